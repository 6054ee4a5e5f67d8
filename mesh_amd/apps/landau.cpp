// mesh_landau: Landau gauge fixing of a stored configuration, the counterpart
// of the reference's `landau` app (src/apps/landau.cpp:8-29): same options,
// the configuration named "<ensemble>/configs/<n>" as readConfig takes it
// (an ensemble directory of mesh/io.h in place of the HDF5 file), geometry
// from the ensemble.
//
// Options: --group su3 --precision 2 (1 = FP32 links, an FP64 file rounded)
//   input.  The reference runs its line-search ascent (Landau::run,
//   src/gauge/landau.h:100-136) for at most 10000 iterations; this build runs
//   even/odd overrelaxation to theta < 1e-12 by default and the reference's
//   solver with --ascent.  Additions: --omega W, --max-iter N.
#include "mesh/gauge.h"
#include "mesh/io.h"

#include <cstdio>
#include <cstdlib>
#include <string>

using namespace mesh;

namespace {
[[noreturn]] void usage(char const *msg)
{
	std::fprintf(stderr,
	             "error: %s\nusage: mesh_landau [--group su3] [--precision 1|2] [--ascent] [--omega W] [--max-iter N]\n"
	             "                   <ensemble>/configs/<n>\n",
	             msg);
	std::exit(2);
}
} // namespace

int main(int argc, char **argv)
{
	std::string group = "su3", configName;
	int precision = 2, max_iter = 10000;
	double omega = 1.7;
	bool ascent = false;
	for (int i = 1; i < argc; ++i)
	{
		std::string a = argv[i];
		auto next = [&]() -> std::string {
			if (i + 1 >= argc)
				usage(("missing value for " + a).c_str());
			return argv[++i];
		};
		if (a == "--group")
			group = next();
		else if (a == "--precision")
			precision = std::atoi(next().c_str());
		else if (a == "--ascent")
			ascent = true;
		else if (a == "--omega")
			omega = std::atof(next().c_str());
		else if (a == "--max-iter")
			max_iter = std::atoi(next().c_str());
		else if (a == "--help" || a == "-h")
			usage("help requested");
		else if (!a.empty() && a[0] == '-')
			usage(("unknown option " + a).c_str());
		else if (configName.empty())
			configName = a;
		else
			usage("more than one input");
	}
	if (configName.empty())
		usage("no input configuration");

	try
	{
		Grid g(configGeometry(configName));
		dispatchByGroup(
		    [&]<typename G>() {
			    typename FieldOf<G>::type U(g);
			    readConfig(configName, U);
			    Landau<G> lan(U);
			    lan.verbose = true;
			    const double F0 = lan(), theta0 = lan.condition() * g.ndim();
			    std::printf("start: functional = %.12f, theta = %.3e\n", F0, theta0);
			    if (ascent)
			    {
				    int it = lan.run_ascent(max_iter);
				    std::printf("landau ascent: %d iterations, condition = %.3e, functional = %.12f\n", it,
				                lan.condition(), lan());
			    }
			    else
			    {
				    int it = lan.run(max_iter, omega);
				    std::printf("landau: %d sweeps, theta = %.3e, functional = %.12f\n", it, lan.last_theta, lan());
			    }
			    // the result as a gauge transformation: g back on the group, U^g rebuilt
			    lan.reunit(lan.g);
			    std::printf("reunitarised g: theta = %.3e, functional = %.12f\n", lan.condition() * g.ndim(), lan());
			    std::printf("plaquette: before %.15f, after %.15f\n", plaquette(U), plaquette(lan.gauge_fixed()));
		    },
		    group, precision);
	}
	catch (std::exception const &e)
	{
		std::fprintf(stderr, "error: %s\n", e.what());
		return 1;
	}
	return 0;
}
