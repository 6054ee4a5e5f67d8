// mesh_hmc: pure-gauge HMC driver on one GPU, the counterpart of the
// reference's `hmc` app (src/apps/hmc.cpp:9-156): same options and defaults,
// same summary; the trajectory runs in libmesh_hip.so.
//
// Options (reference defaults): --group su3 --geom 8 8 8 8 --beta 6
//   --start random|cold|<ensemble>/configs/<n> --seed S --count 100
//   --scheme 4mn|2mn|2lf --epsilon 1 --substeps 8 --precision 2 (1 =
//   FP32-mixed, 2 = FP64) --filename <ensemble> [--force]: every
//   configuration and the histories go to an ensemble directory (the
//   reference's HDF5 file, src/apps/hmc.cpp:63-86, 119-124; format in
//   mesh/io.h); --spacing K writes every K-th configuration (the ensemble's
//   markov_spacing attribute).  A start configuration from an ensemble sets the geometry
//   when --geom is not given and, without --seed, continues that chain's
//   Philox streams, so the run reproduces the uninterrupted chain bit for bit.
// Additions: --therm K (first K trajectories without Metropolis),
//   --landau (Landau gauge fixing of the final configuration by
//   overrelaxation), --landau-ascent (the same by the reference's line-search
//   ascent, src/gauge/landau.h:100-136), --float-sums (with --precision 1:
//   Delta H from FP32 lattice sums of H_old and H_new, as the reference's
//   float build forms it; a comparison baseline).
#include "mesh/gauge.h"
#include "mesh/io.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

using namespace mesh;

namespace {

[[noreturn]] void usage(char const *msg)
{
	std::fprintf(stderr,
	             "error: %s\nusage: mesh_hmc [--group su3|su2|u1] [--geom L0 L1 [L2 [L3]]] [--beta B]\n"
	             "                [--start random|cold|<ensemble>/configs/<n>] [--seed S] [--count N]\n"
	             "                [--scheme 2lf|2mn|4mn] [--epsilon E] [--substeps N] [--precision 1|2]\n"
	             "                [--filename <ensemble> [--force] [--spacing K]] [--therm K] [--landau]\n"
	             "                [--landau-ascent] [--float-sums]\n",
	             msg);
	std::exit(2);
}

bool is_number(std::string const &s)
{
	return !s.empty() && s.find_first_not_of("0123456789") == std::string::npos;
}

} // namespace

int main(int argc, char **argv)
{
	std::string group = "su3", scheme = "4mn", start = "random", seed = "";
	std::vector<int32_t> geom = {8, 8, 8, 8};
	double beta = 6.0, epsilon = 1.0;
	int substeps = 8, count = 100, precision = 2, therm = 0, spacing = 1;
	bool landau = false, ascent = false, geom_given = false, allow_overwrite = false, float_sums = false,
	     therm_given = false;
	std::string filename;

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
		else if (a == "--geom")
		{
			geom.clear();
			geom_given = true;
			while (i + 1 < argc && is_number(argv[i + 1]))
				geom.push_back(std::atoi(argv[++i]));
		}
		else if (a == "--beta")
			beta = std::atof(next().c_str());
		else if (a == "--start")
			start = next();
		else if (a == "--seed")
			seed = next();
		else if (a == "--count")
			count = std::atoi(next().c_str());
		else if (a == "--scheme")
			scheme = next();
		else if (a == "--epsilon")
			epsilon = std::atof(next().c_str());
		else if (a == "--substeps")
			substeps = std::atoi(next().c_str());
		else if (a == "--precision")
			precision = std::atoi(next().c_str());
		else if (a == "--therm")
		{
			therm = std::atoi(next().c_str());
			therm_given = true;
		}
		else if (a == "--landau")
			landau = true;
		else if (a == "--landau-ascent")
			landau = ascent = true;
		else if (a == "--float-sums")
			float_sums = true;
		else if (a == "--plot")
			std::fprintf(stderr, "mesh_hmc: --plot (gnuplot of the plaquette history) is not supported; ignored\n");
		else if (a == "--filename")
			filename = next();
		else if (a == "--force")
			allow_overwrite = true;
		else if (a == "--spacing")
			spacing = std::atoi(next().c_str());
		else if (a == "--help" || a == "-h")
			usage("help requested");
		else
			usage(("unknown option " + a).c_str());
	}
	if (spacing < 1)
		usage("--spacing must be positive");
	const bool seed_given = !seed.empty();
	if (seed.empty())
		seed = std::to_string(std::random_device{}());
	const bool from_file = start != "random" && start != "cold";

	try
	{
		auto deltas = makeDeltas(scheme, epsilon, substeps);
		if (from_file && !geom_given)
			geom = configGeometry(start);
		auto t_total = std::chrono::steady_clock::now();
		mesh_profile_enable(ctx().get(), 1);
		if (float_sums)
			ctx().check(mesh_set_option(ctx().get(), "float_sums", 1));
		dispatchByGroup(
		    [&]<typename G>() {
			    Grid g(geom);
			    Hmc<G> hmc(g);
			    hmc.seed = seed_key(seed);
			    if (start == "random")
				    hmc.randomizeGaugeField();
			    else if (start == "cold")
				    hmc.U.fill_unit();
			    else
			    {
				    readConfig(start, hmc.U);
				    auto [edir, ename] = io::split_config_name(start);
				    auto at = io::Ensemble::open(edir).attributes();
				    if (!seed_given && at.has("philox_key") && is_number(ename))
				    {
					    hmc.seed = std::strtoull(at.get_string("philox_key").c_str(), nullptr, 10);
					    hmc.traj = (uint32_t)(at.get_int("philox_traj0") + std::stoll(ename));
					    std::printf("continuing the chain of %s at trajectory counter %u\n", edir.c_str(), hmc.traj);
					    // a configuration stored inside the thermalisation: the
					    // rest of it again without Metropolis, as the chain had it
					    if (!therm_given && at.has("markov_discard"))
					    {
						    therm = (int)std::max(0LL, at.get_int("markov_discard") - std::stoll(ename));
						    if (therm > 0)
							    std::printf("%d thermalisation trajectories left in that chain\n", therm);
					    }
				    }
				    if (at.has("precision") && at.get_int("precision") != precision)
					    std::fprintf(stderr,
					                 "warning: %s was written at --precision %lld, this run uses %d: the chain "
					                 "does not continue bit for bit\n",
					                 start.c_str(), at.get_int("precision"), precision);
			    }
			    io::Ensemble ens;
			    if (!filename.empty())
			    {
				    ens = io::Ensemble::create(filename, allow_overwrite);
				    ens.set_attribute("group", group);
				    ens.set_attribute("beta", beta);
				    ens.set_attribute("geometry", std::vector<int>(geom.begin(), geom.end()));
				    ens.set_attribute("hmc_scheme", scheme);
				    ens.set_attribute("hmc_epsilon", epsilon);
				    ens.set_attribute("hmc_substeps", substeps);
				    ens.set_attribute("hmc_metropolis", 1);
				    ens.set_attribute("markov_seed", seed);
				    ens.set_attribute("markov_start", start);
				    ens.set_attribute("markov_discard", therm);
				    ens.set_attribute("markov_count", count);
				    ens.set_attribute("markov_spacing", spacing);
				    ens.set_attribute("precision", precision);
				    ens.set_attribute("philox_key", std::to_string(hmc.seed));
				    ens.set_attribute("philox_traj0", (long long)hmc.traj);
			    }
			    for (int iter = 0; iter < count; ++iter)
			    {
				    hmc.runHmcUpdate(beta, deltas, iter >= therm);
				    if (iter == therm - 1)
					    hmc.reset_observables();
				    if (ens && (iter + 1) % spacing == 0)
					    writeConfig(ens, std::to_string(iter + 1), hmc.U);
			    }
			    hmc.print_summary();
			    if (ens)
			    {
				    ens.write_data("plaq_history", hmc.plaq_history);
				    ens.write_data("accept_history", hmc.accept_history);
				    ens.write_data("deltaH_history", hmc.deltaH_history);
				    ens.write_data("time_history", hmc.time_history);
			    }
			    if (landau)
			    {
				    Landau<G> lan(hmc.U);
				    lan.verbose = true;
				    if (ascent)
				    {
					    int it = lan.run_ascent(100000);
					    std::printf("landau ascent: %d iterations, condition = %.3e, functional = %.12f\n", it,
					                lan.condition(), lan());
				    }
				    else
				    {
					    int it = lan.run(100000);
					    std::printf("landau: %d sweeps, theta = %.3e, functional = %.12f\n", it,
					                lan.condition() * g.ndim(), lan());
				    }
			    }
		    },
		    group, precision);
		ctx().synchronize();
		double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_total).count();
		std::printf("========== performance summary ==========\n");
		for (char const *k : {"force_link", "momenta_link", "force_update", "link_update", "delta_h", "momenta",
		                      "landau_or"})
		{
			double ms;
			long long n;
			ctx().check(mesh_profile_read(ctx().get(), k, &ms, &n));
			std::printf("time in %-13s %8.3f s (%5.1f %%, %lld launches)\n", k, ms * 1e-3, 100 * ms * 1e-3 / total,
			            n);
		}
		std::printf("total:                %8.3f s\n", total);
	}
	catch (std::exception const &e)
	{
		std::fprintf(stderr, "error: %s\n", e.what());
		return 1;
	}
	return 0;
}
