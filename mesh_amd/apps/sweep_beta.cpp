// mesh_sweep_beta: pure-gauge HMC over a range of couplings, the counterpart
// of the reference's `sweep_beta` app (src/apps/sweep_beta.cpp:8-97): same
// options and defaults, one chain carried from beta to beta, the mean
// plaquette over the last 90% of each beta's trajectories.
//
// Options (reference defaults): --group su3 --geom 4 4 4 4 --beta-min 0
//   --beta-max 8 --beta-count 30 --seed S --count 50 --scheme 4mn
//   --epsilon 2 --substeps 8 --precision 2 (1 = FP32-mixed).  --plot (gnuplot)
//   is accepted and ignored.
#include "mesh/gauge.h"

#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

using namespace mesh;

namespace {
[[noreturn]] void usage(char const *msg)
{
	std::fprintf(stderr,
	             "error: %s\nusage: mesh_sweep_beta [--group su3|u1] [--geom L0 L1 [L2 [L3]]] [--beta-min B]\n"
	             "                       [--beta-max B] [--beta-count N] [--seed S] [--count N]\n"
	             "                       [--scheme 2lf|2mn|4mn] [--epsilon E] [--substeps N] [--precision 1|2]\n",
	             msg);
	std::exit(2);
}

bool is_number(std::string const &s)
{
	return !s.empty() && s.find_first_not_of("0123456789") == std::string::npos;
}

double mean(std::vector<double> const &v, size_t from)
{
	double s = 0;
	for (size_t i = from; i < v.size(); ++i)
		s += v[i];
	return v.size() > from ? s / (double)(v.size() - from) : 0.0;
}
} // namespace

int main(int argc, char **argv)
{
	std::string group = "su3", scheme = "4mn", seed;
	std::vector<int32_t> geom = {4, 4, 4, 4};
	double beta_min = 0, beta_max = 8, epsilon = 2.0;
	int beta_count = 30, substeps = 8, count = 50, precision = 2;
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
			while (i + 1 < argc && is_number(argv[i + 1]))
				geom.push_back(std::atoi(argv[++i]));
		}
		else if (a == "--beta-min")
			beta_min = std::atof(next().c_str());
		else if (a == "--beta-max")
			beta_max = std::atof(next().c_str());
		else if (a == "--beta-count")
			beta_count = std::atoi(next().c_str());
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
		else if (a == "--plot")
			std::fprintf(stderr, "mesh_sweep_beta: --plot (gnuplot) is not supported; ignored\n");
		else if (a == "--help" || a == "-h")
			usage("help requested");
		else
			usage(("unknown option " + a).c_str());
	}
	if (beta_count < 1 || count < 1)
		usage("--beta-count and --count must be positive");
	if (seed.empty())
		seed = std::to_string(std::random_device{}());

	try
	{
		auto deltas = makeDeltas(scheme, epsilon, substeps);
		dispatchByGroup(
		    [&]<typename G>() {
			    Grid g(geom);
			    Hmc<G> hmc(g);
			    hmc.seed = seed_key(seed);
			    hmc.randomizeGaugeField();
			    for (int iter = 0; iter < beta_count; ++iter)
			    {
				    // the reference divides by beta_count - 1: one beta means beta_min
				    double beta =
				        beta_count > 1 ? beta_min + (beta_max - beta_min) / (beta_count - 1) * iter : beta_min;
				    hmc.reset_observables();
				    for (int i = 0; i < count; ++i)
					    hmc.runHmcUpdate(beta, deltas);
				    double plaq = mean(hmc.plaq_history, hmc.plaq_history.size() / 10);
				    std::printf("beta = %.3f, acc = %.3f, plaq = %.3f\n", beta, mean(hmc.accept_history, 0), plaq);
				    std::fflush(stdout);
			    }
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
