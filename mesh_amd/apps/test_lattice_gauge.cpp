// test_lattice_gauge: the reference's only test (src/tests/lattice_gauge.cpp,
// which prints three ratios without checking them) restated over the C++
// call surface of mesh/gauge.h, with the ratios asserted, plus a plaquette
// and trajectory smoke check.  Exit status 0 iff every check passes.
#include "mesh/gauge.h"

#include <cmath>
#include <cstdio>

using namespace mesh;

namespace {

int failures = 0;

void check(bool ok, char const *what, double value)
{
	std::printf("  %-48s %.12g  %s\n", what, value, ok ? "ok" : "FAIL");
	if (!ok)
		++failures;
}

template <class G> void testImpl(Coordinate geom, uint64_t seed)
{
	std::printf("\n========== testing G=%s%s, Nd=%zu ==========\n", GaugeTraits<G>::name().c_str(),
	            G::prec == MESH_F64 ? "<double>" : "<float>", geom.size());
	Grid g(geom);
	double n = (double)g.size() * g.ndim() * GaugeTraits<G>::dim();
	double tol = 5.0 * std::sqrt(2.0 / n); // 5 sigma of a chi^2 mean

	GaugeField<G> F(g), U(g), tmp(g);
	random_algebra_field(F, seed, 0);
	random_algebra_field(tmp, seed, 1);
	random_gauge_field(U, seed);
	Lattice<G> f(g);
	{
		auto v = tmp[0];
		device_memcpy(f.data(), v.data(), f.bytes());
	}
	// src/tests/lattice_gauge.cpp:23-29
	double r1 = norm2(F) * (1.0 / n) / 0.5;
	check(std::fabs(r1 - 1) < tol, "normalization of norm2(random algebra)", r1);
	double r2 = sum_real_trace_prod(F, F) * (1.0 / n) / (-0.5);
	check(std::fabs(r2 - 1) < tol, "normalization of trace(random algebra^2)", r2);

	// src/tests/lattice_gauge.cpp:31-40
	{
		Landau<G> landau(U);
		random_gauge_field(landau.g, seed + 1);
		double old = landau();
		double eps = G::prec == MESH_F64 ? 1e-4 : 1e-2;
		auto D = landau.deriv();
		double expected = -2.0 * eps * sum_real_trace_prod(f, D);
		evolve(landau.g, f, eps);
		double r3 = (landau() - old) / expected;
		check(std::fabs(r3 - 1) < (G::prec == MESH_F64 ? 1e-3 : 5e-2), "derivative of Landau gauge condition", r3);
	}

	// cshift (lattice.h:366-389) and lattice_sum (lattice.h:270-273): shifting
	// forth and back is the identity, and a shifted field sums to the same
	{
		auto a = cshift(f, g.ndim() - 1, 1);
		auto b = cshift(a, g.ndim() - 1, -1);
		auto s0 = lattice_sum(f), s1 = lattice_sum(a), s2 = lattice_sum(b);
		double d01 = 0, d02 = 0;
		for (int k = 0; k < 9; ++k)
		{
			d01 = std::max(d01, std::abs(s1[k] - s0[k]));
			d02 = std::max(d02, std::abs(s2[k] - s0[k]));
		}
		check(d01 < 1e-6 * g.size(), "lattice_sum invariant under cshift", d01);
		check(d02 == 0.0, "cshift by +1 then -1 gives the field back", d02);
	}

	GaugeField<G> one(g);
	one.fill_unit();
	check(std::fabs(plaquette(one) - 1) < 1e-12, "plaquette of the unit field", plaquette(one));
	double p = plaquette(U);
	check(std::fabs(p) < 5.0 / std::sqrt(g.size() * 18.0 * g.ndim() * (g.ndim() - 1) / 2), "plaquette of Haar links",
	      p);

	Hmc<G> hmc(g);
	hmc.seed = seed;
	hmc.randomizeGaugeField();
	auto d = makeDeltas("2mn", 0.5, 6);
	hmc.runHmcUpdate(6.0, d, false);
	check(std::isfinite(hmc.deltaH_history.back()), "trajectory Delta H is finite", hmc.deltaH_history.back());
	check(hmc.plaq_history.back() > p, "plaquette rises along a hot-start trajectory", hmc.plaq_history.back());
}

} // namespace

int main()
{
	try
	{
		uint64_t seed = 20261202;
		for (auto geom : {Coordinate{8, 8, 8, 8}, Coordinate{8, 8, 8}})
		{
			testImpl<SU3<double>>(geom, seed++);
			testImpl<SU3<float>>(geom, seed++);
		}
	}
	catch (std::exception const &e)
	{
		std::fprintf(stderr, "error: %s\n", e.what());
		return 1;
	}
	std::printf("\n%s (%d failure%s)\n", failures ? "FAILED" : "PASSED", failures, failures == 1 ? "" : "s");
	return failures ? 1 : 0;
}
