// The generic lattice surface (mesh/lattice_apply.h) with krox/mesh's own
// lambdas and lattice expressions, against the library's named kernels and
// against the same lambdas evaluated on host copies of the fields.
//
// Restated from the reference, as a caller of the surface would write them:
//   stapleSum's two lambdas and the cshift sum  (src/gauge/utils.h:110-139)
//   evolve's lambda                             (src/gauge/hmc.h:26-36)
//   reunitize's lambda                          (src/gauge/utils.h:80-84)
//   plaquette as a lattice expression           (src/gauge/utils.h:93-108)
//   wilsonDeriv                                 (src/gauge/utils.h:148-153)
//   Landau functional and condition             (src/gauge/landau.h:23-49)
//   the line search's exp(F * eps) * g          (src/gauge/landau.h:81, 131)
// Prints one PASS/FAIL line per check; exit status = number of failures.
#include "mesh/lattice_apply.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

using namespace mesh;

static int failures = 0;

static void check(bool ok, char const *prec, char const *what, double v)
{
	std::printf("%s %s %s %.3e\n", ok ? "PASS" : "FAIL", prec, what, v);
	if (!ok)
		++failures;
}

template <class G> std::vector<G> host_copy(Field<G> const &f)
{
	auto h = f.copy_to_host();
	static_assert(sizeof(h[0]) == sizeof(G));
	std::vector<G> r(h.size());
	std::memcpy(r.data(), h.data(), h.size() * sizeof(G));
	return r;
}
template <class F, class G, class... Gs> void host_apply(F f, std::vector<G> &a, std::vector<Gs> const &...as)
{
	for (size_t i = 0; i < a.size(); ++i)
		f(a[i], as[i]...);
}
template <class G> bool same_bits(std::vector<G> const &a, std::vector<G> const &b)
{
	return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(G)) == 0;
}
// max |a - b| / max |b| over all entries
template <class G> double rel_diff(Field<G> const &a, Field<G> const &b)
{
	auto x = host_copy(a), y = host_copy(b);
	double d = 0, m = 0;
	for (size_t i = 0; i < x.size(); ++i)
		for (int k = 0; k < 18; ++k)
		{
			d = std::max(d, std::abs((double)x[i].v[k] - (double)y[i].v[k]));
			m = std::max(m, std::abs((double)y[i].v[k]));
		}
	return d / m;
}

// ---- the reference's algorithms, written against the generic surface ----

template <class vG> Lattice<vG> stapleSum_generic(GaugeField<vG> const &U, int mu)
{
	auto S = Lattice<vG>::zeros(U.grid());
	int Nd = U.grid().ndim();
	for (int nu = 0; nu < Nd; ++nu)
	{
		if (nu == mu)
			continue;
		auto tmp = cshift(U[nu], mu, 1);
		lattice_apply([] __host__ __device__(vG & a, vG const &b, vG const &c, vG const &d) { a += b * adj(d * c); },
		              S, tmp, cshift(U[mu], nu, 1), U[nu]);
		lattice_apply([] __host__ __device__(vG & a, vG const &b, vG const &c) { a = adj(b * a) * c; }, tmp, U[mu],
		              U[nu]);
		S += cshift(tmp, nu, -1);
	}
	return S;
}

template <typename G> void evolve_generic(Lattice<G> &U, Lattice<G> const &P, double t)
{
	lattice_apply([t] __host__ __device__(G & a, G const &b) { a = gauge::exp(project_on_algebra(b) * t) * a; }, U,
	              P);
}

template <typename vG> void reunitize_generic(Lattice<vG> &U)
{
	lattice_apply([] __host__ __device__(vG & a) { a = project_on_group_fast(a); }, U);
}

template <typename G> double plaquette_generic(GaugeField<G> const &U)
{
	double vol = U[0].grid().size();
	double s = 0;
	int Nd = U[0].grid().ndim();
	for (int mu = 0; mu < Nd; ++mu)
		for (int nu = mu + 1; nu < Nd; ++nu)
		{
			auto tmp = U[mu] * cshift(U[nu], mu, 1) * adj(cshift(U[mu], nu, 1)) * adj(U[nu]);
			s += sum_real_trace(tmp);
		}
	return s / (vol * Nd * (Nd - 1) * 0.5) / GaugeTraits<G>::Nc();
}

template <typename G> Lattice<G> wilsonDeriv_generic(GaugeField<G> const &U, int mu, double beta)
{
	return project_on_algebra(U[mu] * stapleSum_generic(U, mu) * (beta / (2 * GaugeTraits<G>::Nc())));
}

template <typename G> double landau_functional_generic(GaugeField<G> const &U, Lattice<G> const &h)
{
	int Nd = U.grid().ndim();
	double vol = U.grid().size();
	double s = 0;
	for (int mu = 0; mu < Nd; ++mu)
		s += sum_real_trace(h * U[mu] * adj(cshift(h, mu, 1)));
	return s / (GaugeTraits<G>::Nc() * Nd * vol);
}

template <typename G> double landau_condition_generic(GaugeField<G> const &U, Lattice<G> const &g)
{
	int Nd = U.grid().ndim();
	double vol = U.grid().size();
	auto d = Lattice<G>::zeros(U.grid());
	for (int mu = 0; mu < Nd; ++mu)
	{
		d += U[mu] * adj(cshift(g, mu, 1));
		d += adj(cshift(g * U[mu], mu, -1));
	}
	return norm2(project_on_algebra(g * d)) * (1.0 / (Nd * vol));
}

// ---- checks ----

template <class T> void run(Grid const &geom, uint64_t seed)
{
	using G = SU3<T>;
	char const *pn = sizeof(T) == 8 ? "f64" : "f32";
	const double tol = sizeof(T) == 8 ? 1e-13 : 3e-6;
	GaugeField<G> U(geom), P(geom);
	random_gauge_field(U, seed);
	random_algebra_field(P, seed + 1, 0);

	// the two staple lambdas, on the device and on host copies: same bits
	{
		const int mu = 1, nu = 2;
		auto S = Lattice<G>::zeros(geom);
		auto tmp = cshift(U[nu], mu, 1);
		auto c = cshift(U[mu], nu, 1);
		auto hS = host_copy(S), ht = host_copy(tmp), hc = host_copy(c), hd = host_copy(U[nu]), hu = host_copy(U[mu]);
		auto f1 = [] __host__ __device__(G & a, G const &b, G const &cc, G const &d) { a += b * adj(d * cc); };
		auto f2 = [] __host__ __device__(G & a, G const &b, G const &cc) { a = adj(b * a) * cc; };
		lattice_apply(f1, S, tmp, c, U[nu]);
		host_apply(f1, hS, ht, hc, hd);
		check(same_bits(host_copy(S), hS), pn, "staple lambda 1 device = host (bits)", 0);
		lattice_apply(f2, tmp, U[mu], U[nu]);
		host_apply(f2, ht, hu, hd);
		check(same_bits(host_copy(tmp), ht), pn, "staple lambda 2 device = host (bits)", 0);
	}
	// stapleSum through the generic surface vs the fused staple kernel
	for (int mu = 0; mu < geom.ndim(); ++mu)
	{
		double d = rel_diff(stapleSum_generic(U, mu), stapleSum(U, mu));
		check(d < tol, pn, "stapleSum generic vs mesh_su3_staple", d);
	}
	// evolve's lambda vs mesh_su3_link_update: same bits (small and large t)
	for (double t : {0.07, 3.0})
	{
		Lattice<G> a = U[0], b = U[0];
		evolve_generic(a, P[0], t);
		evolve(b, P[0], t);
		check(same_bits(host_copy(a), host_copy(b)), pn, "evolve lambda vs mesh_su3_link_update (bits)", t);
		auto h = host_copy(U[0]);
		host_apply([t](G &x, G const &p) { x = gauge::exp(project_on_algebra(p) * t) * x; }, h, host_copy(P[0]));
		check(same_bits(host_copy(a), h), pn, "evolve lambda device = host (bits)", t);
	}
	// reunitize's lambda vs mesh_su3_reunit on perturbed links: same bits
	{
		Lattice<G> a = U[1];
		add_mul(a, 1e-3, P[1]);
		Lattice<G> b = a;
		reunitize_generic(a);
		reunitize(b);
		check(same_bits(host_copy(a), host_copy(b)), pn, "reunitize lambda vs mesh_su3_reunit (bits)", 0);
	}
	// plaquette as a lattice expression vs the plaquette kernel
	{
		double pg = plaquette_generic(U), pk = plaquette(U);
		check(std::abs(pg - pk) < (sizeof(T) == 8 ? 1e-14 : 1e-6), pn, "plaquette expression vs mesh_su3_plaquette",
		      std::abs(pg - pk));
	}
	// wilsonDeriv
	{
		const int mu = geom.ndim() - 1;
		double d = rel_diff(wilsonDeriv_generic(U, mu, 6.0), wilsonDeriv(U, mu, 6.0));
		check(d < tol, pn, "wilsonDeriv generic vs mesh_su3_force", d);
	}
	// Landau functional and condition with a random gauge transformation
	{
		Lattice<G> h(geom);
		random_gauge_field(h, seed + 2);
		Landau<G> L(U);
		L.g = h;
		double fg = landau_functional_generic(U, h), fk = L(h);
		check(std::abs(fg - fk) < (sizeof(T) == 8 ? 1e-14 : 1e-6), pn, "Landau functional expression vs kernel",
		      std::abs(fg - fk));
		double cg = landau_condition_generic(U, h), ck = L.condition();
		check(std::abs(cg - ck) < tol * std::abs(ck), pn, "Landau condition expression vs kernel",
		      std::abs(cg - ck) / std::abs(ck));
		// the line search's trial transform exp(F * eps) * g vs the link step kernel
		Lattice<G> F = L.deriv();
		Lattice<G> g1 = exp(F * 0.25) * h;
		Lattice<G> g2 = h;
		evolve(g2, F, 0.25);
		double d = rel_diff(g1, g2);
		check(d < tol, pn, "exp(F * eps) * g vs mesh_su3_link_update", d);
	}
	// lattice_sum_apply and norm2 against host sums
	{
		auto f = [] __host__ __device__(G const &a, G const &b) { return (double)real_trace(a * adj(b)); };
		double s = lattice_sum_apply(f, U[0], U[1]);
		auto x = host_copy(U[0]), y = host_copy(U[1]);
		long double hs = 0;
		for (size_t i = 0; i < x.size(); ++i)
			hs += f(x[i], y[i]);
		check(std::abs(s - (double)hs) < 1e-12 * std::max(1.0, std::abs((double)hs)), pn,
		      "lattice_sum_apply vs host sum", std::abs(s - (double)hs));
		double n1 = norm2(P), n2 = norm2(static_cast<Field<G> const &>(P));
		check(std::abs(n1 - n2) < 1e-12 * n2, pn, "norm2 generic vs mesh_su3_norm2", std::abs(n1 - n2) / n2);
	}
	// LatticeStack views: writing U[mu] writes the stack
	{
		GaugeField<G> V2(geom);
		V2.fill_unit();
		V2[2] = U[2];
		check(same_bits(host_copy(V2[2]), host_copy(U[2])), pn, "assignment to a stack direction", 0);
	}
}

int main()
{
	try
	{
		run<double>(Grid({4, 6, 4, 8}), 31);
		run<float>(Grid({4, 6, 4, 8}), 41);
		run<double>(Grid({6, 4, 10}), 51);
	}
	catch (std::exception const &e)
	{
		std::printf("FAIL exception %s\n", e.what());
		return 100;
	}
	std::printf("%d failures\n", failures);
	return failures;
}
