// CPU restatement of krox/mesh's HMC trajectory with the reference's own pass
// structure: every lattice operation is a separate pass over freshly
// allocated full-lattice temporaries, neighbours come from cshift copies, and
// the force of each direction is staple sum, product, scaling and projection
// in turn (src/gauge/utils.h:93-153, src/gauge/hmc.h:26-59,
// src/gauge/hmc.cpp:75-107, src/lattice/lattice.h:256-389).
//
// TEST INFRASTRUCTURE ONLY: part of oracle/liboracle_mesh.so, used by tests/
// (against the fused restatement of oracle_mesh.cpp) and by bench.py's
// cpu_baseline leg as the reference-structured CPU baseline.  Layout and
// conventions are those of oracle_mesh.cpp: U[Nd][V][3][3] complex, sites
// row-major with dimension 0 slowest, cshift(a, mu, s)[x] = a[x + s mu-hat].
#include "oracle_su3.h"

#include <chrono>
#include <cstring>
#include <vector>

using namespace orc;

namespace {

using M = mat3<double>;
using Lat = std::vector<M>; // one direction / one matrix per site

struct Geo
{
	int nd;
	int L[4];
	int64_t V, stride[4];
	Geo(int const *dims, int nd_) : nd(nd_)
	{
		V = 1;
		for (int i = nd - 1; i >= 0; --i)
		{
			L[i] = dims[i];
			stride[i] = V;
			V *= dims[i];
		}
	}
	int64_t shift(int64_t x, int mu, int s) const
	{
		int c = (int)((x / stride[mu]) % L[mu]);
		int n = ((c + s) % L[mu] + L[mu]) % L[mu];
		return x + (int64_t)(n - c) * stride[mu];
	}
};

// cshift (src/lattice/lattice.h:366-389): a new lattice, r[x] = a[x + s mu]
Lat cshift(Lat const &a, Geo const &g, int mu, int s)
{
	Lat r(a.size());
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < g.V; ++x)
		r[x] = a[g.shift(x, mu, s)];
	return r;
}

// lattice_apply-style passes (src/lattice/lattice.h:256-261, 277-355)
template <class F> Lat map1(Lat const &a, F f)
{
	Lat r(a.size());
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < (int64_t)a.size(); ++x)
		r[x] = f(a[x]);
	return r;
}
template <class F> Lat map2(Lat const &a, Lat const &b, F f)
{
	Lat r(a.size());
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < (int64_t)a.size(); ++x)
		r[x] = f(a[x], b[x]);
	return r;
}
template <class F> void apply(Lat &a, F f)
{
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < (int64_t)a.size(); ++x)
		f(a[x], x);
}
M addm(M const &a, M const &b)
{
	M r;
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
			r.e[i][j] = {a.e[i][j].re + b.e[i][j].re, a.e[i][j].im + b.e[i][j].im};
	return r;
}
M scale(M const &a, double s)
{
	M r;
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
			r.e[i][j] = {a.e[i][j].re * s, a.e[i][j].im * s};
	return r;
}
template <class F> double sum_apply(Lat const &a, F f)
{
	double s = 0;
#pragma omp parallel for schedule(static) reduction(+ : s)
	for (int64_t x = 0; x < (int64_t)a.size(); ++x)
		s += f(a[x]);
	return s;
}

using Field = std::vector<Lat>;

Field load(double const *U, Geo const &g)
{
	Field f(g.nd, Lat(g.V));
	for (int mu = 0; mu < g.nd; ++mu)
		std::memcpy(f[mu].data(), U + (size_t)mu * g.V * 18, g.V * sizeof(M));
	return f;
}
void store(Field const &f, double *U, Geo const &g)
{
	for (int mu = 0; mu < g.nd; ++mu)
		std::memcpy(U + (size_t)mu * g.V * 18, f[mu].data(), g.V * sizeof(M));
}

// plaquette (src/gauge/utils.h:93-108): six products of shifted copies per plane
double plaquette(Field const &U, Geo const &g)
{
	double s = 0;
	for (int mu = 0; mu < g.nd; ++mu)
		for (int nu = mu + 1; nu < g.nd; ++nu)
		{
			Lat a = map2(U[mu], cshift(U[nu], g, mu, 1), mul_nn<double>);
			Lat b = map1(cshift(U[mu], g, nu, 1), adj3<double>);
			Lat c = map2(a, b, mul_nn<double>);
			Lat d = map1(U[nu], adj3<double>);
			Lat t = map2(c, d, mul_nn<double>);
			s += sum_apply(t, [](M const &m) { return (m.e[0][0].re + m.e[1][1].re) + m.e[2][2].re; });
		}
	return s / ((double)g.V * g.nd * (g.nd - 1) * 0.5) / 3.0;
}
double wilson_action(Field const &U, Geo const &g, double beta)
{
	return (1.0 - plaquette(U, g)) * (beta * g.nd * (g.nd - 1) * 0.5 * (double)g.V);
}
double norm2(Field const &P)
{
	double s = 0;
	for (auto const &p : P)
		s += sum_apply(p, [](M const &m) { return norm2_3(m); });
	return s;
}

// stapleSum (src/gauge/utils.h:110-139): zero fill, then per plane a shifted
// copy, two passes and the shifted sum
Lat staple_sum(Field const &U, Geo const &g, int mu)
{
	Lat S(g.V, zero3<double>());
	for (int nu = 0; nu < g.nd; ++nu)
	{
		if (nu == mu)
			continue;
		Lat tmp = cshift(U[nu], g, mu, 1);
		Lat c = cshift(U[mu], g, nu, 1);
		apply(S, [&](M &a, int64_t x) { a = addm(a, mul_nn(tmp[x], adj3(mul_nn(U[nu][x], c[x])))); });
		apply(tmp, [&](M &a, int64_t x) { a = mul_nn(adj3(mul_nn(U[mu][x], a)), U[nu][x]); });
		Lat t2 = cshift(tmp, g, nu, -1);
		apply(S, [&](M &a, int64_t x) { a = addm(a, t2[x]); });
	}
	return S;
}

// P[mu] -= wilsonDeriv(U, mu, beta) * delta for every mu (src/gauge/hmc.h:55-58,
// src/gauge/utils.h:148-153)
void momentum_step(Field const &U, Field &P, Geo const &g, double beta, double delta)
{
	for (int mu = 0; mu < g.nd; ++mu)
	{
		Lat S = staple_sum(U, g, mu);
		Lat US = map2(U[mu], S, mul_nn<double>);
		Lat sc = map1(US, [&](M const &m) { return scale(m, beta / 6.0); });
		Lat F = map1(sc, ta3<double>);
		Lat Fd = map1(F, [&](M const &m) { return scale(m, delta); });
		apply(P[mu], [&](M &a, int64_t x) {
			for (int i = 0; i < 3; ++i)
				for (int j = 0; j < 3; ++j)
					a.e[i][j] = {a.e[i][j].re - Fd[x].e[i][j].re, a.e[i][j].im - Fd[x].e[i][j].im};
		});
	}
}

// evolve for every mu (src/gauge/hmc.h:26-36)
void link_step(Field &U, Field const &P, Geo const &g, double t)
{
	for (int mu = 0; mu < g.nd; ++mu)
		apply(U[mu], [&](M &a, int64_t x) { a = evolve_link(a, P[mu][x], t); });
}

void reunitize(Field &U)
{
	for (auto &u : U)
		apply(u, [](M &a, int64_t) { a = reunit3(a); });
}

Field momenta(Geo const &g, uint64_t seed, uint32_t traj)
{
	Field P(g.nd, Lat(g.V));
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < g.V; ++x)
		for (int mu = 0; mu < g.nd; ++mu)
		{
			double z[8];
			link_normals(seed, TAG_MOM, traj, (uint32_t)x, (uint32_t)mu, 8, z);
			P[mu][x] = algebra_from_normals(z);
		}
	return P;
}

double now()
{
	return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

} // namespace

extern "C" {

// One trajectory in the reference's pass structure (src/gauge/hmc.cpp:75-107):
// momenta, H_old = S(U) + norm2(P), U_new = U, runHmd, reunitize(U_new),
// H_new, accept if u < exp(-Delta H) (the same Philox uniform as the fused
// restatement).  In place on U; out[4] = {Delta H, accepted, H_old, H_new}.
void orc_refpass_trajectory_f64(double *Uh, int const *dims, int nd, double beta, double const *deltas, int nde,
                                uint64_t seed, uint32_t traj, int metropolis, double *out)
{
	Geo g(dims, nd);
	Field U = load(Uh, g);
	Field P = momenta(g, seed, traj);
	double H_old = wilson_action(U, g, beta) + norm2(P);
	Field Un = U;
	for (int i = 0; i < nde; ++i)
		if (i % 2 == 0)
			link_step(Un, P, g, deltas[i]);
		else
			momentum_step(Un, P, g, beta, deltas[i]);
	reunitize(Un);
	double H_new = wilson_action(Un, g, beta) + norm2(P);
	double dH = H_new - H_old;
	uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
	uint32_t ctr[4] = {traj, 0u, 0u, TAG_METRO};
	uint32_t r[4];
	philox4x32_10(ctr, key, r);
	double u = u01_from_bits(r[0], r[1]);
	int acc = metropolis ? (u < std::exp(-dH)) : 1;
	if (acc)
		store(Un, Uh, g);
	out[0] = dH;
	out[1] = acc;
	out[2] = H_old;
	out[3] = H_new;
}

// Wall time of each stage of that trajectory, run once on U (not modified):
// t[0] momenta + H_old, t[1] one momentum step over all directions, t[2] one
// link step over all directions, t[3] reunitize + H_new.  A trajectory with
// N_F momentum and N_L link steps takes t[0] + N_F t[1] + N_L t[2] + t[3].
void orc_refpass_stage_times_f64(double const *Uh, int const *dims, int nd, double beta, double delta, double *t)
{
	Geo g(dims, nd);
	Field U = load(Uh, g);
	double a = now();
	Field P = momenta(g, 1, 0);
	volatile double H = wilson_action(U, g, beta) + norm2(P);
	double b = now();
	momentum_step(U, P, g, beta, delta);
	double c = now();
	link_step(U, P, g, delta);
	double d = now();
	reunitize(U);
	H = wilson_action(U, g, beta) + norm2(P);
	double e = now();
	(void)H;
	t[0] = b - a;
	t[1] = c - b;
	t[2] = d - c;
	t[3] = e - d;
}

} // extern "C"
