// CPU oracle: lattice-level restatement of krox/mesh's pure-gauge HMC,
// Wilson action and Landau gauge functional, plus the Landau overrelaxation
// sweep that the GPU library implements.
//
// TEST INFRASTRUCTURE ONLY: built into oracle/liboracle_mesh.so and loaded
// (through oracle/oracle.py) only by tests/, the smoke check and the
// cpu_baseline leg of bench.py.  The GPU library never links or calls it.
//
// Host layout of a gauge field: double/float U[Nd][V][3][3][2] (one block per
// Lorentz direction, as LatticeStack, src/lattice/lattice.h:159-192), sites
// row-major with dimension 0 slowest (src/lattice/grid.h:39-49), each matrix
// row-major with interleaved (re, im).
//
// Shift convention: shift(x, mu, +1) = x + mu-hat, periodic, i.e.
// cshift(a, mu, +1)[x] = a[x + mu] (src/lattice/lattice.h:366-389,
// src/lattice/device.h:55-57, 82-84).
#include "oracle_su3.h"
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <omp.h>
#include <cstring>
#include <string>
#include <vector>

using namespace orc;

namespace {

struct Geo
{
	int nd;
	int L[4];
	int64_t V;
	int64_t stride[4];

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
	int coord(int64_t x, int mu) const { return (int)((x / stride[mu]) % L[mu]); }
	int64_t shift(int64_t x, int mu, int s) const
	{
		int c = coord(x, mu);
		int n = (c + s) % L[mu];
		if (n < 0)
			n += L[mu];
		return x + (int64_t)(n - c) * stride[mu];
	}
	int parity(int64_t x) const
	{
		int p = 0;
		for (int mu = 0; mu < nd; ++mu)
			p += coord(x, mu);
		return p & 1;
	}
};

template <class T> inline mat3<T> ld(T const *U, int64_t V, int mu, int64_t x)
{
	mat3<T> m;
	std::memcpy(&m, U + (mu * V + x) * 18, sizeof(m));
	return m;
}
template <class T> inline void st(T *U, int64_t V, int mu, int64_t x, mat3<T> const &m)
{
	std::memcpy(U + (mu * V + x) * 18, &m, sizeof(m));
}
template <class T> inline mat3<double> up(mat3<T> const &a)
{
	mat3<double> r;
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
			r.e[i][j] = {(double)a.e[i][j].re, (double)a.e[i][j].im};
	return r;
}
template <class T> inline mat3<T> down(mat3<double> const &a)
{
	mat3<T> r;
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
			r.e[i][j] = {(T)a.e[i][j].re, (T)a.e[i][j].im};
	return r;
}

// staple sum for direction mu at x (src/gauge/utils.h:110-139):
//   sum_{nu != mu} U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
//                + U_nu(x-nu+mu)^+ U_mu(x-nu)^+ U_nu(x-nu)
// For nu < mu the first factor U_nu(x+mu) U_mu(x+nu)^+ is formed as the
// adjoint of U_mu(x+nu) U_nu(x+mu)^+, the first factor of direction nu in
// plane mu, so that both directions of a plane use one rounded product.
template <class T> inline mat3<T> staple_at(T const *U, Geo const &g, int mu, int64_t x)
{
	mat3<T> S = zero3<T>();
	int64_t xpm = g.shift(x, mu, 1);
	for (int nu = 0; nu < g.nd; ++nu)
	{
		if (nu == mu)
			continue;
		int64_t xpn = g.shift(x, nu, 1);
		int64_t xmn = g.shift(x, nu, -1);
		int64_t xpmmn = g.shift(xpm, nu, -1);
		mat3<T> T1 = nu < mu ? adj3(mul_nd(ld(U, g.V, mu, xpn), ld(U, g.V, nu, xpm)))
		                     : mul_nd(ld(U, g.V, nu, xpm), ld(U, g.V, mu, xpn));
		madd_nd(S, T1, ld(U, g.V, nu, x));
		mat3<T> T2 = mul_dd(ld(U, g.V, nu, xpmmn), ld(U, g.V, mu, xmn));
		madd_nn(S, T2, ld(U, g.V, nu, xmn));
	}
	return S;
}

// sum over planes mu<nu of Re tr U_mu(x) U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+,
// evaluated in double from links of type T (src/gauge/utils.h:93-108)
template <class T> inline double plaq_site(T const *U, Geo const &g, int64_t x)
{
	double s = 0.0;
	for (int mu = 0; mu < g.nd; ++mu)
		for (int nu = mu + 1; nu < g.nd; ++nu)
		{
			int64_t xpm = g.shift(x, mu, 1), xpn = g.shift(x, nu, 1);
			mat3<double> A = mul_nn(up(ld(U, g.V, mu, x)), up(ld(U, g.V, nu, xpm)));
			mat3<double> B = mul_nn(up(ld(U, g.V, nu, x)), up(ld(U, g.V, mu, xpn)));
			s += retr_nd(A, B);
		}
	return s;
}

template <class T> inline double kin_site(T const *P, Geo const &g, int64_t x)
{
	double k = 0.0;
	for (int mu = 0; mu < g.nd; ++mu)
		k += norm2_3(up(ld(P, g.V, mu, x)));
	return k;
}

long double lsum(std::vector<double> const &v)
{
	long double s = 0;
	for (double d : v)
		s += d;
	return s;
}

template <class T> void random_gauge(T *U, Geo const &g, uint64_t seed)
{
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < g.V; ++x)
		for (int mu = 0; mu < g.nd; ++mu)
		{
			double z[18];
			link_normals(seed, TAG_HAAR, 0, (uint32_t)x, (uint32_t)mu, 18, z);
			st(U, g.V, mu, x, down<T>(haar_from_normals(z)));
		}
}

template <class T> void momenta(T *P, double *K, Geo const &g, uint64_t seed, uint32_t traj)
{
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < g.V; ++x)
	{
		for (int mu = 0; mu < g.nd; ++mu)
		{
			double z[8];
			link_normals(seed, TAG_MOM, traj, (uint32_t)x, (uint32_t)mu, 8, z);
			st(P, g.V, mu, x, down<T>(algebra_from_normals(z)));
		}
		if (K)
			K[x] = kin_site(P, g, x);
	}
}

template <class T> void force_update(T const *U, T *P, Geo const &g, double beta, double delta)
{
	double k = delta * (beta / 6.0);
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < g.V; ++x)
		for (int mu = 0; mu < g.nd; ++mu)
		{
			mat3<T> S = staple_at(U, g, mu, x);
			st(P, g.V, mu, x, force_link_update(ld(P, g.V, mu, x), ld(U, g.V, mu, x), S, k));
		}
}

template <class T> void link_update(T const *Uin, T *Uout, T const *P, Geo const &g, double t, int reunit)
{
	T tt = (T)t;
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < g.V; ++x)
		for (int mu = 0; mu < g.nd; ++mu)
		{
			mat3<T> u = evolve_link(ld(Uin, g.V, mu, x), ld(P, g.V, mu, x), tt);
			if (reunit)
				u = reunit3(u);
			st(Uout, g.V, mu, x, u);
		}
}

// per-site differenced Delta H (see header comment of orc_su3_delta_h)
template <class T>
void delta_h(T const *Uo, T const *Un, T const *P, double const *Ko, Geo const &g, double beta,
             double *out)
{
	std::vector<double> d(g.V), pso(g.V), psn(g.V), kn(g.V);
	double c = beta / 3.0;
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < g.V; ++x)
	{
		pso[x] = plaq_site(Uo, g, x);
		psn[x] = plaq_site(Un, g, x);
		kn[x] = kin_site(P, g, x);
		double dps = psn[x] - pso[x];
		double dk = kn[x] - Ko[x];
		d[x] = std::fma(-c, dps, dk);
	}
	out[0] = (double)lsum(d);
	out[1] = (double)lsum(pso);
	out[2] = (double)lsum(psn);
	long double ko = 0;
	for (int64_t x = 0; x < g.V; ++x)
		ko += Ko[x];
	out[3] = (double)ko;
	out[4] = (double)lsum(kn);
}

inline double metropolis_uniform(uint64_t seed, uint32_t traj)
{
	uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
	uint32_t ctr[4] = {traj, 0u, 0u, TAG_METRO};
	uint32_t r[4];
	philox4x32_10(ctr, key, r);
	return u01_from_bits(r[0], r[1]);
}

// One trajectory (src/gauge/hmc.cpp:75-107 with src/gauge/hmc.h:38-59):
// refresh P, U_new = U, leapfrog/Omelyan over deltas (even index: link
// update, odd: momentum update), reunitarise U_new (fused into the last link
// update), Delta H, accept if u < exp(-Delta H).
template <class T>
void hmc_trajectory(T *U, Geo const &g, double beta, double const *deltas, int nde, uint64_t seed,
                    uint32_t traj, int metropolis, double *out)
{
	size_t n = (size_t)g.nd * g.V * 18;
	std::vector<T> P(n), Un(n);
	std::vector<double> K(g.V);
	momenta(P.data(), K.data(), g, seed, traj);
	for (int i = 0; i < nde; ++i)
	{
		if (i % 2 == 0)
			link_update(i == 0 ? U : Un.data(), Un.data(), P.data(), g, deltas[i], i == nde - 1);
		else
			force_update(Un.data(), P.data(), g, beta, deltas[i]);
	}
	double dh[5];
	delta_h(U, Un.data(), P.data(), K.data(), g, beta, dh);
	double u = metropolis_uniform(seed, traj);
	int acc = metropolis ? (u < std::exp(-dh[0])) : 1;
	if (acc)
		std::memcpy(U, Un.data(), n * sizeof(T));
	double np = (double)g.V * g.nd * (g.nd - 1) / 2;
	out[0] = dh[0];
	out[1] = acc;
	out[2] = (acc ? dh[2] : dh[1]) / (np * 3.0);
	out[3] = beta * (np - dh[1] / 3.0) + dh[3];
	out[4] = beta * (np - dh[2] / 3.0) + dh[4];
	out[5] = u;
}

////////////////////////////////////////////////////////////////////////////////
// Landau gauge
////////////////////////////////////////////////////////////////////////////////

// F = (1/(Nc Nd V)) sum_{x,mu} Re tr h(x) U_mu(x) h(x+mu)^+   (src/gauge/landau.h:23-34)
// h == nullptr means h = 1.
template <class T> double landau_functional(T const *U, T const *h, Geo const &g)
{
	std::vector<double> s(g.V);
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < g.V; ++x)
	{
		double a = 0.0;
		for (int mu = 0; mu < g.nd; ++mu)
		{
			mat3<double> u = up(ld(U, g.V, mu, x));
			if (h)
			{
				u = mul_nn(up(ld(h, g.V, 0, x)), u);
				u = mul_nd(u, up(ld(h, g.V, 0, g.shift(x, mu, 1))));
			}
			a += (u.e[0][0].re + u.e[1][1].re) + u.e[2][2].re;
		}
		s[x] = a;
	}
	return (double)(lsum(s) / ((long double)3.0 * g.nd * g.V));
}

// gauge-transformed link U^h_mu(x) in double
template <class T> inline mat3<double> ug(T const *U, T const *h, Geo const &g, int mu, int64_t x)
{
	mat3<double> u = up(ld(U, g.V, mu, x));
	if (h)
	{
		u = mul_nn(up(ld(h, g.V, 0, x)), u);
		u = mul_nd(u, up(ld(h, g.V, 0, g.shift(x, mu, 1))));
	}
	return u;
}

// sum_x ||Delta(x)||_F^2 with Delta(x) = sum_mu [TA(U^h_mu(x)) - TA(U^h_mu(x-mu))]
// (src/gauge/landau.h:37-49: condition() = this / (Nd V))
template <class T> double landau_divsum(T const *U, T const *h, Geo const &g)
{
	std::vector<double> s(g.V);
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < g.V; ++x)
	{
		mat3<double> D = zero3<double>();
		for (int mu = 0; mu < g.nd; ++mu)
		{
			mat3<double> a = ta3(ug(U, h, g, mu, x));
			mat3<double> b = ta3(ug(U, h, g, mu, g.shift(x, mu, -1)));
			for (int i = 0; i < 3; ++i)
				for (int j = 0; j < 3; ++j)
				{
					D.e[i][j].re += a.e[i][j].re - b.e[i][j].re;
					D.e[i][j].im += a.e[i][j].im - b.e[i][j].im;
				}
		}
		s[x] = norm2_3(D);
	}
	return (double)lsum(s);
}

// gradient (src/gauge/landau.h:52-65):
//   -0.5/(Nc Nd V) TA(h(x) d(x)), d(x) = sum_mu U_mu(x) h(x+mu)^+ + (h(x-mu) U_mu(x-mu))^+
template <class T> void landau_deriv(T const *U, T const *h, Geo const &g, T *F)
{
	double c = -0.5 / (3.0 * g.nd * (double)g.V);
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < g.V; ++x)
	{
		mat3<double> d = zero3<double>();
		for (int mu = 0; mu < g.nd; ++mu)
		{
			int64_t xp = g.shift(x, mu, 1), xm = g.shift(x, mu, -1);
			mat3<double> a = mul_nd(up(ld(U, g.V, mu, x)), up(ld(h, g.V, 0, xp)));
			mat3<double> b = adj3(mul_nn(up(ld(h, g.V, 0, xm)), up(ld(U, g.V, mu, xm))));
			for (int i = 0; i < 3; ++i)
				for (int j = 0; j < 3; ++j)
				{
					d.e[i][j].re += a.e[i][j].re + b.e[i][j].re;
					d.e[i][j].im += a.e[i][j].im + b.e[i][j].im;
				}
		}
		mat3<double> f = ta3(mul_nn(up(ld(h, g.V, 0, x)), d));
		for (int i = 0; i < 3; ++i)
			for (int j = 0; j < 3; ++j)
			{
				f.e[i][j].re *= c;
				f.e[i][j].im *= c;
			}
		st(F, g.V, 0, x, down<T>(f));
	}
}

// overrelaxation update of one site: K = sum_mu U_mu(x) + U_mu(x-mu)^+,
// three SU(2) subgroups (0,1),(0,2),(1,2), each r = argmax Re tr(r k) raised
// to the power omega by the truncated binomial series and renormalised.
template <class T> inline mat3<T> or_site_transform(mat3<T> K, T omega)
{
	static const int sub[3][2] = {{0, 1}, {0, 2}, {1, 2}};
	mat3<T> G = unit3<T>();
	T c2 = T(0.5) * omega * (omega - T(1));
	for (int s = 0; s < 3; ++s)
	{
		int i = sub[s][0], j = sub[s][1];
		T b0 = T(0.5) * (K.e[i][i].re + K.e[j][j].re);
		T b3 = T(0.5) * (K.e[i][i].im - K.e[j][j].im);
		T b1 = T(0.5) * (K.e[i][j].im + K.e[j][i].im);
		T b2 = T(0.5) * (K.e[i][j].re - K.e[j][i].re);
		T n2 = b0 * b0;
		n2 = std::fma(b1, b1, n2);
		n2 = std::fma(b2, b2, n2);
		n2 = std::fma(b3, b3, n2);
		if (!(n2 > T(0)))
			continue;
		T inv = T(1) / std::sqrt(n2);
		T r0 = b0 * inv, r1 = -(b1 * inv), r2 = -(b2 * inv), r3 = -(b3 * inv);
		T a = r0 - T(1);
		T rv2 = r1 * r1;
		rv2 = std::fma(r2, r2, rv2);
		rv2 = std::fma(r3, r3, rv2);
		T g0 = std::fma(c2, std::fma(a, a, -rv2), std::fma(omega, a, T(1)));
		T fac = std::fma(T(2) * c2, a, omega);
		T g1 = r1 * fac, g2 = r2 * fac, g3 = r3 * fac;
		T gn2 = g0 * g0;
		gn2 = std::fma(g1, g1, gn2);
		gn2 = std::fma(g2, g2, gn2);
		gn2 = std::fma(g3, g3, gn2);
		T gi = T(1) / std::sqrt(gn2);
		g0 *= gi;
		g1 *= gi;
		g2 *= gi;
		g3 *= gi;
		cplx<T> e00 = {g0, g3}, e01 = {g2, g1}, e10 = {-g2, g1}, e11 = {g0, -g3};
		for (int c = 0; c < 3; ++c)
		{
			cplx<T> ki = K.e[i][c], kj = K.e[j][c];
			cplx<T> ni = {T(0), T(0)}, nj = {T(0), T(0)};
			cmadd(ni, e00, ki);
			cmadd(ni, e01, kj);
			cmadd(nj, e10, ki);
			cmadd(nj, e11, kj);
			K.e[i][c] = ni;
			K.e[j][c] = nj;
			cplx<T> gi_ = G.e[i][c], gj_ = G.e[j][c];
			cplx<T> mi = {T(0), T(0)}, mj = {T(0), T(0)};
			cmadd(mi, e00, gi_);
			cmadd(mi, e01, gj_);
			cmadd(mj, e10, gi_);
			cmadd(mj, e11, gj_);
			G.e[i][c] = mi;
			G.e[j][c] = mj;
		}
	}
	return G;
}

template <class T> void landau_or_halfsweep(T *U, T *h, Geo const &g, int parity, double omega)
{
	T om = (T)omega;
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < g.V; ++x)
	{
		if (g.parity(x) != parity)
			continue;
		mat3<T> K = zero3<T>();
		for (int mu = 0; mu < g.nd; ++mu)
		{
			mat3<T> a = ld(U, g.V, mu, x);
			mat3<T> b = ld(U, g.V, mu, g.shift(x, mu, -1));
			for (int i = 0; i < 3; ++i)
				for (int j = 0; j < 3; ++j)
				{
					K.e[i][j].re += a.e[i][j].re;
					K.e[i][j].im += a.e[i][j].im;
					K.e[i][j].re += b.e[j][i].re;
					K.e[i][j].im -= b.e[j][i].im;
				}
		}
		mat3<T> G = or_site_transform(K, om);
		for (int mu = 0; mu < g.nd; ++mu)
		{
			int64_t xm = g.shift(x, mu, -1);
			st(U, g.V, mu, x, mul_nn(G, ld(U, g.V, mu, x)));
			st(U, g.V, mu, xm, mul_nd(ld(U, g.V, mu, xm), G));
		}
		if (h)
			st(h, g.V, 0, x, mul_nn(G, ld(h, g.V, 0, x)));
	}
}

int make_deltas(char const *scheme, double epsilon, int substeps, double *out, int cap)
{
	double delta = std::sqrt(2 * epsilon) / substeps;
	std::vector<double> raw;
	std::string s = scheme;
	if (s == "2lf")
		raw = {0.5, 1.0, 0.5};
	else if (s == "2mn")
	{
		double l = 0.19318332750378361;
		raw = {l, 0.5, 1.0 - 2.0 * l, 0.5, l};
	}
	else if (s == "4mn")
	{
		double rho = 0.1786178958448091, theta = -0.06626458266981843, lam = 0.7123418310626056;
		raw = {rho, lam, theta, 0.5 - lam, 1.0 - 2.0 * (rho + theta), 0.5 - lam, theta, lam, rho};
	}
	else
		return -1;
	std::vector<double> d;
	for (int i = 0; i < substeps; ++i)
	{
		if (i == 0)
			d.insert(d.end(), raw.begin(), raw.end());
		else
		{
			d.back() *= 2;
			d.insert(d.end(), raw.begin() + 1, raw.end());
		}
	}
	for (double &x : d)
		x *= delta;
	if ((int)d.size() > cap)
		return -2;
	std::copy(d.begin(), d.end(), out);
	return (int)d.size();
}

} // namespace

extern "C" {

// integrator step table (src/gauge/hmc.cpp:7-51); returns length, <0 on error
int orc_make_deltas(char const *scheme, double epsilon, int substeps, double *out, int cap)
{
	return make_deltas(scheme, epsilon, substeps, out, cap);
}

// {wilsonAction(U, beta), norm2(P)} of the reference's float build
// (src/gauge/utils.h:94-108, 141-146, 27-41): per-site FP32 values summed as
// device_sum_apply does (src/lattice/device.h:184-210), a 256-wide pair tree
// per block in FP32 and a sequential FP32 sum of the block results; planes
// and directions accumulated in FP64.  P may be NULL.
void orc_su3_float_energy_f32(float const *U, float const *P, int const *dims, int nd, double beta, double *out)
{
	Geo g(dims, nd);
	const int64_t nb = (g.V + 255) / 256;
	std::vector<float> site(g.V);
	auto tree = [&]() {
		float acc = 0.0f;
		for (int64_t b = 0; b < nb; ++b)
		{
			float s[256];
			const int64_t base = b * 256;
			for (int i = 0; i < 256; ++i)
				s[i] = base + i < g.V ? site[base + i] : 0.0f;
			for (int st = 128; st > 0; st >>= 1)
				for (int i = 0; i < st; ++i)
					if (base + i + st < g.V)
						s[i] += s[i + st];
			acc = b == 0 ? s[0] : acc + s[0];
		}
		return acc;
	};
	double sp = 0.0, sk = 0.0;
	for (int mu = 0; mu < nd; ++mu)
		for (int nu = mu + 1; nu < nd; ++nu)
		{
			for (int64_t x = 0; x < g.V; ++x)
			{
				int64_t xpm = g.shift(x, mu, 1), xpn = g.shift(x, nu, 1);
				mat3<float> A = mul_nn(ld(U, g.V, mu, x), ld(U, g.V, nu, xpm));
				mat3<float> B = mul_nn(ld(U, g.V, nu, x), ld(U, g.V, mu, xpn));
				site[x] = retr_nd(A, B);
			}
			sp += tree();
		}
	const double vol = (double)g.V;
	const double plaq = sp / (vol * nd * (nd - 1) * 0.5) / 3.0;
	out[0] = (1.0 - plaq) * (beta * nd * (nd - 1) * 0.5 * vol);
	for (int mu = 0; mu < nd && P; ++mu)
	{
		for (int64_t x = 0; x < g.V; ++x)
			site[x] = norm2_3(ld(P, g.V, mu, x));
		sk += tree();
	}
	out[1] = sk;
}

void orc_philox4x32_10(uint32_t const *ctr, uint32_t const *key, uint32_t *out)
{
	philox4x32_10(ctr, key, out);
}
double orc_det_log(double x) { return det_log(x); }
void orc_det_sincos2pi(double u, double *s, double *c) { det_sincos2pi(u, s, c); }
void orc_link_normals(uint64_t seed, uint32_t tag, uint32_t ctr2, uint32_t site, uint32_t mu, int n,
                      double *z)
{
	link_normals(seed, tag, ctr2, site, mu, n, z);
}
double orc_metropolis_uniform(uint64_t seed, uint32_t traj) { return metropolis_uniform(seed, traj); }

// Philox block of an arbitrary counter -> (u01 in (0,1], z0, z1) as used by
// the U(1) and flow streams
void orc_philox_draw(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, double *out3)
{
	uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
	uint32_t ctr[4] = {c0, c1, c2, c3};
	uint32_t r[4];
	philox4x32_10(ctr, key, r);
	out3[0] = u01_from_bits(r[0], r[1]);
	normal_pair(r, &out3[1], &out3[2]);
}

// element ops on 18-real matrices (double)
void orc_su3_mul_nn(double const *a, double const *b, double *c)
{
	mat3<double> A, B, C;
	std::memcpy(&A, a, sizeof A);
	std::memcpy(&B, b, sizeof B);
	C = mul_nn(A, B);
	std::memcpy(c, &C, sizeof C);
}
void orc_su3_ta(double const *m, double *out)
{
	mat3<double> M;
	std::memcpy(&M, m, sizeof M);
	M = ta3(M);
	std::memcpy(out, &M, sizeof M);
}
void orc_su3_exp(double const *x, double *out)
{
	mat3<double> X;
	std::memcpy(&X, x, sizeof X);
	X = expsu3(X);
	std::memcpy(out, &X, sizeof X);
}
void orc_su3_exp_f32(float const *x, float *out)
{
	mat3<float> X;
	std::memcpy(&X, x, sizeof X);
	X = expsu3(X);
	std::memcpy(out, &X, sizeof X);
}
void orc_su3_reunit(double const *a, double *out)
{
	mat3<double> A;
	std::memcpy(&A, a, sizeof A);
	A = reunit3(A);
	std::memcpy(out, &A, sizeof A);
}
void orc_su3_det(double const *a, double *out)
{
	mat3<double> A;
	std::memcpy(&A, a, sizeof A);
	cplx<double> d = det3(A);
	out[0] = d.re;
	out[1] = d.im;
}
void orc_su3_haar_from_normals(double const *z, double *out)
{
	mat3<double> A = haar_from_normals(z);
	std::memcpy(out, &A, sizeof A);
}
void orc_su3_algebra_from_normals(double const *p, double *out)
{
	mat3<double> A = algebra_from_normals(p);
	std::memcpy(out, &A, sizeof A);
}

// ---- lattice ops; suffix _f64 (double links) / _f32 (float links) ----

void orc_su3_random_gauge_f64(double *U, int const *dims, int nd, uint64_t seed)
{
	random_gauge(U, Geo(dims, nd), seed);
}
void orc_su3_random_gauge_f32(float *U, int const *dims, int nd, uint64_t seed)
{
	random_gauge(U, Geo(dims, nd), seed);
}
void orc_su3_momenta_f64(double *P, double *K, int const *dims, int nd, uint64_t seed, uint32_t traj)
{
	momenta(P, K, Geo(dims, nd), seed, traj);
}
void orc_su3_momenta_f32(float *P, double *K, int const *dims, int nd, uint64_t seed, uint32_t traj)
{
	momenta(P, K, Geo(dims, nd), seed, traj);
}
void orc_su3_plaq_sites_f64(double const *U, int const *dims, int nd, double *out)
{
	Geo g(dims, nd);
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < g.V; ++x)
		out[x] = plaq_site(U, g, x);
}
void orc_su3_plaq_sites_f32(float const *U, int const *dims, int nd, double *out)
{
	Geo g(dims, nd);
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < g.V; ++x)
		out[x] = plaq_site(U, g, x);
}
// normalised average plaquette in [0,1] (src/gauge/utils.h:93-108)
double orc_su3_plaquette_f64(double const *U, int const *dims, int nd)
{
	Geo g(dims, nd);
	std::vector<double> s(g.V);
	orc_su3_plaq_sites_f64(U, dims, nd, s.data());
	return (double)(lsum(s) / ((long double)g.V * g.nd * (g.nd - 1) / 2 * 3));
}
double orc_su3_plaquette_f32(float const *U, int const *dims, int nd)
{
	Geo g(dims, nd);
	std::vector<double> s(g.V);
	orc_su3_plaq_sites_f32(U, dims, nd, s.data());
	return (double)(lsum(s) / ((long double)g.V * g.nd * (g.nd - 1) / 2 * 3));
}
void orc_su3_staple_f64(double const *U, int const *dims, int nd, int mu, double *S)
{
	Geo g(dims, nd);
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < g.V; ++x)
		st(S, g.V, 0, x, staple_at(U, g, mu, x));
}
void orc_su3_staple_f32(float const *U, int const *dims, int nd, int mu, float *S)
{
	Geo g(dims, nd);
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < g.V; ++x)
		st(S, g.V, 0, x, staple_at(U, g, mu, x));
}
// wilsonDeriv (src/gauge/utils.h:148-153): F = (beta/6) TA(U_mu S_mu)
void orc_su3_force_f64(double const *U, int const *dims, int nd, int mu, double beta, double *F)
{
	Geo g(dims, nd);
	double c = beta / 6.0;
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < g.V; ++x)
	{
		mat3<double> A = ta3(mul_nn(ld(U, g.V, mu, x), staple_at(U, g, mu, x)));
		for (int i = 0; i < 3; ++i)
			for (int j = 0; j < 3; ++j)
			{
				A.e[i][j].re *= c;
				A.e[i][j].im *= c;
			}
		st(F, g.V, 0, x, A);
	}
}
void orc_su3_force_update_f64(double const *U, double *P, int const *dims, int nd, double beta, double delta)
{
	force_update(U, P, Geo(dims, nd), beta, delta);
}
void orc_su3_force_update_f32(float const *U, float *P, int const *dims, int nd, double beta, double delta)
{
	force_update(U, P, Geo(dims, nd), beta, delta);
}
void orc_su3_link_update_f64(double const *Uin, double *Uout, double const *P, int const *dims, int nd,
                             double t, int reunit)
{
	link_update(Uin, Uout, P, Geo(dims, nd), t, reunit);
}
void orc_su3_link_update_f32(float const *Uin, float *Uout, float const *P, int const *dims, int nd,
                             double t, int reunit)
{
	link_update(Uin, Uout, P, Geo(dims, nd), t, reunit);
}
void orc_su3_reunit_field_f64(double *U, int64_t nlinks)
{
#pragma omp parallel for schedule(static)
	for (int64_t i = 0; i < nlinks; ++i)
	{
		mat3<double> a;
		std::memcpy(&a, U + 18 * i, sizeof a);
		a = reunit3(a);
		std::memcpy(U + 18 * i, &a, sizeof a);
	}
}
void orc_su3_reunit_field_f32(float *U, int64_t nlinks)
{
#pragma omp parallel for schedule(static)
	for (int64_t i = 0; i < nlinks; ++i)
	{
		mat3<float> a;
		std::memcpy(&a, U + 18 * i, sizeof a);
		a = reunit3(a);
		std::memcpy(U + 18 * i, &a, sizeof a);
	}
}
// h = exp(t TA(F)) h for nlinks matrices (the gauge-transformation update
// exp(F * eps) * g of src/gauge/landau.h:131, with the same link evolution as
// the molecular dynamics)
void orc_su3_exp_mul_field_f64(double *h, double const *F, int64_t nlinks, double t)
{
#pragma omp parallel for schedule(static)
	for (int64_t i = 0; i < nlinks; ++i)
	{
		mat3<double> a, f;
		std::memcpy(&a, h + 18 * i, sizeof a);
		std::memcpy(&f, F + 18 * i, sizeof f);
		a = evolve_link(a, f, t);
		std::memcpy(h + 18 * i, &a, sizeof a);
	}
}
// y = y + (a x) per real, product and sum rounded separately as the
// reference's `F += last * beta` (src/lattice/lattice.h:302-309, 349-355)
void orc_axpy_f64(double *y, double a, double const *x, int64_t n)
{
	for (int64_t i = 0; i < n; ++i)
	{
		double p = x[i] * a;
		y[i] = y[i] + p;
	}
}
void orc_su3_kinetic_sites_f64(double const *P, int const *dims, int nd, double *K)
{
	Geo g(dims, nd);
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < g.V; ++x)
		K[x] = kin_site(P, g, x);
}
void orc_su3_kinetic_sites_f32(float const *P, int const *dims, int nd, double *K)
{
	Geo g(dims, nd);
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < g.V; ++x)
		K[x] = kin_site(P, g, x);
}
// out[5] = {Delta H, sum ps_old, sum ps_new, K_old, K_new}; per site
// d(x) = -(beta/3) (ps_new(x) - ps_old(x)) + (k_new(x) - k_old(x)) summed in
// long double, ps = sum of Re tr over the Nd(Nd-1)/2 plaquettes at x.
void orc_su3_delta_h_f64(double const *Uo, double const *Un, double const *P, double const *Ko,
                         int const *dims, int nd, double beta, double *out)
{
	delta_h(Uo, Un, P, Ko, Geo(dims, nd), beta, out);
}
void orc_su3_delta_h_f32(float const *Uo, float const *Un, float const *P, double const *Ko,
                         int const *dims, int nd, double beta, double *out)
{
	delta_h(Uo, Un, P, Ko, Geo(dims, nd), beta, out);
}
// out[6] = {Delta H, accepted, plaquette after, H_old, H_new, u};
// metropolis == 0 accepts unconditionally (thermalisation)
void orc_su3_hmc_trajectory_f64(double *U, int const *dims, int nd, double beta, double const *deltas,
                                int nde, uint64_t seed, uint32_t traj, int metropolis, double *out)
{
	hmc_trajectory(U, Geo(dims, nd), beta, deltas, nde, seed, traj, metropolis, out);
}
void orc_su3_hmc_trajectory_f32(float *U, int const *dims, int nd, double beta, double const *deltas,
                                int nde, uint64_t seed, uint32_t traj, int metropolis, double *out)
{
	hmc_trajectory(U, Geo(dims, nd), beta, deltas, nde, seed, traj, metropolis, out);
}

// Wall time of each stage of hmc_trajectory, run once on U (not modified):
// t[0] momenta with k(x), t[1] one momentum step, t[2] one link step, t[3]
// the last link step with reunitarisation and Delta H.  A trajectory with
// N_F momentum and N_L link steps takes t[0] + N_F t[1] + (N_L - 1) t[2] + t[3].
void orc_su3_stage_times_f64(double const *U, int const *dims, int nd, double beta, double delta, double *t)
{
	Geo g(dims, nd);
	size_t n = (size_t)g.nd * g.V * 18;
	std::vector<double> P(n), Un(n), K(g.V);
	auto now = [] { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); };
	double a = now();
	momenta(P.data(), K.data(), g, 1, 0);
	double b = now();
	force_update(U, P.data(), g, beta, delta);
	double c = now();
	link_update(U, Un.data(), P.data(), g, delta, 0);
	double d = now();
	link_update(U, Un.data(), P.data(), g, delta, 1);
	double dh[5];
	delta_h(U, Un.data(), P.data(), K.data(), g, beta, dh);
	double e = now();
	t[0] = b - a;
	t[1] = c - b;
	t[2] = d - c;
	t[3] = e - d;
}

// OpenMP threads of the oracle's passes
void orc_set_threads(int n) { omp_set_num_threads(n); }
int orc_max_threads() { return omp_get_max_threads(); }

double orc_landau_functional_f64(double const *U, double const *h, int const *dims, int nd)
{
	return landau_functional(U, h, Geo(dims, nd));
}
double orc_landau_functional_f32(float const *U, float const *h, int const *dims, int nd)
{
	return landau_functional(U, h, Geo(dims, nd));
}
double orc_landau_divsum_f64(double const *U, double const *h, int const *dims, int nd)
{
	return landau_divsum(U, h, Geo(dims, nd));
}
double orc_landau_divsum_f32(float const *U, float const *h, int const *dims, int nd)
{
	return landau_divsum(U, h, Geo(dims, nd));
}
void orc_landau_deriv_f64(double const *U, double const *h, int const *dims, int nd, double *F)
{
	landau_deriv(U, h, Geo(dims, nd), F);
}
void orc_landau_or_halfsweep_f64(double *U, double *h, int const *dims, int nd, int parity, double omega)
{
	landau_or_halfsweep(U, h, Geo(dims, nd), parity, omega);
}
void orc_landau_or_halfsweep_f32(float *U, float *h, int const *dims, int nd, int parity, double omega)
{
	landau_or_halfsweep(U, h, Geo(dims, nd), parity, omega);
}
// random gauge transformation field g(x) (Haar), tag TAG_GAUGE
void orc_su3_random_transform_f64(double *h, int const *dims, int nd, uint64_t seed)
{
	Geo g(dims, nd);
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < g.V; ++x)
	{
		double z[18];
		link_normals(seed, TAG_GAUGE, 0, (uint32_t)x, 0u, 18, z);
		st(h, g.V, 0, x, haar_from_normals(z));
	}
}
// U_mu(x) <- h(x) U_mu(x) h(x+mu)^+
void orc_su3_gauge_transform_f64(double *U, double const *h, int const *dims, int nd)
{
	Geo g(dims, nd);
	size_t n = (size_t)g.nd * g.V * 18;
	std::vector<double> out(n);
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < g.V; ++x)
		for (int mu = 0; mu < g.nd; ++mu)
			st(out.data(), g.V, mu, x, ug(U, h, g, mu, x));
	std::memcpy(U, out.data(), n * sizeof(double));
}

} // extern "C"
