// CPU oracle: SU(2) pure-gauge Wilson HMC (src/gauge/hmc.cpp:75-114 with the
// SU(2) operations of src/gauge/groups.h:110-192), restated per site with the
// operation sequences of the GPU path, so that both precisions agree with it
// bit for bit.
//
// TEST INFRASTRUCTURE ONLY: part of oracle/liboracle_mesh.so, loaded by
// tests/ through oracle/oracle.py.
//
// A link is a unit quaternion (w, x, y, z) = w + x I + y J + z K, the 2x2
// matrix w + i (x, y, z).sigma up to orientation; products are Hamilton
// products, adj the conjugate, Re tr = 2 w.  Host layout U[Nd][V][4], sites
// row-major with dimension 0 slowest.  Momenta (0, p_k / 2), p_k ~ N(0,1)
// (groups.h:171-181), kinetic term 2 (x^2 + y^2 + z^2): the norm that the
// force (beta / 4) project_on_algebra(U Sigma) conserves.
#include "oracle_su3.h"

#include <cmath>
#include <cstring>
#include <vector>

using namespace orc;

namespace {

template <class R> struct qt
{
	R w, x, y, z;
};

template <class R> inline qt<R> qmul(qt<R> a, qt<R> b)
{
	qt<R> r;
	r.w = std::fma(-a.z, b.z, std::fma(-a.y, b.y, std::fma(-a.x, b.x, a.w * b.w)));
	r.x = std::fma(-a.z, b.y, std::fma(a.y, b.z, std::fma(a.x, b.w, a.w * b.x)));
	r.y = std::fma(a.z, b.x, std::fma(a.y, b.w, std::fma(-a.x, b.z, a.w * b.y)));
	r.z = std::fma(a.z, b.w, std::fma(-a.y, b.x, std::fma(a.x, b.y, a.w * b.z)));
	return r;
}
template <class R> inline qt<R> qconj(qt<R> a) { return {a.w, -a.x, -a.y, -a.z}; }
template <class R> inline qt<R> qadd(qt<R> a, qt<R> b) { return {a.w + b.w, a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> inline qt<double> q64(qt<R> a) { return {(double)a.w, (double)a.x, (double)a.y, (double)a.z}; }
template <class R> inline qt<R> qround(qt<double> a) { return {(R)a.w, (R)a.x, (R)a.y, (R)a.z}; }

// exp(X) - 1 of X = (0, v): (cos alpha - 1, v sin(alpha)/alpha) as series in
// alpha^2 to (alpha^2)^11, argument halved while alpha^2 > 1 and squared back
template <class R> inline qt<R> q_expm1(R vx, R vy, R vz)
{
	R a2 = std::fma(vz, vz, std::fma(vy, vy, vx * vx));
	int s = 0;
	R sc = R(1);
	while (a2 > R(1) && s < 64)
	{
		a2 *= R(0.25);
		sc *= R(0.5);
		++s;
	}
	vx *= sc;
	vy *= sc;
	vz *= sc;
	R cm1 = R(0), sn = R(0);
	for (int n = 11; n >= 1; --n)
	{
		const R c2n = R(1) / (R)((2 * n) * (2 * n - 1));
		const R c2n1 = R(1) / (R)((2 * n + 1) * (2 * n));
		cm1 = std::fma(-a2 * c2n, cm1, -a2 * c2n);
		sn = std::fma(-a2 * c2n1, sn, -a2 * c2n1);
	}
	qt<R> D = {cm1, vx + vx * sn, vy + vy * sn, vz + vz * sn};
	for (int k = 0; k < s; ++k)
	{
		qt<R> D2 = qmul(D, D);
		D = {std::fma(R(2), D.w, D2.w), std::fma(R(2), D.x, D2.x), std::fma(R(2), D.y, D2.y),
		     std::fma(R(2), D.z, D2.z)};
	}
	return D;
}

// project_on_group_fast (groups.h:185-189) as a + a (1 - |a|^2)/2, in FP64
template <class R> inline qt<R> q_reunit(qt<R> a0)
{
	qt<double> a = q64(a0);
	double e = std::fma(-a.z, a.z, std::fma(-a.y, a.y, std::fma(-a.x, a.x, std::fma(-a.w, a.w, 1.0))));
	double h = 0.5 * e;
	return qround<R>({std::fma(a.w, h, a.w), std::fma(a.x, h, a.x), std::fma(a.y, h, a.y), std::fma(a.z, h, a.z)});
}

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
		int n = c + s;
		n = n < 0 ? n + L[mu] : (n >= L[mu] ? n - L[mu] : n);
		return x + (int64_t)(n - c) * stride[mu];
	}
};

template <class R> inline qt<R> ld(R const *U, Geo const &g, int mu, int64_t x)
{
	R const *p = U + ((size_t)mu * g.V + x) * 4;
	return {p[0], p[1], p[2], p[3]};
}
template <class R> inline void st(R *U, Geo const &g, int mu, int64_t x, qt<R> q)
{
	R *p = U + ((size_t)mu * g.V + x) * 4;
	p[0] = q.w;
	p[1] = q.x;
	p[2] = q.y;
	p[3] = q.z;
}

template <class R> qt<R> staple(R const *U, Geo const &g, int mu, int64_t x)
{
	qt<R> S = {R(0), R(0), R(0), R(0)};
	const int64_t xpm = g.shift(x, mu, 1);
	for (int nu = 0; nu < g.nd; ++nu)
	{
		if (nu == mu)
			continue;
		const int64_t xpn = g.shift(x, nu, 1), xmn = g.shift(x, nu, -1), xpmmn = g.shift(xpm, nu, -1);
		qt<R> up = qmul(qmul(ld(U, g, nu, xpm), qconj(ld(U, g, mu, xpn))), qconj(ld(U, g, nu, x)));
		qt<R> lo = qmul(qmul(qconj(ld(U, g, nu, xpmmn)), qconj(ld(U, g, mu, xmn))), ld(U, g, nu, xmn));
		S = qadd(qadd(S, up), lo);
	}
	return S;
}

template <class R> double plaq_site(R const *U, Geo const &g, int64_t x)
{
	double s = 0.0;
	for (int mu = 0; mu < g.nd; ++mu)
		for (int nu = mu + 1; nu < g.nd; ++nu)
		{
			qt<double> A = qmul(q64(ld(U, g, mu, x)), q64(ld(U, g, nu, g.shift(x, mu, 1))));
			qt<double> B = qmul(q64(ld(U, g, nu, x)), q64(ld(U, g, mu, g.shift(x, nu, 1))));
			s += std::fma(A.z, B.z, std::fma(A.y, B.y, std::fma(A.x, B.x, A.w * B.w)));
		}
	return s;
}

template <class R> double kin(qt<R> p)
{
	qt<double> q = q64(p);
	return 2.0 * std::fma(q.z, q.z, std::fma(q.y, q.y, q.x * q.x));
}

template <class R> void random_gauge(R *U, Geo const &g, uint64_t seed)
{
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < g.V; ++x)
		for (int mu = 0; mu < g.nd; ++mu)
		{
			double z[4];
			link_normals(seed, TAG_HAAR, 0, (uint32_t)x, (uint32_t)mu, 4, z);
			double n = std::sqrt(std::fma(z[3], z[3], std::fma(z[2], z[2], std::fma(z[1], z[1], z[0] * z[0]))));
			st(U, g, mu, x, qround<R>({z[0] / n, z[1] / n, z[2] / n, z[3] / n}));
		}
}

template <class R> void momenta(R *P, double *K, Geo const &g, uint64_t seed, uint32_t traj)
{
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < g.V; ++x)
	{
		double k = 0.0;
		for (int mu = 0; mu < g.nd; ++mu)
		{
			double z[4];
			link_normals(seed, TAG_MOM, traj, (uint32_t)x, (uint32_t)mu, 4, z);
			qt<R> p = qround<R>({0.0, 0.5 * z[0], 0.5 * z[1], 0.5 * z[2]});
			st(P, g, mu, x, p);
			k += kin(p);
		}
		if (K)
			K[x] = k;
	}
}

template <class R> void force_update(R const *U, R *P, Geo const &g, double beta, double delta)
{
	const double k = delta * (beta / 4.0);
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < g.V; ++x)
		for (int mu = 0; mu < g.nd; ++mu)
		{
			qt<double> M = q64(qmul(ld(U, g, mu, x), staple(U, g, mu, x)));
			qt<double> p = q64(ld(P, g, mu, x));
			st(P, g, mu, x, qround<R>({p.w, std::fma(-k, M.x, p.x), std::fma(-k, M.y, p.y), std::fma(-k, M.z, p.z)}));
		}
}

template <class R> void link_update(R const *Uin, R *Uout, R const *P, Geo const &g, double t, int reunit)
{
	const R tt = (R)t;
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < g.V; ++x)
		for (int mu = 0; mu < g.nd; ++mu)
		{
			qt<R> p = ld(P, g, mu, x), u = ld(Uin, g, mu, x);
			qt<R> D = q_expm1(p.x * tt, p.y * tt, p.z * tt);
			qt<R> un = qadd(u, qmul(D, u));
			if (reunit)
				un = q_reunit(un);
			st(Uout, g, mu, x, un);
		}
}

long double lsum(std::vector<double> const &v)
{
	long double s = 0;
	for (double d : v)
		s += d;
	return s;
}

template <class R>
void hmc_trajectory(R *U, Geo const &g, double beta, double const *deltas, int nde, uint64_t seed, uint32_t traj,
                    int metropolis, double *out)
{
	size_t n = (size_t)g.nd * g.V * 4;
	std::vector<R> P(n), Un(n);
	std::vector<double> K(g.V);
	momenta(P.data(), K.data(), g, seed, traj);
	for (int i = 0; i < nde; ++i)
	{
		if (i % 2 == 0)
			link_update(i == 0 ? U : Un.data(), Un.data(), P.data(), g, deltas[i], i == nde - 1);
		else
			force_update(Un.data(), P.data(), g, beta, deltas[i]);
	}
	std::vector<double> d(g.V), pso(g.V), psn(g.V), kn(g.V);
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < g.V; ++x)
	{
		pso[x] = plaq_site(U, g, x);
		psn[x] = plaq_site(Un.data(), g, x);
		double k = 0.0;
		for (int mu = 0; mu < g.nd; ++mu)
			k += kin(ld(P.data(), g, mu, x));
		kn[x] = k;
		d[x] = std::fma(-beta, psn[x] - pso[x], k - K[x]);
	}
	double dH = (double)lsum(d), so = (double)lsum(pso), sn = (double)lsum(psn), ko = (double)lsum(K),
	       kk = (double)lsum(kn);
	uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
	uint32_t ctr[4] = {traj, 0u, 0u, TAG_METRO};
	uint32_t r[4];
	philox4x32_10(ctr, key, r);
	double u = u01_from_bits(r[0], r[1]);
	int acc = metropolis ? (u < std::exp(-dH)) : 1;
	if (acc)
		std::memcpy(U, Un.data(), n * sizeof(R));
	double np = (double)g.V * g.nd * (g.nd - 1) / 2;
	out[0] = dH;
	out[1] = acc;
	out[2] = (acc ? sn : so) / np;
	out[3] = beta * (np - so) + ko;
	out[4] = beta * (np - sn) + kk;
	out[5] = u;
}

template <class R> double plaquette(R const *U, Geo const &g)
{
	std::vector<double> ps(g.V);
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < g.V; ++x)
		ps[x] = plaq_site(U, g, x);
	return (double)lsum(ps) / ((double)g.V * g.nd * (g.nd - 1) / 2);
}

// ---- Landau gauge (src/gauge/landau.h for SU2<T>); h[V][4], h == NULL is h = 1

// h(x) d(x), d(x) = sum_mu U_mu(x) h(x+mu)^+ + (h(x-mu) U_mu(x-mu))^+ (landau.h:41-46)
template <class R> qt<R> landau_gd(R const *U, R const *h, Geo const &g, int64_t x)
{
	qt<R> d = {R(0), R(0), R(0), R(0)};
	for (int mu = 0; mu < g.nd; ++mu)
	{
		const int64_t xp = g.shift(x, mu, 1), xm = g.shift(x, mu, -1);
		qt<R> a = ld(U, g, mu, x), b = ld(U, g, mu, xm);
		if (h)
		{
			a = qmul(a, qconj(ld(h, g, 0, xp)));
			b = qmul(ld(h, g, 0, xm), b);
		}
		d = qadd(qadd(d, a), qconj(b));
	}
	return h ? qmul(ld(h, g, 0, x), d) : d;
}

// F[h] = sum Re tr (h U h(x+mu)^+) / (Nc Nd V) (landau.h:23-34)
template <class R> double landau_functional(R const *U, R const *h, Geo const &g)
{
	std::vector<double> v(g.V, 0.0);
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < g.V; ++x)
		for (int mu = 0; mu < g.nd; ++mu)
		{
			if (!h)
			{
				v[x] += (double)ld(U, g, mu, x).w;
				continue;
			}
			qt<double> A = q64(qmul(ld(h, g, 0, x), ld(U, g, mu, x)));
			qt<double> B = q64(ld(h, g, 0, g.shift(x, mu, 1)));
			v[x] += std::fma(A.z, B.z, std::fma(A.y, B.y, std::fma(A.x, B.x, A.w * B.w)));
		}
	return (double)lsum(v) / ((double)g.nd * g.V);
}

// sum_x |project_on_algebra(h d)|^2 (landau.h:37-49 without the 1/(Nd V))
template <class R> double landau_divsum(R const *U, R const *h, Geo const &g)
{
	std::vector<double> v(g.V);
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < g.V; ++x)
	{
		qt<double> q = q64(landau_gd(U, h, g, x));
		v[x] = std::fma(q.z, q.z, std::fma(q.y, q.y, q.x * q.x));
	}
	return (double)lsum(v);
}

// -0.5 / (Nc Nd V) project_on_algebra(h d) (landau.h:52-65)
template <class R> void landau_deriv(R const *U, R const *h, Geo const &g, R *F)
{
	const double c = -0.5 / (2.0 * g.nd * (double)g.V);
#pragma omp parallel for schedule(static)
	for (int64_t x = 0; x < g.V; ++x)
	{
		qt<double> q = q64(landau_gd(U, h, g, x));
		st(F, g, 0, x, qround<R>({0.0, c * q.x, c * q.y, c * q.z}));
	}
}

template <class R> void gauge_transform(R const *U, R const *h, Geo const &g, R *Uo)
{
	for (int mu = 0; mu < g.nd; ++mu)
		for (int64_t x = 0; x < g.V; ++x)
			st(Uo, g, mu, x, qmul(qmul(ld(h, g, 0, x), ld(U, g, mu, x)), qconj(ld(h, g, 0, g.shift(x, mu, 1)))));
}

// r = conj(K)/|K|, then 1 + omega (r - 1) + omega (omega - 1)/2 (r - 1)^2 normalised
template <class R> qt<R> or_step(qt<R> K, R omega)
{
	R n2 = std::fma(K.z, K.z, std::fma(K.y, K.y, std::fma(K.x, K.x, K.w * K.w)));
	if (!(n2 > R(0)))
		return {R(1), R(0), R(0), R(0)};
	R inv = R(1) / std::sqrt(n2);
	R r0 = K.w * inv, r1 = -(K.x * inv), r2 = -(K.y * inv), r3 = -(K.z * inv);
	R a = r0 - R(1);
	R rv2 = std::fma(r3, r3, std::fma(r2, r2, r1 * r1));
	const R c2 = R(0.5) * omega * (omega - R(1));
	R g0 = std::fma(c2, std::fma(a, a, -rv2), std::fma(omega, a, R(1)));
	R fac = std::fma(R(2) * c2, a, omega);
	R g1 = r1 * fac, g2 = r2 * fac, g3 = r3 * fac;
	R gi = R(1) / std::sqrt(std::fma(g3, g3, std::fma(g2, g2, std::fma(g1, g1, g0 * g0))));
	return {g0 * gi, g1 * gi, g2 * gi, g3 * gi};
}

// overrelaxation of the sites with (sum of coordinates + parity) even, on the
// transformed links; sites of one parity share no link, so any order is exact
template <class R> void or_half(R *U, R *h, Geo const &g, int parity, R omega)
{
	for (int64_t x = 0; x < g.V; ++x)
	{
		int cs = 0;
		for (int mu = 0; mu < g.nd; ++mu)
			cs += (int)((x / g.stride[mu]) % g.L[mu]);
		if ((cs & 1) != parity)
			continue;
		qt<R> K = {R(0), R(0), R(0), R(0)};
		for (int mu = 0; mu < g.nd; ++mu)
			K = qadd(qadd(K, ld(U, g, mu, x)), qconj(ld(U, g, mu, g.shift(x, mu, -1))));
		qt<R> G = or_step(K, omega);
		for (int mu = 0; mu < g.nd; ++mu)
		{
			const int64_t xm = g.shift(x, mu, -1);
			st(U, g, mu, x, qmul(G, ld(U, g, mu, x)));
			st(U, g, mu, xm, qmul(ld(U, g, mu, xm), qconj(G)));
		}
		if (h)
			st(h, g, 0, x, qmul(G, ld(h, g, 0, x)));
	}
}

} // namespace

extern "C" {

#define ORC_SU2(SFX, R)                                                                                           \
	void orc_su2_random_##SFX(R *U, int const *dims, int nd, uint64_t seed) { random_gauge(U, Geo(dims, nd), seed); } \
	void orc_su2_momenta_##SFX(R *P, double *K, int const *dims, int nd, uint64_t seed, uint32_t traj)          \
	{                                                                                                             \
		momenta(P, K, Geo(dims, nd), seed, traj);                                                                 \
	}                                                                                                             \
	void orc_su2_force_update_##SFX(R const *U, R *P, int const *dims, int nd, double beta, double delta)       \
	{                                                                                                             \
		force_update(U, P, Geo(dims, nd), beta, delta);                                                           \
	}                                                                                                             \
	void orc_su2_link_update_##SFX(R const *Uin, R *Uout, R const *P, int const *dims, int nd, double t,        \
	                               int reunit)                                                                    \
	{                                                                                                             \
		link_update(Uin, Uout, P, Geo(dims, nd), t, reunit);                                                      \
	}                                                                                                             \
	double orc_su2_plaquette_##SFX(R const *U, int const *dims, int nd) { return plaquette(U, Geo(dims, nd)); }   \
	void orc_su2_staple_##SFX(R const *U, int const *dims, int nd, int mu, R *S)                                 \
	{                                                                                                             \
		Geo g(dims, nd);                                                                                          \
		for (int64_t x = 0; x < g.V; ++x)                                                                         \
			st(S, g, 0, x, staple(U, g, mu, x));                                                                  \
	}                                                                                                             \
	void orc_su2_hmc_trajectory_##SFX(R *U, int const *dims, int nd, double beta, double const *deltas, int nde, \
	                                  uint64_t seed, uint32_t traj, int metropolis, double *out)                  \
	{                                                                                                             \
		hmc_trajectory(U, Geo(dims, nd), beta, deltas, nde, seed, traj, metropolis, out);                         \
	}

#define ORC_SU2_LANDAU(SFX, R)                                                                                    \
	double orc_su2_landau_functional_##SFX(R const *U, R const *h, int const *dims, int nd)                      \
	{                                                                                                             \
		return landau_functional(U, h, Geo(dims, nd));                                                           \
	}                                                                                                             \
	double orc_su2_landau_divsum_##SFX(R const *U, R const *h, int const *dims, int nd)                          \
	{                                                                                                             \
		return landau_divsum(U, h, Geo(dims, nd));                                                               \
	}                                                                                                             \
	void orc_su2_landau_deriv_##SFX(R const *U, R const *h, int const *dims, int nd, R *F)                       \
	{                                                                                                             \
		landau_deriv(U, h, Geo(dims, nd), F);                                                                    \
	}                                                                                                             \
	void orc_su2_gauge_transform_##SFX(R const *U, R const *h, int const *dims, int nd, R *Uo)                   \
	{                                                                                                             \
		gauge_transform(U, h, Geo(dims, nd), Uo);                                                                \
	}                                                                                                             \
	void orc_su2_landau_or_half_##SFX(R *U, R *h, int const *dims, int nd, int parity, double omega)             \
	{                                                                                                             \
		or_half(U, h, Geo(dims, nd), parity, (R)omega);                                                          \
	}

ORC_SU2_LANDAU(f64, double)
ORC_SU2_LANDAU(f32, float)
#undef ORC_SU2_LANDAU

ORC_SU2(f64, double)
ORC_SU2(f32, float)
#undef ORC_SU2

// one quaternion product and exp(X) - 1 of an algebra element (element tests)
void orc_su2_qmul(double const *a, double const *b, double *out)
{
	qt<double> r = qmul(qt<double>{a[0], a[1], a[2], a[3]}, qt<double>{b[0], b[1], b[2], b[3]});
	out[0] = r.w;
	out[1] = r.x;
	out[2] = r.y;
	out[3] = r.z;
}
void orc_su2_expm1(double const *v, double *out)
{
	qt<double> r = q_expm1(v[0], v[1], v[2]);
	out[0] = r.w;
	out[1] = r.x;
	out[2] = r.y;
	out[3] = r.z;
}

} // extern "C"
