// SU(2) pure-gauge Wilson-action HMC for gfx950: the reference's
// Hmc<SU2<T>> (src/gauge/hmc.cpp:75-114) with the group operations of
// src/gauge/groups.h:110-192.
//
// A link is a unit quaternion q = (q0, q1, q2, q3), the 2x2 matrix
// q0 - i q_k sigma_k, multiplied as quaternions (Hamilton product); adj is
// the conjugate, Re tr = 2 q0, project_on_algebra keeps (0, q1, q2, q3).
//
// Landau gauge fixing (src/gauge/landau.h, instantiated for SU(2) by
// src/tests/lattice_gauge.cpp:46-53): functional, condition, derivative and
// gauge transformation of the SU(3) path on quaternions, and an even/odd
// overrelaxation half-sweep whose local step is G = (conj(K) / |K|)^omega,
// K(x) = sum_mu U_mu(x) + U_mu(x - mu)^+ of the transformed links.
// Momenta are algebra elements (0, p_k / 2) with p_k ~ N(0,1)
// (random_algebra_element, groups.h:171-181).  The kinetic term is their
// Frobenius norm as 2x2 matrices, 2 (q1^2 + q2^2 + q3^2) = sum p_k^2 / 2: the
// value for which the molecular dynamics with the force
// (beta / (2 Nc)) project_on_algebra(U Sigma) conserves H and whose
// Boltzmann weight those momenta sample (the reference adds the quaternion's
// own norm q1^2 + q2^2 + q3^2, half of it).
//
// Layout: per direction f the sites' quaternions in site order, component c
// of site x at base[4 (f stride + x) + c], stride = V rounded up to 256 (a
// quaternion is one 16-byte vector load in FP32, two in FP64; lanes of a
// wavefront read consecutive quaternions).
//
// Every multiply-add is an explicit fma in a fixed order and cos / sin of the
// exponential are power series in alpha^2 (no library transcendentals), so
// the CPU restatement reproduces these kernels bit for bit in both
// precisions.  The FP32 path keeps the FP32 lessons of the SU(3) path: the
// link step adds an increment to U, the projection onto the group and the
// momentum update run in FP64 and are rounded once.
#include "mesh_internal.h"
#include "rng_dev.h"

#include <cmath>
#include <cstdint>
#include <initializer_list>

// fused momentum + link step with the block's links staged in LDS
// (k_su2_force_link_rows) where the geometry allows it; 0: always the
// thread-per-link kernel
#ifndef MESH_SU2_ROWS
#define MESH_SU2_ROWS 1
#endif

using namespace mesh_dev;

namespace {

template <class R> struct qt
{
	R w, x, y, z; // w + x I + y J + z K
};

template <class R> __device__ __forceinline__ qt<R> qmul(qt<R> a, qt<R> b)
{
	qt<R> r;
	r.w = fma(-a.z, b.z, fma(-a.y, b.y, fma(-a.x, b.x, a.w * b.w)));
	r.x = fma(-a.z, b.y, fma(a.y, b.z, fma(a.x, b.w, a.w * b.x)));
	r.y = fma(a.z, b.x, fma(a.y, b.w, fma(-a.x, b.z, a.w * b.y)));
	r.z = fma(a.z, b.w, fma(-a.y, b.x, fma(a.x, b.y, a.w * b.z)));
	return r;
}
template <class R> __device__ __forceinline__ qt<R> qconj(qt<R> a) { return {a.w, -a.x, -a.y, -a.z}; }
template <class R> __device__ __forceinline__ qt<R> qadd(qt<R> a, qt<R> b)
{
	return {a.w + b.w, a.x + b.x, a.y + b.y, a.z + b.z};
}
template <class R> __device__ __forceinline__ qt<double> q64(qt<R> a)
{
	return {(double)a.w, (double)a.x, (double)a.y, (double)a.z};
}
template <class R> __device__ __forceinline__ qt<R> qround(qt<double> a) { return {(R)a.w, (R)a.x, (R)a.y, (R)a.z}; }

// exp(X) - 1 of an algebra element X = (0, v), alpha = |v|:
// (cos alpha - 1, v sin(alpha) / alpha), both as series in alpha^2 (the
// reference's sin(alpha) / alpha is NaN at alpha = 0, src/gauge/groups.h:
// 159-169); for alpha^2 > 1 the argument is halved and the result squared
// back as D -> 2 D + D^2.  Terms to (alpha^2)^11 / 23!: below 2e-23 at
// alpha = 1.
template <class R> __device__ __forceinline__ qt<R> q_expm1(R vx, R vy, R vz)
{
	R a2 = fma(vz, vz, fma(vy, vy, vx * vx));
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
	// cm1 = sum_{n>=1} (-a2)^n / (2n)!, sinc = sum_{n>=0} (-a2)^n / (2n+1)!, Horner from the top
	R cm1 = R(0), sn = R(0);
#pragma unroll
	for (int n = 11; n >= 1; --n)
	{
		const R c2n = R(1) / (R)((2 * n) * (2 * n - 1));
		const R c2n1 = R(1) / (R)((2 * n + 1) * (2 * n));
		cm1 = fma(-a2 * c2n, cm1, -a2 * c2n);
		sn = fma(-a2 * c2n1, sn, -a2 * c2n1);
	}
	qt<R> D = {cm1, vx + vx * sn, vy + vy * sn, vz + vz * sn};
	for (int k = 0; k < s; ++k)
	{
		qt<R> D2 = qmul(D, D);
		D = {fma(R(2), D.w, D2.w), fma(R(2), D.x, D2.x), fma(R(2), D.y, D2.y), fma(R(2), D.z, D2.z)};
	}
	return D;
}

// project_on_group_fast (src/gauge/groups.h:185-189), a (1.5 - 0.5 |a|^2), as
// a + a (1 - |a|^2) / 2; FP32 links in FP64, rounded once
template <class R> __device__ __forceinline__ qt<R> q_reunit(qt<R> a0)
{
	qt<double> a = q64(a0);
	double e = fma(-a.z, a.z, fma(-a.y, a.y, fma(-a.x, a.x, fma(-a.w, a.w, 1.0))));
	double h = 0.5 * e;
	return qround<R>({fma(a.w, h, a.w), fma(a.x, h, a.x), fma(a.y, h, a.y), fma(a.z, h, a.z)});
}

struct G2
{
	int nd;
	int L[4], st[4];
	long V, stride;
};

// neighbour x + s mu^ (|s| <= L_mu); 32-bit arithmetic (V <= 2^26 sites): a
// 64-bit division is a long emulated sequence on the GPU
__device__ __forceinline__ long nbr(G2 const &g, long x, int mu, int s)
{
	const unsigned xs = (unsigned)x;
	int c = (int)((xs / (unsigned)g.st[mu]) % (unsigned)g.L[mu]);
	int n = c + s;
	n = n < 0 ? n + g.L[mu] : (n >= g.L[mu] ? n - g.L[mu] : n);
	return (long)((int)xs + (n - c) * g.st[mu]);
}

// coordinates of site x (c[mu] = (x / st[mu]) % L[mu], 0 beyond nd), with
// one division per direction, and the neighbour of x along mu from its known
// coordinate c = c[mu]: nbr's site without nbr's division
__device__ __forceinline__ void coords2(G2 const &g, long x, int (&c)[4])
{
	unsigned t = (unsigned)x;
#pragma unroll
	for (int mu = 3; mu >= 0; --mu)
	{
		if (mu < g.nd)
		{
			const unsigned q = t / (unsigned)g.L[mu];
			c[mu] = (int)(t - q * (unsigned)g.L[mu]);
			t = q;
		}
		else
			c[mu] = 0;
	}
}
__device__ __forceinline__ long shc(G2 const &g, long x, int mu, int c, int s)
{
	int n = c + s;
	n = n < 0 ? n + g.L[mu] : (n >= g.L[mu] ? n - g.L[mu] : n);
	return (long)((int)x + (n - c) * g.st[mu]);
}

// quaternion (w, x, y, z) of direction f at site x: four consecutive reals,
// one vector load (two for FP64)
template <class R> using v4 = R __attribute__((ext_vector_type(4)));
template <class R> __device__ __forceinline__ qt<R> ldq(const R *__restrict__ b, G2 const &g, int f, long x)
{
	const v4<R> v = *reinterpret_cast<const v4<R> *>(b + 4 * ((long)f * g.stride + x));
	return {v.x, v.y, v.z, v.w};
}
template <class R> __device__ __forceinline__ void stq(R *__restrict__ b, G2 const &g, int f, long x, qt<R> q)
{
	v4<R> v = {q.w, q.x, q.y, q.z};
	*reinterpret_cast<v4<R> *>(b + 4 * ((long)f * g.stride + x)) = v;
}

// staple sum of direction mu at x (src/gauge/utils.h:110-139), planes in
// ascending nu: U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+ + U_nu(x-nu+mu)^+ U_mu(x-nu)^+ U_nu(x-nu)
template <class R> __device__ __forceinline__ qt<R> staple(G2 const &g, const R *__restrict__ U, int mu, long x)
{
	qt<R> S = {R(0), R(0), R(0), R(0)};
	int c[4];
	coords2(g, x, c);
	const long xpm = shc(g, x, mu, mu == 0 ? c[0] : (mu == 1 ? c[1] : (mu == 2 ? c[2] : c[3])), 1);
#pragma unroll
	for (int nu = 0; nu < 4; ++nu)
	{
		if (nu == mu || nu >= g.nd)
			continue;
		// x + mu has x's coordinate along nu
		const long xpn = shc(g, x, nu, c[nu], 1), xmn = shc(g, x, nu, c[nu], -1), xpmmn = shc(g, xpm, nu, c[nu], -1);
		qt<R> up = qmul(qmul(ldq(U, g, nu, xpm), qconj(ldq(U, g, mu, xpn))), qconj(ldq(U, g, nu, x)));
		qt<R> lo = qmul(qmul(qconj(ldq(U, g, nu, xpmmn)), qconj(ldq(U, g, mu, xmn))), ldq(U, g, nu, xmn));
		S = qadd(qadd(S, up), lo);
	}
	return S;
}

// sum over planes mu < nu of Re P_{mu nu}(x) (the quaternion's real part, Re tr / 2), in FP64;
// c: the coordinates of x (coords2)
template <class R>
__device__ __forceinline__ double plaq_site(G2 const &g, const R *__restrict__ U, long x, int const (&c)[4])
{
	double s = 0.0;
#pragma unroll
	for (int mu = 0; mu < 4; ++mu)
#pragma unroll
		for (int nu = mu + 1; nu < 4; ++nu)
		{
			if (nu >= g.nd)
				continue;
			qt<double> A = qmul(q64(ldq(U, g, mu, x)), q64(ldq(U, g, nu, shc(g, x, mu, c[mu], 1))));
			qt<double> B = qmul(q64(ldq(U, g, nu, x)), q64(ldq(U, g, mu, shc(g, x, nu, c[nu], 1))));
			// Re (A B^+) = A . B
			s += fma(A.z, B.z, fma(A.y, B.y, fma(A.x, B.x, A.w * B.w)));
		}
	return s;
}

template <class R> __device__ __forceinline__ double kin(qt<R> p)
{
	qt<double> q = q64(p);
	return 2.0 * fma(q.z, q.z, fma(q.y, q.y, q.x * q.x));
}

template <class R> __global__ __launch_bounds__(256) void k_su2_random(G2 g, int nf, R *U, uint64_t seed)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= g.V * nf)
		return;
	int f = (int)(i / g.V);
	long x = i - (long)f * g.V;
	double z[4];
	link_normals<4>(seed, TAG_HAAR, 0u, (uint32_t)x, (uint32_t)f, z);
	double n = sqrt(fma(z[3], z[3], fma(z[2], z[2], fma(z[1], z[1], z[0] * z[0]))));
	stq(U, g, f, x, qround<R>({z[0] / n, z[1] / n, z[2] / n, z[3] / n}));
}

template <class R> __global__ __launch_bounds__(256) void k_su2_unit(G2 g, int nf, R *U)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= g.V * nf)
		return;
	int f = (int)(i / g.V);
	stq(U, g, f, i - (long)f * g.V, qt<R>{R(1), R(0), R(0), R(0)});
}

template <class R>
__global__ __launch_bounds__(256) void k_su2_momenta(G2 g, R *P, double *K, uint64_t seed, uint32_t traj)
{
	long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (x >= g.V)
		return;
	double k = 0.0;
	for (int mu = 0; mu < g.nd; ++mu)
	{
		double z[4];
		link_normals<4>(seed, TAG_MOM, traj, (uint32_t)x, (uint32_t)mu, z);
		qt<R> p = qround<R>({0.0, 0.5 * z[0], 0.5 * z[1], 0.5 * z[2]});
		stq(P, g, mu, x, p);
		k += kin(p);
	}
	if (K)
		K[x] = k;
}

// P_mu(x) -= k project_on_algebra(U_mu(x) Sigma_mu(x)), k = delta beta / 4;
// the projection and the update in FP64, P rounded once
template <class R>
__global__ __launch_bounds__(256) void k_su2_force_update(G2 g, const R *__restrict__ U, R *__restrict__ P, double k)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= g.V * g.nd)
		return;
	int mu = (int)(i / g.V);
	long x = i - (long)mu * g.V;
	qt<double> M = q64(qmul(ldq(U, g, mu, x), staple(g, U, mu, x)));
	qt<double> p = q64(ldq(P, g, mu, x));
	stq(P, g, mu, x, qround<R>({p.w, fma(-k, M.x, p.x), fma(-k, M.y, p.y), fma(-k, M.z, p.z)}));
}

// the momentum step of link (mu, x) followed in the same thread by its link
// step Unext_mu(x) = exp(t P_new) U_mu(x) (+ reunitarisation): the arithmetic
// of k_su2_force_update and k_su2_link_update, one pass over the field
// (Unext != U: neighbouring threads still read U)
template <class R>
__global__ __launch_bounds__(256) void k_su2_force_link(G2 g, const R *__restrict__ U, R *__restrict__ P, double k,
                                                        R *__restrict__ Unext, R t, int reunit)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= g.V * g.nd)
		return;
	int mu = (int)((unsigned)i / (unsigned)g.V);
	long x = i - (long)mu * g.V;
	qt<R> u = ldq(U, g, mu, x);
	qt<double> M = q64(qmul(u, staple(g, U, mu, x)));
	qt<double> p = q64(ldq(P, g, mu, x));
	qt<R> pn = qround<R>({p.w, fma(-k, M.x, p.x), fma(-k, M.y, p.y), fma(-k, M.z, p.z)});
	stq(P, g, mu, x, pn);
	qt<R> D = q_expm1(pn.x * t, pn.y * t, pn.z * t);
	qt<R> un = qadd(u, qmul(D, u));
	if (reunit)
		un = q_reunit(un);
	stq(Unext, g, mu, x, un);
}

// k_su2_force_link for 4-d lattices whose x3-rows tile the 64-site blocks
// (V % 64 == 0, 64 % L3 == 0): a workgroup takes one block's 64 sites in
// their four directions (wavefront mu = direction mu), stages the block's
// 4 x 64 links in LDS, and reads there the link itself and the staple links
// at x and at its x3-neighbours (7 of 18); the others come from global memory.
// Workgroups follow the XCD-aware block order of the SU(3) stencil kernels
// (ordered_block_of).  The arithmetic is that of k_su2_force_link (staple():
// the same products in the same order), so the bits are the same.
template <class R>
__global__ __launch_bounds__(256) void k_su2_force_link_rows(G2 g, const R *__restrict__ U, R *__restrict__ P,
                                                             double k, R *__restrict__ Unext, R t, int reunit, int nb,
                                                             int order)
{
	__shared__ v4<R> own[4][64];
	const int lb = ordered_block_of(g, nb, 64, order, blockIdx.x);
	if (lb < 0)
		return;
	const int mu = threadIdx.x >> 6, lane = threadIdx.x & 63;
	const long x = (long)lb * 64 + lane;
	own[mu][lane] = *reinterpret_cast<const v4<R> *>(U + 4 * ((long)mu * g.stride + x));
	__syncthreads();
	auto O = [&](int d, int s) {
		const v4<R> v = own[d][s];
		return qt<R>{v.x, v.y, v.z, v.w};
	};
	// slots of the x3-neighbours (the block starts a row: 64 % L3 == 0)
	const int L3 = g.L[3], c3 = lane % L3, o0 = lane - c3;
	const int sp = o0 + (c3 + 1 == L3 ? 0 : c3 + 1), sm = o0 + (c3 == 0 ? L3 - 1 : c3 - 1);
	qt<R> S = {R(0), R(0), R(0), R(0)};
	int c[4];
	coords2(g, x, c);
	const long xpm = shc(g, x, mu, mu == 0 ? c[0] : (mu == 1 ? c[1] : (mu == 2 ? c[2] : c[3])), 1);
#pragma unroll
	for (int nu = 0; nu < 4; ++nu)
	{
		if (nu == mu)
			continue;
		// x + mu has x's coordinate along nu
		const long xpn = shc(g, x, nu, c[nu], 1), xmn = shc(g, x, nu, c[nu], -1), xpmmn = shc(g, xpm, nu, c[nu], -1);
		const qt<R> a = mu == 3 ? O(nu, sp) : ldq(U, g, nu, xpm); // U_nu(x+mu)
		const qt<R> b = nu == 3 ? O(mu, sp) : ldq(U, g, mu, xpn); // U_mu(x+nu)
		qt<R> up = qmul(qmul(a, qconj(b)), qconj(O(nu, lane)));
		const qt<R> d = nu == 3 ? O(mu, sm) : ldq(U, g, mu, xmn); // U_mu(x-nu)
		const qt<R> e = nu == 3 ? O(nu, sm) : ldq(U, g, nu, xmn); // U_nu(x-nu)
		qt<R> lo = qmul(qmul(qconj(ldq(U, g, nu, xpmmn)), qconj(d)), e);
		S = qadd(qadd(S, up), lo);
	}
	const qt<R> u = O(mu, lane);
	qt<double> M = q64(qmul(u, S));
	qt<double> p = q64(ldq(P, g, mu, x));
	qt<R> pn = qround<R>({p.w, fma(-k, M.x, p.x), fma(-k, M.y, p.y), fma(-k, M.z, p.z)});
	stq(P, g, mu, x, pn);
	qt<R> D = q_expm1(pn.x * t, pn.y * t, pn.z * t);
	qt<R> un = qadd(u, qmul(D, u));
	if (reunit)
		un = q_reunit(un);
	stq(Unext, g, mu, x, un);
}

// Uout = exp(t project_on_algebra(P)) Uin, as Uin + (exp - 1) Uin (+ reunitarisation)
template <class R>
__global__ __launch_bounds__(256) void k_su2_link_update(G2 g, int nf, R *Uout, const R *Uin, const R *__restrict__ P,
                                                         R t, int reunit)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= g.V * nf)
		return;
	int f = (int)(i / g.V);
	long x = i - (long)f * g.V;
	qt<R> p = ldq(P, g, f, x);
	qt<R> u = ldq(Uin, g, f, x);
	qt<R> D = q_expm1(p.x * t, p.y * t, p.z * t);
	qt<R> d = qmul(D, u);
	qt<R> un = qadd(u, d);
	if (reunit)
		un = q_reunit(un);
	stq(Uout, g, f, x, un);
}

template <class R> __global__ __launch_bounds__(256) void k_su2_reunit(G2 g, int nf, R *U)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= g.V * nf)
		return;
	int f = (int)(i / g.V);
	long x = i - (long)f * g.V;
	stq(U, g, f, x, q_reunit(ldq(U, g, f, x)));
}

template <class R> __global__ __launch_bounds__(256) void k_su2_plaq(G2 g, const R *__restrict__ U, double *partials,
                                                                   int nb)
{
	long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
	double v[1] = {0.0};
	if (x < g.V)
	{
		int c[4];
		coords2(g, x, c);
		v[0] = plaq_site(g, U, x, c);
	}
	block_sum_store<1>(v, partials, nb, blockIdx.x);
}

// per-site differenced Delta H: d(x) = -beta (ps_new - ps_old)(x) + (k_new - k_old)(x)
template <class R>
__global__ __launch_bounds__(256) void k_su2_delta_h(G2 g, const R *__restrict__ Uo, const R *__restrict__ Un,
                                                     const R *__restrict__ P, const double *__restrict__ Ko,
                                                     double beta, double *partials, int nb)
{
	long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
	double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
	if (x < g.V)
	{
		int c[4];
		coords2(g, x, c);
		double pso = plaq_site(g, Uo, x, c), psn = plaq_site(g, Un, x, c);
		double kn = 0.0;
		for (int mu = 0; mu < g.nd; ++mu)
			kn += kin(ldq(P, g, mu, x));
		v[0] = fma(-beta, psn - pso, kn - Ko[x]);
		v[1] = pso;
		v[2] = psn;
		v[3] = Ko[x];
		v[4] = kn;
	}
	block_sum_store<5>(v, partials, nb, blockIdx.x);
}

// host order [f][x][4] <-> device [f][stride][4] (directions padded to a
// multiple of 256 sites)
template <class R> __global__ void k_su2_aos(const R *src, R *dst, G2 g, int nf, int to_dev)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= g.V * nf)
		return;
	int f = (int)((unsigned long)i / (unsigned long)g.V);
	long x = i - (long)f * g.V;
	if (to_dev)
		stq(dst, g, f, x, qt<R>{src[4 * i], src[4 * i + 1], src[4 * i + 2], src[4 * i + 3]});
	else
	{
		qt<R> q = ldq(src, g, f, x);
		dst[4 * i] = q.w;
		dst[4 * i + 1] = q.x;
		dst[4 * i + 2] = q.y;
		dst[4 * i + 3] = q.z;
	}
}

// ---- Landau gauge (src/gauge/landau.h) on quaternions

// g(x) d(x), d(x) = sum_mu U_mu(x) h(x+mu)^+ + (h(x-mu) U_mu(x-mu))^+
// (landau.h:41-46); h == NULL means h = 1
template <class R> __device__ __forceinline__ qt<R> landau_gd(G2 const &g, const R *__restrict__ U,
                                                              const R *__restrict__ h, long x)
{
	qt<R> d = {R(0), R(0), R(0), R(0)};
	int c[4];
	coords2(g, x, c);
#pragma unroll
	for (int mu = 0; mu < 4; ++mu)
	{
		if (mu >= g.nd)
			break;
		const long xp = shc(g, x, mu, c[mu], 1), xm = shc(g, x, mu, c[mu], -1);
		qt<R> a = ldq(U, g, mu, x), b = ldq(U, g, mu, xm);
		if (h)
		{
			a = qmul(a, qconj(ldq(h, g, 0, xp)));
			b = qmul(ldq(h, g, 0, xm), b);
		}
		d = qadd(qadd(d, a), qconj(b));
	}
	return h ? qmul(ldq(h, g, 0, x), d) : d;
}

// per site sum_mu Re tr (h U_mu h(x+mu)^+) / 2, as (h U_mu) . h(x+mu) in FP64
template <class R>
__global__ __launch_bounds__(256) void k_su2_lfunc(G2 g, const R *__restrict__ U, const R *__restrict__ h,
                                                   double *partials, int nb)
{
	long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
	double v[1] = {0.0};
	if (x < g.V)
		for (int mu = 0; mu < g.nd; ++mu)
		{
			if (!h)
			{
				v[0] += (double)ldq(U, g, mu, x).w;
				continue;
			}
			qt<double> A = q64(qmul(ldq(h, g, 0, x), ldq(U, g, mu, x)));
			qt<double> B = q64(ldq(h, g, 0, nbr(g, x, mu, 1)));
			v[0] += fma(A.z, B.z, fma(A.y, B.y, fma(A.x, B.x, A.w * B.w)));
		}
	block_sum_store<1>(v, partials, nb, blockIdx.x);
}

// per site |project_on_algebra(g d)|^2 (the quaternion norm, landau.h:48)
template <class R>
__global__ __launch_bounds__(256) void k_su2_ldiv(G2 g, const R *__restrict__ U, const R *__restrict__ h,
                                                  double *partials, int nb)
{
	long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
	double v[1] = {0.0};
	if (x < g.V)
	{
		qt<double> q = q64(landau_gd(g, U, h, x));
		v[0] = fma(q.z, q.z, fma(q.y, q.y, q.x * q.x));
	}
	block_sum_store<1>(v, partials, nb, blockIdx.x);
}

// F(x) = c project_on_algebra(g d), c = -0.5 / (Nc Nd V) (landau.h:52-65)
template <class R>
__global__ __launch_bounds__(256) void k_su2_lderiv(G2 g, const R *__restrict__ U, const R *__restrict__ h, double c,
                                                    R *__restrict__ F)
{
	long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (x >= g.V)
		return;
	qt<double> q = q64(landau_gd(g, U, h, x));
	stq(F, g, 0, x, qround<R>({0.0, c * q.x, c * q.y, c * q.z}));
}

// Uo_mu(x) = h(x) U_mu(x) h(x+mu)^+
template <class R>
__global__ __launch_bounds__(256) void k_su2_gauge_transform(G2 g, R *__restrict__ Uo, const R *__restrict__ U,
                                                             const R *__restrict__ h)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= g.V * g.nd)
		return;
	int mu = (int)(i / g.V);
	long x = i - (long)mu * g.V;
	stq(Uo, g, mu, x, qmul(qmul(ldq(h, g, 0, x), ldq(U, g, mu, x)), qconj(ldq(h, g, 0, nbr(g, x, mu, 1)))));
}

// the local step: r = conj(K) / |K| maximises Re tr (r K); its omega-th power
// to second order in (r - 1), 1 + omega (r - 1) + omega (omega - 1) / 2 (r - 1)^2,
// projected back onto the group
template <class R> __device__ __forceinline__ qt<R> or_step(qt<R> K, R omega)
{
	R n2 = fma(K.z, K.z, fma(K.y, K.y, fma(K.x, K.x, K.w * K.w)));
	if (!(n2 > R(0)))
		return {R(1), R(0), R(0), R(0)};
	R inv = R(1) / sqrt(n2);
	R r0 = K.w * inv, r1 = -(K.x * inv), r2 = -(K.y * inv), r3 = -(K.z * inv);
	R a = r0 - R(1);
	R rv2 = fma(r3, r3, fma(r2, r2, r1 * r1));
	const R c2 = R(0.5) * omega * (omega - R(1));
	R g0 = fma(c2, fma(a, a, -rv2), fma(omega, a, R(1)));
	R fac = fma(R(2) * c2, a, omega);
	R g1 = r1 * fac, g2 = r2 * fac, g3 = r3 * fac;
	R gi = R(1) / sqrt(fma(g3, g3, fma(g2, g2, fma(g1, g1, g0 * g0))));
	return {g0 * gi, g1 * gi, g2 * gi, g3 * gi};
}

// overrelaxation of the sites of one parity (sum of coordinates mod 2) on
// the transformed links U, with h = G h when h is given; the t-th site of the
// parity lies in row t / (L_f / 2) of the fastest dimension f
// Workgroups of 512 consecutive sites (256 of the parity) follow the
// XCD-aware block order of the stencil kernels when V % 512 == 0 (nb > 0);
// each thread reads its site's 8 links once (in this half-sweep no other
// thread touches them: a link joins one site of each parity).
template <class R>
__global__ __launch_bounds__(256) void k_su2_or_half(G2 g, R *U, R *h, int parity, R omega, int nb, int order)
{
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (nb > 0)
	{
		const int lb = ordered_block_of(g, nb, 512, order, blockIdx.x);
		if (lb < 0)
			return;
		t = (long)lb * 256 + threadIdx.x;
	}
	if (t >= (g.V >> 1))
		return;
	const int f = g.nd - 1, hl = g.L[f] >> 1;
	// sites are below 2^31 (nbr's 32-bit arithmetic): 32-bit divisions
	const long row = (long)((unsigned)t / (unsigned)hl);
	const long x0 = row * g.L[f];
	// coordinates of the row's first site in the slower directions, then x's
	// along the row
	int cc[4] = {0, 0, 0, 0}, ps = 0;
#pragma unroll
	for (int mu = 0; mu < 3; ++mu)
		if (mu < f)
		{
			cc[mu] = (int)(((unsigned)x0 / (unsigned)g.st[mu]) % (unsigned)g.L[mu]);
			ps += cc[mu];
		}
	const long x = x0 + 2 * (t - row * hl) + ((parity + ps) & 1);
	qt<R> fw[4], bw[4];
	long xm[4];
	qt<R> K = {R(0), R(0), R(0), R(0)};
#pragma unroll
	for (int mu = 0; mu < 4; ++mu) // unrolled: the arrays stay in registers
		if (mu < g.nd)
		{
			xm[mu] = shc(g, x, mu, mu == f ? (int)(x - x0) : cc[mu], -1);
			fw[mu] = ldq(U, g, mu, x);
			bw[mu] = ldq(U, g, mu, xm[mu]);
			K = qadd(qadd(K, fw[mu]), qconj(bw[mu]));
		}
	qt<R> G = or_step(K, omega);
#pragma unroll
	for (int mu = 0; mu < 4; ++mu)
		if (mu < g.nd)
		{
			stq(U, g, mu, x, qmul(G, fw[mu]));
			stq(U, g, mu, xm[mu], qmul(bw[mu], qconj(G)));
		}
	if (h)
		stq(h, g, 0, x, qmul(G, ldq(h, g, 0, x)));
}

// per site sum over nf of Re tr (A B) = 2 (a.w b.w - a.v . b.v), in FP64
template <class R>
__global__ __launch_bounds__(256) void k_su2_dot(G2 g, int nf, const R *__restrict__ A, const R *__restrict__ B,
                                                 double *partials, int nb)
{
	long x = (long)blockIdx.x * blockDim.x + threadIdx.x;
	double v[1] = {0.0};
	if (x < g.V)
		for (int f = 0; f < nf; ++f)
		{
			qt<double> a = q64(ldq(A, g, f, x)), b = q64(ldq(B, g, f, x));
			v[0] += 2.0 * fma(-a.z, b.z, fma(-a.y, b.y, fma(-a.x, b.x, a.w * b.w)));
		}
	block_sum_store<1>(v, partials, nb, blockIdx.x);
}

// Y += a X
template <class R> __global__ __launch_bounds__(256) void k_su2_axpy(G2 g, int nf, R *Y, R a, const R *__restrict__ X)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= g.V * nf)
		return;
	int f = (int)(i / g.V);
	long x = i - (long)f * g.V;
	qt<R> y = ldq(Y, g, f, x), u = ldq(X, g, f, x);
	stq(Y, g, f, x, qt<R>{fma(a, u.w, y.w), fma(a, u.x, y.x), fma(a, u.y, y.y), fma(a, u.z, y.z)});
}

int make_g2(mesh_ctx *ctx, const mesh_grid *g, G2 *o)
{
	Geo geo;
	if (int rc = mesh_make_geo(ctx, g, &geo))
		return rc;
	o->nd = geo.nd;
	for (int i = 0; i < 4; ++i)
	{
		o->L[i] = geo.L[i];
		o->st[i] = geo.st[i];
	}
	o->V = geo.V;
	o->stride = geo.stride;
	return MESH_OK;
}

inline int nblk(long n) { return (int)((n + 255) / 256); }

template <class R> R *as(void *p) { return static_cast<R *>(p); }
template <class R> const R *as(const void *p) { return static_cast<const R *>(p); }

// f(R{}) with R the element type of the precision in use
template <class F> void by_prec(int prec, F f)
{
	if (prec == MESH_F64)
		f(0.0);
	else
		f(0.0f);
}

int su2_setup(mesh_ctx *ctx, const mesh_grid *g, int prec, G2 *o)
{
	if (int rc = make_g2(ctx, g, o))
		return rc;
	return mesh_check_prec(ctx, prec);
}

int launch_force(mesh_ctx *ctx, G2 const &g, int prec, const void *U, void *P, double beta, double delta)
{
	const double k = delta * (beta / 4.0);
	int ph = mesh_prof_begin(ctx, PROF_FORCE_UPDATE);
	by_prec(prec, [&](auto tag) {
		using R = decltype(tag);
		k_su2_force_update<R><<<nblk(g.V * g.nd), 256, 0, ctx->stream>>>(g, as<R>(U), as<R>(P), k);
	});
	mesh_prof_end(ctx, ph);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int launch_link(mesh_ctx *ctx, G2 const &g, int nf, int prec, void *Uout, const void *Uin, const void *P, double t,
                int reunit)
{
	int ph = mesh_prof_begin(ctx, PROF_LINK_UPDATE);
	by_prec(prec, [&](auto tag) {
		using R = decltype(tag);
		k_su2_link_update<R><<<nblk(g.V * nf), 256, 0, ctx->stream>>>(g, nf, as<R>(Uout), as<R>(Uin), as<R>(P), (R)t,
		                                                             reunit);
	});
	mesh_prof_end(ctx, ph);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int launch_force_link(mesh_ctx *ctx, G2 const &g, int prec, const void *U, void *P, double beta, double delta,
                      void *Unext, double t, int reunit)
{
	const double k = delta * (beta / 4.0);
	int ph = mesh_prof_begin(ctx, PROF_FORCE_LINK);
	const bool rows = MESH_SU2_ROWS && g.nd == 4 && g.V % 64 == 0 && 64 % g.L[3] == 0;
	by_prec(prec, [&](auto tag) {
		using R = decltype(tag);
		if (rows)
		{
			const int nb = (int)(g.V / 64);
			k_su2_force_link_rows<R><<<mesh_grid_blocks(ctx, nb), 256, 0, ctx->stream>>>(
			    g, as<R>(U), as<R>(P), k, as<R>(Unext), (R)t, reunit, nb, mesh_order(ctx));
		}
		else
			k_su2_force_link<R><<<nblk(g.V * g.nd), 256, 0, ctx->stream>>>(g, as<R>(U), as<R>(P), k,
			                                                               as<R>(Unext), (R)t, reunit);
	});
	mesh_prof_end(ctx, ph);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int launch_momenta(mesh_ctx *ctx, G2 const &g, int prec, void *P, double *K, uint64_t seed, uint32_t traj)
{
	int ph = mesh_prof_begin(ctx, PROF_MOMENTA);
	by_prec(prec, [&](auto tag) {
		using R = decltype(tag);
		k_su2_momenta<R><<<nblk(g.V), 256, 0, ctx->stream>>>(g, as<R>(P), K, seed, traj);
	});
	mesh_prof_end(ctx, ph);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

// dh[5] = {Delta H, sum ps_old, sum ps_new, K_old, K_new}
int su2_delta_h(mesh_ctx *ctx, G2 const &g, int prec, const void *Uo, const void *Un, const void *P, const double *Ko,
                double beta, double *dh)
{
	const int nb = nblk(g.V);
	if (int rc = mesh_reserve_partials(ctx, 5 * (size_t)nb))
		return rc;
	int ph = mesh_prof_begin(ctx, PROF_DELTA_H);
	by_prec(prec, [&](auto tag) {
		using R = decltype(tag);
		k_su2_delta_h<R><<<nb, 256, 0, ctx->stream>>>(g, as<R>(Uo), as<R>(Un), as<R>(P), Ko, beta, ctx->partials, nb);
	});
	mesh_prof_end(ctx, ph);
	MESH_LAUNCHED(ctx);
	return mesh_finish_reduction(ctx, nb, 5, dh);
}

double su2_metropolis_uniform(uint64_t seed, uint32_t traj)
{
	u4 c = {traj, 0u, 0u, TAG_METRO};
	u4 r = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
	return u01_from_bits(r.x, r.y);
}

inline double su2_planes(G2 const &g) { return (double)g.V * g.nd * (g.nd - 1) / 2; }

// fields are read with vector loads: every field pointer must be aligned to
// one quaternion (16 B in FP32, 32 B in FP64; mesh_malloc gives 256 B)
int su2_check_aligned(mesh_ctx *ctx, int prec, std::initializer_list<const void *> ps)
{
	const uintptr_t a = prec == MESH_F64 ? 32 : 16;
	for (const void *p : ps)
		if (p && (uintptr_t)p % a)
			return mesh_fail(ctx, MESH_E_ARG, "SU(2) field pointers must be aligned to 16 B (FP32) / 32 B (FP64)");
	return MESH_OK;
}

// one FP64 sum over sites of a kernel's per-site value
template <class K> int su2_site_sum(mesh_ctx *ctx, G2 const &g, K launch, double *out)
{
	const int nb = nblk(g.V);
	if (int rc = mesh_reserve_partials(ctx, nb))
		return rc;
	launch(nb);
	MESH_LAUNCHED(ctx);
	return mesh_finish_reduction(ctx, nb, 1, out);
}


} // namespace

extern "C" {

size_t mesh_su2_field_bytes(const mesh_grid *g, int nf, int prec)
{
	if (!g || g->nd < 2 || g->nd > 4 || nf < 1 || (prec != MESH_F64 && prec != MESH_F32))
		return 0;
	long V = 1;
	for (int i = 0; i < g->nd; ++i)
	{
		if (g->dims[i] < 2)
			return 0;
		V *= g->dims[i];
	}
	return (size_t)nf * 4 * ((V + 255) / 256 * 256) * (prec == MESH_F64 ? 8 : 4);
}

int mesh_su2_transfer(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, void *host, void *dev, int upload)
{
	MESH_CHECK_CTX(ctx);
	G2 geo;
	if (int rc = su2_setup(ctx, g, prec, &geo))
		return rc;
	if (nf < 1 || !host || !dev)
		return mesh_fail(ctx, MESH_E_ARG, "bad SU(2) transfer arguments");
	if (int rc = su2_check_aligned(ctx, prec, {dev}))
		return rc;
	const size_t es = prec == MESH_F64 ? 8 : 4, bytes = (size_t)nf * geo.V * 4 * es;
	void *tmp = nullptr;
	if (hipMallocAsync(&tmp, bytes, ctx->stream) != hipSuccess)
		return mesh_fail(ctx, MESH_E_ALLOC, "cannot allocate transfer staging buffer");
	const int nb = nblk(geo.V * nf);
	if (upload)
	{
		MESH_HIP(ctx, hipMemcpyAsync(tmp, host, bytes, hipMemcpyHostToDevice, ctx->stream));
		by_prec(prec, [&](auto tag) {
			using R = decltype(tag);
			k_su2_aos<R><<<nb, 256, 0, ctx->stream>>>(as<R>(tmp), as<R>(dev), geo, nf, 1);
		});
		MESH_LAUNCHED(ctx);
	}
	else
	{
		by_prec(prec, [&](auto tag) {
			using R = decltype(tag);
			k_su2_aos<R><<<nb, 256, 0, ctx->stream>>>(as<R>(dev), as<R>(tmp), geo, nf, 0);
		});
		MESH_LAUNCHED(ctx);
		MESH_HIP(ctx, hipMemcpyAsync(host, tmp, bytes, hipMemcpyDeviceToHost, ctx->stream));
	}
	MESH_HIP(ctx, hipFreeAsync(tmp, ctx->stream));
	MESH_HIP(ctx, hipStreamSynchronize(ctx->stream));
	return MESH_OK;
}

int mesh_su2_random(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, void *U, uint64_t seed)
{
	MESH_CHECK_CTX(ctx);
	G2 geo;
	if (int rc = su2_setup(ctx, g, prec, &geo))
		return rc;
	if (!U || nf < 1)
		return mesh_fail(ctx, MESH_E_ARG, "bad su2_random arguments");
	if (int rc = su2_check_aligned(ctx, prec, {U}))
		return rc;
	by_prec(prec, [&](auto tag) {
		using R = decltype(tag);
		k_su2_random<R><<<nblk(geo.V * nf), 256, 0, ctx->stream>>>(geo, nf, as<R>(U), seed);
	});
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_su2_unit(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, void *U)
{
	MESH_CHECK_CTX(ctx);
	G2 geo;
	if (int rc = su2_setup(ctx, g, prec, &geo))
		return rc;
	if (!U || nf < 1)
		return mesh_fail(ctx, MESH_E_ARG, "bad su2_unit arguments");
	if (int rc = su2_check_aligned(ctx, prec, {U}))
		return rc;
	by_prec(prec, [&](auto tag) {
		using R = decltype(tag);
		k_su2_unit<R><<<nblk(geo.V * nf), 256, 0, ctx->stream>>>(geo, nf, as<R>(U));
	});
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_su2_momenta(mesh_ctx *ctx, const mesh_grid *g, int prec, void *P, double *K, uint64_t seed, uint32_t traj)
{
	MESH_CHECK_CTX(ctx);
	G2 geo;
	if (int rc = su2_setup(ctx, g, prec, &geo))
		return rc;
	if (!P)
		return mesh_fail(ctx, MESH_E_ARG, "P is NULL");
	if (int rc = su2_check_aligned(ctx, prec, {P}))
		return rc;
	return launch_momenta(ctx, geo, prec, P, K, seed, traj);
}

int mesh_su2_force_update(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U, void *P, double beta,
                          double delta)
{
	MESH_CHECK_CTX(ctx);
	G2 geo;
	if (int rc = su2_setup(ctx, g, prec, &geo))
		return rc;
	if (!U || !P || U == P)
		return mesh_fail(ctx, MESH_E_ARG, "bad su2_force_update arguments");
	if (int rc = su2_check_aligned(ctx, prec, {U, P}))
		return rc;
	return launch_force(ctx, geo, prec, U, P, beta, delta);
}

int mesh_su2_link_update(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, void *Uout, const void *Uin,
                         const void *P, double t, int reunit)
{
	MESH_CHECK_CTX(ctx);
	G2 geo;
	if (int rc = su2_setup(ctx, g, prec, &geo))
		return rc;
	if (!Uout || !Uin || !P || nf < 1)
		return mesh_fail(ctx, MESH_E_ARG, "bad su2_link_update arguments");
	if (int rc = su2_check_aligned(ctx, prec, {Uout, Uin, P}))
		return rc;
	return launch_link(ctx, geo, nf, prec, Uout, Uin, P, t, reunit);
}

int mesh_su2_reunit(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, void *U)
{
	MESH_CHECK_CTX(ctx);
	G2 geo;
	if (int rc = su2_setup(ctx, g, prec, &geo))
		return rc;
	if (!U || nf < 1)
		return mesh_fail(ctx, MESH_E_ARG, "bad su2_reunit arguments");
	if (int rc = su2_check_aligned(ctx, prec, {U}))
		return rc;
	by_prec(prec, [&](auto tag) {
		using R = decltype(tag);
		k_su2_reunit<R><<<nblk(geo.V * nf), 256, 0, ctx->stream>>>(geo, nf, as<R>(U));
	});
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_su2_plaquette(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U, double *out)
{
	MESH_CHECK_CTX(ctx);
	G2 geo;
	if (int rc = su2_setup(ctx, g, prec, &geo))
		return rc;
	if (!U || !out)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to su2_plaquette");
	if (int rc = su2_check_aligned(ctx, prec, {U}))
		return rc;
	const int nb = nblk(geo.V);
	if (int rc = mesh_reserve_partials(ctx, nb))
		return rc;
	by_prec(prec, [&](auto tag) {
		using R = decltype(tag);
		k_su2_plaq<R><<<nb, 256, 0, ctx->stream>>>(geo, as<R>(U), ctx->partials, nb);
	});
	MESH_LAUNCHED(ctx);
	double s;
	if (int rc = mesh_finish_reduction(ctx, nb, 1, &s))
		return rc;
	*out = s / su2_planes(geo);
	return MESH_OK;
}

int mesh_su2_delta_h(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U_old, const void *U_new,
                     const void *P, const double *K_old, double beta, double *out)
{
	MESH_CHECK_CTX(ctx);
	G2 geo;
	if (int rc = su2_setup(ctx, g, prec, &geo))
		return rc;
	if (!U_old || !U_new || !P || !K_old || !out)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to su2_delta_h");
	if (int rc = su2_check_aligned(ctx, prec, {U_old, U_new, P}))
		return rc;
	return su2_delta_h(ctx, geo, prec, U_old, U_new, P, K_old, beta, out);
}

int mesh_su2_hmc_trajectory(mesh_ctx *ctx, const mesh_grid *g, int prec, void **U, void **U_work, void *P, double *K,
                            double beta, const double *deltas, int nd_, uint64_t seed, uint32_t traj, int metropolis,
                            double *out)
{
	MESH_CHECK_CTX(ctx);
	G2 geo;
	if (int rc = su2_setup(ctx, g, prec, &geo))
		return rc;
	if (!U || !U_work || !*U || !*U_work || *U == *U_work || !P || !K || !deltas || !out || nd_ < 1 || nd_ % 2 != 1)
		return mesh_fail(ctx, MESH_E_ARG, "bad su2_hmc_trajectory arguments (deltas must have odd length)");
	if (int rc = su2_check_aligned(ctx, prec, {*U, *U_work, P}))
		return rc;
	int rc;
	if ((rc = launch_momenta(ctx, geo, prec, P, K, seed, traj)))
		return rc;
	const int nfused = (nd_ - 1) / 2;
	if (ctx->fuse_md && nfused > 0)
	{
		// each momentum step fused with the link step after it, ping-pong
		// between the caller's work field and the context's spare, the first
		// link step writing the field that makes the last land in *U_work
		const size_t fb = mesh_su2_field_bytes(g, geo.nd, prec);
		if (fb > ctx->md_spare_bytes)
		{
			if (ctx->md_spare)
			{
				MESH_HIP(ctx, hipStreamSynchronize(ctx->stream));
				MESH_HIP(ctx, hipFree(ctx->md_spare));
				ctx->md_spare = nullptr;
				ctx->md_spare_bytes = 0;
			}
			if (hipMalloc(&ctx->md_spare, fb) != hipSuccess)
				return mesh_fail(ctx, MESH_E_ALLOC, "cannot allocate the spare field of the fused trajectory");
			ctx->md_spare_bytes = fb;
		}
		void *cur = (nfused % 2) ? ctx->md_spare : *U_work;
		void *other = (nfused % 2) ? *U_work : ctx->md_spare;
		if ((rc = launch_link(ctx, geo, geo.nd, prec, cur, *U, P, deltas[0], 0)))
			return rc;
		for (int i = 1; i + 1 < nd_; i += 2)
		{
			if ((rc = launch_force_link(ctx, geo, prec, cur, P, beta, deltas[i], other, deltas[i + 1],
			                            i + 1 == nd_ - 1)))
				return rc;
			void *tmp = cur;
			cur = other;
			other = tmp;
		}
	}
	else
		for (int i = 0; i < nd_; ++i)
		{
			if (i % 2 == 0)
				rc = launch_link(ctx, geo, geo.nd, prec, *U_work, i == 0 ? *U : *U_work, P, deltas[i],
				                 i == nd_ - 1);
			else
				rc = launch_force(ctx, geo, prec, *U_work, P, beta, deltas[i]);
			if (rc)
				return rc;
		}
	double dh[5];
	if ((rc = su2_delta_h(ctx, geo, prec, *U, *U_work, P, K, beta, dh)))
		return rc;
	double u = su2_metropolis_uniform(seed, traj);
	int acc = metropolis ? (u < std::exp(-dh[0])) : 1;
	if (acc)
	{
		void *t = *U;
		*U = *U_work;
		*U_work = t;
	}
	const double np = su2_planes(geo);
	out[0] = dh[0];
	out[1] = acc;
	out[2] = (acc ? dh[2] : dh[1]) / np;
	out[3] = beta * (np - dh[1]) + dh[3];
	out[4] = beta * (np - dh[2]) + dh[4];
	out[5] = u;
	return MESH_OK;
}

// ---- Landau gauge

int mesh_su2_landau_functional(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U, const void *h,
                               double *out)
{
	MESH_CHECK_CTX(ctx);
	G2 geo;
	if (int rc = su2_setup(ctx, g, prec, &geo))
		return rc;
	if (!U || !out)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to su2_landau_functional");
	if (int rc = su2_check_aligned(ctx, prec, {U, h}))
		return rc;
	double s;
	int rc = su2_site_sum(
	    ctx, geo,
	    [&](int nb) {
		    by_prec(prec, [&](auto tag) {
			    using R = decltype(tag);
			    k_su2_lfunc<R><<<nb, 256, 0, ctx->stream>>>(geo, as<R>(U), as<R>(h), ctx->partials, nb);
		    });
	    },
	    &s);
	if (rc)
		return rc;
	*out = s / ((double)geo.nd * geo.V);
	return MESH_OK;
}

int mesh_su2_landau_divsum(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U, const void *h, double *out)
{
	MESH_CHECK_CTX(ctx);
	G2 geo;
	if (int rc = su2_setup(ctx, g, prec, &geo))
		return rc;
	if (!U || !out)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to su2_landau_divsum");
	if (int rc = su2_check_aligned(ctx, prec, {U, h}))
		return rc;
	return su2_site_sum(
	    ctx, geo,
	    [&](int nb) {
		    by_prec(prec, [&](auto tag) {
			    using R = decltype(tag);
			    k_su2_ldiv<R><<<nb, 256, 0, ctx->stream>>>(geo, as<R>(U), as<R>(h), ctx->partials, nb);
		    });
	    },
	    out);
}

int mesh_su2_landau_deriv(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U, const void *h, void *F)
{
	MESH_CHECK_CTX(ctx);
	G2 geo;
	if (int rc = su2_setup(ctx, g, prec, &geo))
		return rc;
	if (!U || !F || F == U || F == h)
		return mesh_fail(ctx, MESH_E_ARG, "bad su2_landau_deriv arguments");
	if (int rc = su2_check_aligned(ctx, prec, {U, h, F}))
		return rc;
	const double c = -0.5 / (2.0 * geo.nd * (double)geo.V);
	by_prec(prec, [&](auto tag) {
		using R = decltype(tag);
		k_su2_lderiv<R><<<nblk(geo.V), 256, 0, ctx->stream>>>(geo, as<R>(U), as<R>(h), c, as<R>(F));
	});
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_su2_gauge_transform(mesh_ctx *ctx, const mesh_grid *g, int prec, void *Uout, const void *U, const void *h)
{
	MESH_CHECK_CTX(ctx);
	G2 geo;
	if (int rc = su2_setup(ctx, g, prec, &geo))
		return rc;
	if (!Uout || !U || !h || Uout == U || Uout == h)
		return mesh_fail(ctx, MESH_E_ARG, "bad su2_gauge_transform arguments");
	if (int rc = su2_check_aligned(ctx, prec, {Uout, U, h}))
		return rc;
	by_prec(prec, [&](auto tag) {
		using R = decltype(tag);
		k_su2_gauge_transform<R><<<nblk(geo.V * geo.nd), 256, 0, ctx->stream>>>(geo, as<R>(Uout), as<R>(U),
		                                                                        as<R>(h));
	});
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_su2_landau_or_halfsweep(mesh_ctx *ctx, const mesh_grid *g, int prec, void *U, void *h, int parity,
                                 double omega)
{
	MESH_CHECK_CTX(ctx);
	G2 geo;
	if (int rc = su2_setup(ctx, g, prec, &geo))
		return rc;
	if (!U || U == h || (parity != 0 && parity != 1))
		return mesh_fail(ctx, MESH_E_ARG, "bad su2_landau_or_halfsweep arguments");
	if (int rc = su2_check_aligned(ctx, prec, {U, h}))
		return rc;
	for (int i = 0; i < geo.nd; ++i)
		if (geo.L[i] % 2)
			return mesh_fail(ctx, MESH_E_SIZE, "checkerboard update needs even extents");
	int ph = mesh_prof_begin(ctx, PROF_LANDAU_OR);
	by_prec(prec, [&](auto tag) {
		using R = decltype(tag);
		const int nb = geo.V % 512 == 0 ? (int)(geo.V / 512) : 0;
		k_su2_or_half<R><<<nb > 0 ? mesh_grid_blocks(ctx, nb) : nblk(geo.V / 2), 256, 0, ctx->stream>>>(
		    geo, as<R>(U), as<R>(h), parity, (R)omega, nb, mesh_order(ctx));
	});
	mesh_prof_end(ctx, ph);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_su2_dot(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, const void *A, const void *B, double *out)
{
	MESH_CHECK_CTX(ctx);
	G2 geo;
	if (int rc = su2_setup(ctx, g, prec, &geo))
		return rc;
	if (!A || !B || !out || nf < 1)
		return mesh_fail(ctx, MESH_E_ARG, "bad su2_dot arguments");
	if (int rc = su2_check_aligned(ctx, prec, {A, B}))
		return rc;
	return su2_site_sum(
	    ctx, geo,
	    [&](int nb) {
		    by_prec(prec, [&](auto tag) {
			    using R = decltype(tag);
			    k_su2_dot<R><<<nb, 256, 0, ctx->stream>>>(geo, nf, as<R>(A), as<R>(B), ctx->partials, nb);
		    });
	    },
	    out);
}

int mesh_su2_axpy(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, void *Y, double a, const void *X)
{
	MESH_CHECK_CTX(ctx);
	G2 geo;
	if (int rc = su2_setup(ctx, g, prec, &geo))
		return rc;
	if (!Y || !X || nf < 1)
		return mesh_fail(ctx, MESH_E_ARG, "bad su2_axpy arguments");
	if (int rc = su2_check_aligned(ctx, prec, {Y, X}))
		return rc;
	by_prec(prec, [&](auto tag) {
		using R = decltype(tag);
		k_su2_axpy<R><<<nblk(geo.V * nf), 256, 0, ctx->stream>>>(geo, nf, as<R>(Y), (R)a, as<R>(X));
	});
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

} // extern "C"
