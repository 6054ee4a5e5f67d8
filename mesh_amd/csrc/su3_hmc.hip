// SU(3) pure-gauge Wilson-action HMC kernels for gfx950.
//
// One kernel per fused step of the trajectory, all over structure-of-arrays
// fields (plane (f,k) = complex entry k of matrix f for every site):
//   k_momenta       P ~ N(0,1) su(3) momenta, k(x) = sum_mu ||P_mu(x)||^2
//   k_force_update  P_mu -= delta beta/6 TA(U_mu Sigma_mu), staples in-kernel
//   k_link_update   U_mu <- exp(t TA(P_mu)) U_mu (+ reunitarisation)
//   k_delta_h       per-site differenced Delta H and both plaquette sums
//   k_plaq          per-site plaquette sum, block-reduced
// Neighbours are addressed in-kernel (no shifted copies as in the
// reference's cshift, src/lattice/lattice.h:366-389).
#include "mesh_internal.h"
#include "rng_dev.h"
#include <cmath>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

using namespace mesh_dev;

// FP64 row-staged staples software-pipelined (staple_rows_pipe; same bits)
#ifndef MESH_PIPE_F64
#define MESH_PIPE_F64 1
#endif
// The global operands of the shared upper-staple factors loaded before the
// staging barrier (share_fetch; same bits), FP32 / FP64; 0: after it.  With
// three FP32 workgroups per CU it gained 1.2% (0.3443 -> 0.3401 ms at 32^4);
// with four (the four-plane mode) it costs 0.4% (0.3417 vs 0.3432 ms, three
// same-box rounds), and in FP64 it is even (0.5626 vs 0.5626 ms).
#ifndef MESH_SHARE_PREFETCH
#define MESH_SHARE_PREFETCH 0
#endif
#ifndef MESH_SHARE_PREFETCH_F64
#define MESH_SHARE_PREFETCH_F64 0
#endif
// waves per SIMD the force kernels are compiled for (0: the compiler's choice)
#ifndef MESH_FORCE_WAVES
#define MESH_FORCE_WAVES 0
#endif
#if MESH_FORCE_WAVES > 0
#define FORCE_OCC __attribute__((amdgpu_waves_per_eu(MESH_FORCE_WAVES)))
#else
#define FORCE_OCC
#endif
// ... for the FP32 wavefront-per-direction kernel only (0: as above)
#ifndef MESH_FORCE_WAVES_F32
#define MESH_FORCE_WAVES_F32 0
#endif
// ... and for the FP64 kernel with the row-staged staple path (0: as above;
// 3 fits in 168 VGPRs but measured slower than two).  The pipelined staples
// would take 260 registers and one wave; 2 holds them at two waves with three
// VGPRs spilled, 1.1% faster than the unpipelined kernel at two waves and 16%
// faster than the pipelined one at one wave (32^4, one box)
#ifndef MESH_FORCE_WAVES_ROWS
#define MESH_FORCE_WAVES_ROWS (MESH_PIPE_F64 ? 2 : 0)
#endif
// upper bound of waves per SIMD for the FP64 kernel with the row-staged staple
// path (8: none): it runs at two either way (170 VGPRs), and with the bound the
// compiler may spend the registers up to 256 on loads in flight
#ifndef MESH_FORCE_MAXW_ROWS
#define MESH_FORCE_MAXW_ROWS 2
#endif
template <class R, int PATH = 0> constexpr int force_max_waves()
{
	return (std::is_same<R, double>::value && PATH == 2) ? MESH_FORCE_MAXW_ROWS : 8;
}
template <class R, int PATH = 0> constexpr int force_min_waves()
{
	return (std::is_same<R, float>::value && MESH_FORCE_WAVES_F32 > 0) ? MESH_FORCE_WAVES_F32
	       : (std::is_same<R, double>::value && PATH == 2 && MESH_FORCE_WAVES_ROWS > 0)
	           ? MESH_FORCE_WAVES_ROWS
	           : (MESH_FORCE_WAVES > 0 ? MESH_FORCE_WAVES : 1);
}

// Wavefront-per-direction kernel: the first upper-staple factor of a plane
// formed once per workgroup and shared through LDS (the same bits either
// way).  MESH_SHARE_UPPER (FP32) / MESH_SHARE_UPPER_F64: 0 each direction
// forms its own, 1 the three planes (0, nu) are shared, 2 all six, 3 the four
// planes other than (0, 1) and (2, 3), one formed by each wavefront (every
// wavefront then forms two factors, as the busiest ones of mode 2 do, in 4 KB
// less LDS per shared plane).  FP64 with all six would need 91 KB of LDS per
// workgroup, one workgroup per CU.
#ifndef MESH_SHARE_UPPER
#define MESH_SHARE_UPPER 3
#endif
#ifndef MESH_SHARE_UPPER_F64
#define MESH_SHARE_UPPER_F64 3
#endif
template <class R> constexpr int share_mode()
{
	return std::is_same<R, float>::value ? MESH_SHARE_UPPER : MESH_SHARE_UPPER_F64;
}
// is plane (p, q), p < q, shared in mode m
__host__ __device__ constexpr bool plane_shared(int m, int p, int q)
{
	return m == 2 || (m == 1 && p == 0) || (m == 3 && p + q != 1 && p + q != 5);
}
// number of shared planes in mode m
__host__ __device__ constexpr int shared_planes(int m) { return m == 2 ? 6 : (m == 3 ? 4 : 3); }

// non-temporal hints on the once-touched fields of the force kernels
#ifndef MESH_STREAM_HINTS
#define MESH_STREAM_HINTS 1
#endif
// waves per SIMD the Delta H kernel is limited to (two fields with neighbours
// in flight: more waves thrash the L2; 0 = the compiler's choice)
// Delta H with the two fields' plaquettes on separate halves of a workgroup
#ifndef MESH_DH_SPLIT
#define MESH_DH_SPLIT 1
#endif
#ifndef MESH_DH_WAVES
#define MESH_DH_WAVES 2
#endif
#if MESH_DH_WAVES > 0
#define DH_OCC __attribute__((amdgpu_waves_per_eu(1, MESH_DH_WAVES)))
#else
#define DH_OCC
#endif
// variant 0 stages the workgroup's own-site links in LDS (4-d, V % 64 == 0)
#ifndef MESH_OWN_LDS
#define MESH_OWN_LDS 1
#endif
// ... and, when the block holds whole x3-rows, the links of the x3-neighbours
#ifndef MESH_OWN_ROWS
#define MESH_OWN_ROWS 1
#endif
// staples with the direction as a compile-time constant (4-d lattices)
#ifndef MESH_STATIC_MU
#define MESH_STATIC_MU 1
#endif

namespace {

template <class R> __device__ __forceinline__ m3<R> load_once(const R *__restrict__ b, int f, Geo const &g, long x)
{
	if constexpr (MESH_STREAM_HINTS != 0)
		return load_stream(b, f, g, x);
	else
		return load(b, f, g, x);
}
template <class R>
__device__ __forceinline__ void store_once(R *__restrict__ b, int f, Geo const &g, long x, m3<R> const &m)
{
	if constexpr (MESH_STREAM_HINTS != 0)
		store_stream(b, f, g, x, m);
	else
		store(b, f, g, x, m);
}

// momenta of the trajectory's fused path: compact antihermitian layout
// (load_alg / store_alg) when PC, else the full matrix layout; streamed
template <bool PC, class R>
__device__ __forceinline__ m3<R> load_mom(const R *__restrict__ P, int f, Geo const &g, long x)
{
	if constexpr (PC)
		return load_alg<R, MESH_STREAM_HINTS != 0>(P, f, g, x);
	else
		return load_once(P, f, g, x);
}
template <bool PC, class R>
__device__ __forceinline__ void store_mom(R *__restrict__ P, int f, Geo const &g, long x, m3<R> const &m)
{
	if constexpr (PC)
		store_alg<R, MESH_STREAM_HINTS != 0>(P, f, g, x, m);
	else
		store_once(P, f, g, x, m);
}

constexpr int SB = 64; // sites per workgroup in the stencil kernels (one wavefront)

// all coordinates of site x: three 32-bit divisions, fastest dimension first
template <class R> __device__ __forceinline__ void coords(Geo const &g, long x, int *c)
{
#if !MESH_COORD32
#pragma unroll
	for (int mu = 0; mu < 4; ++mu)
		c[mu] = mu < g.nd ? coord(g, x, mu) : 0;
	return;
#endif
	unsigned t = (unsigned)x;
#pragma unroll
	for (int mu = 3; mu >= 1; --mu)
	{
		if (mu < g.nd)
		{
			unsigned q = t / (unsigned)g.L[mu];
			c[mu] = (int)(t - q * (unsigned)g.L[mu]);
			t = q;
		}
		else
			c[mu] = 0;
	}
	c[0] = (int)t;
}

// staple sum of direction mu at site x with coordinates c
// (src/gauge/utils.h:110-139)
template <class R> __device__ __forceinline__ m3<R> staple_at(Geo const &g, const R *__restrict__ U, int mu,
                                                            long x, int const *c)
{
	m3<R> S = zero_m3<R>();
	long xpm = shift_c(g, x, mu, c[mu], 1);
#pragma unroll
	for (int nu = 0; nu < 4; ++nu)
	{
		if (nu == mu || nu >= g.nd)
			continue;
		long xpn = shift_c(g, x, nu, c[nu], 1);
		long xmn = shift_c(g, x, nu, c[nu], -1);
		long xpmmn = shift_c(g, xpm, nu, c[nu], -1);
		{
			m3<R> a = load(U, nu, g, xpm), b = load(U, mu, g, xpn);
			m3<R> T1 = nu < mu ? adj_m3(mul_nd(b, a)) : mul_nd(a, b);
			madd_nd(S, T1, load(U, nu, g, x));
		}
		{
			m3<R> T2 = mul_dd(load(U, nu, g, xpmmn), load(U, mu, g, xmn));
			madd_nn(S, T2, load(U, nu, g, xmn));
		}
	}
	return S;
}

// The same for a compile-time direction of a 4-d lattice: no branch between
// the six staples, so the compiler can issue the loads of later staples
// while earlier ones multiply (same operations in the same order).
template <int MU, class R>
__device__ __forceinline__ m3<R> staple_4d(Geo const &g, const R *__restrict__ U, long x, int const *c)
{
	m3<R> S = zero_m3<R>();
	long xpm = shift_c(g, x, MU, c[MU], 1);
#pragma unroll
	for (int nu = 0; nu < 4; ++nu)
	{
		if (nu == MU)
			continue;
		long xpn = shift_c(g, x, nu, c[nu], 1);
		long xmn = shift_c(g, x, nu, c[nu], -1);
		long xpmmn = shift_c(g, xpm, nu, c[nu], -1);
		{
			m3<R> T1 = upper_factor(load(U, nu, g, xpm), load(U, MU, g, xpn), nu < MU);
			madd_nd(S, T1, load(U, nu, g, x));
		}
		{
			m3<R> T2 = mul_dd(load(U, nu, g, xpmmn), load(U, MU, g, xmn));
			madd_nn(S, T2, load(U, nu, g, xmn));
		}
	}
	return S;
}

// The same with links read from the workgroup's 64-site block staged in LDS
// (own[d][k][slot], staged once by the workgroup instead of once per
// direction's wavefront): the four links of site x, and with ROWS (the block
// holds whole x3-rows, L3 | 64) also those of its x3-neighbours.

template <class R> __device__ __forceinline__ m3<R> lds_m3(vec2_t<R> (*a)[9][64], int d, int slot)
{
	m3<R> m;
#pragma unroll
	for (int k = 0; k < 9; ++k)
	{
		vec2_t<R> v = a[d][k][slot];
		m.e[k] = {v.x, v.y};
	}
	return m;
}

// RowsAt of x with the block and its x0, x1 coordinates in scalar registers
// (the same for every lane of the block's wavefront on the rows path)
__device__ __forceinline__ RowsAt rows_at(Geo const &g, long x, int const *c)
{
	RowsAt a = rows_at_lane(g, x, c);
	a.lb = __builtin_amdgcn_readfirstlane(a.lb);
	a.c0 = __builtin_amdgcn_readfirstlane(a.c0);
	a.c1 = __builtin_amdgcn_readfirstlane(a.c1);
	return a;
}

// link f at x + s1 e_a1 + s2 e_a2 (a2 = -1: one shift); xs is that site,
// which the rows path needs only for shifts along x2
template <bool ROWS, class R>
__device__ __forceinline__ m3<R> load_disp(const R *__restrict__ U, int f, Geo const &g, RowsAt const &a, int a1,
                                           int s1, int a2, int s2, long xs)
{
	if constexpr (ROWS)
		if (a1 != 2 && a2 != 2)
		{
			const BlkSlot b = rows_site(g, a, a1, s1, a2, s2);
			const vec2_t<R> *p = reinterpret_cast<const vec2_t<R> *>(U) + blk_off(g, f, b.blk, 0);
			m3<R> m;
#pragma unroll
			for (int k = 0; k < 9; ++k)
			{
				vec2_t<R> v = p[b.l + (k << MESH_SITE_LB)];
				m.e[k] = {v.x, v.y};
			}
			return m;
		}
	return load(U, f, g, xs);
}

// index of plane (p, q), p < q, among the six of a 4-d lattice
__host__ __device__ constexpr int plane_index(int p, int q) { return p == 0 ? q - 1 : (p == 1 ? q + 1 : 5); }
// LDS slot of the shared factor of plane (p, q) in mode m (mode 3: the planes
// (0, 2), (1, 3), (1, 2), (0, 3), formed by wavefronts 0..3)
__host__ __device__ constexpr int share_slot(int m, int p, int q)
{
	return m != 3 ? plane_index(p, q) : (p == 0 ? (q == 2 ? 0 : 3) : (q == 3 ? 1 : 2));
}

// Shared first upper-staple factors: A_pq = U_q(x+p) U_p(x+q)^+ for p < q,
// the factor of direction p in plane q (direction q uses A_pq^+).  Mode 2:
// the six planes split 2, 2, 1, 1 over the wavefronts of directions 0..3;
// mode 1: planes (0, w) formed by wavefront w = 1, 2, 3; mode 3: one plane
// each (share_slot).
// the planes (p, q) whose factors wavefront W forms, f(P, Q, I) with I its
// index among them
template <int W, int MODE, class F> __device__ __forceinline__ void for_shared_planes(F &&f)
{
	using I0 = std::integral_constant<int, 0>;
	using I1 = std::integral_constant<int, 1>;
	using I2 = std::integral_constant<int, 2>;
	using I3 = std::integral_constant<int, 3>;
	if constexpr (MODE == 3)
	{
		if constexpr (W == 0)
			f(I0{}, I2{}, I0{});
		else if constexpr (W == 1)
			f(I1{}, I3{}, I0{});
		else if constexpr (W == 2)
			f(I1{}, I2{}, I0{});
		else
			f(I0{}, I3{}, I0{});
	}
	else if constexpr (MODE == 1)
	{
		if constexpr (W == 1)
			f(I0{}, I1{}, I0{});
		else if constexpr (W == 2)
			f(I0{}, I2{}, I0{});
		else if constexpr (W == 3)
			f(I0{}, I3{}, I0{});
	}
	else if constexpr (W == 0)
	{
		f(I0{}, I1{}, I0{});
		f(I0{}, I2{}, I1{});
	}
	else if constexpr (W == 1)
	{
		f(I1{}, I2{}, I0{});
		f(I1{}, I3{}, I1{});
	}
	else if constexpr (W == 2)
		f(I2{}, I3{}, I0{});
	else
		f(I0{}, I3{}, I0{});
}

// the factors' operands that come from global memory (with ROWS the x3-
// neighbour ones come from the staged block): loaded by share_fetch before
// the staging barrier, since they do not depend on it, so that their latency
// overlaps the staging's
template <class R> struct ShareOps
{
	m3<R> a[2], b[2];
};

template <int W, bool ROWS, int MODE, class R>
__device__ __forceinline__ void share_fetch(Geo const &g, const R *__restrict__ U, long x, int const *c,
                                            ShareOps<R> &F)
{
	RowsAt A{};
	if constexpr (ROWS)
		A = rows_at(g, x, c);
	for_shared_planes<W, MODE>([&](auto P, auto Q, auto I) {
		constexpr int p = decltype(P)::value, q = decltype(Q)::value, i = decltype(I)::value;
		if constexpr (!(ROWS && p == 3))
			F.a[i] = load_disp<ROWS>(U, q, g, A, p, 1, -1, 0, shift_c(g, x, p, c[p], 1));
		if constexpr (!(ROWS && q == 3))
			F.b[i] = load_disp<ROWS>(U, p, g, A, q, 1, -1, 0, shift_c(g, x, q, c[q], 1));
	});
}

template <int W, bool ROWS, int MODE, class R>
__device__ __forceinline__ void share_upper_factors(Geo const &g, long x, int const *c, vec2_t<R> (*own)[9][64],
                                                    vec2_t<R> (*sh)[9][64], int lane, ShareOps<R> const &F)
{
	int sp = 0, sm = 0;
	if constexpr (ROWS)
		x3_slots(g, x, c, sp, sm);
	for_shared_planes<W, MODE>([&](auto P, auto Q, auto I) {
		constexpr int p = decltype(P)::value, q = decltype(Q)::value, i = decltype(I)::value;
		m3<R> a, b;
		if constexpr (ROWS && p == 3)
			a = lds_m3<R>(own, q, sp);
		else
			a = F.a[i];
		if constexpr (ROWS && q == 3)
			b = lds_m3<R>(own, p, sp);
		else
			b = F.b[i];
		m3<R> A = mul_nd(a, b);
#pragma unroll
		for (int k = 0; k < 9; ++k)
			sh[share_slot(MODE, p, q)][k][lane] = vec2_t<R>{A.e[k].re, A.e[k].im};
	});
}

// SH: share mode of the first upper-staple factors (share_upper_factors)
template <int MU, bool ROWS, class R, int SH = 0>
__device__ __forceinline__ m3<R> staple_4d_own(Geo const &g, const R *__restrict__ U, long x, int const *c,
                                               vec2_t<R> (*own)[9][64], int lane, vec2_t<R> (*sh)[9][64] = nullptr)
{
	auto own_at = [&](int d, int slot) { return lds_m3<R>(own, d, slot); };
	// slots of the x3-neighbours x +/- e3 (used with ROWS only)
	RowsAt A{};
	if constexpr (ROWS)
		A = rows_at(g, x, c);
	const int sp = A.sp, sm = A.sm;
	m3<R> S = zero_m3<R>();
	long xpm = shift_c(g, x, MU, c[MU], 1);
#pragma unroll
	for (int nu = 0; nu < 4; ++nu)
	{
		if (nu == MU)
			continue;
		const bool fwd3 = ROWS && MU == 3; // x + mu is an x3-neighbour
		const bool nu3 = ROWS && nu == 3;  // x +/- nu are x3-neighbours
		long xpn = nu3 ? 0 : shift_c(g, x, nu, c[nu], 1);
		long xmn = nu3 ? 0 : shift_c(g, x, nu, c[nu], -1);
		long xpmmn = shift_c(g, xpm, nu, c[nu], -1);
		if (plane_shared(SH, nu < MU ? nu : MU, nu < MU ? MU : nu))
		{
			m3<R> T1 = nu < MU ? adj_m3(lds_m3<R>(sh, share_slot(SH, nu, MU), lane))
			                   : lds_m3<R>(sh, share_slot(SH, MU, nu), lane);
			madd_nd(S, T1, own_at(nu, lane));
		}
		else
		{
			m3<R> a = fwd3 ? own_at(nu, sp) : load_disp<ROWS>(U, nu, g, A, MU, 1, -1, 0, xpm);
			m3<R> b = nu3 ? own_at(MU, sp) : load_disp<ROWS>(U, MU, g, A, nu, 1, -1, 0, xpn);
			m3<R> T1 = upper_factor(a, b, nu < MU);
			madd_nd(S, T1, own_at(nu, lane));
		}
		{
			m3<R> a = load_disp<ROWS>(U, nu, g, A, MU, 1, nu, -1, xpmmn);
			m3<R> b = nu3 ? own_at(MU, sm) : load_disp<ROWS>(U, MU, g, A, nu, -1, -1, 0, xmn);
			m3<R> T2 = mul_dd(a, b);
			madd_nn(S, T2, nu3 ? own_at(nu, sm) : load_disp<ROWS>(U, nu, g, A, nu, -1, -1, 0, xmn));
		}
	}
	return S;
}

// staple_4d_own for whole x3-rows in the block (ROWS) with shared upper
// factors, software-pipelined: every product's phase first issues the global
// loads the next phase needs, and __builtin_amdgcn_sched_barrier keeps them
// ahead of the arithmetic, so that a wavefront multiplies while its next links
// are in flight (the compiler's own schedule waits on one link's loads at a
// time).  The operations and their order are those of staple_4d_own, so the
// bits are the same; at most five links are live.
template <int MU, int SH, class R>
__device__ __forceinline__ m3<R> staple_rows_pipe(Geo const &g, const R *__restrict__ U, long x, int const *c,
                                                  vec2_t<R> (*own)[9][64], int lane, vec2_t<R> (*sh)[9][64])
{
	const RowsAt At = rows_at(g, x, c);
	const int sp = At.sp, sm = At.sm;
	const long xpm = shift_c(g, x, MU, c[MU], 1);
	m3<R> S = zero_m3<R>(), A, B, La, Lb, Lc;
	// operands of the upper staple's first product (planes not shared)
	auto fetch_upper = [&](auto NU) {
		constexpr int nu = decltype(NU)::value;
		if constexpr (!plane_shared(SH, nu < MU ? nu : MU, nu < MU ? MU : nu))
		{
			A = MU == 3 ? lds_m3<R>(own, nu, sp) : load_disp<true>(U, nu, g, At, MU, 1, -1, 0, xpm);
			B = nu == 3 ? lds_m3<R>(own, MU, sp)
			            : load_disp<true>(U, MU, g, At, nu, 1, -1, 0, shift_c(g, x, nu, c[nu], 1));
		}
	};
	// the upper and lower staples of plane (MU, nu); NEXT: the next nu (-1: none)
	auto step = [&](auto NU, auto NEXT) {
		constexpr int nu = decltype(NU)::value, nx = decltype(NEXT)::value;
		const long xmn = nu == 3 ? 0 : shift_c(g, x, nu, c[nu], -1);
		La = load_disp<true>(U, nu, g, At, MU, 1, nu, -1, shift_c(g, xpm, nu, c[nu], -1));
		__builtin_amdgcn_sched_barrier(0);
		m3<R> T1;
		if constexpr (plane_shared(SH, nu < MU ? nu : MU, nu < MU ? MU : nu))
			T1 = nu < MU ? adj_m3(lds_m3<R>(sh, share_slot(SH, nu, MU), lane))
			             : lds_m3<R>(sh, share_slot(SH, MU, nu), lane);
		else
			T1 = upper_factor(A, B, nu < MU);
		Lb = nu == 3 ? lds_m3<R>(own, MU, sm) : load_disp<true>(U, MU, g, At, nu, -1, -1, 0, xmn);
		__builtin_amdgcn_sched_barrier(0);
		madd_nd(S, T1, lds_m3<R>(own, nu, lane));
		Lc = nu == 3 ? lds_m3<R>(own, nu, sm) : load_disp<true>(U, nu, g, At, nu, -1, -1, 0, xmn);
		__builtin_amdgcn_sched_barrier(0);
		m3<R> T2 = mul_dd(La, Lb);
		if constexpr (nx >= 0)
			fetch_upper(std::integral_constant<int, nx>{});
		__builtin_amdgcn_sched_barrier(0);
		madd_nn(S, T2, Lc);
	};
	constexpr int n0 = MU == 0 ? 1 : 0, n1 = MU <= 1 ? 2 : 1, n2 = MU <= 2 ? 3 : 2;
	fetch_upper(std::integral_constant<int, n0>{});
	step(std::integral_constant<int, n0>{}, std::integral_constant<int, n1>{});
	step(std::integral_constant<int, n1>{}, std::integral_constant<int, n2>{});
	step(std::integral_constant<int, n2>{}, std::integral_constant<int, -1>{});
	return S;
}

template <class R>
__device__ __forceinline__ m3<R> staple_any(Geo const &g, const R *__restrict__ U, int mu, long x, int const *c)
{
	if (g.nd != 4)
		return staple_at(g, U, mu, x, c);
	switch (mu)
	{
	case 0:
		return staple_4d<0>(g, U, x, c);
	case 1:
		return staple_4d<1>(g, U, x, c);
	case 2:
		return staple_4d<2>(g, U, x, c);
	default:
		return staple_4d<3>(g, U, x, c);
	}
}

// Re tr of all plaquettes at x, in double
template <class R> __device__ __forceinline__ double plaq_site(Geo const &g, const R *__restrict__ U, long x,
                                                             int const *c)
{
	double s = 0.0;
#pragma unroll
	for (int mu = 0; mu < 4; ++mu)
#pragma unroll
		for (int nu = mu + 1; nu < 4; ++nu)
		{
			if (nu >= g.nd)
				continue;
			long xpm = shift_c(g, x, mu, c[mu], 1), xpn = shift_c(g, x, nu, c[nu], 1);
			m3<double> A = mul_nn(to_f64(load(U, mu, g, x)), to_f64(load(U, nu, g, xpm)));
			m3<double> B = mul_nn(to_f64(load(U, nu, g, x)), to_f64(load(U, mu, g, xpn)));
			s += retr_nd(A, B);
		}
	return s;
}

template <bool PC, class R> __device__ __forceinline__ double kin_site(Geo const &g, const R *__restrict__ P, long x)
{
	double k = 0.0;
	for (int mu = 0; mu < g.nd; ++mu)
		k += norm2_m3(to_f64(PC ? load_alg<R, false>(P, mu, g, x) : load(P, mu, g, x)));
	return k;
}

// ---------------------------------------------------------------- kernels

template <class R> __global__ __launch_bounds__(256) void k_random_gauge(Geo g, int nf, R *U, uint64_t seed,
                                                                        uint32_t tag)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= g.V * nf)
		return;
	int f = (int)(i / g.V);
	long x = lane_site(g, i - (long)f * g.V);
	double z[18];
	link_normals<18>(seed, tag, 0u, (uint32_t)x, (uint32_t)f, z);
	store(U, f, g, x, from_f64<R>(haar_from_normals(z)));
}

template <class R> __global__ __launch_bounds__(256) void k_momenta(Geo g, R *P, double *K, uint64_t seed,
                                                                   uint32_t traj)
{
	long x = lane_site(g, (long)blockIdx.x * blockDim.x + threadIdx.x);
	if (x >= g.V)
		return;
	double k = 0.0;
	for (int mu = 0; mu < g.nd; ++mu)
	{
		double z[8];
		link_normals<8>(seed, TAG_MOM, traj, (uint32_t)x, (uint32_t)mu, z);
		m3<R> p = from_f64<R>(algebra_from_normals(z));
		store(P, mu, g, x, p);
		k += norm2_m3(to_f64(p));
	}
	if (K)
		K[x] = k;
}

// momenta refresh fused with the first link step of the trajectory:
// P_mu(x) drawn as in k_momenta, then Uout_mu(x) = exp(t TA(P_mu(x))) Uin_mu(x)
// as in k_link_update (the same operations, so the same bits); PC: P in the
// compact antihermitian layout
template <class R, bool PC>
__global__ __launch_bounds__(256) void k_momenta_link(Geo g, R *__restrict__ P, double *__restrict__ K, uint64_t seed,
                                                      uint32_t traj, R *__restrict__ Uout, const R *__restrict__ Uin,
                                                      R t, int reunit_flag)
{
	long x = lane_site(g, (long)blockIdx.x * blockDim.x + threadIdx.x);
	if (x >= g.V)
		return;
	double k = 0.0;
	for (int mu = 0; mu < g.nd; ++mu)
	{
		double z[8];
		link_normals<8>(seed, TAG_MOM, traj, (uint32_t)x, (uint32_t)mu, z);
		m3<R> p = from_f64<R>(algebra_from_normals(z));
		store_mom<PC>(P, mu, g, x, p);
		k += norm2_m3(to_f64(p));
		m3<R> X = ta(p);
#pragma unroll
		for (int e = 0; e < 9; ++e)
		{
			X.e[e].re *= t;
			X.e[e].im *= t;
		}
		m3<R> u = exp_mul(X, load_once(Uin, mu, g, x));
		if (reunit_flag)
			u = reunit(u);
		store_once(Uout, mu, g, x, u);
	}
	K[x] = k;
}

template <class R>
__global__ __launch_bounds__(256) void k_plaq(Geo g, const R *__restrict__ U, double *partials, int nb, int remap)
{
	int lb = ordered_block(g, nb, 256, remap);
	if (lb < 0)
		return;
	long x = lane_site(g, (long)lb * 256 + threadIdx.x);
	double v[1] = {0.0};
	if (x < g.V)
	{
		int c[4];
		coords<R>(g, x, c);
		v[0] = plaq_site(g, U, x, c);
	}
	block_sum_store<1>(v, partials, nb, lb);
}

// partial sums of (d(x), ps_old(x), ps_new(x), k_new(x), k_old(x)); PC: P in
// the compact antihermitian layout
template <class R, bool PC>
__global__ __launch_bounds__(256) DH_OCC void k_delta_h(Geo g, const R *__restrict__ Uo, const R *__restrict__ Un,
                                                 const R *__restrict__ P, const double *__restrict__ Ko,
                                                 double c3, double *partials, int nb, int remap)
{
	int lb = ordered_block(g, nb, 256, remap);
	if (lb < 0)
		return;
	long x = lane_site(g, (long)lb * 256 + threadIdx.x);
	double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
	if (x < g.V)
	{
		int c[4];
		coords<R>(g, x, c);
		double pso = plaq_site(g, Uo, x, c);
		double psn = plaq_site(g, Un, x, c);
		double kn = kin_site<PC>(g, P, x);
		double dps = psn - pso;
		double dk = kn - Ko[x];
		v[0] = fma(-c3, dps, dk);
		v[1] = pso;
		v[2] = psn;
		v[3] = kn;
		v[4] = Ko[x];
	}
	block_sum_store<5>(v, partials, nb, lb);
}

// The same with a workgroup of 256 threads over 128 sites: threads 0-127 sum
// the plaquettes of U_old, threads 128-255 those of U_new and the kinetic
// energy, for the same sites; the per-site values are those of k_delta_h (the
// block partials sum them in another order).
template <class R, bool PC>
__global__ __launch_bounds__(256) void k_delta_h_split(Geo g, const R *__restrict__ Uo, const R *__restrict__ Un,
                                                       const R *__restrict__ P, const double *__restrict__ Ko,
                                                       double c3, double *partials, int nb, int remap)
{
	__shared__ double pso_s[128];
	int lb = ordered_block(g, nb, 128, remap);
	if (lb < 0)
		return;
	const bool newer = threadIdx.x >= 128;
	const int t = threadIdx.x & 127;
	long x = lane_site(g, (long)lb * 128 + t);
	double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
	double ps = 0.0, kn = 0.0;
	if (x < g.V)
	{
		int c[4];
		coords<R>(g, x, c);
		if (newer)
		{
			ps = plaq_site(g, Un, x, c);
			kn = kin_site<PC>(g, P, x);
		}
		else
			ps = plaq_site(g, Uo, x, c);
	}
	if (!newer)
		pso_s[t] = ps;
	__syncthreads();
	if (x < g.V)
	{
		if (newer)
		{
			double pso = pso_s[t];
			double dps = ps - pso;
			double dk = kn - Ko[x];
			v[0] = fma(-c3, dps, dk);
			v[2] = ps;
			v[3] = kn;
			v[4] = Ko[x];
		}
		else
			v[1] = ps;
	}
	block_sum_store<5>(v, partials, nb, lb);
}

template <class R>
__global__ __launch_bounds__(256) void k_staple(Geo g, const R *__restrict__ U, int mu, R *__restrict__ S)
{
	long x = lane_site(g, (long)blockIdx.x * blockDim.x + threadIdx.x);
	if (x >= g.V)
		return;
	int c[4];
	coords<R>(g, x, c);
	store(S, 0, g, x, staple_at(g, U, mu, x, c));
}

template <class R>
__global__ __launch_bounds__(256) void k_force(Geo g, const R *__restrict__ U, int mu, R coef, R *__restrict__ F)
{
	long x = lane_site(g, (long)blockIdx.x * blockDim.x + threadIdx.x);
	if (x >= g.V)
		return;
	int c[4];
	coords<R>(g, x, c);
	m3<R> A = ta(mul_nn(load(U, mu, g, x), staple_at(g, U, mu, x, c)));
#pragma unroll
	for (int k = 0; k < 9; ++k)
	{
		A.e[k].re *= coef;
		A.e[k].im *= coef;
	}
	store(F, 0, g, x, A);
}

// the link step that may follow a momentum step in the same thread:
// Unext_mu(x) = exp(tl TA(p)) u (+ reunitarisation), u = U_mu(x), p = new P_mu(x)
template <class R>
__device__ __forceinline__ void fused_link_step(Geo const &g, R *__restrict__ Unext, int mu, long x, m3<R> const &p,
                                                m3<R> const &u, R tl, int reunit_flag)
{
	m3<R> X = ta(p);
#pragma unroll
	for (int k = 0; k < 9; ++k)
	{
		X.e[k].re *= tl;
		X.e[k].im *= tl;
	}
	m3<R> un = exp_mul(X, u);
	if (reunit_flag)
		un = reunit(un);
	store_once(Unext, mu, g, x, un);
}

// fused momentum step: workgroup = 64 sites x 4 directions, wavefront w
// handles direction w of the workgroup's sites.  With Unext != nullptr the
// following link step is fused in: Unext = exp(tl TA(P_new)) U (Unext != U,
// since neighbouring threads still read U), and P is in the compact
// antihermitian layout (the fused kernels run only inside a trajectory).
// Staple paths of k_force_update, one per kernel (a kernel is compiled for
// the registers of its largest path, and the general ones need more):
//   STAPLE_ROWS  4-d, V % 64 == 0, L3 | 64, (L2 L3) % 64 == 0: the block's
//                links staged in LDS,
//                own and x3-neighbour links read there;
//   STAPLE_OWN   4-d, V % 64 == 0: the block's links staged, own links read there;
//   STAPLE_ANY   any geometry: every link from global memory.
enum
{
	STAPLE_ANY = 0,
	STAPLE_OWN = 1,
	STAPLE_ROWS = 2
};
inline int staple_path(Geo const &g)
{
	if (!MESH_OWN_LDS || g.nd != 4 || (g.V & 63) != 0)
		return STAPLE_ANY;
	return (MESH_OWN_ROWS && 64 % g.L[3] == 0 && (g.L[2] * g.L[3]) % 64 == 0) ? STAPLE_ROWS : STAPLE_OWN;
}

template <class R, bool FUSE, bool PC, int PATH>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(force_min_waves<R, PATH>(), force_max_waves<R, PATH>()))) void k_force_update(
    Geo g, const R *__restrict__ U, R *__restrict__ P, double kc, int nb, int remap, R *__restrict__ Unext, R tl,
    int reunit_flag)
{
	int lb = ordered_block(g, nb, SB, remap);
	if (lb < 0)
		return;
	int mu = threadIdx.x >> 6;
	long x = lane_site(g, (long)lb * SB + (threadIdx.x & 63));
	m3<R> S, u;
	if constexpr (PATH != STAPLE_ANY)
	{
		// every thread of the workgroup is a valid (site, direction): stage the
		// block's own links, then the staples read them from LDS
		__shared__ vec2_t<R> own[4][9][64];
		const int lane = threadIdx.x & 63;
		int c[4];
		coords<R>(g, x, c);
		constexpr bool RW = PATH == STAPLE_ROWS; // the block holds whole x3-rows
		constexpr int SH = share_mode<R>();
		ShareOps<R> F;
		constexpr bool PF = std::is_same<R, float>::value ? MESH_SHARE_PREFETCH : MESH_SHARE_PREFETCH_F64;
		if constexpr (SH > 0 && PF)
			switch (mu)
			{
			case 0:
				share_fetch<0, RW, SH>(g, U, x, c, F);
				break;
			case 1:
				share_fetch<1, RW, SH>(g, U, x, c, F);
				break;
			case 2:
				share_fetch<2, RW, SH>(g, U, x, c, F);
				break;
			default:
				share_fetch<3, RW, SH>(g, U, x, c, F);
				break;
			}
		{
			const vec2_t<R> *p = reinterpret_cast<const vec2_t<R> *>(U) + site_off(g, mu, x);
#pragma unroll
			for (int k = 0; k < 9; ++k)
				own[mu][k][lane] = p[k * plane_step(g)];
		}
		__syncthreads();
		if constexpr (SH > 0)
		{
			// planes (0, q) are indices 0..2, so mode 1 needs three
			__shared__ vec2_t<R> sh[shared_planes(SH)][9][64];
			if constexpr (!PF)
				switch (mu)
				{
				case 0:
					share_fetch<0, RW, SH>(g, U, x, c, F);
					break;
				case 1:
					share_fetch<1, RW, SH>(g, U, x, c, F);
					break;
				case 2:
					share_fetch<2, RW, SH>(g, U, x, c, F);
					break;
				default:
					share_fetch<3, RW, SH>(g, U, x, c, F);
					break;
				}
			switch (mu)
			{
			case 0:
				share_upper_factors<0, RW, SH>(g, x, c, own, sh, lane, F);
				break;
			case 1:
				share_upper_factors<1, RW, SH>(g, x, c, own, sh, lane, F);
				break;
			case 2:
				share_upper_factors<2, RW, SH>(g, x, c, own, sh, lane, F);
				break;
			default:
				share_upper_factors<3, RW, SH>(g, x, c, own, sh, lane, F);
				break;
			}
			__syncthreads();
			constexpr bool PIPE = RW && MESH_PIPE_F64 && std::is_same<R, double>::value;
			switch (mu)
			{
			case 0:
				S = PIPE ? staple_rows_pipe<0, SH>(g, U, x, c, own, lane, sh)
				         : staple_4d_own<0, RW, R, SH>(g, U, x, c, own, lane, sh);
				break;
			case 1:
				S = PIPE ? staple_rows_pipe<1, SH>(g, U, x, c, own, lane, sh)
				         : staple_4d_own<1, RW, R, SH>(g, U, x, c, own, lane, sh);
				break;
			case 2:
				S = PIPE ? staple_rows_pipe<2, SH>(g, U, x, c, own, lane, sh)
				         : staple_4d_own<2, RW, R, SH>(g, U, x, c, own, lane, sh);
				break;
			default:
				S = PIPE ? staple_rows_pipe<3, SH>(g, U, x, c, own, lane, sh)
				         : staple_4d_own<3, RW, R, SH>(g, U, x, c, own, lane, sh);
				break;
			}
		}
		else
			switch (mu)
			{
			case 0:
				S = staple_4d_own<0, RW>(g, U, x, c, own, lane);
				break;
			case 1:
				S = staple_4d_own<1, RW>(g, U, x, c, own, lane);
				break;
			case 2:
				S = staple_4d_own<2, RW>(g, U, x, c, own, lane);
				break;
			default:
				S = staple_4d_own<3, RW>(g, U, x, c, own, lane);
				break;
			}
#pragma unroll
		for (int k = 0; k < 9; ++k)
		{
			vec2_t<R> v = own[mu][k][lane];
			u.e[k] = {v.x, v.y};
		}
	}
	else
	{
		if (mu >= g.nd || x >= g.V)
			return;
		int c[4];
		coords<R>(g, x, c);
		S = MESH_STATIC_MU ? staple_any(g, U, mu, x, c) : staple_at(g, U, mu, x, c);
		u = load(U, mu, g, x);
	}
	m3<R> p = mom_step<PC>(load_mom<PC>(P, mu, g, x), mul_nn(u, S), kc);
	store_mom<PC>(P, mu, g, x, p);
	if constexpr (FUSE)
		fused_link_step(g, Unext, mu, x, p, u, tl, reunit_flag);
}

// Fused momentum step with an LDS tile (4-d lattices whose fastest extent
// L3 divides 64 with R = 64 / L3 dividing L2).  A workgroup owns the 64
// sites of R consecutive x2-rows (one site block of the layout) and stages
// all four link directions of those rows plus one halo row on each side in
// LDS; every staple link displaced only along x2/x3 is then read from LDS and
// only the x0/x1-displaced ones from global memory.  Wavefront w handles
// direction w; the arithmetic (and the layout of P) is that of k_force_update.
template <class R, int L3, bool FUSE, bool PC = FUSE>
__global__ __launch_bounds__(256) FORCE_OCC void k_force_update_tile(Geo g, const R *__restrict__ U, R *__restrict__ P,
                                                                     double kc, int nb, int remap, R *__restrict__ Unext,
                                                                     R tl, int reunit_flag)
{
	constexpr int RW = 64 / L3; // rows per tile
	constexpr int NR = RW + 2;  // with halo rows
	// a staged row holds its 4 x 9 planes of L3 entries in the slot order of
	// the field (even x3 first), so the 16-lane groups of a wavefront read
	// consecutive entries; rows are padded by 128 B so that successive rows,
	// read in the same cycle by different lane groups, start on other banks
	constexpr int RS = 36 * L3 + 128 / (int)sizeof(vec2_t<R>);
	__shared__ vec2_t<R> lds[NR * RS];
	auto slot3 = [](int c3) { return MESH_SITE_PAIRS ? ((c3 & 1) * (L3 / 2) + (c3 >> 1)) : c3; };
	int lb = ordered_block(g, nb, 64, remap);
	if (lb < 0)
		return;
	const long base = (long)lb * 64; // first site of the tile (x3 = 0 of its first row)
	const int x2_0 = (int)((base / L3) % g.L[2]);
	const long row0 = base - (long)x2_0 * L3; // site index of (x0, x1, 0, 0)
	// stage rows x2_0 - 1 ... x2_0 + RW (periodic in x2)
	for (int e = threadIdx.x; e < NR * 4 * 9 * L3; e += 256)
	{
		const int j = e % L3; // slot within the row
		const int c = MESH_SITE_PAIRS ? (j < L3 / 2 ? 2 * j : 2 * (j - L3 / 2) + 1) : j;
		int k = (e / L3) % 9;
		int d = (e / (9 * L3)) % 4;
		int lr = e / (36 * L3);
		int x2 = x2_0 - 1 + lr;
		x2 = x2 < 0 ? x2 + g.L[2] : (x2 >= g.L[2] ? x2 - g.L[2] : x2);
		long xs = row0 + (long)x2 * L3 + c;
		lds[lr * RS + (d * 9 + k) * L3 + j] =
		    reinterpret_cast<const vec2_t<R> *>(U)[site_off(g, d, xs) + k * plane_step(g)];
	}
	__syncthreads();

	// the fused kernel gets the direction as a compile-time constant (no branch
	// between staples); the separate momentum step keeps it a wave-uniform
	// value (with constants the compiler spends all registers on it)
	auto compute = [&](auto mu_c) {
		const int mu = (int)mu_c;
		const long x = lane_site(g, base + (threadIdx.x & 63));
		const int r = (int)(x - base) / L3, c3 = (int)(x - base) % L3;
		int c[4];
		coords<R>(g, x, c);
		// matrix of direction d at the site displaced by (dx2, dx3) from x, from LDS
		auto from_lds = [&](int d, int dx2, int dx3) {
			int cc = c3 + dx3;
			cc = cc < 0 ? cc + L3 : (cc >= L3 ? cc - L3 : cc);
			const vec2_t<R> *row = lds + (r + 1 + dx2) * RS + d * 9 * L3 + slot3(cc);
			m3<R> m;
#pragma unroll
			for (int k = 0; k < 9; ++k)
			{
				vec2_t<R> v = row[k * L3];
				m.e[k] = {v.x, v.y};
			}
			return m;
		};
		// displacement of +/-1 along direction a as (dx2, dx3); a < 2 is non-local
		auto d2 = [](int a, int s) { return a == 2 ? s : 0; };
		auto d3 = [](int a, int s) { return a == 3 ? s : 0; };

		m3<R> S = zero_m3<R>();
		const bool mloc = mu >= 2;
		long xpm = shift_c(g, x, mu, c[mu], 1);
#pragma unroll
		for (int nu = 0; nu < 4; ++nu)
		{
			if (nu == mu)
				continue;
			const bool nloc = nu >= 2;
			long xpn = shift_c(g, x, nu, c[nu], 1);
			long xmn = shift_c(g, x, nu, c[nu], -1);
			{
				m3<R> a = mloc ? from_lds(nu, d2(mu, 1), d3(mu, 1)) : load(U, nu, g, xpm);
				m3<R> b = nloc ? from_lds(mu, d2(nu, 1), d3(nu, 1)) : load(U, mu, g, xpn);
				m3<R> T1 = nu < mu ? adj_m3(mul_nd(b, a)) : mul_nd(a, b);
				madd_nd(S, T1, from_lds(nu, 0, 0));
			}
			{
				m3<R> a = (mloc && nloc) ? from_lds(nu, d2(mu, 1) + d2(nu, -1), d3(mu, 1) + d3(nu, -1))
				                         : load(U, nu, g, shift_c(g, xpm, nu, c[nu], -1));
				m3<R> b = nloc ? from_lds(mu, d2(nu, -1), d3(nu, -1)) : load(U, mu, g, xmn);
				m3<R> T2 = mul_dd(a, b);
				madd_nn(S, T2, nloc ? from_lds(nu, d2(nu, -1), d3(nu, -1)) : load(U, nu, g, xmn));
			}
		}
		m3<R> u = from_lds(mu, 0, 0);
		m3<R> p = mom_step<PC>(load_mom<PC>(P, mu, g, x), mul_nn(u, S), kc);
		store_mom<PC>(P, mu, g, x, p);
		if constexpr (FUSE)
			fused_link_step(g, Unext, mu, x, p, u, tl, reunit_flag);
	};
	if constexpr (FUSE && MESH_STATIC_MU)
	{
		switch (threadIdx.x >> 6)
		{
		case 0:
			compute(std::integral_constant<int, 0>{});
			break;
		case 1:
			compute(std::integral_constant<int, 1>{});
			break;
		case 2:
			compute(std::integral_constant<int, 2>{});
			break;
		default:
			compute(std::integral_constant<int, 3>{});
			break;
		}
	}
	else
		compute((int)(threadIdx.x >> 6));
}

// variant: one thread per site looping over the directions
template <class R>
__global__ __launch_bounds__(256) void k_force_update_site(Geo g, const R *__restrict__ U, R *__restrict__ P, double kc,
                                                           int nb, int remap)
{
	int lb = ordered_block(g, nb, 256, remap);
	if (lb < 0)
		return;
	long x = lane_site(g, (long)lb * 256 + threadIdx.x);
	if (x >= g.V)
		return;
	int c[4];
	coords<R>(g, x, c);
	for (int mu = 0; mu < g.nd; ++mu)
	{
		m3<R> S = staple_at(g, U, mu, x, c);
		store(P, mu, g, x, mom_step<false>(load(P, mu, g, x), mul_nn(load(U, mu, g, x), S), kc));
	}
}

// link step over nf*V matrices: Uout = exp(t TA(P)) Uin (+ reunitarise)
template <class R>
__global__ __launch_bounds__(256) void k_link_update(Geo g, int nf, R *Uout, const R *Uin, const R *__restrict__ P,
                                                     R t, int reunit_flag)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= g.V * nf)
		return;
	int f = (int)(i / g.V);
	long x = lane_site(g, i - (long)f * g.V);
	m3<R> X = ta(load(P, f, g, x));
#pragma unroll
	for (int k = 0; k < 9; ++k)
	{
		X.e[k].re *= t;
		X.e[k].im *= t;
	}
	m3<R> u = exp_mul(X, load(Uin, f, g, x));
	if (reunit_flag)
		u = reunit(u);
	store(Uout, f, g, x, u);
}

// dst[x] = src[x + offset e_dir] (periodic) for nf matrices per site
template <class R>
__global__ __launch_bounds__(256) void k_cshift(Geo g, int nf, R *__restrict__ dst, const R *__restrict__ src, int dir,
                                                int offset)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= g.V * nf)
		return;
	int f = (int)(i / g.V);
	long x = lane_site(g, i - (long)f * g.V);
	const int c = coord(g, x, dir), L = g.L[dir];
	int cn = c + offset;
	cn = cn >= L ? cn - L : (cn < 0 ? cn + L : cn);
	store(dst, f, g, x, load(src, f, g, x + (long)(cn - c) * g.st[dir]));
}

// block partial sums of the 18 reals of every matrix (FP64)
template <class R>
__global__ __launch_bounds__(256) void k_sum_m3(Geo g, int nf, const R *__restrict__ src, double *partials, int nb)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	double v[18];
#pragma unroll
	for (int k = 0; k < 18; ++k)
		v[k] = 0.0;
	if (i < g.V * nf)
	{
		int f = (int)(i / g.V);
		long x = lane_site(g, i - (long)f * g.V);
		m3<R> m = load(src, f, g, x);
#pragma unroll
		for (int k = 0; k < 9; ++k)
		{
			v[2 * k] = (double)m.e[k].re;
			v[2 * k + 1] = (double)m.e[k].im;
		}
	}
	block_sum_store<18>(v, partials, nb, blockIdx.x);
}

// Y = Y + (a X) per real, the multiply and the add rounded separately as in
// the reference's lattice expression `Y += X * a`
template <class R>
__global__ __launch_bounds__(256) void k_axpy(Geo g, int nf, R *__restrict__ Y, R a, const R *__restrict__ X)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= g.V * nf)
		return;
	int f = (int)(i / g.V);
	long x = lane_site(g, i - (long)f * g.V);
	m3<R> y = load(Y, f, g, x);
	m3<R> b = load(X, f, g, x);
#pragma unroll
	for (int k = 0; k < 9; ++k)
	{
		R pr = b.e[k].re * a, pi = b.e[k].im * a;
		y.e[k].re = y.e[k].re + pr;
		y.e[k].im = y.e[k].im + pi;
	}
	store(Y, f, g, x, y);
}

template <class R> __global__ __launch_bounds__(256) void k_reunit(Geo g, int nf, R *U)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= g.V * nf)
		return;
	int f = (int)(i / g.V);
	long x = lane_site(g, i - (long)f * g.V);
	store(U, f, g, x, reunit(load(U, f, g, x)));
}

// sum over f,x of ||A||^2 (B == nullptr) or Re tr(A B), partials per block
template <class R>
__global__ __launch_bounds__(256) void k_trace_sum(Geo g, int nf, const R *__restrict__ A, const R *__restrict__ B,
                                                   double *partials, int nb)
{
	long x = lane_site(g, (long)blockIdx.x * 256 + threadIdx.x);
	double v[1] = {0.0};
	if (x < g.V)
	{
		double s = 0.0;
		for (int f = 0; f < nf; ++f)
		{
			m3<double> a = to_f64(load(A, f, g, x));
			if (!B)
				s += norm2_m3(a);
			else
			{
				m3<double> p = mul_nn(a, to_f64(load(B, f, g, x)));
				s += (p.e[0].re + p.e[4].re) + p.e[8].re;
			}
		}
		v[0] = s;
	}
	block_sum_store<1>(v, partials, nb, blockIdx.x);
}

// ---- the reference's float build (precision 1) energies: a per-site FP32
// value summed as device_sum_apply does (src/lattice/device.h:184-210), a
// 256-wide shared-memory pair tree per workgroup in FP32 over the sites in
// lexicographic order; the workgroup results are summed on the host in FP32.
// mode 0: Re tr of the (mu, nu) plaquette, mode 1: ||P_mu||^2 (PC: compact P).
template <bool PC>
__global__ __launch_bounds__(256) void k_float_tree(Geo g, const float *__restrict__ A, int mode, int mu, int nu,
                                                    float *__restrict__ partials)
{
	__shared__ float s[256];
	const long gi = (long)blockIdx.x * 256 + threadIdx.x;
	const int li = threadIdx.x;
	if (gi < g.V)
	{
		if (mode == 0)
		{
			const long xpm = shift_c(g, gi, mu, coord(g, gi, mu), 1), xpn = shift_c(g, gi, nu, coord(g, gi, nu), 1);
			m3<float> P1 = mul_nn(load(A, mu, g, gi), load(A, nu, g, xpm));
			m3<float> P2 = mul_nn(load(A, nu, g, gi), load(A, mu, g, xpn));
			s[li] = retr_nd(P1, P2);
		}
		else
			s[li] = norm2_m3(PC ? load_alg<float, false>(A, mu, g, gi) : load(A, mu, g, gi));
	}
	__syncthreads();
	for (int st = 128; st > 0; st >>= 1)
	{
		if (li < st && gi + st < g.V)
			s[li] += s[li + st];
		__syncthreads();
	}
	if (li == 0)
		partials[blockIdx.x] = s[0];
}

// ---------------------------------------------------------------- host helpers

struct Setup
{
	Geo geo;
	int rc;
};

Setup setup(mesh_ctx *ctx, const mesh_grid *g, int prec)
{
	Setup s;
	s.rc = mesh_make_geo(ctx, g, &s.geo);
	if (!s.rc)
		s.rc = mesh_check_prec(ctx, prec);
	return s;
}

inline int nblk(long n, int b) { return (int)((n + b - 1) / b); }

int plaq_sum(mesh_ctx *ctx, Geo const &geo, int prec, const void *U, double *out)
{
	int nb = nblk(geo.V, 256);
	if (int rc = mesh_reserve_partials(ctx, nb))
		return rc;
	int gb = mesh_grid_blocks(ctx, nb);
	int ph = mesh_prof_begin(ctx, PROF_PLAQ);
	if (prec == MESH_F64)
		k_plaq<double><<<gb, 256, 0, ctx->stream>>>(geo, (const double *)U, ctx->partials, nb, mesh_order(ctx));
	else
		k_plaq<float><<<gb, 256, 0, ctx->stream>>>(geo, (const float *)U, ctx->partials, nb, mesh_order(ctx));
	mesh_prof_end(ctx, ph);
	MESH_LAUNCHED(ctx);
	return mesh_finish_reduction(ctx, nb, 1, out);
}

double n_plaquettes(Geo const &geo) { return (double)geo.V * geo.nd * (geo.nd - 1) / 2; }

int launch_link_update(mesh_ctx *ctx, Geo const &geo, int nf, int prec, void *Uout, const void *Uin, const void *P,
                       double t, int reunit_flag);

// force kernel variant in use: 0 wavefront per direction (the default in both
// precisions), 1 thread per site, 2 LDS tile (falls back to 0 where the tile
// does not fit)
int force_variant_of(mesh_ctx *ctx, int)
{
	return ctx->force_variant >= 0 ? ctx->force_variant : 0;
}

// momentum step P -= delta beta/6 TA(U Sigma); with Unext != nullptr the
// link step Unext = exp(tl TA(P)) U (+ reunit) follows, fused into the same
// kernel where the variant supports it (Unext must differ from U; with
// `compact` P is in the compact antihermitian layout, which only the fused
// kernels read: force_variant_of != 1)
int launch_force_update(mesh_ctx *ctx, Geo const &geo, int prec, const void *U, void *P, double beta, double delta,
                        void *Unext = nullptr, double tl = 0.0, int reunit_flag = 0, bool compact = false)
{
	double k = delta * (beta / 6.0);
	int ph = mesh_prof_begin(ctx, Unext ? PROF_FORCE_LINK : PROF_FORCE_UPDATE);
	const int L3 = geo.L[3], RW = L3 > 0 ? 64 / L3 : 0;
	const bool tile_ok = geo.nd == 4 && (L3 == 16 || L3 == 32 || L3 == 64) && geo.L[2] % RW == 0;
	bool fused = Unext != nullptr;
	const int variant = force_variant_of(ctx, prec);
	if (variant == 2 && tile_ok)
	{
		int nb = (int)(geo.V / 64);
		int gb = mesh_grid_blocks(ctx, nb);
		auto launch = [&](auto tag, auto l3) {
			using Rt = decltype(tag);
			constexpr int l3v = decltype(l3)::value;
			if (fused && compact)
				k_force_update_tile<Rt, l3v, true, true><<<gb, 256, 0, ctx->stream>>>(
				    geo, (const Rt *)U, (Rt *)P, k, nb, mesh_order(ctx), (Rt *)Unext, (Rt)tl, reunit_flag);
			else if (fused)
				k_force_update_tile<Rt, l3v, true, false><<<gb, 256, 0, ctx->stream>>>(
				    geo, (const Rt *)U, (Rt *)P, k, nb, mesh_order(ctx), (Rt *)Unext, (Rt)tl, reunit_flag);
			else
				k_force_update_tile<Rt, l3v, false><<<gb, 256, 0, ctx->stream>>>(
				    geo, (const Rt *)U, (Rt *)P, k, nb, mesh_order(ctx), nullptr, (Rt)0, 0);
		};
		using I16 = std::integral_constant<int, 16>;
		using I32 = std::integral_constant<int, 32>;
		using I64 = std::integral_constant<int, 64>;
		if (prec == MESH_F64)
			L3 == 16 ? launch(0.0, I16{}) : L3 == 32 ? launch(0.0, I32{}) : launch(0.0, I64{});
		else
			L3 == 16 ? launch(0.0f, I16{}) : L3 == 32 ? launch(0.0f, I32{}) : launch(0.0f, I64{});
	}
	else if (variant == 1)
	{
		int nb = nblk(geo.V, 256);
		int gb = mesh_grid_blocks(ctx, nb);
		if (prec == MESH_F64)
			k_force_update_site<double><<<gb, 256, 0, ctx->stream>>>(geo, (const double *)U, (double *)P, k, nb,
			                                                          ctx->xcd_remap);
		else
			k_force_update_site<float><<<gb, 256, 0, ctx->stream>>>(geo, (const float *)U, (float *)P, k,
			                                                         nb, ctx->xcd_remap);
		fused = false;
	}
	else
	{
		int nb = nblk(geo.V, SB);
		int gb = mesh_grid_blocks(ctx, nb);
		auto launch = [&](auto tag, auto path) {
			using Rt = decltype(tag);
			constexpr int PT = decltype(path)::value;
			if (fused && compact)
				k_force_update<Rt, true, true, PT><<<gb, 256, 0, ctx->stream>>>(
				    geo, (const Rt *)U, (Rt *)P, k, nb, mesh_order(ctx), (Rt *)Unext, (Rt)tl, reunit_flag);
			else if (fused)
				k_force_update<Rt, true, false, PT><<<gb, 256, 0, ctx->stream>>>(
				    geo, (const Rt *)U, (Rt *)P, k, nb, mesh_order(ctx), (Rt *)Unext, (Rt)tl, reunit_flag);
			else
				k_force_update<Rt, false, false, PT><<<gb, 256, 0, ctx->stream>>>(
				    geo, (const Rt *)U, (Rt *)P, k, nb, mesh_order(ctx), nullptr, (Rt)0, 0);
		};
		auto by_path = [&](auto tag) {
			switch (staple_path(geo))
			{
			case STAPLE_ROWS:
				launch(tag, std::integral_constant<int, STAPLE_ROWS>{});
				break;
			case STAPLE_OWN:
				launch(tag, std::integral_constant<int, STAPLE_OWN>{});
				break;
			default:
				launch(tag, std::integral_constant<int, STAPLE_ANY>{});
				break;
			}
		};
		prec == MESH_F64 ? by_path(0.0) : by_path(0.0f);
	}
	mesh_prof_end(ctx, ph);
	MESH_LAUNCHED(ctx);
	if (Unext && !fused)
	{
		if (compact)
			return mesh_fail(ctx, MESH_E_ARG, "compact momenta need a fused force variant");
		return launch_link_update(ctx, geo, geo.nd, prec, Unext, U, P, tl, reunit_flag);
	}
	return MESH_OK;
}

int launch_link_update(mesh_ctx *ctx, Geo const &geo, int nf, int prec, void *Uout, const void *Uin, const void *P,
                       double t, int reunit_flag)
{
	int nb = nblk(geo.V * nf, 256);
	int ph = mesh_prof_begin(ctx, PROF_LINK_UPDATE);
	if (prec == MESH_F64)
		k_link_update<double><<<nb, 256, 0, ctx->stream>>>(geo, nf, (double *)Uout, (const double *)Uin,
		                                                   (const double *)P, t, reunit_flag);
	else
		k_link_update<float><<<nb, 256, 0, ctx->stream>>>(geo, nf, (float *)Uout, (const float *)Uin,
		                                                  (const float *)P, (float)t, reunit_flag);
	mesh_prof_end(ctx, ph);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int launch_momenta(mesh_ctx *ctx, Geo const &geo, int prec, void *P, double *K, uint64_t seed, uint32_t traj)
{
	int nb = nblk(geo.V, 256);
	int ph = mesh_prof_begin(ctx, PROF_MOMENTA);
	if (prec == MESH_F64)
		k_momenta<double><<<nb, 256, 0, ctx->stream>>>(geo, (double *)P, K, seed, traj);
	else
		k_momenta<float><<<nb, 256, 0, ctx->stream>>>(geo, (float *)P, K, seed, traj);
	mesh_prof_end(ctx, ph);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int launch_momenta_link(mesh_ctx *ctx, Geo const &geo, int prec, void *P, double *K, uint64_t seed, uint32_t traj,
                        void *Uout, const void *Uin, double t, int reunit_flag, bool compact)
{
	int nb = nblk(geo.V, 256);
	int ph = mesh_prof_begin(ctx, PROF_MOMENTA_LINK);
	auto launch = [&](auto tag, auto pc) {
		using Rt = decltype(tag);
		k_momenta_link<Rt, decltype(pc)::value><<<nb, 256, 0, ctx->stream>>>(
		    geo, (Rt *)P, K, seed, traj, (Rt *)Uout, (const Rt *)Uin, (Rt)t, reunit_flag);
	};
	using PC1 = std::true_type;
	using PC0 = std::false_type;
	if (prec == MESH_F64)
		compact ? launch(0.0, PC1{}) : launch(0.0, PC0{});
	else
		compact ? launch(0.0f, PC1{}) : launch(0.0f, PC0{});
	mesh_prof_end(ctx, ph);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

// out = {wilsonAction(U, beta), norm2(P)} of the reference's float build
// (src/gauge/utils.h:94-108, 141-146, 27-41): each plane's and each
// direction's FP32 sum, accumulated in FP64
int float_energy(mesh_ctx *ctx, Geo const &geo, const float *U, const float *P, bool compact, double beta,
                 double *out)
{
	const int nb = nblk(geo.V, 256);
	if (int rc = mesh_reserve_partials(ctx, nb))
		return rc;
	float *dp = reinterpret_cast<float *>(ctx->partials);
	std::vector<float> h(nb);
	auto tree = [&](const float *A, int mode, int mu, int nu, bool pc, float *r) -> int {
		if (pc)
			k_float_tree<true><<<nb, 256, 0, ctx->stream>>>(geo, A, mode, mu, nu, dp);
		else
			k_float_tree<false><<<nb, 256, 0, ctx->stream>>>(geo, A, mode, mu, nu, dp);
		MESH_LAUNCHED(ctx);
		MESH_HIP(ctx, hipMemcpyAsync(h.data(), dp, nb * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
		MESH_HIP(ctx, hipStreamSynchronize(ctx->stream));
		float acc = h[0];
		for (int i = 1; i < nb; ++i)
			acc += h[i];
		*r = acc;
		return MESH_OK;
	};
	const int nd = geo.nd;
	double sp = 0.0, sk = 0.0;
	for (int mu = 0; mu < nd; ++mu)
		for (int nu = mu + 1; nu < nd; ++nu)
		{
			float r;
			if (int rc = tree(U, 0, mu, nu, false, &r))
				return rc;
			sp += r;
		}
	const double vol = (double)geo.V;
	const double plaq = sp / (vol * nd * (nd - 1) * 0.5) / 3.0;
	out[0] = (1.0 - plaq) * (beta * nd * (nd - 1) * 0.5 * vol);
	for (int mu = 0; mu < nd && P; ++mu)
	{
		float r;
		if (int rc = tree(P, 1, mu, 0, compact, &r))
			return rc;
		sk += r;
	}
	out[1] = sk;
	return MESH_OK;
}

int delta_h(mesh_ctx *ctx, Geo const &geo, int prec, const void *Uo, const void *Un, const void *P, const double *Ko,
            double beta, double *out, bool compact = false)
{
	const int sbk = MESH_DH_SPLIT ? 128 : 256; // sites per workgroup
	int nb = nblk(geo.V, sbk);
	if (int rc = mesh_reserve_partials(ctx, 5 * (size_t)nb))
		return rc;
	int gb = mesh_grid_blocks(ctx, nb);
	double c3 = beta / 3.0;
	int ph = mesh_prof_begin(ctx, PROF_DELTA_H);
	auto launch = [&](auto tag, auto pc) {
		using Rt = decltype(tag);
		if (MESH_DH_SPLIT)
			k_delta_h_split<Rt, decltype(pc)::value><<<gb, 256, 0, ctx->stream>>>(
			    geo, (const Rt *)Uo, (const Rt *)Un, (const Rt *)P, Ko, c3, ctx->partials, nb, mesh_order(ctx));
		else
			k_delta_h<Rt, decltype(pc)::value><<<gb, 256, 0, ctx->stream>>>(
			    geo, (const Rt *)Uo, (const Rt *)Un, (const Rt *)P, Ko, c3, ctx->partials, nb, mesh_order(ctx));
	};
	using PC1 = std::true_type;
	using PC0 = std::false_type;
	if (prec == MESH_F64)
		compact ? launch(0.0, PC1{}) : launch(0.0, PC0{});
	else
		compact ? launch(0.0f, PC1{}) : launch(0.0f, PC0{});
	mesh_prof_end(ctx, ph);
	MESH_LAUNCHED(ctx);
	double r[5];
	if (int rc = mesh_finish_reduction(ctx, nb, 5, r))
		return rc;
	out[0] = r[0];
	out[1] = r[1];
	out[2] = r[2];
	out[3] = r[4];
	out[4] = r[3];
	return MESH_OK;
}

// Metropolis uniform of trajectory traj: Philox(seed; traj, 0, 0, METRO)
double metropolis_uniform(uint64_t seed, uint32_t traj)
{
	u4 c = {traj, 0u, 0u, TAG_METRO};
	u4 r = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
	return u01_from_bits(r.x, r.y);
}

} // namespace

extern "C" {

int mesh_su3_random_gauge(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, void *U, uint64_t seed,
                          uint32_t stream)
{
	MESH_CHECK_CTX(ctx);
	Setup s = setup(ctx, g, prec);
	if (s.rc)
		return s.rc;
	if (!U || nf < 1)
		return mesh_fail(ctx, MESH_E_ARG, "bad random_gauge arguments");
	uint32_t tag = stream ? stream : TAG_HAAR;
	int nb = nblk(s.geo.V * nf, 256);
	if (prec == MESH_F64)
		k_random_gauge<double><<<nb, 256, 0, ctx->stream>>>(s.geo, nf, (double *)U, seed, tag);
	else
		k_random_gauge<float><<<nb, 256, 0, ctx->stream>>>(s.geo, nf, (float *)U, seed, tag);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_su3_momenta(mesh_ctx *ctx, const mesh_grid *g, int prec, void *P, double *K, uint64_t seed, uint32_t traj)
{
	MESH_CHECK_CTX(ctx);
	Setup s = setup(ctx, g, prec);
	if (s.rc)
		return s.rc;
	if (!P)
		return mesh_fail(ctx, MESH_E_ARG, "P is NULL");
	return launch_momenta(ctx, s.geo, prec, P, K, seed, traj);
}

int mesh_su3_plaquette(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U, double *out)
{
	MESH_CHECK_CTX(ctx);
	Setup s = setup(ctx, g, prec);
	if (s.rc)
		return s.rc;
	if (!U || !out)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to plaquette");
	double sum;
	if (int rc = plaq_sum(ctx, s.geo, prec, U, &sum))
		return rc;
	*out = sum / (n_plaquettes(s.geo) * 3.0);
	return MESH_OK;
}

int mesh_su3_wilson_action(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U, double beta, double *out)
{
	MESH_CHECK_CTX(ctx);
	Setup s = setup(ctx, g, prec);
	if (s.rc)
		return s.rc;
	if (!U || !out)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to wilson_action");
	double sum;
	if (int rc = plaq_sum(ctx, s.geo, prec, U, &sum))
		return rc;
	*out = beta * (n_plaquettes(s.geo) - sum / 3.0);
	return MESH_OK;
}

static int trace_sum(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, const void *A, const void *B, double *out)
{
	MESH_CHECK_CTX(ctx);
	Setup s = setup(ctx, g, prec);
	if (s.rc)
		return s.rc;
	if (!A || !out || nf < 1)
		return mesh_fail(ctx, MESH_E_ARG, "bad trace-sum arguments");
	int nb = nblk(s.geo.V, 256);
	if (int rc = mesh_reserve_partials(ctx, nb))
		return rc;
	if (prec == MESH_F64)
		k_trace_sum<double><<<nb, 256, 0, ctx->stream>>>(s.geo, nf, (const double *)A, (const double *)B,
		                                                 ctx->partials, nb);
	else
		k_trace_sum<float><<<nb, 256, 0, ctx->stream>>>(s.geo, nf, (const float *)A, (const float *)B,
		                                                ctx->partials, nb);
	MESH_LAUNCHED(ctx);
	return mesh_finish_reduction(ctx, nb, 1, out);
}

int mesh_su3_norm2(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, const void *P, double *out)
{
	return trace_sum(ctx, g, nf, prec, P, nullptr, out);
}

int mesh_su3_sum_real_trace_prod(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, const void *A, const void *B,
                                 double *out)
{
	if (!B)
		return mesh_fail(ctx, MESH_E_ARG, "B is NULL");
	return trace_sum(ctx, g, nf, prec, A, B, out);
}

int mesh_su3_staple(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U, int mu, void *S)
{
	MESH_CHECK_CTX(ctx);
	Setup s = setup(ctx, g, prec);
	if (s.rc)
		return s.rc;
	if (!U || !S || mu < 0 || mu >= s.geo.nd)
		return mesh_fail(ctx, MESH_E_ARG, "bad staple arguments");
	int nb = nblk(s.geo.V, 256);
	if (prec == MESH_F64)
		k_staple<double><<<nb, 256, 0, ctx->stream>>>(s.geo, (const double *)U, mu, (double *)S);
	else
		k_staple<float><<<nb, 256, 0, ctx->stream>>>(s.geo, (const float *)U, mu, (float *)S);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_su3_force(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U, int mu, double beta, void *F)
{
	MESH_CHECK_CTX(ctx);
	Setup s = setup(ctx, g, prec);
	if (s.rc)
		return s.rc;
	if (!U || !F || mu < 0 || mu >= s.geo.nd)
		return mesh_fail(ctx, MESH_E_ARG, "bad force arguments");
	int nb = nblk(s.geo.V, 256);
	double c = beta / 6.0;
	if (prec == MESH_F64)
		k_force<double><<<nb, 256, 0, ctx->stream>>>(s.geo, (const double *)U, mu, c, (double *)F);
	else
		k_force<float><<<nb, 256, 0, ctx->stream>>>(s.geo, (const float *)U, mu, (float)c, (float *)F);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_su3_force_update(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U, void *P, double beta,
                          double delta)
{
	MESH_CHECK_CTX(ctx);
	Setup s = setup(ctx, g, prec);
	if (s.rc)
		return s.rc;
	if (!U || !P)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to force_update");
	return launch_force_update(ctx, s.geo, prec, U, P, beta, delta);
}

int mesh_su3_link_update(mesh_ctx *ctx, const mesh_grid *g, int prec, void *Uout, const void *Uin, const void *P,
                         double t, int reunit_flag)
{
	MESH_CHECK_CTX(ctx);
	Setup s = setup(ctx, g, prec);
	if (s.rc)
		return s.rc;
	if (!Uout || !Uin || !P)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to link_update");
	return launch_link_update(ctx, s.geo, s.geo.nd, prec, Uout, Uin, P, t, reunit_flag);
}

int mesh_su3_exp_mul(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, void *U, const void *F, double t)
{
	MESH_CHECK_CTX(ctx);
	Setup s = setup(ctx, g, prec);
	if (s.rc)
		return s.rc;
	if (!U || !F || nf < 1)
		return mesh_fail(ctx, MESH_E_ARG, "bad exp_mul arguments");
	return launch_link_update(ctx, s.geo, nf, prec, U, U, F, t, 0);
}

int mesh_su3_cshift(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, void *dst, const void *src, int dir,
                    int offset)
{
	MESH_CHECK_CTX(ctx);
	Setup s = setup(ctx, g, prec);
	if (s.rc)
		return s.rc;
	if (!dst || !src || nf < 1)
		return mesh_fail(ctx, MESH_E_ARG, "bad cshift arguments");
	if (dst == src)
		return mesh_fail(ctx, MESH_E_ARG, "inplace cshift not supported");
	if (dir < 0 || dir >= s.geo.nd || offset == 0 || offset <= -s.geo.L[dir] || offset >= s.geo.L[dir])
		return mesh_fail(ctx, MESH_E_ARG, "cshift needs 0 <= dir < nd and 0 < |offset| < L[dir]");
	int nb = nblk(s.geo.V * nf, 256);
	if (prec == MESH_F64)
		k_cshift<double><<<nb, 256, 0, ctx->stream>>>(s.geo, nf, (double *)dst, (const double *)src, dir, offset);
	else
		k_cshift<float><<<nb, 256, 0, ctx->stream>>>(s.geo, nf, (float *)dst, (const float *)src, dir, offset);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_su3_sum(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, const void *src, double *out)
{
	MESH_CHECK_CTX(ctx);
	Setup s = setup(ctx, g, prec);
	if (s.rc)
		return s.rc;
	if (!src || !out || nf < 1)
		return mesh_fail(ctx, MESH_E_ARG, "bad sum arguments");
	int nb = nblk(s.geo.V * nf, 256);
	if (int rc = mesh_reserve_partials(ctx, 18 * (size_t)nb))
		return rc;
	if (prec == MESH_F64)
		k_sum_m3<double><<<nb, 256, 0, ctx->stream>>>(s.geo, nf, (const double *)src, ctx->partials, nb);
	else
		k_sum_m3<float><<<nb, 256, 0, ctx->stream>>>(s.geo, nf, (const float *)src, ctx->partials, nb);
	MESH_LAUNCHED(ctx);
	return mesh_finish_reduction(ctx, nb, 18, out);
}

int mesh_su3_axpy(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, void *Y, double a, const void *X)
{
	MESH_CHECK_CTX(ctx);
	Setup s = setup(ctx, g, prec);
	if (s.rc)
		return s.rc;
	if (!Y || !X || nf < 1)
		return mesh_fail(ctx, MESH_E_ARG, "bad axpy arguments");
	int nb = nblk(s.geo.V * nf, 256);
	if (prec == MESH_F64)
		k_axpy<double><<<nb, 256, 0, ctx->stream>>>(s.geo, nf, (double *)Y, a, (const double *)X);
	else
		k_axpy<float><<<nb, 256, 0, ctx->stream>>>(s.geo, nf, (float *)Y, (float)a, (const float *)X);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_su3_reunit(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, void *U)
{
	MESH_CHECK_CTX(ctx);
	Setup s = setup(ctx, g, prec);
	if (s.rc)
		return s.rc;
	if (!U || nf < 1)
		return mesh_fail(ctx, MESH_E_ARG, "bad reunit arguments");
	int nb = nblk(s.geo.V * nf, 256);
	if (prec == MESH_F64)
		k_reunit<double><<<nb, 256, 0, ctx->stream>>>(s.geo, nf, (double *)U);
	else
		k_reunit<float><<<nb, 256, 0, ctx->stream>>>(s.geo, nf, (float *)U);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_su3_delta_h(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *Uo, const void *Un, const void *P,
                     const double *Ko, double beta, double *out)
{
	MESH_CHECK_CTX(ctx);
	Setup s = setup(ctx, g, prec);
	if (s.rc)
		return s.rc;
	if (!Uo || !Un || !P || !Ko || !out)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to delta_h");
	return delta_h(ctx, s.geo, prec, Uo, Un, P, Ko, beta, out);
}

int mesh_make_deltas(const char *scheme, double epsilon, int substeps, double *out, int cap)
{
	if (!scheme || !out || substeps < 1 || !(epsilon > 0))
		return MESH_E_ARG;
	std::vector<double> raw;
	if (!std::strcmp(scheme, "2lf"))
		raw = {0.5, 1.0, 0.5};
	else if (!std::strcmp(scheme, "2mn"))
	{
		double l = 0.19318332750378361;
		raw = {l, 0.5, 1.0 - 2.0 * l, 0.5, l};
	}
	else if (!std::strcmp(scheme, "4mn"))
	{
		double rho = 0.1786178958448091, theta = -0.06626458266981843, lam = 0.7123418310626056;
		raw = {rho, lam, theta, 0.5 - lam, 1.0 - 2.0 * (rho + theta), 0.5 - lam, theta, lam, rho};
	}
	else
		return MESH_E_ARG;
	// Langevin-time epsilon to MD step: delta = sqrt(2 epsilon) / substeps;
	// consecutive substeps share their boundary position step
	double delta = std::sqrt(2 * epsilon) / substeps;
	std::vector<double> d;
	for (int i = 0; i < substeps; ++i)
	{
		if (i > 0)
			d.back() *= 2;
		d.insert(d.end(), raw.begin() + (i > 0 ? 1 : 0), raw.end());
	}
	if ((int)d.size() > cap)
		return MESH_E_SIZE;
	for (size_t i = 0; i < d.size(); ++i)
		out[i] = d[i] * delta;
	return (int)d.size();
}

int mesh_su3_hmc_trajectory(mesh_ctx *ctx, const mesh_grid *g, int prec, void **U, void **U_work, void *P,
                            double *K, double beta, const double *deltas, int nd_, uint64_t seed, uint32_t traj,
                            int metropolis, double *out)
{
	MESH_CHECK_CTX(ctx);
	Setup s = setup(ctx, g, prec);
	if (s.rc)
		return s.rc;
	if (!U || !U_work || !*U || !*U_work || !P || !K || !deltas || !out)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to hmc_trajectory");
	if (nd_ < 1 || nd_ % 2 != 1)
		return mesh_fail(ctx, MESH_E_ARG, "deltas must have odd length (src/gauge/hmc.h:44)");
	if (*U == *U_work)
		return mesh_fail(ctx, MESH_E_ARG, "U and U_work must differ");
	int rc;
	const int nfused = (nd_ - 1) / 2;
	// option "float_sums": Delta H as the reference's float build forms it
	const bool fsum = ctx->float_sums && prec == MESH_F32;
	double h_old[2] = {0.0, 0.0};
	// the fused kernels keep P in the compact antihermitian layout (half the
	// bytes of a matrix field, the same bits; option "compact_momenta" = 0
	// keeps the 18-real matrix layout)
	const bool compact = ctx->compact_p && ctx->fuse_md && nfused > 0 && force_variant_of(ctx, prec) != 1;
	if (!(ctx->fuse_md && nfused > 0) && (rc = launch_momenta(ctx, s.geo, prec, P, K, seed, traj)))
		return rc;
	if (ctx->fuse_md && nfused > 0)
	{
		// momentum step i fused with link step i+1, ping-pong between the
		// caller's work field and a spare; the first link step writes the
		// field that makes the last one land in *U_work
		size_t fb = mesh_su3_field_bytes(g, s.geo.nd, prec);
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
		if ((rc = launch_momenta_link(ctx, s.geo, prec, P, K, seed, traj, cur, *U, deltas[0], 0, compact)))
			return rc;
		// the link step only reads P: P is still the trajectory's initial momenta
		if (fsum && (rc = float_energy(ctx, s.geo, (const float *)*U, (const float *)P, compact, beta, h_old)))
			return rc;
		for (int i = 1; i + 1 < nd_; i += 2)
		{
			if ((rc = launch_force_update(ctx, s.geo, prec, cur, P, beta, deltas[i], other, deltas[i + 1],
			                              i + 1 == nd_ - 1, compact)))
				return rc;
			void *t = cur;
			cur = other;
			other = t;
		}
	}
	else
		for (int i = 0; i < nd_; ++i)
		{
			if (i == 0 && fsum &&
			    (rc = float_energy(ctx, s.geo, (const float *)*U, (const float *)P, false, beta, h_old)))
				return rc;
			if (i % 2 == 0)
				rc = launch_link_update(ctx, s.geo, s.geo.nd, prec, *U_work, i == 0 ? *U : *U_work, P, deltas[i],
				                        i == nd_ - 1);
			else
				rc = launch_force_update(ctx, s.geo, prec, *U_work, P, beta, deltas[i]);
			if (rc)
				return rc;
		}
	double dh[5];
	if ((rc = delta_h(ctx, s.geo, prec, *U, *U_work, P, K, beta, dh, compact)))
		return rc;
	double np = n_plaquettes(s.geo);
	double H_old = beta * (np - dh[1] / 3.0) + dh[3], H_new = beta * (np - dh[2] / 3.0) + dh[4];
	if (fsum)
	{
		// the reference's H_new - H_old of separately summed FP32 energies (hmc.cpp:85-92)
		double h_new[2];
		if ((rc = float_energy(ctx, s.geo, (const float *)*U_work, (const float *)P, compact, beta, h_new)))
			return rc;
		H_old = h_old[0] + h_old[1];
		H_new = h_new[0] + h_new[1];
		dh[0] = H_new - H_old;
	}
	double u = metropolis_uniform(seed, traj);
	int acc = metropolis ? (u < std::exp(-dh[0])) : 1;
	if (acc)
	{
		void *t = *U;
		*U = *U_work;
		*U_work = t;
	}
	out[0] = dh[0];
	out[1] = acc;
	out[2] = (acc ? dh[2] : dh[1]) / (np * 3.0);
	out[3] = H_old;
	out[4] = H_new;
	out[5] = u;
	return MESH_OK;
}

int mesh_su3_float_energy(mesh_ctx *ctx, const mesh_grid *g, const float *U, const float *P, int p_compact,
                          double beta, double *out)
{
	MESH_CHECK_CTX(ctx);
	Setup s = setup(ctx, g, MESH_F32);
	if (s.rc)
		return s.rc;
	if (!U || !out)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to su3_float_energy");
	return float_energy(ctx, s.geo, U, P, p_compact != 0, beta, out);
}

} // extern "C"

// ---- plaquette without allocation -------------------------------------------
//
// mesh_su3_plaquette_noalloc keeps its per-block partial sums in this
// file-scope array instead of the context's buffer (which grows with
// hipMalloc).  In host code the array's name is not a device address:
// mesh_scratch_setup resolves the address once per device with
// hipGetSymbolAddress, and launches take that pointer.  One partial per 256
// sites: 65536 doubles (512 KiB per GPU, held until the process ends) cover
// lattices up to 2^24 sites (64^4).
constexpr int NOALLOC_PARTIALS = 65536;
static __device__ __align__(256) double scratch[NOALLOC_PARTIALS];

namespace {

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex scratch_mu;
double *scratch_dev[SCRATCH_MAX_DEVICES] = {};

__global__ void k_div1(double *o, double d)
{
	if (threadIdx.x == 0)
		o[0] = o[0] / d;
}

} // namespace

int mesh_scratch_setup(mesh_ctx *ctx)
{
	std::lock_guard<std::mutex> lock(scratch_mu);
	if (ctx->device < 0 || ctx->device >= SCRATCH_MAX_DEVICES)
		return mesh_fail(ctx, MESH_E_ARG, "device index beyond the scratch table");
	if (!scratch_dev[ctx->device])
	{
		void *p = nullptr;
		hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !p)
			return mesh_hip_fail(ctx, e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
		scratch_dev[ctx->device] = static_cast<double *>(p);
	}
	ctx->scratch = scratch_dev[ctx->device];
	return MESH_OK;
}

extern "C" {

int mesh_su3_plaquette_noalloc(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U, double *out_dev)
{
	MESH_CHECK_CTX(ctx);
	Setup s = setup(ctx, g, prec);
	if (s.rc)
		return s.rc;
	const int nb = nblk(s.geo.V, 256);
	if (nb > NOALLOC_PARTIALS)
		return mesh_fail(ctx, MESH_E_SIZE,
		                 "plaquette_noalloc: lattice exceeds 2^24 sites (the fixed scratch holds 65536 block sums)");
	if (!U || !out_dev)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to plaquette_noalloc");
	if (!ctx->scratch)
		return mesh_fail(ctx, MESH_E_STATE, "plaquette_noalloc: scratch address not resolved");
	const int gb = mesh_grid_blocks(ctx, nb);
	if (prec == MESH_F64)
		k_plaq<double><<<gb, 256, 0, ctx->stream>>>(s.geo, (const double *)U, ctx->scratch, nb, mesh_order(ctx));
	else
		k_plaq<float><<<gb, 256, 0, ctx->stream>>>(s.geo, (const float *)U, ctx->scratch, nb, mesh_order(ctx));
	MESH_LAUNCHED(ctx);
	if (int rc = mesh_reduce_to_device(ctx, ctx->scratch, nb, out_dev))
		return rc;
	k_div1<<<1, 64, 0, ctx->stream>>>(out_dev, n_plaquettes(s.geo) * 3.0);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

} // extern "C"
