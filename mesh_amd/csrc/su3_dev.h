// Device-side SU(3) algebra for gfx950.
//
// All multiply-adds are explicit fma() and the library is compiled with
// -ffp-contract=off, so each function performs one fixed sequence of IEEE
// operations.  The sequences are chosen to be the natural minimal ones
// (4 fma per complex multiply-add, k-loops in ascending order) and are the
// same sequences the CPU restatement used by the tests performs, which makes
// FP64 molecular dynamics reproducible bit for bit across CPU and GPU.
//
// Reference semantics: src/gauge/groups.h:198-268 (adj, real_trace,
// project_on_algebra, project_on_group_fast, exp).
#pragma once
#include <cmath>
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <utility>

namespace mesh_dev {

// float and double overloads of fma also on the host side of hipcc
using std::fma;

template <class R> struct vec2;
template <> struct vec2<double>
{
	using type = double2;
};
template <> struct vec2<float>
{
	using type = float2;
};
template <class R> using vec2_t = typename vec2<R>::type;

template <class R> struct cpx
{
	R re, im;
};

template <class R> struct m3
{
	cpx<R> e[9]; // row-major
};

#define MD __device__ __forceinline__
// the element algebra also compiles for the host side of a hipcc translation
// unit (the generic lattice operations evaluate the same lambdas on the host)
#define MHD __host__ __device__ __forceinline__

template <class R> MHD cpx<R> c0() { return {R(0), R(0)}; }

// acc += a*b
template <class R> MHD void cmadd(cpx<R> &acc, cpx<R> a, cpx<R> b)
{
	acc.re = fma(a.re, b.re, acc.re);
	acc.re = fma(-a.im, b.im, acc.re);
	acc.im = fma(a.re, b.im, acc.im);
	acc.im = fma(a.im, b.re, acc.im);
}
// acc -= a*b
template <class R> MHD void cmsub(cpx<R> &acc, cpx<R> a, cpx<R> b)
{
	acc.re = fma(-a.re, b.re, acc.re);
	acc.re = fma(a.im, b.im, acc.re);
	acc.im = fma(-a.re, b.im, acc.im);
	acc.im = fma(-a.im, b.re, acc.im);
}
// acc += a*conj(b)
template <class R> MHD void cmadd_cb(cpx<R> &acc, cpx<R> a, cpx<R> b)
{
	acc.re = fma(a.re, b.re, acc.re);
	acc.re = fma(a.im, b.im, acc.re);
	acc.im = fma(a.im, b.re, acc.im);
	acc.im = fma(-a.re, b.im, acc.im);
}
// acc += conj(a)*b
template <class R> MHD void cmadd_ca(cpx<R> &acc, cpx<R> a, cpx<R> b)
{
	acc.re = fma(a.re, b.re, acc.re);
	acc.re = fma(a.im, b.im, acc.re);
	acc.im = fma(a.re, b.im, acc.im);
	acc.im = fma(-a.im, b.re, acc.im);
}
// acc += conj(a)*conj(b)
template <class R> MHD void cmadd_cc(cpx<R> &acc, cpx<R> a, cpx<R> b)
{
	acc.re = fma(a.re, b.re, acc.re);
	acc.re = fma(-a.im, b.im, acc.re);
	acc.im = fma(-a.re, b.im, acc.im);
	acc.im = fma(-a.im, b.re, acc.im);
}
template <class R> MHD cpx<R> cmul(cpx<R> a, cpx<R> b)
{
	cpx<R> r = c0<R>();
	cmadd(r, a, b);
	return r;
}

// C = A B
template <class R> MHD m3<R> mul_nn(m3<R> const &A, m3<R> const &B)
{
	m3<R> C;
#pragma unroll
	for (int i = 0; i < 3; ++i)
#pragma unroll
		for (int j = 0; j < 3; ++j)
		{
			cpx<R> acc = c0<R>();
#pragma unroll
			for (int k = 0; k < 3; ++k)
				cmadd(acc, A.e[3 * i + k], B.e[3 * k + j]);
			C.e[3 * i + j] = acc;
		}
	return C;
}
// C = A B^+
template <class R> MHD m3<R> mul_nd(m3<R> const &A, m3<R> const &B)
{
	m3<R> C;
#pragma unroll
	for (int i = 0; i < 3; ++i)
#pragma unroll
		for (int j = 0; j < 3; ++j)
		{
			cpx<R> acc = c0<R>();
#pragma unroll
			for (int k = 0; k < 3; ++k)
				cmadd_cb(acc, A.e[3 * i + k], B.e[3 * j + k]);
			C.e[3 * i + j] = acc;
		}
	return C;
}
// C = A^+ B^+
template <class R> MHD m3<R> mul_dd(m3<R> const &A, m3<R> const &B)
{
	m3<R> C;
#pragma unroll
	for (int i = 0; i < 3; ++i)
#pragma unroll
		for (int j = 0; j < 3; ++j)
		{
			cpx<R> acc = c0<R>();
#pragma unroll
			for (int k = 0; k < 3; ++k)
				cmadd_cc(acc, A.e[3 * k + i], B.e[3 * j + k]);
			C.e[3 * i + j] = acc;
		}
	return C;
}
// S += A B
template <class R> MHD void madd_nn(m3<R> &S, m3<R> const &A, m3<R> const &B)
{
#pragma unroll
	for (int i = 0; i < 3; ++i)
#pragma unroll
		for (int j = 0; j < 3; ++j)
		{
			cpx<R> acc = S.e[3 * i + j];
#pragma unroll
			for (int k = 0; k < 3; ++k)
				cmadd(acc, A.e[3 * i + k], B.e[3 * k + j]);
			S.e[3 * i + j] = acc;
		}
}
// S += A B^+
template <class R> MHD void madd_nd(m3<R> &S, m3<R> const &A, m3<R> const &B)
{
#pragma unroll
	for (int i = 0; i < 3; ++i)
#pragma unroll
		for (int j = 0; j < 3; ++j)
		{
			cpx<R> acc = S.e[3 * i + j];
#pragma unroll
			for (int k = 0; k < 3; ++k)
				cmadd_cb(acc, A.e[3 * i + k], B.e[3 * j + k]);
			S.e[3 * i + j] = acc;
		}
}

// A^+ (exact: a transpose and a sign)
template <class R> MHD m3<R> adj_m3(m3<R> const &A)
{
	m3<R> r;
#pragma unroll
	for (int i = 0; i < 3; ++i)
#pragma unroll
		for (int j = 0; j < 3; ++j)
			r.e[3 * i + j] = {A.e[3 * j + i].re, -A.e[3 * j + i].im};
	return r;
}
// First factor of the upper staple of direction mu in plane nu,
// U_nu(x+mu) U_mu(x+nu)^+ from a = U_nu(x+mu), b = U_mu(x+nu).  For nu < mu
// it is the adjoint of b a^+, the factor of direction nu in plane mu, so the
// two directions of a plane use one rounded product.
template <class R> MHD m3<R> upper_factor(m3<R> const &a, m3<R> const &b, bool nu_below_mu)
{
	return nu_below_mu ? adj_m3(mul_nd(b, a)) : mul_nd(a, b);
}
template <class R> MHD m3<R> zero_m3()
{
	m3<R> r;
#pragma unroll
	for (int k = 0; k < 9; ++k)
		r.e[k] = c0<R>();
	return r;
}
template <class R> MHD m3<R> unit_m3()
{
	m3<R> r = zero_m3<R>();
	r.e[0].re = R(1);
	r.e[4].re = R(1);
	r.e[8].re = R(1);
	return r;
}

// Re tr(A B^+)
template <class R> MHD R retr_nd(m3<R> const &A, m3<R> const &B)
{
	R s = R(0);
#pragma unroll
	for (int k = 0; k < 9; ++k)
	{
		s = fma(A.e[k].re, B.e[k].re, s);
		s = fma(A.e[k].im, B.e[k].im, s);
	}
	return s;
}
template <class R> MHD R norm2_m3(m3<R> const &A)
{
	R s = R(0);
#pragma unroll
	for (int k = 0; k < 9; ++k)
	{
		s = fma(A.e[k].re, A.e[k].re, s);
		s = fma(A.e[k].im, A.e[k].im, s);
	}
	return s;
}

// traceless antihermitian part (M - M^+)/2 - tr(M - M^+)/6
template <class R> MHD m3<R> ta(m3<R> const &M)
{
	m3<R> A;
	R t = ((M.e[0].im + M.e[4].im) + M.e[8].im) * (R(1) / R(3));
#pragma unroll
	for (int i = 0; i < 3; ++i)
	{
		A.e[4 * i] = {R(0), M.e[4 * i].im - t};
#pragma unroll
		for (int j = 0; j < 3; ++j)
			if (j != i)
			{
				A.e[3 * i + j].re = R(0.5) * (M.e[3 * i + j].re - M.e[3 * j + i].re);
				A.e[3 * i + j].im = R(0.5) * (M.e[3 * i + j].im + M.e[3 * j + i].im);
			}
	}
	return A;
}

template <class R> MHD cpx<R> det(m3<R> const &X)
{
	cpx<R> m0 = c0<R>(), m1 = c0<R>(), m2 = c0<R>();
	cmadd(m0, X.e[4], X.e[8]);
	cmsub(m0, X.e[5], X.e[7]);
	cmadd(m1, X.e[3], X.e[8]);
	cmsub(m1, X.e[5], X.e[6]);
	cmadd(m2, X.e[3], X.e[7]);
	cmsub(m2, X.e[4], X.e[6]);
	cpx<R> d = c0<R>();
	cmadd(d, X.e[0], m0);
	cmsub(d, X.e[1], m1);
	cmadd(d, X.e[2], m2);
	return d;
}

template <class R> constexpr int exp_terms() { return sizeof(R) == 8 ? 18 : 10; }
// for ||X||_F^2 <= EXP_SMALL_R2 the first exp_terms_small() terms suffice:
// the remainder is below 0.2^7/7! = 2.5e-9 (FP32) or 0.2^12/12! = 8.6e-18
// (FP64), under half an ulp of 1 as with the full count at ||X||_F = 1
constexpr double EXP_SMALL_R2 = 0.04;
template <class R> constexpr int exp_terms_small() { return sizeof(R) == 8 ? 11 : 6; }

// X^2 of an exactly antihermitian X (every t TA(P) is one: TA's roundings are
// antisymmetric): hermitian, so the upper triangle is formed and the lower
// one is its conjugate, and the diagonal is the real -sum_k |X_ik|^2
template <class R> MHD m3<R> sq_ah(m3<R> const &X)
{
	m3<R> Q;
#pragma unroll
	for (int i = 0; i < 3; ++i)
	{
		R s = R(0);
#pragma unroll
		for (int k = 0; k < 3; ++k)
		{
			s = fma(X.e[3 * i + k].re, X.e[3 * i + k].re, s);
			s = fma(X.e[3 * i + k].im, X.e[3 * i + k].im, s);
		}
		Q.e[4 * i] = {-s, R(0)};
	}
#pragma unroll
	for (int i = 0; i < 3; ++i)
#pragma unroll
		for (int j = i + 1; j < 3; ++j)
		{
			cpx<R> acc = c0<R>();
#pragma unroll
			for (int k = 0; k < 3; ++k)
				cmadd(acc, X.e[3 * i + k], X.e[3 * k + j]);
			Q.e[3 * i + j] = acc;
			Q.e[3 * j + i] = {acc.re, -acc.im};
		}
	return Q;
}

// Im det X of an exactly antihermitian X (its determinant is imaginary):
// with H = -i X hermitian, det X = -i det H and
// det H = h0 h1 h2 + 2 Re(H01 H12 H20) - h0 |H12|^2 - h1 |H02|^2 - h2 |H01|^2
template <class R> MHD R imdet_ah(m3<R> const &X)
{
	const R h0 = X.e[0].im, h1 = X.e[4].im, h2 = X.e[8].im;
	const cpx<R> a = {X.e[1].im, -X.e[1].re}; // H01
	const cpx<R> b = {X.e[5].im, -X.e[5].re}; // H12
	const cpx<R> c = {X.e[2].im, -X.e[2].re}; // H02 (H20 = conj H02)
	const cpx<R> p = cmul(a, b);
	const R t = fma(p.re, c.re, p.im * c.im);
	const R n01 = fma(a.re, a.re, a.im * a.im), n12 = fma(b.re, b.re, b.im * b.im),
	        n02 = fma(c.re, c.re, c.im * c.im);
	R d = (h0 * h1) * h2;
	d = fma(R(2), t, d);
	d = fma(-h0, n12, d);
	d = fma(-h1, n02, d);
	d = fma(-h2, n01, d);
	return -d;
}

// one term of the Cayley-Hamilton series; b = i bi is the (imaginary) det X
template <class R, int N>
MHD void exp_series_term(R a, R bi, cpx<R> &al, cpx<R> &be, cpx<R> &ga, cpx<R> &f0, cpx<R> &f1, cpx<R> &f2)
{
	const R inv = R(1) / R(N + 1);
	cpx<R> gb = {-ga.im * bi, ga.re * bi};
	cpx<R> nal = {gb.re * inv, gb.im * inv};
	cpx<R> nbe = {fma(ga.re, a, al.re) * inv, fma(ga.im, a, al.im) * inv};
	cpx<R> nga = {be.re * inv, be.im * inv};
	al = nal;
	be = nbe;
	ga = nga;
	f0.re += al.re;
	f0.im += al.im;
	f1.re += be.re;
	f1.im += be.im;
	f2.re += ga.re;
	f2.im += ga.im;
}
template <class R, int N0, int... N>
MHD void exp_series(R a, R bi, cpx<R> &al, cpx<R> &be, cpx<R> &ga, cpx<R> &f0, cpx<R> &f1, cpx<R> &f2,
                   std::integer_sequence<int, N...>)
{
	(exp_series_term<R, N0 + N>(a, bi, al, be, ga, f0, f1, f2), ...);
}

// exp(X) - 1 of an exactly antihermitian, traceless matrix (X = t TA(P)):
// Cayley-Hamilton with power-series coefficients (X^3 = a X + b with
// a = tr X^2 / 2, b = det X, imaginary for such X),
// exp X = f0 1 + f1 X + f2 X^2, argument halved until ||X||_F <= 1 and the
// result squared back as D -> 2 D + D^2; the series stops after
// exp_terms_small() terms where ||X||_F^2 <= EXP_SMALL_R2.
//
// The 1 is never added: f0 - 1 is summed from its small terms and every entry
// of D is formed at its own magnitude.  Rounding exp X itself to FP32 puts a
// diagonal near 1 on a grid of 6e-8 and makes |exp X| systematically 1.4e-9
// too small per link step (measured over 4e5 random links); in U + D U only
// the final sum is rounded at the magnitude of U, which is unbiased.
template <class R> MHD m3<R> expm1su3(m3<R> X)
{
	m3<R> X2 = sq_ah(X);
	R trx2 = (X2.e[0].re + X2.e[4].re) + X2.e[8].re;
	R r2 = -trx2;
	int s = 0;
	R sc = R(1);
	// bounded: an infinite or NaN argument must not loop forever
	while (r2 > R(1) && s < 64)
	{
		r2 *= R(0.25);
		sc *= R(0.5);
		++s;
	}
	if (s > 0)
	{
		R sc2 = sc * sc;
#pragma unroll
		for (int k = 0; k < 9; ++k)
		{
			X.e[k].re *= sc;
			X.e[k].im *= sc;
			X2.e[k].re *= sc2;
			X2.e[k].im *= sc2;
		}
		trx2 *= sc2;
	}
	R a = R(0.5) * trx2;
	const R bi = imdet_ah(X);
	cpx<R> al = {R(1), R(0)}, be = c0<R>(), ga = c0<R>();
	cpx<R> g0 = c0<R>(), f1 = c0<R>(), f2 = c0<R>();
	constexpr int NS = exp_terms_small<R>(), NF = exp_terms<R>();
	exp_series<R, 0>(a, bi, al, be, ga, g0, f1, f2, std::make_integer_sequence<int, NS>{});
	if (r2 > R(EXP_SMALL_R2))
		exp_series<R, NS>(a, bi, al, be, ga, g0, f1, f2, std::make_integer_sequence<int, NF - NS>{});
	m3<R> D;
#pragma unroll
	for (int k = 0; k < 9; ++k)
	{
		cpx<R> acc = (k % 4 == 0) ? g0 : c0<R>();
		cmadd(acc, f1, X.e[k]);
		cmadd(acc, f2, X2.e[k]);
		D.e[k] = acc;
	}
	for (int k = 0; k < s; ++k)
	{
		m3<R> N;
#pragma unroll
		for (int i = 0; i < 3; ++i)
#pragma unroll
			for (int j = 0; j < 3; ++j)
			{
				cpx<R> acc = {D.e[3 * i + j].re + D.e[3 * i + j].re, D.e[3 * i + j].im + D.e[3 * i + j].im};
#pragma unroll
				for (int l = 0; l < 3; ++l)
					cmadd(acc, D.e[3 * i + l], D.e[3 * l + j]);
				N.e[3 * i + j] = acc;
			}
		D = N;
	}
	return D;
}

// U + D U (D = exp X - 1): the product summed from zero, then one rounding at
// the magnitude of U
template <class R> MHD m3<R> add_mul(m3<R> const &D, m3<R> const &U)
{
	m3<R> r;
#pragma unroll
	for (int i = 0; i < 3; ++i)
#pragma unroll
		for (int j = 0; j < 3; ++j)
		{
			cpx<R> acc = c0<R>();
#pragma unroll
			for (int k = 0; k < 3; ++k)
				cmadd(acc, D.e[3 * i + k], U.e[3 * k + j]);
			r.e[3 * i + j] = {U.e[3 * i + j].re + acc.re, U.e[3 * i + j].im + acc.im};
		}
	return r;
}

// link evolution exp(X) U (src/gauge/hmc.h:26-36) as U + (exp X - 1) U
template <class R> MHD m3<R> exp_mul(m3<R> const &X, m3<R> const &U) { return add_mul(expm1su3(X), U); }

// exp of a traceless antihermitian matrix, 1 + (exp X - 1)
template <class R> MHD m3<R> expsu3(m3<R> const &X)
{
	m3<R> E = expm1su3(X);
	E.e[0].re += R(1);
	E.e[4].re += R(1);
	E.e[8].re += R(1);
	return E;
}

// project_on_group_fast (src/gauge/groups.h:237-243): r = 1.5 a - 0.5 a a^+ a,
// then r *= 4/3 - det(r)/3, evaluated as the same quantities in increment
// form, r = a + (1 - a a^+) a / 2 and r += r (1 - det r) / 3: the constants
// 4/3 and 1/3 would be rounded up in FP32 (each by 3e-8 relative) and scale
// every link up by 1.3e-8; here 1/3 only multiplies a difference of order the
// non-unitarity.
template <class R> MHD m3<R> reunit_inc(m3<R> const &A)
{
	m3<R> M; // 1 - A A^+
#pragma unroll
	for (int i = 0; i < 3; ++i)
#pragma unroll
		for (int j = 0; j < 3; ++j)
		{
			cpx<R> acc = {i == j ? R(1) : R(0), R(0)};
#pragma unroll
			for (int k = 0; k < 3; ++k)
			{
				cpx<R> a = A.e[3 * i + k], b = A.e[3 * j + k];
				acc.re = fma(-a.re, b.re, acc.re);
				acc.re = fma(-a.im, b.im, acc.re);
				acc.im = fma(-a.im, b.re, acc.im);
				acc.im = fma(a.re, b.im, acc.im);
			}
			M.e[3 * i + j] = acc;
		}
	m3<R> r;
#pragma unroll
	for (int i = 0; i < 3; ++i)
#pragma unroll
		for (int j = 0; j < 3; ++j)
		{
			cpx<R> acc = c0<R>();
#pragma unroll
			for (int k = 0; k < 3; ++k)
				cmadd(acc, M.e[3 * i + k], A.e[3 * k + j]);
			r.e[3 * i + j] = {fma(R(0.5), acc.re, A.e[3 * i + j].re), fma(R(0.5), acc.im, A.e[3 * i + j].im)};
		}
	cpx<R> d = det(r);
	cpx<R> c = {(R(1) - d.re) * (R(1) / R(3)), -d.im * (R(1) / R(3))};
#pragma unroll
	for (int k = 0; k < 9; ++k)
	{
		cpx<R> q = cmul(r.e[k], c);
		r.e[k] = {r.e[k].re + q.re, r.e[k].im + q.im};
	}
	return r;
}

template <class R> MHD m3<double> to_f64(m3<R> const &a)
{
	m3<double> r;
#pragma unroll
	for (int k = 0; k < 9; ++k)
		r.e[k] = {(double)a.e[k].re, (double)a.e[k].im};
	return r;
}
template <class R> MHD m3<R> from_f64(m3<double> const &a)
{
	m3<R> r;
#pragma unroll
	for (int k = 0; k < 9; ++k)
		r.e[k] = {(R)a.e[k].re, (R)a.e[k].im};
	return r;
}

// Reunitarisation of a stored link.  FP32 links are projected in FP64 and
// rounded once: in FP32 the projection's own roundings (1 - a a^+ is formed
// from terms near 1) leave a mean ||U||^2/3 - 1 of -3.9e-9, against 1e-11
// when rounded once from FP64 (both over 4e5 random links).
template <class R> MHD m3<R> reunit(m3<R> const &A)
{
	if constexpr (sizeof(R) == 8)
		return reunit_inc(A);
	else
		return from_f64<R>(reunit_inc(to_f64(A)));
}

// momentum step P <- P - k TA(M), M = U S (src/gauge/hmc.h:57 with
// src/gauge/utils.h:148-153): TA and the update in FP64, P rounded once.  A
// force rounded to FP32 before the update (a function of the links alone,
// unlike P) raises the trajectory's mean Delta H by 4.6e-8 per site against
// the FP64 force (2mn, 12 substeps, thermalised 12^4 at beta = 6); formed in
// FP64 from the FP32 product the difference is at the noise floor.  With AH
// the caller's P is exactly antihermitian (the compact layout): the strict
// lower triangle is then written as -conj of the upper one, which is the
// value the full update would round to.
template <bool AH, class R> MHD m3<R> mom_step(m3<R> const &p, m3<R> const &M, double k)
{
	m3<double> A = ta(to_f64(M));
	m3<R> r;
#pragma unroll
	for (int i = 0; i < 3; ++i)
#pragma unroll
		for (int j = 0; j < 3; ++j)
		{
			const int e = 3 * i + j;
			if (AH && j < i)
				continue;
			r.e[e].re = (R)fma(-k, A.e[e].re, (double)p.e[e].re);
			r.e[e].im = (R)fma(-k, A.e[e].im, (double)p.e[e].im);
		}
	if constexpr (AH)
	{
#pragma unroll
		for (int i = 0; i < 3; ++i)
#pragma unroll
			for (int j = 0; j < i; ++j)
				r.e[3 * i + j] = {-r.e[3 * j + i].re, r.e[3 * j + i].im};
	}
	return r;
}

// ---- geometry ---------------------------------------------------------------

struct Geo
{
	int nd;
	int L[4];
	int st[4]; // site strides, st[nd-1] = 1
	long V;
	long stride; // sites per direction block, V rounded up to a multiple of 256
};

// ---- field access -------------------------------------------------------------
//
// Matrix f of site x, complex entry k:
//   base[f * 9 * stride + (x / B) * 9 * B + k * B + slot(x % B)]
// i.e. each direction is split into blocks of B = 2^MESH_SITE_LB consecutive
// sites and a block stores its 9 complex planes back to back
// (MESH_SITE_LB = 0: plain SoA, one block of `stride` sites).  Within a
// plane the even sites of the block come first, then the odd ones:
// slot(l) = (l & 1) * B/2 + l / 2 (MESH_SITE_PAIRS = 1), so that with an even
// fastest extent each checkerboard parity of a row is a contiguous run (the
// Landau half-sweep reads and writes whole cache lines); a wavefront of B
// consecutive sites still covers the same contiguous B-site plane segment.
#ifndef MESH_SITE_LB
#define MESH_SITE_LB 6
#endif
#ifndef MESH_SITE_PAIRS
#define MESH_SITE_PAIRS 1
#endif

MHD long site_off(Geo const &g, int f, long x)
{
	long o = (long)(f * 9) * g.stride;
	if (MESH_SITE_LB)
	{
		long l = x & ((1L << MESH_SITE_LB) - 1);
		if (MESH_SITE_PAIRS)
			l = ((l & 1) << (MESH_SITE_LB - 1)) | (l >> 1);
		return o + (((x >> MESH_SITE_LB) * 9) << MESH_SITE_LB) + l;
	}
	return o + x;
}
MD long plane_step(Geo const &g) { return MESH_SITE_LB ? (1L << MESH_SITE_LB) : g.stride; }

// Site of the linear thread index t in kernels with one thread per site: the
// inverse of the slot order within each B-site group, so that lane l of a
// wavefront touches plane slot l and a wavefront's accesses stay contiguous.
// Identity unless V is a multiple of B (the permutation must map [0, V) onto
// itself).
MD long lane_site(Geo const &g, long t)
{
#if MESH_SITE_PAIRS && MESH_SITE_LB
	constexpr long B = 1L << MESH_SITE_LB, hb = B / 2;
	if ((g.V & (B - 1)) == 0)
	{
		long l = t & (B - 1);
		l = l < hb ? 2 * l : 2 * (l - hb) + 1;
		return (t & ~(B - 1)) | l;
	}
#endif
	return t;
}

template <class R> MD m3<R> load(const R *__restrict__ base, int f, Geo const &g, long x)
{
	const vec2_t<R> *p = reinterpret_cast<const vec2_t<R> *>(base) + site_off(g, f, x);
	const long ks = plane_step(g);
	m3<R> m;
#pragma unroll
	for (int k = 0; k < 9; ++k)
	{
		vec2_t<R> v = p[k * ks];
		m.e[k] = {v.x, v.y};
	}
	return m;
}
template <class R> MD void store(R *__restrict__ base, int f, Geo const &g, long x, m3<R> const &m)
{
	vec2_t<R> *p = reinterpret_cast<vec2_t<R> *>(base) + site_off(g, f, x);
	const long ks = plane_step(g);
#pragma unroll
	for (int k = 0; k < 9; ++k)
	{
		vec2_t<R> v;
		v.x = m.e[k].re;
		v.y = m.e[k].im;
		p[k * ks] = v;
	}
}

// The same with non-temporal hints, for fields a kernel reads or writes once
// (momenta, the next link field): they should not evict the links that
// neighbouring workgroups still read through the L2.
template <class R> using nt2_t = R __attribute__((ext_vector_type(2)));
template <class R> MD m3<R> load_stream(const R *__restrict__ base, int f, Geo const &g, long x)
{
	const nt2_t<R> *p = reinterpret_cast<const nt2_t<R> *>(base) + site_off(g, f, x);
	const long ks = plane_step(g);
	m3<R> m;
#pragma unroll
	for (int k = 0; k < 9; ++k)
	{
		nt2_t<R> v = __builtin_nontemporal_load(p + k * ks);
		m.e[k] = {v.x, v.y};
	}
	return m;
}
template <class R> MD void store_stream(R *__restrict__ base, int f, Geo const &g, long x, m3<R> const &m)
{
	nt2_t<R> *p = reinterpret_cast<nt2_t<R> *>(base) + site_off(g, f, x);
	const long ks = plane_step(g);
#pragma unroll
	for (int k = 0; k < 9; ++k)
	{
		nt2_t<R> v = {m.e[k].re, m.e[k].im};
		__builtin_nontemporal_store(v, p + k * ks);
	}
}

// Compact field of exactly antihermitian matrices (the momenta inside one HMC
// trajectory): 9 reals per matrix instead of 18.  Per block of B sites, in
// the slot order of site_off: planes 0-2 hold the complex upper entries
// (0,1), (0,2), (1,2), plane 3 the pair (Im m00, Im m11), then a half plane
// of Im m22; direction f starts at 9 * f * S reals.  Unpacking sets the lower
// entries to -conj of the upper ones and the diagonal real parts to zero, so a
// matrix with these symmetries in every bit (momenta from
// algebra_from_normals, updated by fma with ta-projected forces, which keep
// them) comes back bit for bit, up to the sign of zero entries.
MD long alg_block(Geo const &g, int f, long x)
{
	long o = (long)(f * 9) * g.stride;
	if (MESH_SITE_LB)
		o += ((x >> MESH_SITE_LB) * 9) << MESH_SITE_LB;
	return o;
}
MD long alg_slot(long x)
{
	if (!MESH_SITE_LB)
		return x;
	long l = x & ((1L << MESH_SITE_LB) - 1);
	return MESH_SITE_PAIRS ? (((l & 1) << (MESH_SITE_LB - 1)) | (l >> 1)) : l;
}
template <class R, bool NT> MD m3<R> load_alg(const R *__restrict__ base, int f, Geo const &g, long x)
{
	const R *b = base + alg_block(g, f, x);
	const long l = alg_slot(x), ks = plane_step(g);
	const nt2_t<R> *p = reinterpret_cast<const nt2_t<R> *>(b) + l;
	nt2_t<R> v[4];
	R d2;
	if constexpr (NT)
	{
#pragma unroll
		for (int k = 0; k < 4; ++k)
			v[k] = __builtin_nontemporal_load(p + k * ks);
		d2 = __builtin_nontemporal_load(b + 8 * ks + l);
	}
	else
	{
#pragma unroll
		for (int k = 0; k < 4; ++k)
			v[k] = p[k * ks];
		d2 = b[8 * ks + l];
	}
	m3<R> m;
	m.e[1] = {v[0].x, v[0].y};
	m.e[3] = {-v[0].x, v[0].y};
	m.e[2] = {v[1].x, v[1].y};
	m.e[6] = {-v[1].x, v[1].y};
	m.e[5] = {v[2].x, v[2].y};
	m.e[7] = {-v[2].x, v[2].y};
	m.e[0] = {R(0), v[3].x};
	m.e[4] = {R(0), v[3].y};
	m.e[8] = {R(0), d2};
	return m;
}
template <class R, bool NT> MD void store_alg(R *__restrict__ base, int f, Geo const &g, long x, m3<R> const &m)
{
	R *b = base + alg_block(g, f, x);
	const long l = alg_slot(x), ks = plane_step(g);
	nt2_t<R> *p = reinterpret_cast<nt2_t<R> *>(b) + l;
	const nt2_t<R> v[4] = {{m.e[1].re, m.e[1].im},
	                       {m.e[2].re, m.e[2].im},
	                       {m.e[5].re, m.e[5].im},
	                       {m.e[0].im, m.e[4].im}};
	if constexpr (NT)
	{
#pragma unroll
		for (int k = 0; k < 4; ++k)
			__builtin_nontemporal_store(v[k], p + k * ks);
		__builtin_nontemporal_store(m.e[8].im, b + 8 * ks + l);
	}
	else
	{
#pragma unroll
		for (int k = 0; k < 4; ++k)
			p[k * ks] = v[k];
		b[8 * ks + l] = m.e[8].im;
	}
}

// site indices are < 2^26 (mesh_make_geo), so 32-bit division suffices (the
// 64-bit one is a long emulated sequence)
#ifndef MESH_COORD32
#define MESH_COORD32 1
#endif
#if MESH_COORD32
MD int coord(Geo const &g, long x, int mu) { return (int)(((unsigned)x / (unsigned)g.st[mu]) % (unsigned)g.L[mu]); }
#else
MD int coord(Geo const &g, long x, int mu) { return (int)((x / g.st[mu]) % g.L[mu]); }
#endif
MHD long shift_c(Geo const &g, long x, int mu, int c, int s)
{
	int n = c + s;
	n = n >= g.L[mu] ? n - g.L[mu] : (n < 0 ? n + g.L[mu] : n);
	return x + (long)(n - c) * g.st[mu];
}

// shift_c in whole blocks, for a direction whose site stride is a multiple
// of the block length (g.st[mu] % 2^MESH_SITE_LB == 0): the blocks a site
// moves by, which leaves its position within the block (and so its slot) as
// it is
MHD int blk_shift(Geo const &g, int mu, int c, int s)
{
	int n = c + s;
	n = n >= g.L[mu] ? n - g.L[mu] : (n < 0 ? n + g.L[mu] : n);
	return (n - c) * (g.st[mu] >> MESH_SITE_LB);
}
// element offset of matrix f, block blk, plane 0, slot l (site_off with the
// site's block and slot given)
MHD long blk_off(Geo const &g, int f, int blk, int l)
{
	return (long)(f * 9) * g.stride + (((long)blk * 9) << MESH_SITE_LB) + l;
}

// plane slot of offset o within a 64-site block (site_off's slot order)
MHD int block_slot(int o) { return MESH_SITE_PAIRS ? (((o & 1) << 5) | (o >> 1)) : o; }

// slots of the x3-neighbours x +/- e3 in x's block (the block holds whole
// x3-rows: L3 | 64)
MHD void x3_slots(Geo const &g, long x, int const *c, int &sp, int &sm)
{
	const int o = (int)(x & 63), c3 = c[3], L3 = g.L[3], o0 = o - c3;
	sp = block_slot(o0 + (c3 + 1 == L3 ? 0 : c3 + 1));
	sm = block_slot(o0 + (c3 == 0 ? L3 - 1 : c3 - 1));
}

// Rows path of the fused MD step (4-d, V % 64 == 0, L3 | 64 and
// (L2 L3) % 64 == 0): a 64-site block holds whole x3-rows and lies within one
// (x0, x1) plane.  A shift along x0 or x1 then moves every site of a block by
// the same number of blocks and keeps its slot, and a shift along x3 stays in
// the block, so a link displaced only along x0, x1 and x3 is found from the
// block index, uniform over the block's wavefront, and a slot: its address is
// formed in scalar registers but for the lane's slot.  RowsAt holds, for site
// x, its block lb, its coordinates c0 and c1 (the block's), its slot and the
// slots of x +/- e3.
struct RowsAt
{
	int lb, c0, c1, slot, sp, sm;
};
MHD RowsAt rows_at_lane(Geo const &g, long x, int const *c)
{
	RowsAt a;
	a.lb = (int)(x >> MESH_SITE_LB);
	a.c0 = c[0];
	a.c1 = c[1];
	a.slot = block_slot((int)(x & 63));
	x3_slots(g, x, c, a.sp, a.sm);
	return a;
}
struct BlkSlot
{
	int blk, l;
};
// block and slot of the site x + s1 e_a1 + s2 e_a2 (a1, a2 != 2; a2 = -1: one
// shift; |s| = 1)
MHD BlkSlot rows_site(Geo const &g, RowsAt const &a, int a1, int s1, int a2, int s2)
{
	static_assert(MESH_SITE_LB == 6, "the rows path assumes 64-site blocks");
	const int d0 = (a1 == 0 ? s1 : 0) + (a2 == 0 ? s2 : 0), d1 = (a1 == 1 ? s1 : 0) + (a2 == 1 ? s2 : 0);
	const int d3 = (a1 == 3 ? s1 : 0) + (a2 == 3 ? s2 : 0);
	BlkSlot b;
	b.blk = a.lb + (d0 ? blk_shift(g, 0, a.c0, d0) : 0) + (d1 ? blk_shift(g, 1, a.c1, d1) : 0);
	b.l = d3 == 0 ? a.slot : (d3 > 0 ? a.sp : a.sm);
	return b;
}

#undef MD

} // namespace mesh_dev
