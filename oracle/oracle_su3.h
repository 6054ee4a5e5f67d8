// CPU oracle: element-level SU(3) / su(3) algebra.
//
// TEST INFRASTRUCTURE ONLY.  This header belongs to the CPU restatement of
// krox/mesh's lattice arithmetic that the GPU library is checked against.  It
// is compiled into oracle/liboracle_mesh.so, which only tests/, the smoke check
// and bench.py's cpu_baseline leg may load.  Nothing in mesh_amd/ includes it.
//
// Every operation is written with an explicit order of IEEE operations (all
// multiply-adds through std::fma, -ffp-contract=off), so that a GPU kernel that
// performs the same sequence produces bit-identical FP64 / FP32 results.
//
// Reference semantics followed (paths relative to the reference root):
//   adj, real_trace              src/gauge/groups.h:214-222
//   project_on_algebra (TA)      src/gauge/groups.h:224-228  (util's
//                                antihermitian_traceless; restated from the
//                                definition (M - M^+)/2 - tr(M - M^+)/(2 Nc))
//   project_on_group_fast        src/gauge/groups.h:237-243
//   exp (algebra -> group)       src/gauge/groups.h:263-268  (util::exp is not
//                                available; restated as the exact matrix
//                                exponential, evaluated by Cayley-Hamilton with
//                                power-series coefficients)
//   random_group_element         src/gauge/groups.h:255-261  (Gram-Schmidt of a
//                                complex Gaussian matrix, row 0 divided by det)
//   random_algebra_element       src/gauge/groups.h:245-253  (same distribution:
//                                P = sum_a p_a i T^a, p_a ~ N(0,1))
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace orc {

template <class T> struct cplx
{
	T re, im;
};

template <class T> struct mat3
{
	cplx<T> e[3][3];
};

// acc += a*b
template <class T> inline void cmadd(cplx<T> &acc, cplx<T> a, cplx<T> b)
{
	acc.re = std::fma(a.re, b.re, acc.re);
	acc.re = std::fma(-a.im, b.im, acc.re);
	acc.im = std::fma(a.re, b.im, acc.im);
	acc.im = std::fma(a.im, b.re, acc.im);
}
// acc -= a*b
template <class T> inline void cmsub(cplx<T> &acc, cplx<T> a, cplx<T> b)
{
	acc.re = std::fma(-a.re, b.re, acc.re);
	acc.re = std::fma(a.im, b.im, acc.re);
	acc.im = std::fma(-a.re, b.im, acc.im);
	acc.im = std::fma(-a.im, b.re, acc.im);
}
// acc += a*conj(b)
template <class T> inline void cmadd_cb(cplx<T> &acc, cplx<T> a, cplx<T> b)
{
	acc.re = std::fma(a.re, b.re, acc.re);
	acc.re = std::fma(a.im, b.im, acc.re);
	acc.im = std::fma(a.im, b.re, acc.im);
	acc.im = std::fma(-a.re, b.im, acc.im);
}
// acc += conj(a)*b
template <class T> inline void cmadd_ca(cplx<T> &acc, cplx<T> a, cplx<T> b)
{
	acc.re = std::fma(a.re, b.re, acc.re);
	acc.re = std::fma(a.im, b.im, acc.re);
	acc.im = std::fma(a.re, b.im, acc.im);
	acc.im = std::fma(-a.im, b.re, acc.im);
}
// acc += conj(a)*conj(b)
template <class T> inline void cmadd_cc(cplx<T> &acc, cplx<T> a, cplx<T> b)
{
	acc.re = std::fma(a.re, b.re, acc.re);
	acc.re = std::fma(-a.im, b.im, acc.re);
	acc.im = std::fma(-a.re, b.im, acc.im);
	acc.im = std::fma(-a.im, b.re, acc.im);
}
template <class T> inline cplx<T> cmul(cplx<T> a, cplx<T> b)
{
	cplx<T> r = {T(0), T(0)};
	cmadd(r, a, b);
	return r;
}

// C = A*B
template <class T> inline mat3<T> mul_nn(mat3<T> const &A, mat3<T> const &B)
{
	mat3<T> C;
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
		{
			cplx<T> acc = {T(0), T(0)};
			for (int k = 0; k < 3; ++k)
				cmadd(acc, A.e[i][k], B.e[k][j]);
			C.e[i][j] = acc;
		}
	return C;
}
// C = A*B^+
template <class T> inline mat3<T> mul_nd(mat3<T> const &A, mat3<T> const &B)
{
	mat3<T> C;
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
		{
			cplx<T> acc = {T(0), T(0)};
			for (int k = 0; k < 3; ++k)
				cmadd_cb(acc, A.e[i][k], B.e[j][k]);
			C.e[i][j] = acc;
		}
	return C;
}
// C = A^+*B
template <class T> inline mat3<T> mul_dn(mat3<T> const &A, mat3<T> const &B)
{
	mat3<T> C;
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
		{
			cplx<T> acc = {T(0), T(0)};
			for (int k = 0; k < 3; ++k)
				cmadd_ca(acc, A.e[k][i], B.e[k][j]);
			C.e[i][j] = acc;
		}
	return C;
}
// C = A^+*B^+
template <class T> inline mat3<T> mul_dd(mat3<T> const &A, mat3<T> const &B)
{
	mat3<T> C;
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
		{
			cplx<T> acc = {T(0), T(0)};
			for (int k = 0; k < 3; ++k)
				cmadd_cc(acc, A.e[k][i], B.e[j][k]);
			C.e[i][j] = acc;
		}
	return C;
}
// S += A*B (accumulating into the existing entries of S)
template <class T> inline void madd_nn(mat3<T> &S, mat3<T> const &A, mat3<T> const &B)
{
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
		{
			cplx<T> acc = S.e[i][j];
			for (int k = 0; k < 3; ++k)
				cmadd(acc, A.e[i][k], B.e[k][j]);
			S.e[i][j] = acc;
		}
}
// S += A*B^+
template <class T> inline void madd_nd(mat3<T> &S, mat3<T> const &A, mat3<T> const &B)
{
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
		{
			cplx<T> acc = S.e[i][j];
			for (int k = 0; k < 3; ++k)
				cmadd_cb(acc, A.e[i][k], B.e[j][k]);
			S.e[i][j] = acc;
		}
}

template <class T> inline mat3<T> zero3()
{
	mat3<T> r;
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
			r.e[i][j] = {T(0), T(0)};
	return r;
}
template <class T> inline mat3<T> unit3()
{
	mat3<T> r = zero3<T>();
	for (int i = 0; i < 3; ++i)
		r.e[i][i].re = T(1);
	return r;
}
template <class T> inline mat3<T> adj3(mat3<T> const &A)
{
	mat3<T> r;
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
			r.e[i][j] = {A.e[j][i].re, -A.e[j][i].im};
	return r;
}

// Re tr(A * B^+) = sum_ij Re(A_ij conj(B_ij)), row-major fma chain.
template <class T> inline T retr_nd(mat3<T> const &A, mat3<T> const &B)
{
	T s = T(0);
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
		{
			s = std::fma(A.e[i][j].re, B.e[i][j].re, s);
			s = std::fma(A.e[i][j].im, B.e[i][j].im, s);
		}
	return s;
}

// ||A||_F^2, row-major fma chain.
template <class T> inline T norm2_3(mat3<T> const &A)
{
	T s = T(0);
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
		{
			s = std::fma(A.e[i][j].re, A.e[i][j].re, s);
			s = std::fma(A.e[i][j].im, A.e[i][j].im, s);
		}
	return s;
}

// Traceless antihermitian part: (M - M^+)/2 - tr(M - M^+)/(2*3) * 1.
template <class T> inline mat3<T> ta3(mat3<T> const &M)
{
	mat3<T> A;
	T t = ((M.e[0][0].im + M.e[1][1].im) + M.e[2][2].im) * (T(1) / T(3));
	for (int i = 0; i < 3; ++i)
	{
		A.e[i][i] = {T(0), M.e[i][i].im - t};
		for (int j = 0; j < 3; ++j)
			if (j != i)
			{
				A.e[i][j].re = T(0.5) * (M.e[i][j].re - M.e[j][i].re);
				A.e[i][j].im = T(0.5) * (M.e[i][j].im + M.e[j][i].im);
			}
	}
	return A;
}

// determinant, cofactor expansion along row 0
template <class T> inline cplx<T> det3(mat3<T> const &X)
{
	cplx<T> m0 = {T(0), T(0)}, m1 = {T(0), T(0)}, m2 = {T(0), T(0)};
	cmadd(m0, X.e[1][1], X.e[2][2]);
	cmsub(m0, X.e[1][2], X.e[2][1]);
	cmadd(m1, X.e[1][0], X.e[2][2]);
	cmsub(m1, X.e[1][2], X.e[2][0]);
	cmadd(m2, X.e[1][0], X.e[2][1]);
	cmsub(m2, X.e[1][1], X.e[2][0]);
	cplx<T> d = {T(0), T(0)};
	cmadd(d, X.e[0][0], m0);
	cmsub(d, X.e[0][1], m1);
	cmadd(d, X.e[0][2], m2);
	return d;
}

// number of power-series terms in expsu3 (remainder < 1/(N+1)! for ||X||_F <= 1);
// for ||X||_F^2 <= exp_small_r2 the first exp_terms_small() terms (remainder
// below 0.2^7/7! = 2.5e-9 FP32, 0.2^12/12! = 8.6e-18 FP64)
template <class T> constexpr int exp_terms() { return sizeof(T) == 8 ? 18 : 10; }
template <class T> constexpr int exp_terms_small() { return sizeof(T) == 8 ? 11 : 6; }
constexpr double exp_small_r2 = 0.04;

// X^2 of an exactly antihermitian X: hermitian, upper triangle formed, lower
// one its conjugate, real diagonal -sum_k |X_ik|^2
template <class T> inline mat3<T> sq_ah(mat3<T> const &X)
{
	mat3<T> Q;
	for (int i = 0; i < 3; ++i)
	{
		T s = T(0);
		for (int k = 0; k < 3; ++k)
		{
			s = std::fma(X.e[i][k].re, X.e[i][k].re, s);
			s = std::fma(X.e[i][k].im, X.e[i][k].im, s);
		}
		Q.e[i][i] = {-s, T(0)};
	}
	for (int i = 0; i < 3; ++i)
		for (int j = i + 1; j < 3; ++j)
		{
			cplx<T> acc = {T(0), T(0)};
			for (int k = 0; k < 3; ++k)
				cmadd(acc, X.e[i][k], X.e[k][j]);
			Q.e[i][j] = acc;
			Q.e[j][i] = {acc.re, -acc.im};
		}
	return Q;
}

// Im det X of an exactly antihermitian X: with H = -i X hermitian,
// det X = -i det H, det H = h0 h1 h2 + 2 Re(H01 H12 H20) - h0 |H12|^2
// - h1 |H02|^2 - h2 |H01|^2
template <class T> inline T imdet_ah(mat3<T> const &X)
{
	const T h0 = X.e[0][0].im, h1 = X.e[1][1].im, h2 = X.e[2][2].im;
	const cplx<T> a = {X.e[0][1].im, -X.e[0][1].re};
	const cplx<T> b = {X.e[1][2].im, -X.e[1][2].re};
	const cplx<T> c = {X.e[0][2].im, -X.e[0][2].re};
	const cplx<T> p = cmul(a, b);
	const T t = std::fma(p.re, c.re, p.im * c.im);
	const T n01 = std::fma(a.re, a.re, a.im * a.im), n12 = std::fma(b.re, b.re, b.im * b.im),
	        n02 = std::fma(c.re, c.re, c.im * c.im);
	T d = (h0 * h1) * h2;
	d = std::fma(T(2), t, d);
	d = std::fma(-h0, n12, d);
	d = std::fma(-h1, n02, d);
	d = std::fma(-h2, n01, d);
	return -d;
}

// exp(X) - 1 for an exactly antihermitian, traceless X (X = t TA(P)).
// Cayley-Hamilton: X^3 = a X + b 1 with a = tr(X^2)/2, b = det X, so that
// X^n = alpha_n 1 + beta_n X + gamma_n X^2 with
//   alpha_{n+1} = gamma_n b, beta_{n+1} = alpha_n + gamma_n a, gamma_{n+1} = beta_n
// and exp X = f0 1 + f1 X + f2 X^2, f_k = sum_n (coefficient_n)/n!.  For
// ||X||_F > 1 the argument is first divided by 2^s (exact) and the result
// squared s times (D -> 2 D + D^2); the (scaled) ||X||_F^2 sets the number of
// terms.  The 1 of exp X is never added: f0 - 1 is summed from its small terms
// (g0 below), so every entry of D is rounded at its own magnitude.
template <class T> inline mat3<T> expm1su3(mat3<T> X)
{
	mat3<T> X2 = sq_ah(X);
	T trx2 = (X2.e[0][0].re + X2.e[1][1].re) + X2.e[2][2].re;
	T r2 = -trx2; // ||X||_F^2
	int s = 0;
	T sc = T(1);
	while (r2 > T(1) && s < 64)
	{
		r2 *= T(0.25);
		sc *= T(0.5);
		++s;
	}
	if (s > 0)
	{
		T sc2 = sc * sc;
		for (int i = 0; i < 3; ++i)
			for (int j = 0; j < 3; ++j)
			{
				X.e[i][j].re *= sc;
				X.e[i][j].im *= sc;
				X2.e[i][j].re *= sc2;
				X2.e[i][j].im *= sc2;
			}
		trx2 *= sc2;
	}
	T a = T(0.5) * trx2;
	const T bi = imdet_ah(X);
	cplx<T> al = {T(1), T(0)}, be = {T(0), T(0)}, ga = {T(0), T(0)};
	cplx<T> g0 = {T(0), T(0)}, f1 = {T(0), T(0)}, f2 = {T(0), T(0)};
	const int nterms = r2 > T(exp_small_r2) ? exp_terms<T>() : exp_terms_small<T>();
	for (int n = 0; n < nterms; ++n)
	{
		T inv = T(1) / T(n + 1);
		cplx<T> gb = {-ga.im * bi, ga.re * bi};
		cplx<T> nal = {gb.re * inv, gb.im * inv};
		cplx<T> nbe = {std::fma(ga.re, a, al.re) * inv,
		               std::fma(ga.im, a, al.im) * inv};
		cplx<T> nga = {be.re * inv, be.im * inv};
		al = nal;
		be = nbe;
		ga = nga;
		g0.re += al.re;
		g0.im += al.im;
		f1.re += be.re;
		f1.im += be.im;
		f2.re += ga.re;
		f2.im += ga.im;
	}
	mat3<T> D;
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
		{
			cplx<T> acc = (i == j) ? g0 : cplx<T>{T(0), T(0)};
			cmadd(acc, f1, X.e[i][j]);
			cmadd(acc, f2, X2.e[i][j]);
			D.e[i][j] = acc;
		}
	for (int k = 0; k < s; ++k)
	{
		mat3<T> N;
		for (int i = 0; i < 3; ++i)
			for (int j = 0; j < 3; ++j)
			{
				cplx<T> acc = {D.e[i][j].re + D.e[i][j].re, D.e[i][j].im + D.e[i][j].im};
				for (int l = 0; l < 3; ++l)
					cmadd(acc, D.e[i][l], D.e[l][j]);
				N.e[i][j] = acc;
			}
		D = N;
	}
	return D;
}

// exp(X) = 1 + (exp X - 1)
template <class T> inline mat3<T> expsu3(mat3<T> const &X)
{
	mat3<T> E = expm1su3(X);
	for (int i = 0; i < 3; ++i)
		E.e[i][i].re += T(1);
	return E;
}

// U + D U: the product summed from zero, then one rounding at U's magnitude
template <class T> inline mat3<T> add_mul(mat3<T> const &D, mat3<T> const &U)
{
	mat3<T> r;
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
		{
			cplx<T> acc = {T(0), T(0)};
			for (int k = 0; k < 3; ++k)
				cmadd(acc, D.e[i][k], U.e[k][j]);
			r.e[i][j] = {U.e[i][j].re + acc.re, U.e[i][j].im + acc.im};
		}
	return r;
}

template <class A, class B> inline mat3<B> convert3(mat3<A> const &a)
{
	mat3<B> r;
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
			r.e[i][j] = {(B)a.e[i][j].re, (B)a.e[i][j].im};
	return r;
}

// project_on_group_fast (src/gauge/groups.h:237-243): r = 1.5 a - 0.5 a a^+ a
// and r *= 4/3 - det(r)/3, evaluated in increment form:
// r = a + (1 - a a^+) a / 2, then r += r (1 - det r) / 3.
template <class T> inline mat3<T> reunit_inc(mat3<T> const &A)
{
	mat3<T> M;
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
		{
			cplx<T> acc = {i == j ? T(1) : T(0), T(0)};
			for (int k = 0; k < 3; ++k)
			{
				cplx<T> a = A.e[i][k], b = A.e[j][k];
				acc.re = std::fma(-a.re, b.re, acc.re);
				acc.re = std::fma(-a.im, b.im, acc.re);
				acc.im = std::fma(-a.im, b.re, acc.im);
				acc.im = std::fma(a.re, b.im, acc.im);
			}
			M.e[i][j] = acc;
		}
	mat3<T> r;
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
		{
			cplx<T> acc = {T(0), T(0)};
			for (int k = 0; k < 3; ++k)
				cmadd(acc, M.e[i][k], A.e[k][j]);
			r.e[i][j] = {std::fma(T(0.5), acc.re, A.e[i][j].re), std::fma(T(0.5), acc.im, A.e[i][j].im)};
		}
	cplx<T> d = det3(r);
	cplx<T> c = {(T(1) - d.re) * (T(1) / T(3)), -d.im * (T(1) / T(3))};
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
		{
			cplx<T> q = cmul(r.e[i][j], c);
			r.e[i][j] = {r.e[i][j].re + q.re, r.e[i][j].im + q.im};
		}
	return r;
}

// reunitarisation of a stored link: FP32 links are projected in FP64 and
// rounded once
template <class T> inline mat3<T> reunit3(mat3<T> const &A)
{
	if constexpr (sizeof(T) == 8)
		return reunit_inc(A);
	else
		return convert3<double, T>(reunit_inc(convert3<T, double>(A)));
}

// link update U <- exp(t * TA(P)) * U, as U + (exp(t TA(P)) - 1) U
template <class T> inline mat3<T> evolve_link(mat3<T> const &U, mat3<T> const &P, T t)
{
	mat3<T> X = ta3(P);
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
		{
			X.e[i][j].re *= t;
			X.e[i][j].im *= t;
		}
	return add_mul(expm1su3(X), U);
}

// momentum update P <- P - k * TA(U * S): the product in T, TA and the update
// in FP64, P rounded once
template <class T> inline mat3<T> force_link_update(mat3<T> const &P, mat3<T> const &U,
                                                   mat3<T> const &S, double k)
{
	mat3<double> A = ta3(convert3<T, double>(mul_nn(U, S)));
	mat3<T> r;
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
		{
			r.e[i][j].re = (T)std::fma(-k, A.e[i][j].re, (double)P.e[i][j].re);
			r.e[i][j].im = (T)std::fma(-k, A.e[i][j].im, (double)P.e[i][j].im);
		}
	return r;
}

////////////////////////////////////////////////////////////////////////////////
// Random numbers: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and a
// deterministic Box-Muller transform built from +,*,fma,/ and sqrt only, so
// that CPU and GPU produce bit-identical normals.
////////////////////////////////////////////////////////////////////////////////

inline void philox4x32_10(uint32_t const ctr_in[4], uint32_t const key_in[2], uint32_t out[4])
{
	uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
	uint32_t k0 = key_in[0], k1 = key_in[1];
	for (int r = 0; r < 10; ++r)
	{
		uint64_t p0 = (uint64_t)0xD2511F53u * c0;
		uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
		uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
		uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
		uint32_t n0 = hi1 ^ c1 ^ k0;
		uint32_t n1 = lo1;
		uint32_t n2 = hi0 ^ c3 ^ k1;
		uint32_t n3 = lo0;
		c0 = n0;
		c1 = n1;
		c2 = n2;
		c3 = n3;
		k0 += 0x9E3779B9u;
		k1 += 0xBB67AE85u;
	}
	out[0] = c0;
	out[1] = c1;
	out[2] = c2;
	out[3] = c3;
}

// (hi, lo) -> uniform double in (0, 1], 53 random bits
inline double u01_from_bits(uint32_t hi, uint32_t lo)
{
	uint64_t w = ((((uint64_t)hi) << 32) | lo) >> 11;
	return (double)(w + 1) * (1.0 / 9007199254740992.0);
}

// natural log of a positive normal double: x = m 2^e, m in [sqrt(1/2), sqrt 2),
// log m = 2 atanh(s), s = (m-1)/(m+1), series through s^23.
inline double det_log(double x)
{
	uint64_t bits;
	std::memcpy(&bits, &x, 8);
	int e = (int)((bits >> 52) & 0x7ff) - 1023;
	bits = (bits & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
	double m;
	std::memcpy(&m, &bits, 8);
	if (m > 1.4142135623730951)
	{
		m *= 0.5;
		e += 1;
	}
	double s = (m - 1.0) / (m + 1.0);
	double s2 = s * s;
	double p = 1.0 / 23.0;
	p = std::fma(p, s2, 1.0 / 21.0);
	p = std::fma(p, s2, 1.0 / 19.0);
	p = std::fma(p, s2, 1.0 / 17.0);
	p = std::fma(p, s2, 1.0 / 15.0);
	p = std::fma(p, s2, 1.0 / 13.0);
	p = std::fma(p, s2, 1.0 / 11.0);
	p = std::fma(p, s2, 1.0 / 9.0);
	p = std::fma(p, s2, 1.0 / 7.0);
	p = std::fma(p, s2, 1.0 / 5.0);
	p = std::fma(p, s2, 1.0 / 3.0);
	p = std::fma(p, s2, 1.0);
	double lm = 2.0 * s * p;
	return std::fma((double)e, 0.6931471805599453094, lm);
}

// sin(2 pi u), cos(2 pi u) for u in [0, 1]: quadrant reduction (exact) and
// Taylor polynomials on [-pi/4, pi/4].
inline void det_sincos2pi(double u, double *sn, double *cs)
{
	double q = std::floor(4.0 * u + 0.5);
	double f = u - q * 0.25;
	double th = f * 6.283185307179586477;
	double t2 = th * th;
	double ps = -1.0 / 121645100408832000.0;             // -1/19!
	ps = std::fma(ps, t2, 1.0 / 355687428096000.0);      // 1/17!
	ps = std::fma(ps, t2, -1.0 / 1307674368000.0);       // -1/15!
	ps = std::fma(ps, t2, 1.0 / 6227020800.0);           // 1/13!
	ps = std::fma(ps, t2, -1.0 / 39916800.0);            // -1/11!
	ps = std::fma(ps, t2, 1.0 / 362880.0);               // 1/9!
	ps = std::fma(ps, t2, -1.0 / 5040.0);                // -1/7!
	ps = std::fma(ps, t2, 1.0 / 120.0);                  // 1/5!
	ps = std::fma(ps, t2, -1.0 / 6.0);                   // -1/3!
	ps = std::fma(ps, t2, 1.0);
	double S = th * ps;
	double pc = 1.0 / 2432902008176640000.0;             // 1/20!
	pc = std::fma(pc, t2, -1.0 / 6402373705728000.0);    // -1/18!
	pc = std::fma(pc, t2, 1.0 / 20922789888000.0);       // 1/16!
	pc = std::fma(pc, t2, -1.0 / 87178291200.0);         // -1/14!
	pc = std::fma(pc, t2, 1.0 / 479001600.0);            // 1/12!
	pc = std::fma(pc, t2, -1.0 / 3628800.0);             // -1/10!
	pc = std::fma(pc, t2, 1.0 / 40320.0);                // 1/8!
	pc = std::fma(pc, t2, -1.0 / 720.0);                 // -1/6!
	pc = std::fma(pc, t2, 1.0 / 24.0);                   // 1/4!
	pc = std::fma(pc, t2, -0.5);
	double C = std::fma(pc, t2, 1.0);
	int qi = ((int)q) & 3;
	if (qi == 0)
	{
		*sn = S;
		*cs = C;
	}
	else if (qi == 1)
	{
		*sn = C;
		*cs = -S;
	}
	else if (qi == 2)
	{
		*sn = -S;
		*cs = -C;
	}
	else
	{
		*sn = -C;
		*cs = S;
	}
}

// two standard normals from one Philox block
inline void normal_pair(uint32_t const r[4], double *z0, double *z1)
{
	double u1 = u01_from_bits(r[0], r[1]);
	double u2 = u01_from_bits(r[2], r[3]);
	double R = std::sqrt(-2.0 * det_log(u1));
	double sn, cs;
	det_sincos2pi(u2, &sn, &cs);
	*z0 = R * cs;
	*z1 = R * sn;
}

// Philox stream tags (counter word 3)
enum : uint32_t
{
	TAG_HAAR = 0x48414152u,    // initial gauge field
	TAG_MOM = 0x4d4f4d30u,     // HMC momenta
	TAG_METRO = 0x4d455452u,   // Metropolis uniforms
	TAG_GAUGE = 0x47415547u,   // random gauge transformations
	TAG_ALG = 0x414c4730u,     // random algebra fields
};

// n normals for link (site, mu) of a stream; counter = (site, mu, draw | ctr2<<5, tag)
inline void link_normals(uint64_t seed, uint32_t tag, uint32_t ctr2, uint32_t site, uint32_t mu,
                         int n, double *z)
{
	uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
	for (int d = 0; d < (n + 1) / 2; ++d)
	{
		uint32_t ctr[4] = {site, mu, (uint32_t)d | (ctr2 << 5), tag};
		uint32_t r[4];
		philox4x32_10(ctr, key, r);
		double z0, z1;
		normal_pair(r, &z0, &z1);
		z[2 * d] = z0;
		if (2 * d + 1 < n)
			z[2 * d + 1] = z1;
	}
}

// Haar-random SU(3) from 18 normals: rows of a complex Gaussian matrix are
// Gram-Schmidt orthonormalised, then row 0 is divided by the determinant.
inline mat3<double> haar_from_normals(double const *z)
{
	mat3<double> a;
	for (int i = 0; i < 3; ++i)
		for (int j = 0; j < 3; ++j)
			a.e[i][j] = {z[6 * i + 2 * j], z[6 * i + 2 * j + 1]};
	for (int i = 0; i < 3; ++i)
	{
		// projection done twice (re-orthogonalisation) for full accuracy
		for (int pass = 0; pass < 2; ++pass)
			for (int k = 0; k < i; ++k)
			{
				// c = <row_k, row_i> = sum_j conj(a_kj) a_ij
				cplx<double> c = {0.0, 0.0};
				for (int j = 0; j < 3; ++j)
					cmadd_ca(c, a.e[k][j], a.e[i][j]);
				for (int j = 0; j < 3; ++j)
					cmsub(a.e[i][j], c, a.e[k][j]);
			}
		double n2 = 0.0;
		for (int j = 0; j < 3; ++j)
		{
			n2 = std::fma(a.e[i][j].re, a.e[i][j].re, n2);
			n2 = std::fma(a.e[i][j].im, a.e[i][j].im, n2);
		}
		double inv = 1.0 / std::sqrt(n2);
		for (int j = 0; j < 3; ++j)
		{
			a.e[i][j].re *= inv;
			a.e[i][j].im *= inv;
		}
	}
	cplx<double> d = det3(a);
	// a_0j / d = a_0j * conj(d) / |d|^2
	double dn = std::fma(d.re, d.re, d.im * d.im);
	double id = 1.0 / dn;
	for (int j = 0; j < 3; ++j)
	{
		cplx<double> t = {0.0, 0.0};
		cmadd_cb(t, a.e[0][j], d);
		a.e[0][j] = {t.re * id, t.im * id};
	}
	return a;
}

// su(3) element sum_a p_a i T^a, T^a = lambda^a / 2 (Gell-Mann), from 8 normals
inline mat3<double> algebra_from_normals(double const *p)
{
	mat3<double> m = zero3<double>();
	const double h = 0.5;
	const double r3 = 0.28867513459481288225; // 1/(2 sqrt 3)
	// i*lambda1/2: (0,1),(1,0) entries i/2
	// i*lambda2/2: (0,1) -> i*(-i)/2 = 1/2 ; (1,0) -> i*(i)/2 = -1/2
	m.e[0][1] = {h * p[1], h * p[0]};
	m.e[1][0] = {-(h * p[1]), h * p[0]};
	m.e[0][2] = {h * p[4], h * p[3]};
	m.e[2][0] = {-(h * p[4]), h * p[3]};
	m.e[1][2] = {h * p[6], h * p[5]};
	m.e[2][1] = {-(h * p[6]), h * p[5]};
	m.e[0][0] = {0.0, std::fma(h, p[2], r3 * p[7])};
	m.e[1][1] = {0.0, std::fma(-h, p[2], r3 * p[7])};
	m.e[2][2] = {0.0, -2.0 * (r3 * p[7])};
	return m;
}

} // namespace orc
