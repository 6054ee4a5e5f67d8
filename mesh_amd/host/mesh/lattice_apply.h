// Generic lattice operations for hipcc-compiled callers: lattice_apply,
// lattice_sum_apply and the lattice expressions of krox/mesh
// (src/lattice/lattice.h:256-355, src/lattice/device.h:139-265,
// src/gauge/utils.h:27-33), over the library's blocked structure-of-arrays
// field layout.
//
// A device lambda receives whole elements, SU3<T> (mesh/lattice.h: one 3x3
// complex matrix, row-major), one per lattice argument, for the site its
// thread handles: the kernel gathers each matrix from its 9 SoA planes into
// registers, calls the lambda and scatters the first argument back.  The
// element algebra (products, adj, traces, project_on_algebra, exp,
// project_on_group_fast) is the library's own (su3_dev.h), so a lambda that
// performs the library's operations produces its bits: the reference's evolve
// lambda (src/gauge/hmc.h:31-35) gives mesh_su3_link_update's field, and
// gauge::exp(X) * U is evaluated as U + (exp(X) - 1) U like the library's
// link step.  Element operations are __host__ __device__, so the same lambda
// also runs on host copies of the fields.
//
// Sums (lattice_sum_apply, sum_real_trace, norm2) are deterministic: each
// workgroup sums its 256 sites in a fixed order, one more pass sums the
// workgroup partials in a fixed order, all in FP64 (the reference sums in the
// lambda's return type and adds the block partials on the host,
// src/lattice/device.h:246-265).
//
// Element types other than SU3<float> / SU3<double> are not provided: the
// U(1) fields of this library store angles in their own batched layout.
//
// The including translation unit is compiled with hipcc for gfx950 and the
// library's layout macros (MESH_SITE_LB, MESH_SITE_PAIRS) at their defaults;
// it links against libmesh_hip.so.
#pragma once

#ifndef __HIPCC__
#error "mesh/lattice_apply.h contains device code: compile the including translation unit with hipcc"
#endif

#include "../../csrc/su3_dev.h"
#include "mesh/gauge.h"

#include <hip/hip_runtime.h>

#include <cstring>
#include <type_traits>
#include <utility>

namespace mesh {

// ---------------------------------------------------------------- elements

namespace detail {

template <class T> __host__ __device__ __forceinline__ mesh_dev::m3<T> m3_of(SU3<T> const &a)
{
	mesh_dev::m3<T> m;
#pragma unroll
	for (int k = 0; k < 9; ++k)
		m.e[k] = {a.v[2 * k], a.v[2 * k + 1]};
	return m;
}
template <class T> __host__ __device__ __forceinline__ SU3<T> su3_of(mesh_dev::m3<T> const &m)
{
	SU3<T> a;
#pragma unroll
	for (int k = 0; k < 9; ++k)
	{
		a.v[2 * k] = m.e[k].re;
		a.v[2 * k + 1] = m.e[k].im;
	}
	return a;
}

} // namespace detail

#define MESH_HD __host__ __device__ __forceinline__

template <class T> MESH_HD SU3<T> operator*(SU3<T> const &a, SU3<T> const &b)
{
	return detail::su3_of(mesh_dev::mul_nn(detail::m3_of(a), detail::m3_of(b)));
}
template <class T> MESH_HD SU3<T> operator+(SU3<T> const &a, SU3<T> const &b)
{
	SU3<T> r;
#pragma unroll
	for (int k = 0; k < 18; ++k)
		r.v[k] = a.v[k] + b.v[k];
	return r;
}
template <class T> MESH_HD SU3<T> operator-(SU3<T> const &a, SU3<T> const &b)
{
	SU3<T> r;
#pragma unroll
	for (int k = 0; k < 18; ++k)
		r.v[k] = a.v[k] - b.v[k];
	return r;
}
template <class T> MESH_HD SU3<T> operator-(SU3<T> const &a)
{
	SU3<T> r;
#pragma unroll
	for (int k = 0; k < 18; ++k)
		r.v[k] = -a.v[k];
	return r;
}
// a * s for a real scalar, entry by entry in T (util's Matrix * double)
template <class T, class S>
    requires std::is_arithmetic_v<S>
MESH_HD SU3<T> operator*(SU3<T> const &a, S s)
{
	SU3<T> r;
	const T t = (T)s;
#pragma unroll
	for (int k = 0; k < 18; ++k)
		r.v[k] = a.v[k] * t;
	return r;
}
template <class T, class S>
    requires std::is_arithmetic_v<S>
MESH_HD SU3<T> operator*(S s, SU3<T> const &a)
{
	return a * s;
}
template <class T> MESH_HD SU3<T> &operator+=(SU3<T> &a, SU3<T> const &b) { return a = a + b; }
template <class T> MESH_HD SU3<T> &operator-=(SU3<T> &a, SU3<T> const &b) { return a = a - b; }
template <class T> MESH_HD SU3<T> &operator*=(SU3<T> &a, SU3<T> const &b) { return a = a * b; }
template <class T, class S>
    requires std::is_arithmetic_v<S>
MESH_HD SU3<T> &operator*=(SU3<T> &a, S s)
{
	return a = a * s;
}

template <class T> MESH_HD SU3<T> adj(SU3<T> const &a) { return detail::su3_of(mesh_dev::adj_m3(detail::m3_of(a))); }
// Re tr a (src/gauge/groups.h:214-222)
template <class T> MESH_HD T real_trace(SU3<T> const &a) { return (a.v[0] + a.v[8]) + a.v[16]; }
// ||a||_F^2
template <class T> MESH_HD T norm2(SU3<T> const &a) { return mesh_dev::norm2_m3(detail::m3_of(a)); }
// traceless antihermitian part (src/gauge/groups.h:224-228)
template <class T> MESH_HD SU3<T> project_on_algebra(SU3<T> const &a)
{
	return detail::su3_of(mesh_dev::ta(detail::m3_of(a)));
}
// project_on_group_fast (src/gauge/groups.h:237-243), in the library's
// increment form (FP32 elements projected in FP64 and rounded once)
template <class T> MESH_HD SU3<T> project_on_group_fast(SU3<T> const &a)
{
	return detail::su3_of(mesh_dev::reunit(detail::m3_of(a)));
}
// exp of an algebra element, kept as D = exp(X) - 1: multiplied from the left
// onto a matrix it gives U + D U (the library's link step); anywhere else it
// becomes the matrix 1 + D
template <class T> struct ExpSU3
{
	SU3<T> D;
	MESH_HD operator SU3<T>() const
	{
		SU3<T> r = D;
		r.v[0] += T(1);
		r.v[8] += T(1);
		r.v[16] += T(1);
		return r;
	}
};
template <class T> MESH_HD ExpSU3<T> exp(SU3<T> const &X)
{
	return {detail::su3_of(mesh_dev::expm1su3(detail::m3_of(X)))};
}
template <class T> MESH_HD SU3<T> operator*(ExpSU3<T> const &e, SU3<T> const &U)
{
	return detail::su3_of(mesh_dev::add_mul(detail::m3_of(e.D), detail::m3_of(U)));
}
template <class T> MESH_HD SU3<T> operator*(SU3<T> const &U, ExpSU3<T> const &e) { return U * SU3<T>(e); }

namespace gauge {
using mesh::adj;
using mesh::exp;
using mesh::project_on_algebra;
using mesh::project_on_group_fast;
using mesh::real_trace;
} // namespace gauge

// ---------------------------------------------------------------- kernels

namespace detail {

template <class G> struct is_su3 : std::false_type
{};
template <class T> struct is_su3<SU3<T>> : std::true_type
{};

// geometry of a lattice in the library's layout (as mesh_make_geo builds it)
inline mesh_dev::Geo geo_of(Grid const &g)
{
	mesh_dev::Geo geo{};
	geo.nd = g.ndim();
	long s = 1;
	for (int i = 3; i >= 0; --i)
	{
		if (i >= geo.nd)
		{
			geo.L[i] = 1;
			geo.st[i] = 0;
			continue;
		}
		geo.L[i] = g.shape(i);
		geo.st[i] = (int)s;
		s *= g.shape(i);
	}
	geo.V = s;
	geo.stride = (s + 255) / 256 * 256;
	return geo;
}

template <class T> __device__ __forceinline__ SU3<T> load_su3(const T *base, mesh_dev::Geo const &g, long x)
{
	return su3_of(mesh_dev::load(base, 0, g, x));
}
template <class T> __device__ __forceinline__ void store_su3(T *base, mesh_dev::Geo const &g, long x, SU3<T> const &a)
{
	mesh_dev::store(base, 0, g, x, m3_of(a));
}

// element-wise map; LOAD = false skips reading the first argument (it starts
// as zero), for the out-of-place operators whose lambdas only assign it
template <bool LOAD, class F, class T, class... Ts>
__global__ __launch_bounds__(256) void k_apply(F f, mesh_dev::Geo g, T *a, const Ts *... as)
{
	const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= g.V)
		return;
	const long x = mesh_dev::lane_site(g, t);
	SU3<T> e;
	if constexpr (LOAD)
		e = load_su3(a, g, x);
	else
		e = su3_of(mesh_dev::zero_m3<T>());
	f(e, load_su3(as, g, x)...);
	store_su3(a, g, x, e);
}

// fixed-order sum of 256 doubles of a workgroup: wavefront tree by shuffles,
// then the four wavefront sums in order
__device__ __forceinline__ double block_sum(double v, double *lds)
{
#pragma unroll
	for (int o = 32; o > 0; o >>= 1)
		v += __shfl_down(v, o, 64);
	if ((threadIdx.x & 63) == 0)
		lds[threadIdx.x >> 6] = v;
	__syncthreads();
	return ((lds[0] + lds[1]) + lds[2]) + lds[3];
}

template <class F, class T, class... Ts>
__global__ __launch_bounds__(256) void k_sum_apply(F f, mesh_dev::Geo g, double *partials, const T *a,
                                                    const Ts *... as)
{
	__shared__ double lds[4];
	const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	double v = 0.0;
	if (t < g.V)
	{
		const long x = mesh_dev::lane_site(g, t);
		v = (double)f(load_su3(a, g, x), load_su3(as, g, x)...);
	}
	const double s = block_sum(v, lds);
	if (threadIdx.x == 0)
		partials[blockIdx.x] = s;
}

template <int = 0> __global__ __launch_bounds__(256) void k_sum_partials(const double *p, int n, double *out)
{
	__shared__ double lds[4];
	double v = 0.0;
	for (int i = threadIdx.x; i < n; i += 256)
		v += p[i];
	const double s = block_sum(v, lds);
	if (threadIdx.x == 0)
		out[0] = s;
}

inline void hip_check(hipError_t e, char const *what)
{
	if (e != hipSuccess)
		throw std::runtime_error(std::string("hip error: ") + what + ": " + hipGetErrorString(e));
}

// the library context's stream: the generic kernels are ordered with the
// named operations of mesh/gauge.h
inline hipStream_t stream() { return (hipStream_t)mesh_stream(Context::global().get()); }

template <class G> void check_conformable(Lattice<G> const &a) { (void)a; }
template <class G, class H, class... Hs>
void check_conformable(Lattice<G> const &a, Lattice<H> const &b, Lattice<Hs> const &...rest)
{
	if (!(a.grid() == b.grid()))
		throw std::runtime_error("lattice_apply: lattices on different grids");
	check_conformable(a, rest...);
}

// device buffer of n doubles for the workgroup partials of a sum, grown as
// needed and kept (one per host thread, like the reference's per-call
// temporaries but without an allocation per call)
inline double *partials(size_t n)
{
	struct Buf
	{
		double *p = nullptr;
		size_t n = 0;
		~Buf()
		{
			if (p)
				(void)hipFree(p);
		}
	};
	thread_local Buf b;
	if (n > b.n)
	{
		if (b.p)
		{
			hip_check(hipStreamSynchronize(stream()), "synchronise");
			hip_check(hipFree(b.p), "free");
			b.p = nullptr;
			b.n = 0;
		}
		hip_check(hipMalloc(&b.p, n * sizeof(double)), "allocate sum partials");
		b.n = n;
	}
	return b.p;
}

template <bool LOAD, class F, class G, class... Gs> void apply(F f, Lattice<G> &a, Lattice<Gs> const &...as)
{
	static_assert(is_su3<G>::value && (is_su3<Gs>::value && ...), "lattice_apply: SU3<T> lattices only");
	check_conformable(a, as...);
	const mesh_dev::Geo g = geo_of(a.grid());
	const int nb = (int)((g.V + 255) / 256);
	k_apply<LOAD><<<nb, 256, 0, stream()>>>(f, g, static_cast<typename G::real *>(a.data()),
	                                         static_cast<const typename Gs::real *>(as.data())...);
	hip_check(hipGetLastError(), "lattice_apply launch");
}

} // namespace detail

// ---------------------------------------------------------------- lattice_apply

// f(a[x], as[x]...) for every site x; a is read and written, the others are
// read (src/lattice/lattice.h:256-261).  f is a __device__ (or __host__
// __device__) lambda taking SU3<T>& and SU3<T> const&.  Asynchronous on the
// library's stream.
template <typename F, typename G, typename... Gs> void lattice_apply(F f, Lattice<G> &a, Lattice<Gs> const &...as)
{
	detail::apply<true>(f, a, as...);
}

// sum over sites of f(a[x], as[x]...) for f returning a real number
// (src/lattice/lattice.h:263-268); the result has f's return type, the sum is
// formed in FP64
template <typename F, typename G, typename... Gs>
auto lattice_sum_apply(F f, Lattice<G> const &a, Lattice<Gs> const &...as)
    -> decltype(f(std::declval<G const &>(), std::declval<Gs const &>()...))
{
	using R = decltype(f(std::declval<G const &>(), std::declval<Gs const &>()...));
	static_assert(std::is_arithmetic_v<R>, "lattice_sum_apply: f must return a real number");
	static_assert(detail::is_su3<G>::value && (detail::is_su3<Gs>::value && ...),
	              "lattice_sum_apply: SU3<T> lattices only");
	detail::check_conformable(a, as...);
	const mesh_dev::Geo g = detail::geo_of(a.grid());
	const int nb = (int)((g.V + 255) / 256);
	double *p = detail::partials((size_t)nb + 1);
	detail::k_sum_apply<<<nb, 256, 0, detail::stream()>>>(f, g, p + 1,
	                                                       static_cast<const typename G::real *>(a.data()),
	                                                       static_cast<const typename Gs::real *>(as.data())...);
	detail::hip_check(hipGetLastError(), "lattice_sum_apply launch");
	detail::k_sum_partials<><<<1, 256, 0, detail::stream()>>>(p + 1, nb, p);
	detail::hip_check(hipGetLastError(), "lattice_sum_apply launch");
	double s = 0.0;
	detail::hip_check(hipMemcpyAsync(&s, p, sizeof(double), hipMemcpyDeviceToHost, detail::stream()), "copy");
	detail::hip_check(hipStreamSynchronize(detail::stream()), "synchronise");
	return (R)s;
}

// ---------------------------------------------------------------- expressions

// a op b, a op s, a op= b, a op= s (src/lattice/lattice.h:277-355)
#define MESH_LATTICE_BINARY(op)                                                                                  \
	template <class T> Lattice<SU3<T>> operator op(Lattice<SU3<T>> const &lhs, Lattice<SU3<T>> const &rhs)      \
	{                                                                                                            \
		Lattice<SU3<T>> r(lhs.grid());                                                                           \
		detail::apply<false>([] __device__(SU3<T> & a, SU3<T> const &b, SU3<T> const &c) { a = b op c; }, r, lhs, \
		                     rhs);                                                                               \
		return r;                                                                                                \
	}                                                                                                            \
	template <class T> Lattice<SU3<T>> &operator op##=(Lattice<SU3<T>> &lhs, Lattice<SU3<T>> const &rhs)        \
	{                                                                                                            \
		detail::apply<true>([] __device__(SU3<T> & a, SU3<T> const &b) { a = a op b; }, lhs, rhs);                \
		return lhs;                                                                                              \
	}                                                                                                            \
	template <class T> Lattice<SU3<T>> operator op(Lattice<SU3<T>> const &lhs, SU3<T> const &rhs)               \
	{                                                                                                            \
		Lattice<SU3<T>> r(lhs.grid());                                                                           \
		detail::apply<false>([rhs] __device__(SU3<T> & a, SU3<T> const &b) { a = b op rhs; }, r, lhs);            \
		return r;                                                                                                \
	}                                                                                                            \
	template <class T> Lattice<SU3<T>> &operator op##=(Lattice<SU3<T>> &lhs, SU3<T> const &rhs)                 \
	{                                                                                                            \
		detail::apply<true>([rhs] __device__(SU3<T> & a) { a = a op rhs; }, lhs);                                 \
		return lhs;                                                                                              \
	}

MESH_LATTICE_BINARY(+)
MESH_LATTICE_BINARY(-)
MESH_LATTICE_BINARY(*)
#undef MESH_LATTICE_BINARY

// a * s, s * a, a *= s for a real scalar (src/lattice/lattice.h:330-336)
template <class T> Lattice<SU3<T>> operator*(Lattice<SU3<T>> const &lhs, double s)
{
	Lattice<SU3<T>> r(lhs.grid());
	detail::apply<false>([s] __device__(SU3<T> & a, SU3<T> const &b) { a = b * s; }, r, lhs);
	return r;
}
template <class T> Lattice<SU3<T>> operator*(double s, Lattice<SU3<T>> const &rhs) { return rhs * s; }
template <class T> Lattice<SU3<T>> &operator*=(Lattice<SU3<T>> &lhs, double s)
{
	detail::apply<true>([s] __device__(SU3<T> & a) { a = a * s; }, lhs);
	return lhs;
}

// element-wise functions (src/lattice/lattice.h:278-288, src/gauge/utils.h:30-33)
template <class T> Lattice<SU3<T>> adj(Lattice<SU3<T>> const &arg)
{
	Lattice<SU3<T>> r(arg.grid());
	detail::apply<false>([] __device__(SU3<T> & a, SU3<T> const &b) { a = adj(b); }, r, arg);
	return r;
}
template <class T> Lattice<SU3<T>> project_on_algebra(Lattice<SU3<T>> const &arg)
{
	Lattice<SU3<T>> r(arg.grid());
	detail::apply<false>([] __device__(SU3<T> & a, SU3<T> const &b) { a = project_on_algebra(b); }, r, arg);
	return r;
}

// exp of an algebra lattice, kept lazy: exp(X) * U is evaluated per site as
// U + (exp(X) - 1) U (the library's link step); converted to a lattice it is
// 1 + (exp(X) - 1).  It refers to X, so it is used within the expression that
// forms it, as in the reference's exp(F * eps) * g (src/gauge/landau.h:81, 131).
template <class T> struct LatticeExp
{
	Lattice<SU3<T>> const &X;
	operator Lattice<SU3<T>>() const
	{
		Lattice<SU3<T>> r(X.grid());
		detail::apply<false>([] __device__(SU3<T> & a, SU3<T> const &b) { a = SU3<T>(exp(b)); }, r, X);
		return r;
	}
};
template <class T> LatticeExp<T> exp(Lattice<SU3<T>> const &X) { return {X}; }
template <class T> Lattice<SU3<T>> operator*(LatticeExp<T> const &e, Lattice<SU3<T>> const &U)
{
	Lattice<SU3<T>> r(U.grid());
	detail::apply<false>([] __device__(SU3<T> & a, SU3<T> const &x, SU3<T> const &u) { a = exp(x) * u; }, r, e.X,
	                     U);
	return r;
}

// sums (src/gauge/utils.h:28-29, MESH_DEFINE_LATTICE_SUM); the per-site
// values are formed in FP64 from the elements, as the library's reductions do
template <class T> double sum_real_trace(Lattice<SU3<T>> const &arg)
{
	return lattice_sum_apply(
	    [] __device__(SU3<T> const &a) { return ((double)a.v[0] + (double)a.v[8]) + (double)a.v[16]; }, arg);
}
template <class T> double norm2(Lattice<SU3<T>> const &arg)
{
	return lattice_sum_apply(
	    [] __device__(SU3<T> const &a) { return mesh_dev::norm2_m3(mesh_dev::to_f64(detail::m3_of(a))); }, arg);
}
template <class T> double norm2(LatticeStack<SU3<T>> const &U)
{
	double s = 0;
	for (size_t mu = 0; mu < U.size(); ++mu)
		s += norm2(U[mu]);
	return s;
}

// a += s * b for a scalar s (src/lattice/lattice.h:338-345)
template <class T, class S>
    requires std::is_arithmetic_v<S>
void add_mul(Lattice<SU3<T>> &lhs, S s, Lattice<SU3<T>> const &rhs)
{
	detail::apply<true>([s] __device__(SU3<T> & a, SU3<T> const &b) { a = a + b * s; }, lhs, rhs);
}

#undef MESH_HD

} // namespace mesh
