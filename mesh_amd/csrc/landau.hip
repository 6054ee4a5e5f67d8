// Landau gauge fixing on gfx950: functional, gauge condition, gradient and
// the even/odd overrelaxation half-sweep.
//
// Reference entry: src/gauge/landau.h (Landau<G>::operator(), condition,
// deriv, run).  The reference maximises the functional by gradient ascent
// with a line search (landau.h:77-136); this library provides the
// checkerboard overrelaxation update (Mandula & Ogilvie, Phys. Lett. B 248
// (1990) 156) over the three SU(2) subgroups (Cabibbo & Marinari, Phys.
// Lett. B 119 (1982) 387), with r^omega taken from the binomial series to
// second order and renormalised.  Sites of one parity share no link, so a
// half-sweep updates all of them at once; each link is read and written once
// per half-sweep.
#include "mesh_internal.h"

// FP64 half-sweep: forward links with non-temporal loads and stores
#ifndef MESH_OR_STREAM
#define MESH_OR_STREAM 1
#endif

using namespace mesh_dev;

namespace {

// U^h_mu(x) = h(x) U_mu(x) h(x+mu)^+ in double (h == nullptr: identity)
template <class R>
__device__ __forceinline__ m3<double> ug(Geo const &g, const R *__restrict__ U, const R *__restrict__ h, int mu,
                                         long x, long xp)
{
	m3<double> u = to_f64(load(U, mu, g, x));
	if (h)
	{
		u = mul_nn(to_f64(load(h, 0, g, x)), u);
		u = mul_nd(u, to_f64(load(h, 0, g, xp)));
	}
	return u;
}

template <class R>
__global__ __launch_bounds__(256) void k_functional(Geo g, const R *__restrict__ U, const R *__restrict__ h,
                                                    double *partials, int nb)
{
	long x = lane_site(g, (long)blockIdx.x * 256 + threadIdx.x);
	double v[1] = {0.0};
	if (x < g.V)
	{
		double a = 0.0;
		for (int mu = 0; mu < g.nd; ++mu)
		{
			m3<double> u = ug(g, U, h, mu, x, shift_c(g, x, mu, coord(g, x, mu), 1));
			a += (u.e[0].re + u.e[4].re) + u.e[8].re;
		}
		v[0] = a;
	}
	block_sum_store<1>(v, partials, nb, blockIdx.x);
}

template <class R>
__global__ __launch_bounds__(256) void k_divsum(Geo g, const R *__restrict__ U, const R *__restrict__ h,
                                                double *partials, int nb)
{
	long x = lane_site(g, (long)blockIdx.x * 256 + threadIdx.x);
	double v[1] = {0.0};
	if (x < g.V)
	{
		m3<double> D = zero_m3<double>();
		for (int mu = 0; mu < g.nd; ++mu)
		{
			int c = coord(g, x, mu);
			long xm = shift_c(g, x, mu, c, -1);
			m3<double> a = ta(ug(g, U, h, mu, x, shift_c(g, x, mu, c, 1)));
			m3<double> b = ta(ug(g, U, h, mu, xm, x));
#pragma unroll
			for (int k = 0; k < 9; ++k)
			{
				D.e[k].re += a.e[k].re - b.e[k].re;
				D.e[k].im += a.e[k].im - b.e[k].im;
			}
		}
		v[0] = norm2_m3(D);
	}
	block_sum_store<1>(v, partials, nb, blockIdx.x);
}

template <class R>
__global__ __launch_bounds__(256) void k_deriv(Geo g, const R *__restrict__ U, const R *__restrict__ h, double c,
                                               R *__restrict__ F)
{
	long x = lane_site(g, (long)blockIdx.x * 256 + threadIdx.x);
	if (x >= g.V)
		return;
	m3<double> d = zero_m3<double>();
	for (int mu = 0; mu < g.nd; ++mu)
	{
		int cm = coord(g, x, mu);
		long xp = shift_c(g, x, mu, cm, 1), xm = shift_c(g, x, mu, cm, -1);
		m3<double> a = mul_nd(to_f64(load(U, mu, g, x)), to_f64(load(h, 0, g, xp)));
		m3<double> b = mul_nn(to_f64(load(h, 0, g, xm)), to_f64(load(U, mu, g, xm)));
		// d += a + b^+
#pragma unroll
		for (int i = 0; i < 3; ++i)
#pragma unroll
			for (int j = 0; j < 3; ++j)
			{
				d.e[3 * i + j].re += a.e[3 * i + j].re + b.e[3 * j + i].re;
				d.e[3 * i + j].im += a.e[3 * i + j].im + -b.e[3 * j + i].im;
			}
	}
	m3<double> f = ta(mul_nn(to_f64(load(h, 0, g, x)), d));
#pragma unroll
	for (int k = 0; k < 9; ++k)
	{
		f.e[k].re *= c;
		f.e[k].im *= c;
	}
	store(F, 0, g, x, from_f64<R>(f));
}

// G in SU(3) raising Re tr(G K), built from three SU(2) subgroup steps with
// overrelaxation parameter omega
template <class R> __device__ __forceinline__ m3<R> or_transform(m3<R> K, R omega)
{
	m3<R> G = unit_m3<R>();
	const R c2 = R(0.5) * omega * (omega - R(1));
#pragma unroll
	for (int s = 0; s < 3; ++s)
	{
		const int i = s == 2 ? 1 : 0;
		const int j = s == 0 ? 1 : 2;
		R b0 = R(0.5) * (K.e[4 * i].re + K.e[4 * j].re);
		R b3 = R(0.5) * (K.e[4 * i].im - K.e[4 * j].im);
		R b1 = R(0.5) * (K.e[3 * i + j].im + K.e[3 * j + i].im);
		R b2 = R(0.5) * (K.e[3 * i + j].re - K.e[3 * j + i].re);
		R n2 = b0 * b0;
		n2 = fma(b1, b1, n2);
		n2 = fma(b2, b2, n2);
		n2 = fma(b3, b3, n2);
		if (!(n2 > R(0)))
			continue;
		R inv = R(1) / sqrt(n2);
		R r0 = b0 * inv, r1 = -(b1 * inv), r2 = -(b2 * inv), r3 = -(b3 * inv);
		R a = r0 - R(1);
		R rv2 = r1 * r1;
		rv2 = fma(r2, r2, rv2);
		rv2 = fma(r3, r3, rv2);
		R g0 = fma(c2, fma(a, a, -rv2), fma(omega, a, R(1)));
		R fac = fma(R(2) * c2, a, omega);
		R g1 = r1 * fac, g2 = r2 * fac, g3 = r3 * fac;
		R gn2 = g0 * g0;
		gn2 = fma(g1, g1, gn2);
		gn2 = fma(g2, g2, gn2);
		gn2 = fma(g3, g3, gn2);
		R gi = R(1) / sqrt(gn2);
		g0 *= gi;
		g1 *= gi;
		g2 *= gi;
		g3 *= gi;
		cpx<R> e00 = {g0, g3}, e01 = {g2, g1}, e10 = {-g2, g1}, e11 = {g0, -g3};
#pragma unroll
		for (int c = 0; c < 3; ++c)
		{
			cpx<R> ki = K.e[3 * i + c], kj = K.e[3 * j + c];
			cpx<R> ni = c0<R>(), nj = c0<R>();
			cmadd(ni, e00, ki);
			cmadd(ni, e01, kj);
			cmadd(nj, e10, ki);
			cmadd(nj, e11, kj);
			K.e[3 * i + c] = ni;
			K.e[3 * j + c] = nj;
			cpx<R> gi_ = G.e[3 * i + c], gj_ = G.e[3 * j + c];
			cpx<R> mi = c0<R>(), mj = c0<R>();
			cmadd(mi, e00, gi_);
			cmadd(mi, e01, gj_);
			cmadd(mj, e10, gi_);
			cmadd(mj, e11, gj_);
			G.e[3 * i + c] = mi;
			G.e[3 * j + c] = mj;
		}
	}
	return G;
}

// overrelaxation at the t-th site of the given parity; rows along the
// fastest dimension (even extent) hold half their sites in each parity
template <class R> __device__ __forceinline__ void or_site(Geo const &g, R *U, R *h, int parity, R omega, long t)
{
	int f = g.nd - 1;
	int hl = g.L[f] >> 1;
	// sites are below 2^26: 32-bit divisions
	long row = (long)((unsigned)t / (unsigned)hl);
	int j = (int)(t - row * hl);
	long x0 = row * g.L[f];
	// x's coordinates: the row start's in the slower directions, 2 j + the
	// parity bit along the row (one division per direction in all)
	int cc[4] = {0, 0, 0, 0}, ps = 0;
#pragma unroll
	for (int mu = 0; mu < 3; ++mu)
		if (mu < f)
		{
			cc[mu] = coord(g, x0, mu);
			ps += cc[mu];
		}
	long x = x0 + 2 * j + ((parity + ps) & 1);
	m3<R> K = zero_m3<R>();
	long xm[4];
	// The links are kept in registers for the update, or the L2 has evicted
	// them by then and the field is read twice per half-sweep: FP32 keeps all
	// 2 Nd, FP64 the Nd forward ones (all would cost 288 registers)
	constexpr bool keep = sizeof(R) == 4;
	m3<R> la[4], lb[keep ? 4 : 1];
#pragma unroll
	for (int mu = 0; mu < 4; ++mu)
	{
		if (mu >= g.nd)
			break;
		xm[mu] = shift_c(g, x, mu, mu == f ? (int)(x - x0) : cc[mu], -1);
		// FP64: the forward links stream past the L2 so that the backward
		// ones, read again after the transformation, are still there
		m3<R> a = (!keep && MESH_OR_STREAM) ? load_stream(U, mu, g, x) : load(U, mu, g, x);
		m3<R> b = load(U, mu, g, xm[mu]);
#pragma unroll
		for (int i = 0; i < 3; ++i)
#pragma unroll
			for (int jj = 0; jj < 3; ++jj)
			{
				K.e[3 * i + jj].re += a.e[3 * i + jj].re;
				K.e[3 * i + jj].im += a.e[3 * i + jj].im;
				K.e[3 * i + jj].re += b.e[3 * jj + i].re;
				K.e[3 * i + jj].im -= b.e[3 * jj + i].im;
			}
		la[mu] = a;
		if constexpr (keep)
			lb[mu] = b;
	}
	m3<R> G = or_transform(K, omega);
#pragma unroll
	for (int mu = 0; mu < 4; ++mu)
	{
		if (mu >= g.nd)
			break;
		if constexpr (keep)
		{
			store(U, mu, g, x, mul_nn(G, la[mu]));
			store(U, mu, g, xm[mu], mul_nd(lb[mu], G));
		}
		else if (MESH_OR_STREAM)
		{
			store_stream(U, mu, g, x, mul_nn(G, la[mu]));
			store(U, mu, g, xm[mu], mul_nd(load(U, mu, g, xm[mu]), G));
		}
		else
		{
			store(U, mu, g, x, mul_nn(G, la[mu]));
			store(U, mu, g, xm[mu], mul_nd(load(U, mu, g, xm[mu]), G));
		}
	}
	if (h)
		store(h, 0, g, x, mul_nn(G, load(h, 0, g, x)));
}

template <class R> __global__ __launch_bounds__(256) void k_or_half(Geo g, R *U, R *h, int parity, R omega)
{
	long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (t < (g.V >> 1))
		or_site(g, U, h, parity, omega, t);
}

// The same with a workgroup per 256-site span (whole rows: L_fast | 256), the
// spans in the block order of the force kernels, so that a span and the spans
// holding its backward neighbours run close together and share the L2 lines
// that each half-sweep touches half of.
template <class R>
__global__ __launch_bounds__(128) void k_or_half_span(Geo g, R *U, R *h, int parity, R omega, int nb, int order)
{
	int lb = ordered_block(g, nb, 256, order);
	if (lb < 0)
		return;
	or_site(g, U, h, parity, omega, (long)lb * 128 + threadIdx.x);
}

template <class R>
__global__ __launch_bounds__(256) void k_gauge_transform(Geo g, R *__restrict__ Uo, const R *__restrict__ U,
                                                         const R *__restrict__ h)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= g.V * g.nd)
		return;
	int mu = (int)(i / g.V);
	long x = lane_site(g, i - (long)mu * g.V);
	long xp = shift_c(g, x, mu, coord(g, x, mu), 1);
	m3<R> u = mul_nn(load(h, 0, g, x), load(U, mu, g, x));
	store(Uo, mu, g, x, mul_nd(u, load(h, 0, g, xp)));
}

int setup(mesh_ctx *ctx, const mesh_grid *g, int prec, Geo *geo)
{
	if (int rc = mesh_make_geo(ctx, g, geo))
		return rc;
	return mesh_check_prec(ctx, prec);
}

inline int nblk(long n) { return (int)((n + 255) / 256); }

} // namespace

extern "C" {

int mesh_landau_functional(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U, const void *h, double *out)
{
	MESH_CHECK_CTX(ctx);
	Geo geo;
	if (int rc = setup(ctx, g, prec, &geo))
		return rc;
	if (!U || !out)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to landau_functional");
	int nb = nblk(geo.V);
	if (int rc = mesh_reserve_partials(ctx, nb))
		return rc;
	if (prec == MESH_F64)
		k_functional<double><<<nb, 256, 0, ctx->stream>>>(geo, (const double *)U, (const double *)h, ctx->partials, nb);
	else
		k_functional<float><<<nb, 256, 0, ctx->stream>>>(geo, (const float *)U, (const float *)h, ctx->partials, nb);
	MESH_LAUNCHED(ctx);
	double s;
	if (int rc = mesh_finish_reduction(ctx, nb, 1, &s))
		return rc;
	*out = s / (3.0 * geo.nd * (double)geo.V);
	return MESH_OK;
}

int mesh_landau_divsum(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U, const void *h, double *out)
{
	MESH_CHECK_CTX(ctx);
	Geo geo;
	if (int rc = setup(ctx, g, prec, &geo))
		return rc;
	if (!U || !out)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to landau_divsum");
	int nb = nblk(geo.V);
	if (int rc = mesh_reserve_partials(ctx, nb))
		return rc;
	if (prec == MESH_F64)
		k_divsum<double><<<nb, 256, 0, ctx->stream>>>(geo, (const double *)U, (const double *)h, ctx->partials, nb);
	else
		k_divsum<float><<<nb, 256, 0, ctx->stream>>>(geo, (const float *)U, (const float *)h, ctx->partials, nb);
	MESH_LAUNCHED(ctx);
	return mesh_finish_reduction(ctx, nb, 1, out);
}

int mesh_landau_deriv(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U, const void *h, void *F)
{
	MESH_CHECK_CTX(ctx);
	Geo geo;
	if (int rc = setup(ctx, g, prec, &geo))
		return rc;
	if (!U || !h || !F)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to landau_deriv");
	double c = -0.5 / (3.0 * geo.nd * (double)geo.V);
	int nb = nblk(geo.V);
	if (prec == MESH_F64)
		k_deriv<double><<<nb, 256, 0, ctx->stream>>>(geo, (const double *)U, (const double *)h, c, (double *)F);
	else
		k_deriv<float><<<nb, 256, 0, ctx->stream>>>(geo, (const float *)U, (const float *)h, c, (float *)F);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_landau_or_halfsweep(mesh_ctx *ctx, const mesh_grid *g, int prec, void *U, void *h, int parity, double omega)
{
	MESH_CHECK_CTX(ctx);
	Geo geo;
	if (int rc = setup(ctx, g, prec, &geo))
		return rc;
	if (!U || (parity != 0 && parity != 1))
		return mesh_fail(ctx, MESH_E_ARG, "bad landau_or_halfsweep arguments");
	for (int i = 0; i < geo.nd; ++i)
		if (geo.L[i] % 2)
			return mesh_fail(ctx, MESH_E_SIZE, "checkerboard update needs even extents");
	int ph = mesh_prof_begin(ctx, PROF_LANDAU_OR);
	const int lf = geo.L[geo.nd - 1];
	if (geo.V % 256 == 0 && 256 % lf == 0)
	{
		int nb = (int)(geo.V / 256);
		int gb = mesh_grid_blocks(ctx, nb);
		if (prec == MESH_F64)
			k_or_half_span<double><<<gb, 128, 0, ctx->stream>>>(geo, (double *)U, (double *)h, parity, omega, nb,
			                                                    mesh_order(ctx));
		else
			k_or_half_span<float><<<gb, 128, 0, ctx->stream>>>(geo, (float *)U, (float *)h, parity, (float)omega,
			                                                   nb, mesh_order(ctx));
	}
	else
	{
		int nb = nblk(geo.V / 2);
		if (prec == MESH_F64)
			k_or_half<double><<<nb, 256, 0, ctx->stream>>>(geo, (double *)U, (double *)h, parity, omega);
		else
			k_or_half<float><<<nb, 256, 0, ctx->stream>>>(geo, (float *)U, (float *)h, parity, (float)omega);
	}
	mesh_prof_end(ctx, ph);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_su3_gauge_transform(mesh_ctx *ctx, const mesh_grid *g, int prec, void *Uout, const void *U, const void *h)
{
	MESH_CHECK_CTX(ctx);
	Geo geo;
	if (int rc = setup(ctx, g, prec, &geo))
		return rc;
	if (!Uout || !U || !h || Uout == U)
		return mesh_fail(ctx, MESH_E_ARG, "bad gauge_transform arguments");
	int nb = nblk(geo.V * geo.nd);
	if (prec == MESH_F64)
		k_gauge_transform<double><<<nb, 256, 0, ctx->stream>>>(geo, (double *)Uout, (const double *)U,
		                                                       (const double *)h);
	else
		k_gauge_transform<float><<<nb, 256, 0, ctx->stream>>>(geo, (float *)Uout, (const float *)U,
		                                                      (const float *)h);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

} // extern "C"
