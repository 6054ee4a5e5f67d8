// Plaquette flow for 2-d U(1): a learned independence proposal built on the
// structure of the 2-d theory (no counterpart in the reference; U(1)
// conventions of u1.hip).
//
// On an L0 x L1 torus the V plaquette angles of a link configuration are
// independent variables except for one constraint (their sum is 0 mod 2 pi),
// and the links are fixed by V - 1 plaquettes, two holonomies and a gauge
// transformation.  The proposal:
//   * every plaquette but the last, P(L0-1, L1-1), is theta_P = h(u), u
//     uniform on [-pi, pi), with the learned circle map
//     h(u) = g(wrap(u - t)) + t,  g(v) = sum_k w_k 2 atan(s_k tan(v / 2))
//     (NK = 8 non-compact projections; s_k = exp(clamp(y_k)), w = softmax,
//     t = atan2(y_{2K+1}, y_{2K}));
//   * maximal-tree gauge: theta_1(x0, x1) = 0 for x1 < L1 - 1 and
//     theta_0(x0, L1 - 1) = 0 for x0 < L0 - 1; then
//       theta_0(x0, x1) = sum_{k = x1}^{L1 - 2} P(x0, k)   (+ h0 on row L0 - 1)
//       theta_1(0, L1 - 1) = h1,
//       theta_1(x0 + 1, L1 - 1) = theta_1(x0, L1 - 1) + P(x0, L1 - 1)
//                                 - theta_0(x0, L1 - 1) + theta_0(x0, 0),
//     with holonomies h0, h1 uniform;
//   * a uniform gauge transformation theta_mu(x) += a(x) - a(x + mu).
// The map (V - 1 plaquettes, 2 holonomies, V - 1 gauge angles) -> 2 V link
// angles is integer-linear and unimodular, so
//   log q(U) = sum_{P != last} log q1(theta_P(U)) - (V + 1) log 2 pi,
//   q1(theta) = 1 / (2 pi g'(v)),  v = g^-1(wrap(theta - t)),
// a function of the plaquettes of U only.  The last plaquette follows from
// the others; its Boltzmann weight is what the Metropolis step corrects.
#include "mesh_internal.h"
#include "rng_dev.h"
#include "u1_plaq.h"

// log q: table-guided inverse of the circle map (0: Newton from the target)
#ifndef MESH_PFLOW_TABLE
#define MESH_PFLOW_TABLE 1
#endif

using namespace mesh_dev;
using u1_plaq::block_sum;
using u1_plaq::PGeo;
using u1_plaq::plaqp;
using u1_plaq::site;
using u1_plaq::wrapd;

namespace {

constexpr int K = 8;             // components
constexpr int NPF = 2 * K + 2;   // parameters
constexpr int THREADS = 256;
constexpr uint32_t TAG_PFLOW = 0x50464c57u;
constexpr float PI_F = 3.14159265358979323846f;
constexpr float TWO_PI_F = 6.28318530717958647692f;
constexpr double LOG_2PI = 1.8378770664093454836;

__device__ __forceinline__ float wrapf(float a)
{
	a = a - TWO_PI_F * floorf((a + PI_F) * (1.0f / TWO_PI_F));
	return a >= PI_F ? a - TWO_PI_F : (a < -PI_F ? a + TWO_PI_F : a);
}

struct Map
{
	float s[K], w[K], t, ta, tb;
};

__device__ __forceinline__ Map map_params(const float *__restrict__ y)
{
	Map p;
	float m = y[K];
	for (int k = 1; k < K; ++k)
		m = fmaxf(m, y[K + k]);
	float z = 0.f;
	for (int k = 0; k < K; ++k)
	{
		p.s[k] = expf(fminf(fmaxf(y[k], -8.f), 8.f));
		p.w[k] = expf(y[K + k] - m);
		z += p.w[k];
	}
	for (int k = 0; k < K; ++k)
		p.w[k] /= z;
	p.ta = y[2 * K];
	p.tb = y[2 * K + 1];
	p.t = atan2f(p.tb, p.ta);
	return p;
}

__device__ __forceinline__ void g_eval(Map const &p, float v, float &G, float &Gp)
{
	float c = cosf(0.5f * v), sn = sinf(0.5f * v);
	G = 0.f;
	Gp = 0.f;
	for (int k = 0; k < K; ++k)
	{
		float d = fmaf(c, c, p.s[k] * p.s[k] * sn * sn);
		G = fmaf(p.w[k], 2.f * atan2f(p.s[k] * sn, c), G);
		Gp = fmaf(p.w[k], p.s[k] / d, Gp);
	}
}

// v = g^-1(target) on [lo, hi] (a bracket of the root) from the start u:
// Newton inside a shrinking bracket, bisection when a step leaves it, until a
// step or the bracket is below 2e-7 max(1, |v|) -- an FP32 ulp of v is up to
// 2.4e-7 on [-pi, pi], so an absolute bound would let an iterate that
// alternates between neighbouring floats run to the cap (at most 40
// evaluations); returns g'(v) in Gp
__device__ __forceinline__ float g_inverse(Map const &p, float target, float &Gp, float lo = -PI_F, float hi = PI_F,
                                           float u = 2.f * PI_F)
{
	if (u > PI_F)
		u = target;
	float G;
	for (int b = 0; b < 40; ++b)
	{
		g_eval(p, u, G, Gp);
		float f = G - target;
		if (f == 0.f)
			break;
		if (f < 0.f)
			lo = u;
		else
			hi = u;
		if (hi - lo <= 2e-7f * fmaxf(1.f, fabsf(u)))
			break;
		float un = u - f / Gp;
		if (un > lo && un < hi)
		{
			bool done = fabsf(un - u) <= 2e-7f * fmaxf(1.f, fabsf(un));
			u = un;
			if (done)
				break;
		}
		else
			u = 0.5f * (lo + hi);
	}
	g_eval(p, u, G, Gp);
	return u;
}

// g on the grid v_j = -pi + j 2 pi / NTAB, j = 0..NTAB, one value per thread
// of the workgroup: G is monotone, so a table lookup brackets the root of
// g(v) = target within three cells and interpolation starts Newton within
// about (2 pi / NTAB)^2 of it
constexpr int NTAB = 256;
__device__ __forceinline__ void g_table(Map const &p, float *tab)
{
	for (int j = threadIdx.x; j <= NTAB; j += THREADS)
	{
		float G, Gp;
		g_eval(p, fmaf((float)j, TWO_PI_F / NTAB, -PI_F), G, Gp);
		tab[j] = G;
	}
}
__device__ __forceinline__ float g_inverse_tab(Map const &p, const float *tab, float target, float &Gp)
{
	int a = 0, b = NTAB;
	while (b - a > 1)
	{
		int m = (a + b) >> 1;
		if (tab[m] <= target)
			a = m;
		else
			b = m;
	}
	const float h = TWO_PI_F / NTAB;
	const float va = fmaf((float)a, h, -PI_F);
	const float dG = tab[a + 1] - tab[a];
	float u = dG > 0.f ? fmaf((target - tab[a]) / dG, h, va) : va + 0.5f * h;
	// one cell of margin either side: the table holds g rounded to FP32
	const float lo = a > 0 ? va - h : -PI_F, hi = a + 2 <= NTAB ? va + 2.f * h : PI_F;
	u = fminf(fmaxf(u, lo), hi);
	return g_inverse(p, target, Gp, lo, hi, u);
}

__device__ __forceinline__ float uniform_angle(uint64_t seed, uint32_t step, uint32_t a, uint32_t b)
{
	return u1_plaq::uniform_angle(seed, step, a, b, TAG_PFLOW);
}

// Proposals: links of configuration c into theta_out[c][2V]; with out,
// out[2c + {0,1}] = {log q, S/beta} of the proposal, log q from the draws
// (-log g' of each free plaquette, the inverse map's value up to rounding)
// and S from the links' plaquettes.
// Draws: plaquette x (x != V-1) counter (x, 2c, step<<5, tag); holonomies
// (V, 2c, ..) and (V+1, 2c, ..); gauge angle a(x) (x, 2c+1, ..).
// LDS: 4V floats.
__device__ void pflow_sample(PGeo const &g, const float *__restrict__ params, uint64_t seed, uint32_t step,
                             float *__restrict__ theta_out, double *out, int c, float *lds, double (*red)[4])
{
	float *P = lds;          // V plaquette angles (the last unused)
	float *t = lds + g.V;    // 2V links
	float *al = lds + 3 * g.V; // V gauge angles
	const Map p = map_params(params);
	const uint32_t cc = (uint32_t)c << 1;
	double v[1] = {0.0};
	for (int x = threadIdx.x; x < g.V; x += THREADS)
	{
		if (x < g.V - 1)
		{
			float u = wrapf(uniform_angle(seed, step, (uint32_t)x, cc) - p.t), G, Gp;
			g_eval(p, u, G, Gp);
			P[x] = wrapf(G + p.t);
			v[0] -= (double)logf(Gp);
		}
		al[x] = uniform_angle(seed, step, (uint32_t)x, cc | 1u);
	}
	__syncthreads();
	const float h0 = uniform_angle(seed, step, (uint32_t)g.V, cc);
	const float h1 = uniform_angle(seed, step, (uint32_t)g.V + 1u, cc);
	float *o = theta_out + (long)c * 2 * g.V;
	u1_plaq::links_from_plaquettes(g, P, h0, h1, al, t, o, THREADS);
	if (!out)
		return;
	__syncthreads();
	double w[2] = {v[0], 0.0};
	for (int x = threadIdx.x; x < g.V; x += THREADS)
	{
		const int x0 = x / g.L1, x1 = x - x0 * g.L1;
		w[1] += 1.0 - cos((double)plaqp(g, o, x0, x1));
	}
	block_sum<2>(w, red);
	if (threadIdx.x == 0)
	{
		out[2 * c + 0] = w[0] - (double)(g.V + 1) * LOG_2PI - (double)(g.V - 1) * LOG_2PI;
		out[2 * c + 1] = w[1];
	}
}

// out[2c + {0,1}] = {log q(U), S(U)/beta} of configurations theta[c][2V];
// LDS: 2V floats
__device__ void pflow_logq(PGeo const &g, const float *__restrict__ params, const float *__restrict__ theta,
                           double *out, int c, float *t, double (*red)[4], float *tab)
{
	const Map p = map_params(params);
	for (int i = threadIdx.x; i < 2 * g.V; i += THREADS)
		t[i] = theta[(long)c * 2 * g.V + i];
	if (MESH_PFLOW_TABLE)
		g_table(p, tab);
	__syncthreads();
	double v[2] = {0.0, 0.0};
	for (int x = threadIdx.x; x < g.V; x += THREADS)
	{
		int x0 = x / g.L1, x1 = x - x0 * g.L1;
		float tp = plaqp(g, t, x0, x1);
		v[1] += 1.0 - cos((double)tp);
		if (x < g.V - 1)
		{
			float Gp;
			if (MESH_PFLOW_TABLE)
				g_inverse_tab(p, tab, wrapf(tp - p.t), Gp);
			else
				g_inverse(p, wrapf(tp - p.t), Gp);
			v[0] -= (double)logf(Gp);
		}
	}
	block_sum<2>(v, red);
	if (threadIdx.x == 0)
	{
		out[2 * c + 0] = v[0] - (double)(g.V + 1) * LOG_2PI - (double)(g.V - 1) * LOG_2PI;
		out[2 * c + 1] = v[1];
	}
}

__global__ __launch_bounds__(THREADS) void k_pflow_sample(PGeo g, const float *__restrict__ params, uint64_t seed,
                                                          uint32_t step, float *__restrict__ theta_out)
{
	extern __shared__ float lds[];
	pflow_sample(g, params, seed, step, theta_out, nullptr, blockIdx.x, lds, nullptr);
}

__global__ __launch_bounds__(THREADS) void k_pflow_logq(PGeo g, const float *__restrict__ params,
                                                        const float *__restrict__ theta, double *out)
{
	extern __shared__ float t[]; // 2V
	__shared__ double red[2][4];
	__shared__ float tab[NTAB + 1];
	pflow_logq(g, params, theta, out, blockIdx.x, t, red, tab);
}

// one MH proposal round in one launch: workgroups 0..batch-1 take log q of
// the current configurations, batch..2 batch-1 draw the proposals with their
// log q and S/beta, side by side on the CUs
__global__ __launch_bounds__(THREADS) void k_pflow_propose(PGeo g, const float *__restrict__ params, int batch,
                                                           const float *__restrict__ theta, double *state,
                                                           uint64_t seed, uint32_t step, float *theta_prop,
                                                           double *prop_out)
{
	extern __shared__ float lds[];
	__shared__ double red[2][4];
	__shared__ float tab[NTAB + 1];
	if ((int)blockIdx.x < batch)
		pflow_logq(g, params, theta, state, blockIdx.x, lds, red, tab);
	else
		pflow_sample(g, params, seed, step, theta_prop, prop_out, blockIdx.x - batch, lds, red);
}

// Reverse-KL gradient of one configuration: loss = log q + beta S with
// theta_P = h(u_P) for the free plaquettes and theta_last = -sum theta_P;
// partial[c][NPF] (FP64), out[2c + {0,1}] = {log q, S/beta}.
__global__ __launch_bounds__(THREADS) void k_pflow_grad(PGeo g, const float *__restrict__ params, uint64_t seed,
                                                        uint32_t step, double beta, double *__restrict__ partial,
                                                        double *out)
{
	__shared__ double red[NPF + 3][4];
	__shared__ double s_last;
	const int c = blockIdx.x;
	const float *y = params;
	const Map p = map_params(y);
	const uint32_t cc = (uint32_t)c << 1;
	// pass 1: sum of the free plaquettes (fixes the last one)
	double v1[1] = {0.0};
	for (int x = threadIdx.x; x < g.V - 1; x += THREADS)
	{
		float v = wrapf(uniform_angle(seed, step, (uint32_t)x, cc) - p.t), G, Gp;
		g_eval(p, v, G, Gp);
		v1[0] += (double)wrapf(G + p.t);
	}
	block_sum<1>(v1, red);
	if (threadIdx.x == 0)
		s_last = v1[0];
	__syncthreads();
	const double th_last = wrapd(-s_last);
	const float sin_last = (float)sin(th_last);
	// pass 2: per-plaquette adjoints and parameter gradients
	double acc[NPF + 2];
	for (int i = 0; i < NPF + 2; ++i)
		acc[i] = 0.0;
	for (int x = threadIdx.x; x < g.V - 1; x += THREADS)
	{
		float v = wrapf(uniform_angle(seed, step, (uint32_t)x, cc) - p.t);
		float cs = cosf(0.5f * v), sn = sinf(0.5f * v);
		float G = 0.f, Gp = 0.f, Gpp = 0.f, gk[K], gpk[K], dgs[K], dgps[K];
		for (int k = 0; k < K; ++k)
		{
			float s2 = p.s[k] * p.s[k];
			float d = fmaf(cs, cs, s2 * sn * sn);
			gk[k] = 2.f * atan2f(p.s[k] * sn, cs);
			gpk[k] = p.s[k] / d;
			dgs[k] = 2.f * sn * cs / d;
			dgps[k] = (cs * cs - s2 * sn * sn) / (d * d);
			G = fmaf(p.w[k], gk[k], G);
			Gp = fmaf(p.w[k], gpk[k], Gp);
			Gpp = fmaf(p.w[k], -p.s[k] * (s2 - 1.f) * sn * cs / (d * d), Gpp);
		}
		float th = wrapf(G + p.t);
		acc[NPF] += -(double)logf(Gp);
		acc[NPF + 1] += 1.0 - cos((double)th);
		// dL/dtheta_P (through beta S and the last plaquette)
		float A = (float)beta * (sinf(th) - sin_last);
		float iG = 1.f / Gp;
		float dLdt = A * (1.f - Gp) + Gpp * iG;
		float dLdw[K], sw = 0.f;
		for (int k = 0; k < K; ++k)
		{
			dLdw[k] = A * gk[k] - gpk[k] * iG;
			sw = fmaf(p.w[k], dLdw[k], sw);
			float dLds = p.w[k] * (A * dgs[k] - dgps[k] * iG);
			acc[k] += (y[k] > -8.f && y[k] < 8.f) ? (double)(p.s[k] * dLds) : 0.0;
		}
		for (int k = 0; k < K; ++k)
			acc[K + k] += (double)(p.w[k] * (dLdw[k] - sw));
		float r2 = fmaf(p.ta, p.ta, p.tb * p.tb) + 1e-20f;
		acc[2 * K] += (double)(dLdt * (-p.tb / r2));
		acc[2 * K + 1] += (double)(dLdt * (p.ta / r2));
	}
	block_sum<NPF + 2>(acc, red);
	if (threadIdx.x == 0)
	{
		for (int i = 0; i < NPF; ++i)
			partial[(long)c * NPF + i] = acc[i];
		out[2 * c + 0] = acc[NPF] - (double)(g.V + 1) * LOG_2PI - (double)(g.V - 1) * LOG_2PI;
		out[2 * c + 1] = acc[NPF + 1] + (1.0 - cos(th_last));
	}
}

__global__ void k_pflow_reduce(const double *__restrict__ partial, int batch, double scale, float *__restrict__ grad)
{
	int i = threadIdx.x;
	if (i >= NPF)
		return;
	double s = 0.0;
	for (int c = 0; c < batch; ++c)
		s += partial[(long)c * NPF + i];
	grad[i] = (float)(s * scale);
}

// log s_k spread over [-3, 1], w uniform, t = 0.  Distinct components: with
// equal ones every component would receive the same gradient for ever.
__global__ void k_pflow_init(float *p)
{
	int i = threadIdx.x;
	if (i < K)
		p[i] = -3.f + 4.f * (float)i / (float)(K - 1);
	else if (i < NPF)
		p[i] = i == 2 * K ? 1.f : 0.f;
}

int pflow_setup(mesh_ctx *ctx, const mesh_grid *g, int batch, PGeo *pg)
{
	if (!g || g->nd != 2 || g->dims[0] < 2 || g->dims[1] < 2)
		return mesh_fail(ctx, MESH_E_SIZE, "the plaquette flow needs a 2-d lattice with extents >= 2");
	long V = (long)g->dims[0] * g->dims[1];
	if (V > 4096)
		return mesh_fail(ctx, MESH_E_SIZE, "the plaquette flow holds a configuration in LDS: volume <= 4096");
	if (batch < 1)
		return mesh_fail(ctx, MESH_E_ARG, "batch must be >= 1");
	pg->L0 = g->dims[0];
	pg->L1 = g->dims[1];
	pg->V = (int)V;
	return MESH_OK;
}

} // namespace

extern "C" {

int mesh_pflow_num_params(void) { return NPF; }

int mesh_pflow_init(mesh_ctx *ctx, float *params)
{
	MESH_CHECK_CTX(ctx);
	if (!params)
		return mesh_fail(ctx, MESH_E_ARG, "params is NULL");
	k_pflow_init<<<1, 64, 0, ctx->stream>>>(params);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_pflow_sample(mesh_ctx *ctx, const mesh_grid *g, int batch, const float *params, uint64_t seed, uint32_t step,
                      float *theta_out)
{
	MESH_CHECK_CTX(ctx);
	PGeo pg;
	if (int rc = pflow_setup(ctx, g, batch, &pg))
		return rc;
	if (!params || !theta_out)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to pflow_sample");
	size_t lds = 4 * (size_t)pg.V * sizeof(float);
	MESH_HIP(ctx, hipFuncSetAttribute((const void *)k_pflow_sample, hipFuncAttributeMaxDynamicSharedMemorySize,
	                                  (int)lds));
	k_pflow_sample<<<batch, THREADS, lds, ctx->stream>>>(pg, params, seed, step, theta_out);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_pflow_propose(mesh_ctx *ctx, const mesh_grid *g, int batch, const float *params, const float *theta,
                       double *state_dev, uint64_t seed, uint32_t step, float *theta_prop, double *prop_dev)
{
	MESH_CHECK_CTX(ctx);
	PGeo pg;
	if (int rc = pflow_setup(ctx, g, batch, &pg))
		return rc;
	if (!params || !theta || !state_dev || !theta_prop || !prop_dev)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to pflow_propose");
	if (theta == theta_prop || state_dev == prop_dev)
		return mesh_fail(ctx, MESH_E_ARG, "pflow_propose: the proposals and their log q need their own buffers");
	size_t lds = 4 * (size_t)pg.V * sizeof(float);
	MESH_HIP(ctx, hipFuncSetAttribute((const void *)k_pflow_propose, hipFuncAttributeMaxDynamicSharedMemorySize,
	                                  (int)lds));
	k_pflow_propose<<<2 * batch, THREADS, lds, ctx->stream>>>(pg, params, batch, theta, state_dev, seed, step,
	                                                          theta_prop, prop_dev);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_pflow_logq(mesh_ctx *ctx, const mesh_grid *g, int batch, const float *params, const float *theta,
                    double *out_dev)
{
	MESH_CHECK_CTX(ctx);
	PGeo pg;
	if (int rc = pflow_setup(ctx, g, batch, &pg))
		return rc;
	if (!params || !theta || !out_dev)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to pflow_logq");
	size_t lds = 2 * (size_t)pg.V * sizeof(float);
	k_pflow_logq<<<batch, THREADS, lds, ctx->stream>>>(pg, params, theta, out_dev);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

size_t mesh_pflow_workspace_bytes(const mesh_grid *g, int batch)
{
	PGeo pg;
	if (pflow_setup(nullptr, g, batch, &pg))
		return 0;
	return (size_t)batch * NPF * sizeof(double);
}

int mesh_pflow_gradient(mesh_ctx *ctx, const mesh_grid *g, int batch, const float *params, uint64_t seed,
                        uint32_t step, double beta, void *workspace, float *grad, double *out_dev)
{
	MESH_CHECK_CTX(ctx);
	PGeo pg;
	if (int rc = pflow_setup(ctx, g, batch, &pg))
		return rc;
	if (!params || !workspace || !grad || !out_dev)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to pflow_gradient");
	double *part = (double *)workspace;
	k_pflow_grad<<<batch, THREADS, 0, ctx->stream>>>(pg, params, seed, step, beta, part, out_dev);
	MESH_LAUNCHED(ctx);
	k_pflow_reduce<<<1, 64, 0, ctx->stream>>>(part, batch, 1.0 / batch, grad);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_pflow_adam(mesh_ctx *ctx, float *params, const float *grad, float *m, float *v, int t, double lr)
{
	MESH_CHECK_CTX(ctx);
	if (!params || !grad || !m || !v || t < 1)
		return mesh_fail(ctx, MESH_E_ARG, "bad pflow_adam arguments");
	return mesh_launch_adam(ctx, NPF, params, grad, m, v, t, lr);
}

} // extern "C"
