// U(1) pure-gauge theory on a 2-d L0 x L1 periodic lattice, batched over
// independent configurations ("chains"), FP32 angles.
//
// Reference semantics: U(1) element algebra src/gauge/groups.h:29-108
// (exp(i a) = (cos a, sin a); momenta a.im = N(0,1)/sqrt 2, i.e. K = sum q^2;
// force (beta/2) Im(U S)), HMC driver src/gauge/hmc.cpp:75-107 instantiated
// for U1 at hmc.cpp:109-110, Wilson action src/gauge/utils.h:141-146.
//
// Conventions: angle theta[c][mu][x], x = x0 * L1 + x1, mu = 0 along x0;
// theta_P(x) = theta_0(x) + theta_1(x+0) - theta_0(x+1) - theta_1(x);
// S = beta sum_x (1 - cos theta_P(x)); Q = round(sum_x wrap(theta_P(x)) / 2 pi).
//
// Design: a configuration (2 V floats, 8 KB at 32^2) lives in LDS for the
// whole trajectory; one workgroup of 256 threads owns one configuration, so a
// trajectory is one launch over the batch with no HBM traffic between steps.
#include "mesh_internal.h"
#include "rng_dev.h"

using namespace mesh_dev;

namespace {

constexpr int U1_THREADS = 256;
constexpr uint32_t TAG_U1MOM = 0x55314d4fu;   // momenta
constexpr uint32_t TAG_U1METRO = 0x55314d45u; // Metropolis
constexpr uint32_t TAG_U1RAND = 0x55315244u;  // random start
constexpr float PI_F = 3.14159265358979323846f;
constexpr float TWO_PI_F = 6.28318530717958647692f;

struct U1Geo
{
	int L0, L1, V;
};

__device__ __forceinline__ float wrap_pi(float a)
{
	// to [-pi, pi)
	a = a - TWO_PI_F * floorf((a + PI_F) * (1.0f / TWO_PI_F));
	return a >= PI_F ? a - TWO_PI_F : (a < -PI_F ? a + TWO_PI_F : a);
}

__device__ __forceinline__ int xp0(U1Geo const &g, int x) { return x + g.L1 >= g.V ? x + g.L1 - g.V : x + g.L1; }
__device__ __forceinline__ int xm0(U1Geo const &g, int x) { return x - g.L1 < 0 ? x - g.L1 + g.V : x - g.L1; }
__device__ __forceinline__ int xp1(U1Geo const &g, int x)
{
	int c = x % g.L1;
	return c + 1 == g.L1 ? x + 1 - g.L1 : x + 1;
}
__device__ __forceinline__ int xm1(U1Geo const &g, int x)
{
	int c = x % g.L1;
	return c == 0 ? x - 1 + g.L1 : x - 1;
}

// plaquette angle at x from an LDS configuration t[2*V] (unwrapped)
__device__ __forceinline__ float plaq(U1Geo const &g, const float *t, int x)
{
	return t[x] + t[g.V + xp0(g, x)] - t[xp1(g, x)] - t[g.V + x];
}

// block sum of NQ doubles over 256 threads; result valid in thread 0
template <int NQ> __device__ __forceinline__ void bsum(double (&v)[NQ], double (*red)[4])
{
	int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
	for (int q = 0; q < NQ; ++q)
	{
		double s = v[q];
#pragma unroll
		for (int o = 32; o >= 1; o >>= 1)
			s += __shfl_xor(s, o, 64);
		if (lane == 0)
			red[q][w] = s;
	}
	__syncthreads();
	if (threadIdx.x == 0)
#pragma unroll
		for (int q = 0; q < NQ; ++q)
			v[q] = ((red[q][0] + red[q][1]) + red[q][2]) + red[q][3];
	__syncthreads();
}

// one normal for link (c, mu, x) of a stream
__device__ __forceinline__ double u1_normal(uint64_t seed, uint32_t tag, uint32_t step, int c, int mu, int x)
{
	u4 ctr = {(uint32_t)x, ((uint32_t)c << 1) | (uint32_t)mu, step << 5, tag};
	u4 r = philox4x32_10(ctr, (uint32_t)seed, (uint32_t)(seed >> 32));
	double z0, z1;
	normal_pair(r, &z0, &z1);
	return z0;
}

// uniform in [0,1) for (c, mu, x) of a stream
__device__ __forceinline__ double u1_uniform(uint64_t seed, uint32_t tag, uint32_t step, int c, int mu, int x)
{
	u4 ctr = {(uint32_t)x, ((uint32_t)c << 1) | (uint32_t)mu, step << 5, tag};
	u4 r = philox4x32_10(ctr, (uint32_t)seed, (uint32_t)(seed >> 32));
	return u01_from_bits(r.x, r.y) - 0x1p-53;
}

// Haar-random (uniform) angles
__global__ __launch_bounds__(256) void k_u1_random(U1Geo g, int batch, float *th, uint64_t seed)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= (long)batch * 2 * g.V)
		return;
	int c = (int)(i / (2 * g.V));
	int r = (int)(i - (long)c * 2 * g.V);
	int mu = r / g.V, x = r - mu * g.V;
	double u = u1_uniform(seed, TAG_U1RAND, 0u, c, mu, x);
	th[i] = (float)(6.283185307179586477 * u - 3.141592653589793238);
}

// observables per configuration: out[c*4 + {0,1,2,3}] = {S/beta = sum(1-cos),
// Q (integer), sum cos theta_P, sum wrap(theta_P)}
__global__ __launch_bounds__(U1_THREADS) void k_u1_obs(U1Geo g, const float *__restrict__ th, double *out)
{
	extern __shared__ float lds[];
	__shared__ double red[4][4];
	int c = blockIdx.x;
	const float *t0 = th + (long)c * 2 * g.V;
	for (int i = threadIdx.x; i < 2 * g.V; i += U1_THREADS)
		lds[i] = t0[i];
	__syncthreads();
	double v[3] = {0.0, 0.0, 0.0};
	for (int x = threadIdx.x; x < g.V; x += U1_THREADS)
	{
		float p = wrap_pi(plaq(g, lds, x));
		double cp = cos((double)p);
		v[0] += 1.0 - cp;
		v[1] += cp;
		v[2] += (double)p;
	}
	bsum<3>(v, red);
	if (threadIdx.x == 0)
	{
		out[4 * c + 0] = v[0];
		out[4 * c + 1] = rint(v[2] * (1.0 / 6.283185307179586477));
		out[4 * c + 2] = v[1];
		out[4 * c + 3] = v[2];
	}
}

// One HMC trajectory per configuration, entirely in LDS.
// out[c*4 + {0,1,2,3}] = {Delta H, accepted, u, S_new/beta}
__global__ __launch_bounds__(U1_THREADS) void k_u1_hmc(U1Geo g, float *th, const double *__restrict__ deltas, int nd,
                                                      double beta, uint64_t seed, uint32_t traj, int metropolis,
                                                      double *out)
{
	extern __shared__ float lds[];
	float *t = lds;               // angles, 2V
	float *q = lds + 2 * g.V;     // momenta, 2V
	float *sp = lds + 4 * g.V;    // sin theta_P, V
	__shared__ double red[4][4];
	__shared__ int acc_s;
	int c = blockIdx.x;
	float *tg = th + (long)c * 2 * g.V;
	const float hb = (float)(0.5 * beta);
	double eold = 0.0;
	for (int i = threadIdx.x; i < 2 * g.V; i += U1_THREADS)
	{
		t[i] = tg[i];
		int mu = i >= g.V, x = i - mu * g.V;
		float qq = (float)(u1_normal(seed, TAG_U1MOM, traj, c, mu, x) * 0.70710678118654752440);
		q[i] = qq;
	}
	__syncthreads();
	// per-site energy before: beta (1 - cos theta_P(x)) + q_0(x)^2 + q_1(x)^2
	for (int x = threadIdx.x; x < g.V; x += U1_THREADS)
	{
		double k = (double)q[x] * q[x] + (double)q[g.V + x] * q[g.V + x];
		eold += beta * (1.0 - cos((double)plaq(g, t, x))) + k;
	}
	for (int s = 0; s < nd; ++s)
	{
		const float d = (float)deltas[s];
		if (s % 2 == 0)
		{
			for (int i = threadIdx.x; i < 2 * g.V; i += U1_THREADS)
				t[i] = wrap_pi(fmaf(d, q[i], t[i]));
		}
		else
		{
			for (int x = threadIdx.x; x < g.V; x += U1_THREADS)
				sp[x] = sinf(plaq(g, t, x));
			__syncthreads();
			const float dh = d * hb;
			for (int x = threadIdx.x; x < g.V; x += U1_THREADS)
			{
				float s0 = sp[x] - sp[xm1(g, x)];
				float s1 = sp[xm0(g, x)] - sp[x];
				q[x] = fmaf(-dh, s0, q[x]);
				q[g.V + x] = fmaf(-dh, s1, q[g.V + x]);
			}
		}
		__syncthreads();
	}
	double v[2] = {0.0, 0.0};
	for (int x = threadIdx.x; x < g.V; x += U1_THREADS)
	{
		double k = (double)q[x] * q[x] + (double)q[g.V + x] * q[g.V + x];
		double sn = 1.0 - cos((double)plaq(g, t, x));
		v[0] += beta * sn + k;
		v[1] += sn;
	}
	double dd[2] = {v[0], v[1]};
	double eo[1] = {eold};
	bsum<2>(dd, red);
	bsum<1>(eo, red);
	if (threadIdx.x == 0)
	{
		double dH = dd[0] - eo[0];
		u4 ctr = {traj, (uint32_t)c, 0u, TAG_U1METRO};
		u4 r = philox4x32_10(ctr, (uint32_t)seed, (uint32_t)(seed >> 32));
		double u = u01_from_bits(r.x, r.y);
		int a = metropolis ? (u < exp(-dH)) : 1;
		acc_s = a;
		out[4 * c + 0] = dH;
		out[4 * c + 1] = a;
		out[4 * c + 2] = u;
		out[4 * c + 3] = dd[1];
	}
	__syncthreads();
	if (acc_s)
		for (int i = threadIdx.x; i < 2 * g.V; i += U1_THREADS)
			tg[i] = t[i];
}

int u1_setup(mesh_ctx *ctx, const mesh_grid *g, int batch, U1Geo *ug)
{
	if (!g || g->nd != 2 || g->dims[0] < 2 || g->dims[1] < 2)
		return mesh_fail(ctx, MESH_E_SIZE, "U(1) lattices are 2-dimensional with extents >= 2");
	long V = (long)g->dims[0] * g->dims[1];
	if (V > 8192)
		return mesh_fail(ctx, MESH_E_SIZE, "U(1) lattice volume exceeds 8192 sites (LDS-resident kernels)");
	if (batch < 1 || batch > (1 << 24))
		return mesh_fail(ctx, MESH_E_ARG, "batch must be in [1, 2^24]");
	ug->L0 = g->dims[0];
	ug->L1 = g->dims[1];
	ug->V = (int)V;
	return MESH_OK;
}

} // namespace

extern "C" {

int mesh_u1_random(mesh_ctx *ctx, const mesh_grid *g, int batch, float *theta, uint64_t seed)
{
	MESH_CHECK_CTX(ctx);
	U1Geo ug;
	if (int rc = u1_setup(ctx, g, batch, &ug))
		return rc;
	if (!theta)
		return mesh_fail(ctx, MESH_E_ARG, "theta is NULL");
	long n = (long)batch * 2 * ug.V;
	k_u1_random<<<(int)((n + 255) / 256), 256, 0, ctx->stream>>>(ug, batch, theta, seed);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_u1_observables(mesh_ctx *ctx, const mesh_grid *g, int batch, const float *theta, double *out_dev)
{
	MESH_CHECK_CTX(ctx);
	U1Geo ug;
	if (int rc = u1_setup(ctx, g, batch, &ug))
		return rc;
	if (!theta || !out_dev)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to u1_observables");
	k_u1_obs<<<batch, U1_THREADS, 2 * ug.V * sizeof(float), ctx->stream>>>(ug, theta, out_dev);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_u1_hmc(mesh_ctx *ctx, const mesh_grid *g, int batch, float *theta, const double *deltas, int ndeltas,
                double beta, uint64_t seed, uint32_t traj, int metropolis, double *out_dev)
{
	MESH_CHECK_CTX(ctx);
	U1Geo ug;
	if (int rc = u1_setup(ctx, g, batch, &ug))
		return rc;
	if (!theta || !deltas || !out_dev || ndeltas < 1 || ndeltas % 2 != 1)
		return mesh_fail(ctx, MESH_E_ARG, "bad u1_hmc arguments (deltas must have odd length)");
	if (ug.V > 4096)
		return mesh_fail(ctx, MESH_E_SIZE, "U(1) HMC keeps a configuration in LDS: volume <= 4096");
	if (int rc = mesh_stage_doubles(ctx, deltas, ndeltas))
		return rc;
	size_t lds = 5 * (size_t)ug.V * sizeof(float);
	MESH_HIP(ctx, hipFuncSetAttribute((const void *)k_u1_hmc, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
	k_u1_hmc<<<batch, U1_THREADS, lds, ctx->stream>>>(ug, theta, ctx->aux, ndeltas, beta, seed, traj, metropolis,
	                                                  out_dev);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

} // extern "C"
