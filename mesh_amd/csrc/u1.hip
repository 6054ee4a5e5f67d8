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

// Philox counter of link (c, mu, x) of a stream: (x, c << 1 | mu bit 0,
// step << 5 | mu bit 1 << 4, tag) -- on 2-d lattices (mu < 2) simply
// (x, c << 1 | mu, step << 5, tag)
__device__ __forceinline__ u4 u1_ctr(uint32_t tag, uint32_t step, int c, int mu, int x)
{
	return {(uint32_t)x, ((uint32_t)c << 1) | (uint32_t)(mu & 1), (step << 5) | ((uint32_t)(mu >> 1) << 4), tag};
}

// one normal for link (c, mu, x) of a stream
__device__ __forceinline__ double u1_normal(uint64_t seed, uint32_t tag, uint32_t step, int c, int mu, int x)
{
	u4 ctr = u1_ctr(tag, step, c, mu, x);
	u4 r = philox4x32_10(ctr, (uint32_t)seed, (uint32_t)(seed >> 32));
	double z0, z1;
	normal_pair(r, &z0, &z1);
	return z0;
}

// uniform in [0,1) for (c, mu, x) of a stream
__device__ __forceinline__ double u1_uniform(uint64_t seed, uint32_t tag, uint32_t step, int c, int mu, int x)
{
	u4 ctr = u1_ctr(tag, step, c, mu, x);
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
	// every wave has read the start angles before the first link step moves them
	__syncthreads();
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

// ---- any dimension 2 <= nd <= 4, one thread per (configuration, site) -------
//
// Same conventions and arithmetic per link as the LDS kernels above: theta and
// momenta in FP32, theta_{mu nu}(x) = theta_mu(x) + theta_nu(x+mu) -
// theta_mu(x+nu) - theta_nu(x) summed left to right, link step theta = wrap(theta
// + t q), momentum step q_mu(x) -= t (beta/2) sum_{nu != mu} [sin theta_{mu nu}(x)
// - sin theta_{mu nu}(x - nu)] (nu ascending; theta_{nu mu} = -theta_{mu nu}),
// energies per site in FP64, differenced per site for Delta H.  On a 2-d lattice
// every link therefore gets the same bits as from the LDS kernels.

struct U1NGeo
{
	int nd, V, np;   // dimensions, sites, planes nd (nd - 1) / 2
	int L[4], st[4]; // extents, row-major strides (dimension 0 slowest)
};

__device__ __forceinline__ int n_fwd(U1NGeo const &g, int x, int mu)
{
	int c = (x / g.st[mu]) % g.L[mu];
	return c + 1 == g.L[mu] ? x - (g.L[mu] - 1) * g.st[mu] : x + g.st[mu];
}
__device__ __forceinline__ int n_bwd(U1NGeo const &g, int x, int mu)
{
	int c = (x / g.st[mu]) % g.L[mu];
	return c == 0 ? x + (g.L[mu] - 1) * g.st[mu] : x - g.st[mu];
}
// plane index of (mu, nu), mu < nu, in the order (0,1), (0,2), ..., (nd-2, nd-1)
__device__ __forceinline__ int plane_of(U1NGeo const &g, int mu, int nu)
{
	return mu * (2 * g.nd - mu - 1) / 2 + (nu - mu - 1);
}
// plaquette angle of plane (mu, nu) at x of one configuration t[nd V] (unwrapped)
__device__ __forceinline__ float plaq_n(U1NGeo const &g, const float *t, int mu, int nu, int x)
{
	return t[mu * g.V + x] + t[nu * g.V + n_fwd(g, x, mu)] - t[mu * g.V + n_fwd(g, x, nu)] - t[nu * g.V + x];
}

__global__ __launch_bounds__(256) void k_u1n_random(U1NGeo g, long n, float *th, uint64_t seed)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const long per = (long)g.nd * g.V;
	int c = (int)(i / per);
	int r = (int)(i - c * per);
	int mu = r / g.V, x = r - mu * g.V;
	double u = u1_uniform(seed, TAG_U1RAND, 0u, c, mu, x);
	th[i] = (float)(6.283185307179586477 * u - 3.141592653589793238);
}

// momenta, a copy of the start configuration, per-site energy before
__global__ __launch_bounds__(256) void k_u1n_start(U1NGeo g, int batch, const float *__restrict__ th,
                                                   float *__restrict__ th_old, float *__restrict__ q,
                                                   double *__restrict__ e_old, double beta, uint64_t seed,
                                                   uint32_t traj)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= (long)batch * g.V)
		return;
	int c = (int)(i / g.V), x = (int)(i - (long)c * g.V);
	const long o = (long)c * g.nd * g.V;
	const float *t = th + o;
	double e = 0.0;
	for (int mu = 0; mu < g.nd; ++mu)
		for (int nu = mu + 1; nu < g.nd; ++nu)
			e += beta * (1.0 - cos((double)plaq_n(g, t, mu, nu, x)));
	for (int mu = 0; mu < g.nd; ++mu)
	{
		float qq = (float)(u1_normal(seed, TAG_U1MOM, traj, c, mu, x) * 0.70710678118654752440);
		q[o + mu * g.V + x] = qq;
		e += (double)qq * qq;
		th_old[o + mu * g.V + x] = t[mu * g.V + x];
	}
	e_old[i] = e;
}

__global__ __launch_bounds__(256) void k_u1n_link(long n, float *__restrict__ th, const float *__restrict__ q, float d)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		th[i] = wrap_pi(fmaf(d, q[i], th[i]));
}

// sin theta_P of every plane at every site
__global__ __launch_bounds__(256) void k_u1n_sinp(U1NGeo g, int batch, const float *__restrict__ th,
                                                  float *__restrict__ sp)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= (long)batch * g.V)
		return;
	int c = (int)(i / g.V), x = (int)(i - (long)c * g.V);
	const float *t = th + (long)c * g.nd * g.V;
	float *s = sp + (long)c * g.np * g.V;
	for (int mu = 0; mu < g.nd; ++mu)
		for (int nu = mu + 1; nu < g.nd; ++nu)
			s[plane_of(g, mu, nu) * g.V + x] = sinf(plaq_n(g, t, mu, nu, x));
}

__global__ __launch_bounds__(256) void k_u1n_mom(U1NGeo g, int batch, const float *__restrict__ sp,
                                                 float *__restrict__ q, float dh)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= (long)batch * g.V)
		return;
	int c = (int)(i / g.V), x = (int)(i - (long)c * g.V);
	const float *s = sp + (long)c * g.np * g.V;
	float *qc = q + (long)c * g.nd * g.V;
	for (int mu = 0; mu < g.nd; ++mu)
	{
		float f = 0.0f;
		bool first = true;
		for (int nu = 0; nu < g.nd; ++nu)
		{
			if (nu == mu)
				continue;
			const int xb = n_bwd(g, x, nu);
			float a, b;
			if (mu < nu)
			{
				const float *sv = s + plane_of(g, mu, nu) * g.V;
				a = sv[x];
				b = sv[xb];
			}
			else
			{
				const float *sv = s + plane_of(g, nu, mu) * g.V;
				a = -sv[x];
				b = -sv[xb];
			}
			const float term = a - b;
			f = first ? term : f + term;
			first = false;
		}
		qc[mu * g.V + x] = fmaf(-dh, f, qc[mu * g.V + x]);
	}
}

// block partial sums per configuration of (e_new - e_old, S_new / beta);
// grid = batch * nbk blocks
__global__ __launch_bounds__(256) void k_u1n_energy(U1NGeo g, int nbk, const float *__restrict__ th,
                                                    const float *__restrict__ q, const double *__restrict__ e_old,
                                                    double beta, double *__restrict__ partials)
{
	__shared__ double red[4][4];
	const int c = blockIdx.x / nbk, b = blockIdx.x - c * nbk;
	const int x = b * 256 + threadIdx.x;
	double v[2] = {0.0, 0.0};
	if (x < g.V)
	{
		const long o = (long)c * g.nd * g.V;
		const float *t = th + o;
		double sn = 0.0, k = 0.0;
		for (int mu = 0; mu < g.nd; ++mu)
			for (int nu = mu + 1; nu < g.nd; ++nu)
				sn += 1.0 - cos((double)plaq_n(g, t, mu, nu, x));
		for (int mu = 0; mu < g.nd; ++mu)
		{
			double qq = q[o + mu * g.V + x];
			k += qq * qq;
		}
		v[0] = (beta * sn + k) - e_old[(long)c * g.V + x];
		v[1] = sn;
	}
	bsum<2>(v, red);
	if (threadIdx.x == 0)
	{
		partials[2 * (long)blockIdx.x + 0] = v[0];
		partials[2 * (long)blockIdx.x + 1] = v[1];
	}
}

// per configuration: Delta H, Metropolis (the uniform of the LDS kernel)
__global__ __launch_bounds__(256) void k_u1n_accept(int batch, int nbk, const double *__restrict__ partials,
                                                    uint64_t seed, uint32_t traj, int metropolis, double *out)
{
	int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= batch)
		return;
	double dH = 0.0, sn = 0.0;
	for (int b = 0; b < nbk; ++b)
	{
		dH += partials[2 * ((long)c * nbk + b) + 0];
		sn += partials[2 * ((long)c * nbk + b) + 1];
	}
	u4 ctr = {traj, (uint32_t)c, 0u, TAG_U1METRO};
	u4 r = philox4x32_10(ctr, (uint32_t)seed, (uint32_t)(seed >> 32));
	double u = u01_from_bits(r.x, r.y);
	int a = metropolis ? (u < exp(-dH)) : 1;
	out[4 * c + 0] = dH;
	out[4 * c + 1] = a;
	out[4 * c + 2] = u;
	out[4 * c + 3] = sn;
}

// rejected configurations get their start angles back
__global__ __launch_bounds__(256) void k_u1n_restore(long n, long per, float *__restrict__ th,
                                                     const float *__restrict__ th_old, const double *__restrict__ out)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	if (out[4 * (i / per) + 1] == 0.0)
		th[i] = th_old[i];
}

// block partial sums per configuration of (sum (1 - cos), sum cos, sum wrap theta_P)
__global__ __launch_bounds__(256) void k_u1n_obs(U1NGeo g, int nbk, const float *__restrict__ th,
                                                 double *__restrict__ partials)
{
	__shared__ double red[4][4];
	const int c = blockIdx.x / nbk, b = blockIdx.x - c * nbk;
	const int x = b * 256 + threadIdx.x;
	double v[3] = {0.0, 0.0, 0.0};
	if (x < g.V)
	{
		const float *t = th + (long)c * g.nd * g.V;
		for (int mu = 0; mu < g.nd; ++mu)
			for (int nu = mu + 1; nu < g.nd; ++nu)
			{
				float p = wrap_pi(plaq_n(g, t, mu, nu, x));
				double cp = cos((double)p);
				v[0] += 1.0 - cp;
				v[1] += cp;
				v[2] += (double)p;
			}
	}
	bsum<3>(v, red);
	if (threadIdx.x == 0)
		for (int k = 0; k < 3; ++k)
			partials[3 * (long)blockIdx.x + k] = v[k];
}

// out[4c + {0,1,2,3}] as mesh_u1_observables; Q only on 2-d lattices
__global__ __launch_bounds__(256) void k_u1n_obs_final(int batch, int nbk, int nd, const double *__restrict__ partials,
                                                       double *out)
{
	int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= batch)
		return;
	double v[3] = {0.0, 0.0, 0.0};
	for (int b = 0; b < nbk; ++b)
		for (int k = 0; k < 3; ++k)
			v[k] += partials[3 * ((long)c * nbk + b) + k];
	out[4 * c + 0] = v[0];
	out[4 * c + 1] = nd == 2 ? rint(v[2] * (1.0 / 6.283185307179586477)) : 0.0;
	out[4 * c + 2] = v[1];
	out[4 * c + 3] = v[2];
}

int u1n_setup(mesh_ctx *ctx, const mesh_grid *g, int batch, U1NGeo *ug)
{
	if (!g || g->nd < 2 || g->nd > 4)
		return mesh_fail(ctx, MESH_E_SIZE, "U(1) lattices have 2 to 4 dimensions");
	long V = 1;
	for (int i = 0; i < g->nd; ++i)
	{
		if (g->dims[i] < 2)
			return mesh_fail(ctx, MESH_E_SIZE, "U(1) lattice extents must be >= 2");
		V *= g->dims[i];
	}
	if (batch < 1 || (long)batch * g->nd * V >= (1L << 31))
		return mesh_fail(ctx, MESH_E_SIZE, "U(1) batch x links exceeds 2^31");
	ug->nd = g->nd;
	ug->V = (int)V;
	ug->np = g->nd * (g->nd - 1) / 2;
	int s = 1;
	for (int i = g->nd - 1; i >= 0; --i)
	{
		ug->L[i] = g->dims[i];
		ug->st[i] = s;
		s *= g->dims[i];
	}
	for (int i = g->nd; i < 4; ++i)
		ug->L[i] = 1, ug->st[i] = (int)V;
	return MESH_OK;
}

// the LDS-resident kernels serve 2-d lattices up to `vmax` sites, unless the
// context asks for the general kernels (option "u1_general")
bool u1_lds(mesh_ctx *ctx, const mesh_grid *g, long vmax)
{
	return !ctx->u1_general && g && g->nd == 2 && (long)g->dims[0] * g->dims[1] <= vmax;
}

// context workspace of the general kernels (grown as needed)
int u1n_workspace(mesh_ctx *ctx, size_t bytes)
{
	if (bytes <= ctx->u1_ws_bytes)
		return MESH_OK;
	if (ctx->u1_ws)
	{
		MESH_HIP(ctx, hipStreamSynchronize(ctx->stream));
		MESH_HIP(ctx, hipFree(ctx->u1_ws));
		ctx->u1_ws = nullptr;
		ctx->u1_ws_bytes = 0;
	}
	if (hipMalloc(&ctx->u1_ws, bytes) != hipSuccess)
		return mesh_fail(ctx, MESH_E_ALLOC, "cannot allocate the U(1) workspace");
	ctx->u1_ws_bytes = bytes;
	return MESH_OK;
}

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

int u1n_hmc(mesh_ctx *ctx, U1NGeo const &ug, int batch, float *theta, const double *deltas, int ndeltas, double beta,
            uint64_t seed, uint32_t traj, int metropolis, double *out)
{
	const long nl = (long)batch * ug.nd * ug.V, ns = (long)batch * ug.V;
	const int nbk = (ug.V + 255) / 256;
	const size_t b_q = align256(nl * sizeof(float)), b_sp = align256((size_t)batch * ug.np * ug.V * sizeof(float)),
	             b_e = align256(ns * sizeof(double)), b_p = align256((size_t)batch * nbk * 2 * sizeof(double));
	if (int rc = u1n_workspace(ctx, 2 * b_q + b_sp + b_e + b_p))
		return rc;
	char *w = (char *)ctx->u1_ws;
	float *q = (float *)w, *th_old = (float *)(w + b_q), *sp = (float *)(w + 2 * b_q);
	double *e_old = (double *)(w + 2 * b_q + b_sp), *partials = (double *)(w + 2 * b_q + b_sp + b_e);
	const int gl = (int)((nl + 255) / 256), gs = (int)((ns + 255) / 256);
	k_u1n_start<<<gs, 256, 0, ctx->stream>>>(ug, batch, theta, th_old, q, e_old, beta, seed, traj);
	const float hb = (float)(0.5 * beta);
	for (int s = 0; s < ndeltas; ++s)
	{
		const float d = (float)deltas[s];
		if (s % 2 == 0)
			k_u1n_link<<<gl, 256, 0, ctx->stream>>>(nl, theta, q, d);
		else
		{
			k_u1n_sinp<<<gs, 256, 0, ctx->stream>>>(ug, batch, theta, sp);
			k_u1n_mom<<<gs, 256, 0, ctx->stream>>>(ug, batch, sp, q, d * hb);
		}
	}
	k_u1n_energy<<<batch * nbk, 256, 0, ctx->stream>>>(ug, nbk, theta, q, e_old, beta, partials);
	k_u1n_accept<<<(batch + 255) / 256, 256, 0, ctx->stream>>>(batch, nbk, partials, seed, traj, metropolis, out);
	k_u1n_restore<<<gl, 256, 0, ctx->stream>>>(nl, (long)ug.nd * ug.V, theta, th_old, out);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
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

// ---- Landau gauge (src/gauge/landau.h for U1<T>) on angles: the gauge
// field g(x) = exp(i phi(x)); every value in FP64 from the FP32 angles and
// rounded once

__device__ __forceinline__ double wrap_pi_d(double a)
{
	// to [-pi, pi)
	const double tp = 6.283185307179586;
	a = a - tp * floor((a + 3.141592653589793) * (1.0 / tp));
	return a >= 3.141592653589793 ? a - tp : (a < -3.141592653589793 ? a + tp : a);
}

// angle of g(x) U_mu(x) g(x+mu)^*: (phi(x) + theta_mu(x)) - phi(x+mu)
__device__ __forceinline__ double u1_tang(U1NGeo const &g, const float *th, const float *ph, int mu, int x)
{
	const double a = (double)th[mu * g.V + x];
	return ph ? ((double)ph[x] + a) - (double)ph[n_fwd(g, x, mu)] : a;
}

// Im(g(x) d(x)) = sum_mu sin(angle at (mu, x)) - sin(angle at (mu, x - mu)) (landau.h:41-48)
__device__ __forceinline__ double u1_lgd(U1NGeo const &g, const float *th, const float *ph, int x)
{
	double s = 0.0;
	for (int mu = 0; mu < g.nd; ++mu)
	{
		s += sin(u1_tang(g, th, ph, mu, x));
		s -= sin(u1_tang(g, th, ph, mu, n_bwd(g, x, mu)));
	}
	return s;
}

// per site: mode 0 sum_mu cos(angle) (the functional), mode 1 Im(g d)^2 (the condition)
__global__ __launch_bounds__(256) void k_u1_lsum(U1NGeo g, const float *__restrict__ th,
                                                 const float *__restrict__ ph, int mode, double *partials, int nb)
{
	int x = blockIdx.x * blockDim.x + threadIdx.x;
	double v[1] = {0.0};
	if (x < g.V)
	{
		if (mode == 0)
			for (int mu = 0; mu < g.nd; ++mu)
				v[0] += cos(u1_tang(g, th, ph, mu, x));
		else
		{
			double s = u1_lgd(g, th, ph, x);
			v[0] = s * s;
		}
	}
	block_sum_store<1>(v, partials, nb, blockIdx.x);
}

// F(x) = c Im(g d), c = -0.5 / (Nd V) (landau.h:52-65; the algebra element i F)
__global__ __launch_bounds__(256) void k_u1_lderiv(U1NGeo g, const float *__restrict__ th,
                                                   const float *__restrict__ ph, double c, float *__restrict__ F)
{
	int x = blockIdx.x * blockDim.x + threadIdx.x;
	if (x < g.V)
		F[x] = (float)(c * u1_lgd(g, th, ph, x));
}

// theta^g_mu(x) = wrap(phi(x) + theta_mu(x) - phi(x+mu))
__global__ __launch_bounds__(256) void k_u1_gauge_transform(U1NGeo g, float *__restrict__ tho,
                                                            const float *__restrict__ th,
                                                            const float *__restrict__ ph)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= g.nd * g.V)
		return;
	int mu = i / g.V;
	tho[i] = (float)wrap_pi_d(u1_tang(g, th, ph, mu, i - mu * g.V));
}

// overrelaxation of one parity on the transformed angles: K(x) = sum_mu
// exp(i theta_mu(x)) + exp(-i theta_mu(x - mu)), the local maximum of
// Re(exp(i a) K) is a = -arg K, the step is omega times that angle
__global__ __launch_bounds__(256) void k_u1_or_half(U1NGeo g, float *th, float *ph, int parity, double omega)
{
	int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= (g.V >> 1))
		return;
	const int f = g.nd - 1, hl = g.L[f] >> 1;
	const int row = t / hl;
	const int x0 = row * g.L[f];
	int ps = 0;
	for (int mu = 0; mu < f; ++mu)
		ps += (x0 / g.st[mu]) % g.L[mu];
	const int x = x0 + 2 * (t - row * hl) + ((parity + ps) & 1);
	double kr = 0.0, ki = 0.0;
	for (int mu = 0; mu < g.nd; ++mu)
	{
		const double a = (double)th[mu * g.V + x], b = (double)th[mu * g.V + n_bwd(g, x, mu)];
		kr += cos(a);
		ki += sin(a);
		kr += cos(b);
		ki -= sin(b);
	}
	if (kr == 0.0 && ki == 0.0)
		return;
	const double al = -omega * atan2(ki, kr);
	for (int mu = 0; mu < g.nd; ++mu)
	{
		const int xm = n_bwd(g, x, mu);
		th[mu * g.V + x] = (float)wrap_pi_d((double)th[mu * g.V + x] + al);
		th[mu * g.V + xm] = (float)wrap_pi_d((double)th[mu * g.V + xm] - al);
	}
	if (ph)
		ph[x] = (float)wrap_pi_d((double)ph[x] + al);
}

// phi_out = wrap(phi_in + t F): exp(t i F) g (landau.h:85)
__global__ __launch_bounds__(256) void k_u1_gauge_step(int n, float *out, const float *in, const float *__restrict__ F,
                                                       double t)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
	{
		const double d = t * (double)F[i];
		out[i] = (float)wrap_pi_d(d + (double)in[i]);
	}
}

// sum Re tr (i a i b) = -sum a b, in FP64
__global__ __launch_bounds__(256) void k_u1_dot(long n, const float *__restrict__ A, const float *__restrict__ B,
                                                double *partials, int nb)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	double v[1] = {i < n ? -((double)A[i] * (double)B[i]) : 0.0};
	block_sum_store<1>(v, partials, nb, blockIdx.x);
}

__global__ __launch_bounds__(256) void k_u1_axpy(long n, float *Y, float a, const float *__restrict__ X)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		Y[i] = fmaf(a, X[i], Y[i]);
}

int u1_lsum(mesh_ctx *ctx, const mesh_grid *g, const float *th, const float *ph, int mode, double *out)
{
	U1NGeo ng;
	if (int rc = u1n_setup(ctx, g, 1, &ng))
		return rc;
	if (!th || !out)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to the U(1) Landau sums");
	const int nb = (ng.V + 255) / 256;
	if (int rc = mesh_reserve_partials(ctx, nb))
		return rc;
	k_u1_lsum<<<nb, 256, 0, ctx->stream>>>(ng, th, ph, mode, ctx->partials, nb);
	MESH_LAUNCHED(ctx);
	double s;
	if (int rc = mesh_finish_reduction(ctx, nb, 1, &s))
		return rc;
	*out = mode == 0 ? s / ((double)ng.nd * ng.V) : s;
	return MESH_OK;
}

} // namespace

extern "C" {

int mesh_u1_landau_functional(mesh_ctx *ctx, const mesh_grid *g, const float *theta, const float *phi, double *out)
{
	MESH_CHECK_CTX(ctx);
	return u1_lsum(ctx, g, theta, phi, 0, out);
}

int mesh_u1_landau_divsum(mesh_ctx *ctx, const mesh_grid *g, const float *theta, const float *phi, double *out)
{
	MESH_CHECK_CTX(ctx);
	return u1_lsum(ctx, g, theta, phi, 1, out);
}

int mesh_u1_landau_deriv(mesh_ctx *ctx, const mesh_grid *g, const float *theta, const float *phi, float *F)
{
	MESH_CHECK_CTX(ctx);
	U1NGeo ng;
	if (int rc = u1n_setup(ctx, g, 1, &ng))
		return rc;
	if (!theta || !F || F == theta || F == phi)
		return mesh_fail(ctx, MESH_E_ARG, "bad u1_landau_deriv arguments");
	k_u1_lderiv<<<(ng.V + 255) / 256, 256, 0, ctx->stream>>>(ng, theta, phi, -0.5 / ((double)ng.nd * ng.V), F);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_u1_gauge_transform(mesh_ctx *ctx, const mesh_grid *g, float *theta_out, const float *theta, const float *phi)
{
	MESH_CHECK_CTX(ctx);
	U1NGeo ng;
	if (int rc = u1n_setup(ctx, g, 1, &ng))
		return rc;
	if (!theta_out || !theta || !phi || theta_out == theta || theta_out == phi)
		return mesh_fail(ctx, MESH_E_ARG, "bad u1_gauge_transform arguments");
	k_u1_gauge_transform<<<(ng.nd * ng.V + 255) / 256, 256, 0, ctx->stream>>>(ng, theta_out, theta, phi);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_u1_landau_or_halfsweep(mesh_ctx *ctx, const mesh_grid *g, float *theta, float *phi, int parity,
                                double omega)
{
	MESH_CHECK_CTX(ctx);
	U1NGeo ng;
	if (int rc = u1n_setup(ctx, g, 1, &ng))
		return rc;
	if (!theta || theta == phi || (parity != 0 && parity != 1))
		return mesh_fail(ctx, MESH_E_ARG, "bad u1_landau_or_halfsweep arguments");
	for (int i = 0; i < ng.nd; ++i)
		if (ng.L[i] % 2)
			return mesh_fail(ctx, MESH_E_SIZE, "checkerboard update needs even extents");
	int ph = mesh_prof_begin(ctx, PROF_LANDAU_OR);
	k_u1_or_half<<<(ng.V / 2 + 255) / 256, 256, 0, ctx->stream>>>(ng, theta, phi, parity, omega);
	mesh_prof_end(ctx, ph);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_u1_gauge_step(mesh_ctx *ctx, const mesh_grid *g, float *phi_out, const float *phi_in, const float *F,
                       double t)
{
	MESH_CHECK_CTX(ctx);
	U1NGeo ng;
	if (int rc = u1n_setup(ctx, g, 1, &ng))
		return rc;
	if (!phi_out || !phi_in || !F)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to u1_gauge_step");
	k_u1_gauge_step<<<(ng.V + 255) / 256, 256, 0, ctx->stream>>>(ng.V, phi_out, phi_in, F, t);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_u1_dot(mesh_ctx *ctx, long n, const float *A, const float *B, double *out)
{
	MESH_CHECK_CTX(ctx);
	if (n < 1 || n >= (1L << 31) || !A || !B || !out)
		return mesh_fail(ctx, MESH_E_ARG, "bad u1_dot arguments");
	const int nb = (int)((n + 255) / 256);
	if (int rc = mesh_reserve_partials(ctx, nb))
		return rc;
	k_u1_dot<<<nb, 256, 0, ctx->stream>>>(n, A, B, ctx->partials, nb);
	MESH_LAUNCHED(ctx);
	return mesh_finish_reduction(ctx, nb, 1, out);
}

int mesh_u1_axpy(mesh_ctx *ctx, long n, float *Y, double a, const float *X)
{
	MESH_CHECK_CTX(ctx);
	if (n < 1 || n >= (1L << 31) || !Y || !X)
		return mesh_fail(ctx, MESH_E_ARG, "bad u1_axpy arguments");
	k_u1_axpy<<<(int)((n + 255) / 256), 256, 0, ctx->stream>>>(n, Y, (float)a, X);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_u1_random(mesh_ctx *ctx, const mesh_grid *g, int batch, float *theta, uint64_t seed)
{
	MESH_CHECK_CTX(ctx);
	if (!u1_lds(ctx, g, 8192))
	{
		U1NGeo ng;
		if (int rc = u1n_setup(ctx, g, batch, &ng))
			return rc;
		if (!theta)
			return mesh_fail(ctx, MESH_E_ARG, "theta is NULL");
		long n = (long)batch * ng.nd * ng.V;
		k_u1n_random<<<(int)((n + 255) / 256), 256, 0, ctx->stream>>>(ng, n, theta, seed);
		MESH_LAUNCHED(ctx);
		return MESH_OK;
	}
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
	if (!u1_lds(ctx, g, 8192))
	{
		U1NGeo ng;
		if (int rc = u1n_setup(ctx, g, batch, &ng))
			return rc;
		if (!theta || !out_dev)
			return mesh_fail(ctx, MESH_E_ARG, "NULL argument to u1_observables");
		const int nbk = (ng.V + 255) / 256;
		if (int rc = mesh_reserve_partials(ctx, 3 * (size_t)batch * nbk))
			return rc;
		k_u1n_obs<<<batch * nbk, 256, 0, ctx->stream>>>(ng, nbk, theta, ctx->partials);
		k_u1n_obs_final<<<(batch + 255) / 256, 256, 0, ctx->stream>>>(batch, nbk, ng.nd, ctx->partials, out_dev);
		MESH_LAUNCHED(ctx);
		return MESH_OK;
	}
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
	if (!u1_lds(ctx, g, 4096))
	{
		U1NGeo ng;
		if (int rc = u1n_setup(ctx, g, batch, &ng))
			return rc;
		if (!theta || !deltas || !out_dev || ndeltas < 1 || ndeltas % 2 != 1)
			return mesh_fail(ctx, MESH_E_ARG, "bad u1_hmc arguments (deltas must have odd length)");
		return u1n_hmc(ctx, ng, batch, theta, deltas, ndeltas, beta, seed, traj, metropolis, out_dev);
	}
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
