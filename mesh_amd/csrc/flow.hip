// Gauge-equivariant coupling-layer flow for 2-d U(1), used as an
// independence proposal beside HMC (no counterpart in the reference; the
// U(1) conventions are those of u1.hip).
//
// Layer l (8 by default, any count up to 64): direction mu = l % 2, offset
// a = (l / 2) % 4, nu = 1 - mu (the pattern repeats every 8 layers).
//  * active plaquettes P(y) with y_nu = a (mod 4); each contains exactly one
//    updated link: U_0(y) for mu = 0 (bottom edge, sign +), U_1(y + 0) for
//    mu = 1 (right edge, sign +);
//  * frozen plaquettes: y_nu = a+1, a+2 (mu = 0) or a-1, a-2 (mu = 1) mod 4;
//    they contain no updated link;
//  * context of P(y): (cos, sin) of the frozen plaquettes y + i mu^ + j nu^,
//    j in {+1,+2} (mu = 0) / {-1,-2} (mu = 1), i in {-1,0,1}; feature
//    2 (3 jj + i + 1) + {0: cos, 1: sin};
//  * MLP 12 -> 64 -> 64 -> 10, tanh, outputs y: s_k = exp(clamp(y_k)),
//    w = softmax(y_{4..7}), t = atan2(y_9, y_8);
//  * h(theta) = g(wrap(theta - t)) + t,  g(u) = sum_k w_k 2 atan(s_k tan(u/2))
//    (mixture of non-compact projections: a circle diffeomorphism);
//  * updated link theta_l += h(theta_P) - theta_P;  log det += log g'(u).
// The map is triangular (the context never changes within a layer) and
// gauge-equivariant (it moves links by gauge-invariant amounts).
//
// Kernels: one workgroup (256 threads) per configuration, angles in LDS;
// the dense layers of the MLP run on FP32 MFMA (v_mfma_f32_16x16x4_f32) in
// 16-row tiles, one tile per wavefront.
#include "flow_mlp.h"
#include "mesh_internal.h"
#include "rng_dev.h"

using namespace mesh_dev;
using namespace flow_mlp;

namespace {

constexpr int MAX_LAYERS = 64; // layers: 8 by default, up to 64
constexpr int NI = 12;         // features
constexpr int NP_LAYER = Layout<NI>::NP; // 5642 parameters per layer
constexpr int THREADS = 256;
constexpr uint32_t TAG_FLOWZ = 0x464c575au; // prior samples
constexpr uint32_t TAG_FLOWI = 0x464c5749u; // initial weights

struct FGeo
{
	int L0, L1, V, NA;
};

__device__ __forceinline__ int site(FGeo const &g, int x0, int x1)
{
	x0 %= g.L0;
	x0 += x0 < 0 ? g.L0 : 0;
	x1 %= g.L1;
	x1 += x1 < 0 ? g.L1 : 0;
	return x0 * g.L1 + x1;
}

__device__ __forceinline__ float plaqf(FGeo const &g, const float *t, int x0, int x1)
{
	return t[site(g, x0, x1)] + t[g.V + site(g, x0 + 1, x1)] - t[site(g, x0, x1 + 1)] - t[g.V + site(g, x0, x1)];
}

// active plaquette r of layer (mu, a) -> its corner (y0, y1)
__device__ __forceinline__ void active_corner(FGeo const &g, int mu, int a, int r, int &y0, int &y1)
{
	if (mu == 0)
	{
		int m = r / g.L0;
		y0 = r - m * g.L0;
		y1 = a + 4 * m;
	}
	else
	{
		int m = r / g.L1;
		y1 = r - m * g.L1;
		y0 = a + 4 * m;
	}
}

// frozen neighbour f (0..5) of the active plaquette at (y0, y1)
__device__ __forceinline__ void frozen_corner(int mu, int f, int y0, int y1, int &z0, int &z1)
{
	int jj = f / 3, i = f % 3 - 1;
	if (mu == 0)
	{
		z0 = y0 + i;
		z1 = y1 + 1 + jj;
	}
	else
	{
		z1 = y1 + i;
		z0 = y0 - 1 - jj;
	}
}

__device__ __forceinline__ int updated_link(FGeo const &g, int mu, int y0, int y1)
{
	return mu == 0 ? site(g, y0, y1) : g.V + site(g, y0 + 1, y1);
}

using LWF = LW<NI>;
using TileF = Tile<NI>;

// features of the 16 active plaquettes r0..r0+15 (rows beyond NA are zero)
__device__ void features(FGeo const &g, const float *t, int mu, int a, int r0, TileF &T)
{
	int l = threadIdx.x & 63;
	for (int e = l; e < 16 * 16; e += 64)
	{
		int row = e / 16, col = e % 16;
		float v = 0.f;
		int r = r0 + row;
		if (r < g.NA && col < NI)
		{
			int y0, y1, z0, z1;
			active_corner(g, mu, a, r, y0, y1);
			frozen_corner(mu, col / 2, y0, y1, z0, z1);
			float p = plaqf(g, t, z0, z1);
			v = (col & 1) ? sinf(p) : cosf(p);
		}
		T.x[e] = v;
	}
}

// ---- kernels ------------------------------------------------------------------

struct FlowShared
{
	LWF lw;
	TileF tile[4];
	double red[4][4];
};

// MLP outputs of tiles it .. it+15 (four per wavefront) into Ys[tile][row*16+k]
__device__ void mlp_round(FGeo const &g, const float *th, int mu, int a, int it, int NT, FlowShared &sh,
                          float (*Ys)[256])
{
	const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
	TileF &T = sh.tile[w];
	for (int q = 0; q < 4; ++q)
	{
		const int tl = q * 4 + w, tix = it + tl;
		if (tix >= NT)
			break;
		features(g, th, mu, a, tix * 16, T);
		wave_sync();
		mlp_forward<NI>(sh.lw, T);
		for (int e = l; e < 256; e += 64)
			Ys[tl][e] = T.y()[e];
		wave_sync();
	}
}

// z ~ uniform -> U.  theta_layers (optional): input of every layer, for the
// backward pass.  out[c*2 + {0,1}] = {log q(U), S(U)/beta = sum(1 - cos)}.
__global__ __launch_bounds__(THREADS) void k_flow_forward(FGeo g, int nl, const float *__restrict__ params,
                                                          uint64_t seed, uint32_t step, float *theta_out,
                                                          float *theta_layers, double *out)
{
	extern __shared__ float th[]; // 2V
	__shared__ FlowShared sh;
	__shared__ float Ys[16][256];
	const int c = blockIdx.x, w = threadIdx.x >> 6, l = threadIdx.x & 63;
	for (int i = threadIdx.x; i < 2 * g.V; i += THREADS)
	{
		int mu = i >= g.V, x = i - mu * g.V;
		u4 ctr = {(uint32_t)x, ((uint32_t)c << 1) | (uint32_t)mu, step << 5, TAG_FLOWZ};
		u4 r = philox4x32_10(ctr, (uint32_t)seed, (uint32_t)(seed >> 32));
		double u = u01_from_bits(r.x, r.y) - 0x1p-53;
		th[i] = (float)(6.283185307179586477 * u - 3.141592653589793238);
	}
	double logdet = 0.0;
	const int NT = (g.NA + 15) / 16;
	for (int layer = 0; layer < nl; ++layer)
	{
		const int mu = layer & 1, a = (layer >> 1) & 3;
		__syncthreads();
		if (theta_layers)
			for (int i = threadIdx.x; i < 2 * g.V; i += THREADS)
				theta_layers[((long)c * nl + layer) * 2 * g.V + i] = th[i];
		load_weights<NI>(sh.lw, params + layer * NP_LAYER);
		__syncthreads();
		// within a layer the MLP inputs (frozen plaquettes) never change, so
		// all MLPs of a round run first and then every thread moves one link
		for (int it = 0; it < NT; it += 16)
		{
			mlp_round(g, th, mu, a, it, NT, sh, Ys);
			__syncthreads();
			const int r = it * 16 + threadIdx.x;
			if (r < g.NA)
			{
				int y0, y1;
				active_corner(g, mu, a, r, y0, y1);
				NCP p = ncp_params(&Ys[threadIdx.x >> 4][(threadIdx.x & 15) * 16]);
				float tp = plaqf(g, th, y0, y1);
				float u = wrapf(tp - p.t), G, Gp;
				ncp_eval(p, u, G, Gp);
				int ul = updated_link(g, mu, y0, y1);
				th[ul] = wrapf(th[ul] + (G + p.t - tp));
				logdet += (double)logf(Gp);
			}
			__syncthreads();
		}
	}
	__syncthreads();
	double v[2] = {logdet, 0.0};
	for (int x = threadIdx.x; x < g.V; x += THREADS)
		v[1] += 1.0 - cos((double)plaqf(g, th, x / g.L1, x % g.L1));
	for (int q = 0; q < 2; ++q)
	{
		double s = v[q];
		for (int o = 32; o >= 1; o >>= 1)
			s += __shfl_xor(s, o, 64);
		if (l == 0)
			sh.red[q][w] = s;
	}
	__syncthreads();
	if (threadIdx.x == 0)
	{
		double ld = ((sh.red[0][0] + sh.red[0][1]) + sh.red[0][2]) + sh.red[0][3];
		double s = ((sh.red[1][0] + sh.red[1][1]) + sh.red[1][2]) + sh.red[1][3];
		out[2 * c + 0] = -2.0 * g.V * 1.8378770664093454836 /* log 2 pi */ - ld;
		out[2 * c + 1] = s;
	}
	for (int i = threadIdx.x; i < 2 * g.V; i += THREADS)
		theta_out[(long)c * 2 * g.V + i] = th[i];
}

// U -> z and log q(U), layers in reverse with g^-1 by safeguarded Newton.
// out[c*2 + {0,1}] = {log q(U), S(U)/beta}
__global__ __launch_bounds__(THREADS) void k_flow_inverse(FGeo g, int nl, const float *__restrict__ params,
                                                          const float *__restrict__ theta_in, float *z_out, double *out)
{
	extern __shared__ float th[];
	__shared__ FlowShared sh;
	__shared__ float Ys[16][256];
	const int c = blockIdx.x, w = threadIdx.x >> 6, l = threadIdx.x & 63;
	for (int i = threadIdx.x; i < 2 * g.V; i += THREADS)
		th[i] = theta_in[(long)c * 2 * g.V + i];
	double v[2] = {0.0, 0.0};
	__syncthreads();
	for (int x = threadIdx.x; x < g.V; x += THREADS)
		v[1] += 1.0 - cos((double)plaqf(g, th, x / g.L1, x % g.L1));
	const int NT = (g.NA + 15) / 16;
	for (int layer = nl - 1; layer >= 0; --layer)
	{
		const int mu = layer & 1, a = (layer >> 1) & 3;
		__syncthreads();
		load_weights<NI>(sh.lw, params + layer * NP_LAYER);
		__syncthreads();
		for (int it = 0; it < NT; it += 16)
		{
			mlp_round(g, th, mu, a, it, NT, sh, Ys);
			__syncthreads();
			const int r = it * 16 + threadIdx.x;
			if (r < g.NA)
			{
				int y0, y1;
				active_corner(g, mu, a, r, y0, y1);
				NCP p = ncp_params(&Ys[threadIdx.x >> 4][(threadIdx.x & 15) * 16]);
				float tpn = plaqf(g, th, y0, y1);
				float Gp;
				float u = ncp_inverse(p, wrapf(tpn - p.t), Gp);
				int ul = updated_link(g, mu, y0, y1);
				th[ul] = wrapf(th[ul] - tpn + u + p.t);
				v[0] += (double)logf(Gp);
			}
			__syncthreads();
		}
	}
	__syncthreads();
	for (int q = 0; q < 2; ++q)
	{
		double s = v[q];
		for (int o = 32; o >= 1; o >>= 1)
			s += __shfl_xor(s, o, 64);
		if (l == 0)
			sh.red[q][w] = s;
	}
	__syncthreads();
	if (threadIdx.x == 0)
	{
		double ld = ((sh.red[0][0] + sh.red[0][1]) + sh.red[0][2]) + sh.red[0][3];
		double s = ((sh.red[1][0] + sh.red[1][1]) + sh.red[1][2]) + sh.red[1][3];
		out[2 * c + 0] = -2.0 * g.V * 1.8378770664093454836 - ld;
		out[2 * c + 1] = s;
	}
	if (z_out)
		for (int i = threadIdx.x; i < 2 * g.V; i += THREADS)
			z_out[(long)c * 2 * g.V + i] = th[i];
}

// gradient of the per-configuration loss  beta sum_x (1 - cos theta_P) - log det
// with respect to every weight; partial[c][layer * NP_LAYER + i].
// theta_layers: input of each layer (from k_flow_forward); theta_final: output;
// dxs: scratch of NA * NI floats per configuration.
__global__ __launch_bounds__(THREADS) void k_flow_backward(FGeo g, int nl, const float *__restrict__ params,
                                                           const float *__restrict__ theta_layers,
                                                           const float *__restrict__ theta_final, float beta,
                                                           float *__restrict__ dxs, float *__restrict__ partial)
{
	extern __shared__ float dyn[];
	float *th = dyn;                // 2V: input of the current layer
	float *ab = dyn + 2 * g.V;      // 2V: adjoint w.r.t. the current layer's output
	float *gp = dyn + 4 * g.V;      // V: plaquette gradients of the layer
	float *gact = dyn + 5 * g.V;    // NA: dL/dtheta_P of the active plaquettes
	__shared__ FlowShared sh;
	const int c = blockIdx.x, w = threadIdx.x >> 6, l = threadIdx.x & 63;
	float *dx = dxs + (long)c * g.NA * NI;
	float *part = partial + (long)c * nl * NP_LAYER;
	// adjoint of the final links: dS/dtheta
	for (int i = threadIdx.x; i < 2 * g.V; i += THREADS)
		th[i] = theta_final[(long)c * 2 * g.V + i];
	__syncthreads();
	for (int x = threadIdx.x; x < g.V; x += THREADS)
		gp[x] = sinf(plaqf(g, th, x / g.L1, x % g.L1));
	__syncthreads();
	for (int x = threadIdx.x; x < g.V; x += THREADS)
	{
		int x0 = x / g.L1, x1 = x % g.L1;
		ab[x] = beta * (gp[x] - gp[site(g, x0, x1 - 1)]);
		ab[g.V + x] = beta * (gp[site(g, x0 - 1, x1)] - gp[x]);
	}
	const int NT = (g.NA + 15) / 16;
	for (int layer = nl - 1; layer >= 0; --layer)
	{
		const int mu = layer & 1, a = (layer >> 1) & 3;
		__syncthreads();
		for (int i = threadIdx.x; i < 2 * g.V; i += THREADS)
			th[i] = theta_layers[((long)c * nl + layer) * 2 * g.V + i];
		load_weights<NI>(sh.lw, params + layer * NP_LAYER);
		__syncthreads();
		Grad<NI> gr;
		gr.zero();
		for (int it = 0; it < (NT + 3) / 4; ++it)
		{
			const int tix = it * 4 + w;
			TileF &T = sh.tile[w];
			if (tix < NT)
			{
				features(g, th, mu, a, tix * 16, T);
				wave_sync();
				mlp_forward<NI>(sh.lw, T);
				// output adjoints of the 16 rows (lanes 0..15 own one row each):
				// the updated link moves by g(u) + t - theta_P, so its adjoint A
				// reaches theta_P as A (g' - 1) plus the log-det term
				if (l < 16)
				{
					float dy[16];
					for (int k = 0; k < 16; ++k)
						dy[k] = 0.f;
					int r = tix * 16 + l;
					if (r < g.NA)
					{
						int y0, y1;
						active_corner(g, mu, a, r, y0, y1);
						const float *y = T.y() + l * 16;
						float tp = plaqf(g, th, y0, y1);
						float A = ab[updated_link(g, mu, y0, y1)];
						gact[r] = ncp_adjoint(y, wrapf(tp - ncp_params(y).t), A, dy) - A;
					}
					for (int k = 0; k < 16; ++k)
						T.y()[l * 16 + k] = dy[k];
				}
				wave_sync();
				mlp_backward<NI>(sh.lw, T, gr, dx, tix * 16, g.NA);
			}
			__syncthreads();
		}
		// per-configuration weight gradients of this layer, waves in order
		store_grad<NI>(gr, part + layer * NP_LAYER);
		// plaquette gradients: active from gact, frozen from the feature adjoints
		for (int x = threadIdx.x; x < g.V; x += THREADS)
		{
			int x0 = x / g.L1, x1 = x % g.L1;
			int cn = mu == 0 ? x1 : x0;
			int rel = ((cn - a) % 4 + 4) % 4;
			float v = 0.f;
			if (rel == 0)
			{
				int m = (cn - a) / 4;
				v = gact[mu == 0 ? m * g.L0 + x0 : m * g.L1 + x1];
			}
			else if ((mu == 0 && (rel == 1 || rel == 2)) || (mu == 1 && (rel == 3 || rel == 2)))
			{
				int jj = mu == 0 ? rel - 1 : (rel == 3 ? 0 : 1);
				float pa = plaqf(g, th, x0, x1);
				float cs = cosf(pa), sn = sinf(pa);
				for (int i = -1; i <= 1; ++i)
				{
					int y0, y1;
					if (mu == 0)
					{
						y0 = x0 - i;
						y1 = x1 - 1 - jj;
					}
					else
					{
						y1 = x1 - i;
						y0 = x0 + 1 + jj;
					}
					y0 = ((y0 % g.L0) + g.L0) % g.L0;
					y1 = ((y1 % g.L1) + g.L1) % g.L1;
					int m = mu == 0 ? (y1 - a) / 4 : (y0 - a) / 4;
					int r = mu == 0 ? m * g.L0 + y0 : m * g.L1 + y1;
					int f = jj * 3 + (i + 1);
					v = fmaf(-sn, dx[(long)r * NI + 2 * f], v);
					v = fmaf(cs, dx[(long)r * NI + 2 * f + 1], v);
				}
			}
			gp[x] = v;
		}
		__syncthreads();
		for (int x = threadIdx.x; x < g.V; x += THREADS)
		{
			int x0 = x / g.L1, x1 = x % g.L1;
			ab[x] += gp[x] - gp[site(g, x0, x1 - 1)];
			ab[g.V + x] += gp[site(g, x0 - 1, x1)] - gp[x];
		}
	}
}

// grad[i] = scale * sum_c partial[c][i] (fixed order)
__global__ void k_flow_reduce(int n, const float *__restrict__ partial, int batch, float scale,
                              float *__restrict__ grad)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	double s = 0.0;
	for (int c = 0; c < batch; ++c)
		s += partial[(long)c * n + i];
	grad[i] = (float)(s * scale);
}

// Adam (Kingma & Ba 2015)
__global__ void k_adam(int n, float *__restrict__ p, const float *__restrict__ gr, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	float gi = gr[i];
	float mi = fmaf(b1, m[i], (1.f - b1) * gi);
	float vi = fmaf(b2, v[i], (1.f - b2) * gi * gi);
	m[i] = mi;
	v[i] = vi;
	p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// initial weights (flow_mlp::init_param: near the identity transform)
__global__ void k_flow_init(int n, float *p, uint64_t seed)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		init_param<NI>(i, p, seed, TAG_FLOWI);
}

// Metropolis-Hastings with an independence proposal, per chain c:
// accept U' with probability min(1, exp[(-beta S' - log q') - (-beta S - log q)]).
// state[c*2 + {0,1}] = {log q, S/beta} of the current U (updated on acceptance),
// prop[c*2 + {0,1}] the same for the proposal; acc[c] = 1 if accepted.
__global__ __launch_bounds__(THREADS) void k_flow_mh(FGeo g, float *theta, double *state, const float *prop_theta,
                                                     const double *prop, double beta, uint64_t seed, uint32_t step,
                                                     double *acc)
{
	__shared__ int a_s;
	const int c = blockIdx.x;
	if (threadIdx.x == 0)
	{
		double lw_new = -beta * prop[2 * c + 1] - prop[2 * c + 0];
		double lw_old = -beta * state[2 * c + 1] - state[2 * c + 0];
		u4 ctr = {step, (uint32_t)c, 1u, 0x464c4d48u /* MH */};
		u4 r = philox4x32_10(ctr, (uint32_t)seed, (uint32_t)(seed >> 32));
		double u = u01_from_bits(r.x, r.y);
		int a = u < exp(lw_new - lw_old);
		a_s = a;
		acc[c] = a;
		if (a)
		{
			state[2 * c + 0] = prop[2 * c + 0];
			state[2 * c + 1] = prop[2 * c + 1];
		}
	}
	__syncthreads();
	if (a_s)
		for (int i = threadIdx.x; i < 2 * g.V; i += THREADS)
			theta[(long)c * 2 * g.V + i] = prop_theta[(long)c * 2 * g.V + i];
}

int flow_setup(mesh_ctx *ctx, const mesh_grid *g, int batch, int layers, FGeo *fg)
{
	if (layers < 1 || layers > MAX_LAYERS)
		return mesh_fail(ctx, MESH_E_ARG, "the flow has 1 to 64 layers");
	if (!g || g->nd != 2 || g->dims[0] % 4 || g->dims[1] % 4 || g->dims[0] < 4 || g->dims[1] < 4)
		return mesh_fail(ctx, MESH_E_SIZE, "the flow needs a 2-d lattice with extents divisible by 4");
	long V = (long)g->dims[0] * g->dims[1];
	if (V > 4096)
		return mesh_fail(ctx, MESH_E_SIZE, "the flow kernels hold a configuration in LDS: volume <= 4096");
	if (batch < 1)
		return mesh_fail(ctx, MESH_E_ARG, "batch must be >= 1");
	fg->L0 = g->dims[0];
	fg->L1 = g->dims[1];
	fg->V = (int)V;
	fg->NA = (int)(V / 4);
	return MESH_OK;
}

} // namespace

extern "C" {

int mesh_flow_num_params(int layers) { return layers >= 1 && layers <= MAX_LAYERS ? layers * NP_LAYER : 0; }

int mesh_flow_init(mesh_ctx *ctx, int layers, float *params, uint64_t seed)
{
	MESH_CHECK_CTX(ctx);
	if (!params || layers < 1 || layers > MAX_LAYERS)
		return mesh_fail(ctx, MESH_E_ARG, "bad flow_init arguments");
	int n = layers * NP_LAYER;
	k_flow_init<<<(n + 255) / 256, 256, 0, ctx->stream>>>(n, params, seed);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_flow_sample(mesh_ctx *ctx, const mesh_grid *g, int batch, int layers, const float *params, uint64_t seed,
                     uint32_t step, float *theta_out, float *theta_layers, double *out_dev)
{
	MESH_CHECK_CTX(ctx);
	FGeo fg;
	if (int rc = flow_setup(ctx, g, batch, layers, &fg))
		return rc;
	if (!params || !theta_out || !out_dev)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to flow_sample");
	MESH_HIP(ctx, hipFuncSetAttribute((const void *)k_flow_forward, hipFuncAttributeMaxDynamicSharedMemorySize,
	                                  (int)(2 * fg.V * sizeof(float))));
	k_flow_forward<<<batch, THREADS, 2 * fg.V * sizeof(float), ctx->stream>>>(fg, layers, params, seed, step,
	                                                                          theta_out, theta_layers, out_dev);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_flow_logq(mesh_ctx *ctx, const mesh_grid *g, int batch, int layers, const float *params, const float *theta,
                   float *z_out, double *out_dev)
{
	MESH_CHECK_CTX(ctx);
	FGeo fg;
	if (int rc = flow_setup(ctx, g, batch, layers, &fg))
		return rc;
	if (!params || !theta || !out_dev)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to flow_logq");
	MESH_HIP(ctx, hipFuncSetAttribute((const void *)k_flow_inverse, hipFuncAttributeMaxDynamicSharedMemorySize,
	                                  (int)(2 * fg.V * sizeof(float))));
	k_flow_inverse<<<batch, THREADS, 2 * fg.V * sizeof(float), ctx->stream>>>(fg, layers, params, theta, z_out,
	                                                                          out_dev);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

size_t mesh_flow_workspace_bytes(const mesh_grid *g, int batch, int layers)
{
	FGeo fg;
	if (flow_setup(nullptr, g, batch, layers, &fg))
		return 0;
	size_t n = (size_t)batch * layers * 2 * fg.V       // theta per layer
	           + (size_t)batch * fg.NA * NI            // feature adjoints
	           + (size_t)batch * layers * NP_LAYER     // per-configuration gradients
	           + (size_t)batch * 2 * fg.V;         // samples
	return n * sizeof(float) + (size_t)batch * 2 * sizeof(double);
}

int mesh_flow_gradient(mesh_ctx *ctx, const mesh_grid *g, int batch, int layers, const float *params, uint64_t seed,
                       uint32_t step, double beta, void *workspace, float *grad, double *out_dev)
{
	MESH_CHECK_CTX(ctx);
	FGeo fg;
	if (int rc = flow_setup(ctx, g, batch, layers, &fg))
		return rc;
	if (!params || !workspace || !grad || !out_dev)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to flow_gradient");
	float *tl = (float *)workspace;
	float *dxs = tl + (size_t)batch * layers * 2 * fg.V;
	float *part = dxs + (size_t)batch * fg.NA * NI;
	float *samp = part + (size_t)batch * layers * NP_LAYER;
	MESH_HIP(ctx, hipFuncSetAttribute((const void *)k_flow_forward, hipFuncAttributeMaxDynamicSharedMemorySize,
	                                  (int)(2 * fg.V * sizeof(float))));
	k_flow_forward<<<batch, THREADS, 2 * fg.V * sizeof(float), ctx->stream>>>(fg, layers, params, seed, step, samp,
	                                                                          tl, out_dev);
	MESH_LAUNCHED(ctx);
	size_t lds = (5 * (size_t)fg.V + fg.NA) * sizeof(float);
	MESH_HIP(ctx, hipFuncSetAttribute((const void *)k_flow_backward, hipFuncAttributeMaxDynamicSharedMemorySize,
	                                  (int)lds));
	k_flow_backward<<<batch, THREADS, lds, ctx->stream>>>(fg, layers, params, tl, samp, (float)beta, dxs, part);
	MESH_LAUNCHED(ctx);
	int n = layers * NP_LAYER;
	k_flow_reduce<<<(n + 255) / 256, 256, 0, ctx->stream>>>(n, part, batch, 1.0f / batch, grad);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_flow_mh(mesh_ctx *ctx, const mesh_grid *g, int batch, float *theta, double *state, const float *prop_theta,
                 const double *prop, double beta, uint64_t seed, uint32_t step, double *acc_dev)
{
	MESH_CHECK_CTX(ctx);
	// any 2-d lattice: the step only copies configurations (proposals of
	// either flow)
	if (!g || g->nd != 2 || g->dims[0] < 2 || g->dims[1] < 2 || batch < 1)
		return mesh_fail(ctx, MESH_E_SIZE, "flow_mh needs a 2-d lattice and batch >= 1");
	FGeo fg;
	fg.L0 = g->dims[0];
	fg.L1 = g->dims[1];
	fg.V = g->dims[0] * g->dims[1];
	fg.NA = fg.V / 4;
	if (!theta || !state || !prop_theta || !prop || !acc_dev)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to flow_mh");
	k_flow_mh<<<batch, THREADS, 0, ctx->stream>>>(fg, theta, state, prop_theta, prop, beta, seed, step, acc_dev);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_flow_adam(mesh_ctx *ctx, int layers, float *params, const float *grad, float *m, float *v, int t, double lr)
{
	MESH_CHECK_CTX(ctx);
	if (!params || !grad || !m || !v || t < 1 || layers < 1 || layers > MAX_LAYERS)
		return mesh_fail(ctx, MESH_E_ARG, "bad flow_adam arguments");
	return mesh_launch_adam(ctx, layers * NP_LAYER, params, grad, m, v, t, lr);
}

} // extern "C"

int mesh_launch_adam(mesh_ctx *ctx, int n, float *params, const float *grad, float *m, float *v, int t, double lr)
{
	const double b1 = 0.9, b2 = 0.999;
	float c1 = (float)(1.0 / (1.0 - std::pow(b1, t))), c2 = (float)(1.0 / (1.0 - std::pow(b2, t)));
	k_adam<<<(n + 255) / 256, 256, 0, ctx->stream>>>(n, params, grad, m, v, (float)lr, (float)b1, (float)b2, 1e-8f,
	                                                 c1, c2);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}
