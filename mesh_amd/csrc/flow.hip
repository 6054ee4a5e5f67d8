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
#include "mesh_internal.h"
#include "rng_dev.h"

using namespace mesh_dev;

namespace {

constexpr int MAX_LAYERS = 64;                          // layers: 8 by default, up to 64
constexpr int NK = 4;                                   // NCP components
constexpr int NI = 12;                                  // features
constexpr int NH = 64;                                  // hidden width
constexpr int NO = 2 * NK + 2;                          // outputs
constexpr int OW1 = 0, OB1 = OW1 + NI * NH, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
              OB3 = OW3 + NH * NO, NP_LAYER = OB3 + NO; // 5642 parameters per layer
constexpr int THREADS = 256;
constexpr uint32_t TAG_FLOWZ = 0x464c575au;  // prior samples
constexpr uint32_t TAG_FLOWI = 0x464c5749u;  // initial weights
constexpr float PI_F = 3.14159265358979323846f;
constexpr float TWO_PI_F = 6.28318530717958647692f;

typedef float f4 __attribute__((ext_vector_type(4)));

// orders the LDS accesses of one wavefront (no cross-wavefront wait)
__device__ __forceinline__ void wave_sync()
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct FGeo
{
	int L0, L1, V, NA;
};

__device__ __forceinline__ float wrapf(float a)
{
	a = a - TWO_PI_F * floorf((a + PI_F) * (1.0f / TWO_PI_F));
	return a >= PI_F ? a - TWO_PI_F : (a < -PI_F ? a + TWO_PI_F : a);
}

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

// ---- MFMA tile product -----------------------------------------------------
// D(16x16) = sum_k A(i,k) B(k,n), A(i,k) = A[i*sai + k*sak], B(k,n) = B[k*sbk + n*sbn],
// k < K (multiple of 4).  Lane l supplies A(l%16, k0 + l/16) and
// B(k0 + l/16, l%16); it receives D(4*(l/16) + j, l%16), j < 4.
__device__ __forceinline__ f4 tile_mm(const float *A, int sai, int sak, const float *B, int sbk, int sbn, int K)
{
	int l = threadIdx.x & 63, i = l & 15, kk = l >> 4;
	f4 acc = {0.f, 0.f, 0.f, 0.f};
	for (int k0 = 0; k0 < K; k0 += 4)
		acc = __builtin_amdgcn_mfma_f32_16x16x4f32(A[i * sai + (k0 + kk) * sak], B[(k0 + kk) * sbk + i * sbn], acc,
		                                           0, 0, 0);
	return acc;
}

// LDS image of one layer's weights (W3/b3 padded to 16 columns)
struct LW
{
	float w1[NI * NH], b1[NH], w2[NH * NH], b2[NH], w3[NH * 16], b3[16];
};

__device__ void load_weights(LW &lw, const float *__restrict__ p)
{
	for (int i = threadIdx.x; i < NI * NH; i += THREADS)
		lw.w1[i] = p[OW1 + i];
	for (int i = threadIdx.x; i < NH * NH; i += THREADS)
		lw.w2[i] = p[OW2 + i];
	for (int i = threadIdx.x; i < NH * 16; i += THREADS)
	{
		int k = i / 16, n = i % 16;
		lw.w3[i] = n < NO ? p[OW3 + k * NO + n] : 0.f;
	}
	for (int i = threadIdx.x; i < NH; i += THREADS)
	{
		lw.b1[i] = p[OB1 + i];
		lw.b2[i] = p[OB2 + i];
	}
	for (int i = threadIdx.x; i < 16; i += THREADS)
		lw.b3[i] = i < NO ? p[OB3 + i] : 0.f;
}

// per-wavefront tile scratch: features, two hidden layers, outputs
struct Tile
{
	float x[16 * 16], h1[16 * NH], h2[16 * NH], y[16 * 16];
};

// features of the 16 active plaquettes r0..r0+15 (rows beyond NA are zero)
__device__ void features(FGeo const &g, const float *t, int mu, int a, int r0, Tile &T)
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

// MLP forward of one tile: h1 = tanh(x W1 + b1), h2 = tanh(h1 W2 + b2), y = h2 W3 + b3
__device__ void mlp_forward(LW const &lw, Tile &T)
{
	int l = threadIdx.x & 63, c = l & 15, rq = (l >> 4) * 4;
	for (int n0 = 0; n0 < NH; n0 += 16)
	{
		f4 d = tile_mm(T.x, 16, 1, lw.w1 + n0, NH, 1, NI);
		for (int j = 0; j < 4; ++j)
			T.h1[(rq + j) * NH + n0 + c] = tanhf(d[j] + lw.b1[n0 + c]);
	}
	wave_sync();
	for (int n0 = 0; n0 < NH; n0 += 16)
	{
		f4 d = tile_mm(T.h1, NH, 1, lw.w2 + n0, NH, 1, NH);
		for (int j = 0; j < 4; ++j)
			T.h2[(rq + j) * NH + n0 + c] = tanhf(d[j] + lw.b2[n0 + c]);
	}
	wave_sync();
	f4 d = tile_mm(T.h2, NH, 1, lw.w3, 16, 1, NH);
	for (int j = 0; j < 4; ++j)
		T.y[(rq + j) * 16 + c] = d[j] + lw.b3[c];
	wave_sync();
}

// transform parameters from the MLP outputs
struct NCP
{
	float s[NK], w[NK], t, ta, tb;
};

__device__ __forceinline__ NCP ncp_params(const float *y)
{
	NCP p;
	float m = y[NK];
	for (int k = 1; k < NK; ++k)
		m = fmaxf(m, y[NK + k]);
	float z = 0.f;
	for (int k = 0; k < NK; ++k)
	{
		p.s[k] = expf(fminf(fmaxf(y[k], -8.f), 8.f));
		p.w[k] = expf(y[NK + k] - m);
		z += p.w[k];
	}
	for (int k = 0; k < NK; ++k)
		p.w[k] /= z;
	p.ta = y[2 * NK];
	p.tb = y[2 * NK + 1];
	p.t = atan2f(p.tb, p.ta);
	return p;
}

// g(u), g'(u) of the mixture
__device__ __forceinline__ void ncp_eval(NCP const &p, float u, float &G, float &Gp)
{
	float c = cosf(0.5f * u), sn = sinf(0.5f * u);
	G = 0.f;
	Gp = 0.f;
	for (int k = 0; k < NK; ++k)
	{
		float d = fmaf(c, c, p.s[k] * p.s[k] * sn * sn);
		G = fmaf(p.w[k], 2.f * atan2f(p.s[k] * sn, c), G);
		Gp = fmaf(p.w[k], p.s[k] / d, Gp);
	}
}

// ---- kernels ------------------------------------------------------------------

struct FlowShared
{
	LW lw;
	Tile tile[4];
	double red[4][4];
};

// MLP outputs of tiles it .. it+15 (four per wavefront) into Ys[tile][row*16+k]
__device__ void mlp_round(FGeo const &g, const float *th, int mu, int a, int it, int NT, FlowShared &sh,
                          float (*Ys)[256])
{
	const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
	Tile &T = sh.tile[w];
	for (int q = 0; q < 4; ++q)
	{
		const int tl = q * 4 + w, tix = it + tl;
		if (tix >= NT)
			break;
		features(g, th, mu, a, tix * 16, T);
		wave_sync();
		mlp_forward(sh.lw, T);
		for (int e = l; e < 256; e += 64)
			Ys[tl][e] = T.y[e];
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
		load_weights(sh.lw, params + layer * NP_LAYER);
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
		load_weights(sh.lw, params + layer * NP_LAYER);
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
				float target = wrapf(tpn - p.t);
				// g is increasing on [-pi, pi] with g(+-pi) = +-pi: Newton steps
				// inside a bracket that every evaluation narrows, bisection when a
				// step leaves it, until a step or the bracket is below
				// 2e-7 max(1, |u|) (an FP32 ulp of u is up to 2.4e-7 here, so an
				// absolute bound could not stop an iterate that alternates
				// between neighbouring floats; at most 40 evaluations)
				float lo = -PI_F, hi = PI_F, u = target, G, Gp;
				for (int b = 0; b < 40; ++b)
				{
					ncp_eval(p, u, G, Gp);
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
				ncp_eval(p, u, G, Gp);
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
	const int lc = l & 15, rq = (l >> 4) * 4;
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
		load_weights(sh.lw, params + layer * NP_LAYER);
		__syncthreads();
		f4 aw1[4], aw2[16], aw3[4];
		for (int q = 0; q < 4; ++q)
			aw1[q] = aw3[q] = f4{0.f, 0.f, 0.f, 0.f};
		for (int q = 0; q < 16; ++q)
			aw2[q] = f4{0.f, 0.f, 0.f, 0.f};
		float ab1 = 0.f, ab2 = 0.f, ab3 = 0.f; // bias gradients, lane = column
		for (int it = 0; it < (NT + 3) / 4; ++it)
		{
			const int tix = it * 4 + w;
			Tile &T = sh.tile[w];
			const bool on = tix < NT;
			if (on)
			{
				features(g, th, mu, a, tix * 16, T);
				wave_sync();
				mlp_forward(sh.lw, T);
				// output adjoints of the 16 rows (lanes 0..15 own one row each)
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
						const float *y = T.y + l * 16;
						NCP p = ncp_params(y);
						float tp = plaqf(g, th, y0, y1);
						float u = wrapf(tp - p.t);
						float cc = cosf(0.5f * u), sn = sinf(0.5f * u);
						float G = 0.f, Gp = 0.f, Gpp = 0.f, gk[NK], gpk[NK], dgs[NK], dgps[NK];
						for (int k = 0; k < NK; ++k)
						{
							float s2 = p.s[k] * p.s[k];
							float d = fmaf(cc, cc, s2 * sn * sn);
							gk[k] = 2.f * atan2f(p.s[k] * sn, cc);
							gpk[k] = p.s[k] / d;
							dgs[k] = 2.f * sn * cc / d;
							dgps[k] = (cc * cc - s2 * sn * sn) / (d * d);
							G = fmaf(p.w[k], gk[k], G);
							Gp = fmaf(p.w[k], gpk[k], Gp);
							Gpp = fmaf(p.w[k], -p.s[k] * (s2 - 1.f) * sn * cc / (d * d), Gpp);
						}
						float A = ab[updated_link(g, mu, y0, y1)];
						float iG = 1.f / Gp;
						gact[r] = A * (Gp - 1.f) - Gpp * iG;
						float dLdt = A * (1.f - Gp) + Gpp * iG;
						float dLdw[NK], sw = 0.f;
						for (int k = 0; k < NK; ++k)
						{
							dLdw[k] = A * gk[k] - gpk[k] * iG;
							sw = fmaf(p.w[k], dLdw[k], sw);
							float dLds = p.w[k] * (A * dgs[k] - dgps[k] * iG);
							dy[k] = (y[k] > -8.f && y[k] < 8.f) ? p.s[k] * dLds : 0.f;
						}
						for (int k = 0; k < NK; ++k)
							dy[NK + k] = p.w[k] * (dLdw[k] - sw);
						float r2 = fmaf(p.ta, p.ta, p.tb * p.tb) + 1e-20f;
						dy[2 * NK] = dLdt * (-p.tb / r2);
						dy[2 * NK + 1] = dLdt * (p.ta / r2);
					}
					for (int k = 0; k < 16; ++k)
						T.y[l * 16 + k] = dy[k];
				}
				wave_sync();
				// dW3 += H2^T dY, db3
				for (int i0 = 0; i0 < NH; i0 += 16)
				{
					f4 d = tile_mm(T.h2 + i0, 1, NH, T.y, 16, 1, 16);
					aw3[i0 / 16] += d;
				}
				if (l < 16)
					for (int rr = 0; rr < 16; ++rr)
						ab3 += T.y[rr * 16 + l];
				// dA2 = (dY W3^T) * (1 - h2^2), into h2
				f4 dh2[4];
				for (int n0 = 0; n0 < NH; n0 += 16)
					dh2[n0 / 16] = tile_mm(T.y, 16, 1, sh.lw.w3 + n0 * 16, 1, 16, 16);
				wave_sync();
				for (int n0 = 0; n0 < NH; n0 += 16)
					for (int j = 0; j < 4; ++j)
					{
						float hv = T.h2[(rq + j) * NH + n0 + lc];
						T.h2[(rq + j) * NH + n0 + lc] = dh2[n0 / 16][j] * (1.f - hv * hv);
					}
				wave_sync();
				// dW2 += H1^T dA2, db2
				for (int i0 = 0; i0 < NH; i0 += 16)
					for (int n0 = 0; n0 < NH; n0 += 16)
						aw2[(i0 / 16) * 4 + n0 / 16] += tile_mm(T.h1 + i0, 1, NH, T.h2 + n0, NH, 1, 16);
				for (int rr = 0; rr < 16; ++rr)
					ab2 += T.h2[rr * NH + l];
				// dA1 = (dA2 W2^T) * (1 - h1^2), into h1
				f4 dh1[4];
				for (int n0 = 0; n0 < NH; n0 += 16)
					dh1[n0 / 16] = tile_mm(T.h2, NH, 1, sh.lw.w2 + n0 * NH, 1, NH, NH);
				wave_sync();
				for (int n0 = 0; n0 < NH; n0 += 16)
					for (int j = 0; j < 4; ++j)
					{
						float hv = T.h1[(rq + j) * NH + n0 + lc];
						T.h1[(rq + j) * NH + n0 + lc] = dh1[n0 / 16][j] * (1.f - hv * hv);
					}
				wave_sync();
				// dW1 += X^T dA1, db1
				for (int n0 = 0; n0 < NH; n0 += 16)
					aw1[n0 / 16] += tile_mm(T.x, 1, 16, T.h1 + n0, NH, 1, 16);
				for (int rr = 0; rr < 16; ++rr)
					ab1 += T.h1[rr * NH + l];
				// dX = dA1 W1^T (feature columns 0..11)
				f4 dxt = tile_mm(T.h1, NH, 1, sh.lw.w1, 1, NH, NH);
				for (int j = 0; j < 4; ++j)
				{
					int r = tix * 16 + rq + j;
					if (lc < NI && r < g.NA)
						dx[(long)r * NI + lc] = dxt[j];
				}
				wave_sync();
			}
			__syncthreads();
		}
		// per-configuration weight gradients of this layer, waves in order
		float *pl = part + layer * NP_LAYER;
		for (int ww = 0; ww < 4; ++ww)
		{
			if (w == ww)
			{
				for (int n0 = 0; n0 < NH; n0 += 16)
					for (int j = 0; j < 4; ++j)
					{
						int k = rq + j, n = n0 + lc;
						if (k < NI)
						{
							float *dst = pl + OW1 + k * NH + n;
							*dst = (ww ? *dst : 0.f) + aw1[n0 / 16][j];
						}
					}
				for (int i0 = 0; i0 < NH; i0 += 16)
					for (int n0 = 0; n0 < NH; n0 += 16)
						for (int j = 0; j < 4; ++j)
						{
							float *dst = pl + OW2 + (i0 + rq + j) * NH + n0 + lc;
							*dst = (ww ? *dst : 0.f) + aw2[(i0 / 16) * 4 + n0 / 16][j];
						}
				for (int i0 = 0; i0 < NH; i0 += 16)
					for (int j = 0; j < 4; ++j)
						if (lc < NO)
						{
							float *dst = pl + OW3 + (i0 + rq + j) * NO + lc;
							*dst = (ww ? *dst : 0.f) + aw3[i0 / 16][j];
						}
				pl[OB1 + l] = (ww ? pl[OB1 + l] : 0.f) + ab1;
				pl[OB2 + l] = (ww ? pl[OB2 + l] : 0.f) + ab2;
				if (l < NO)
					pl[OB3 + l] = (ww ? pl[OB3 + l] : 0.f) + ab3;
			}
			__syncthreads();
		}
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

// initial weights: W ~ N(0, 1/fan_in) for the hidden layers, the last layer
// 1e-2 of that (identical components would receive identical gradients for
// ever) with the bias of the identity transform (s_k = 1, w uniform, t = 0)
__global__ void k_flow_init(int n, float *p, uint64_t seed)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	int o = i % NP_LAYER;
	float v = 0.f;
	double sd = 0.0;
	if (o >= OW1 && o < OB1)
		sd = 1.0 / sqrt((double)NI);
	else if (o >= OW2 && o < OB2)
		sd = 1.0 / sqrt((double)NH);
	else if (o >= OW3 && o < OB3)
		sd = 1e-2 / sqrt((double)NH); // small: breaks the symmetry of the NK components
	if (sd > 0)
	{
		u4 ctr = {(uint32_t)i, 0u, 0u, TAG_FLOWI};
		u4 r = philox4x32_10(ctr, (uint32_t)seed, (uint32_t)(seed >> 32));
		double z0, z1;
		normal_pair(r, &z0, &z1);
		v = (float)(z0 * sd);
	}
	if (o == OB3 + 2 * NK)
		v = 1.f;
	p[i] = v;
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
