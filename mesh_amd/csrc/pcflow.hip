// Plaquette coupling flow for 2-d U(1): gauge-equivariant coupling layers
// acting on the plaquettes, used as an independence proposal beside HMC (no
// counterpart in the reference; U(1) conventions of u1.hip).
//
// Variables: the V - 1 free plaquettes of an L0 x L1 torus (all but the last,
// P(L0-1, L1-1), which the others fix); the links follow by the integration
// of u1_plaq.h with uniform holonomies and gauge, so the flow acts on
// gauge-invariant variables only and its density over links is gauge
// invariant.
//  * prior: every free plaquette uniform on [-pi, pi);
//  * layer l: mask class m = l % 4; the plaquettes with
//    (x0 mod 2, x1 mod 2) = (m % 2, m / 2) are active (the last excluded),
//    every other one is frozen.  Each active plaquette a moves to
//    h(a) = g(wrap(a - t)) + t with the mixture g of flow_mlp.h, its
//    parameters from an MLP 40 -> 64 -> 64 -> 10 of the frozen 3x3
//    neighbourhood: (cos, sin) of the 8 surrounding plaquettes (1x1 loops),
//    of the 8 1x2 loops formed by neighbouring pairs of them and of the 4
//    L-shaped loops that each 2x2 square of the neighbourhood leaves when its
//    (active) centre is taken out (a loop's angle is the sum of its
//    plaquettes); the last plaquette enters the context as 0 (it is not a
//    variable of the flow);
//  * log q(U) = -sum over layers and active plaquettes of log g'
//    - (V - 1) log 2 pi (prior) - (V + 1) log 2 pi (holonomies, gauge).
// Within a layer the active plaquettes see only frozen ones, so the map is
// triangular with the exact log-determinant above, and it does not perturb
// any plaquette it has not transformed (a link-space coupling layer moves a
// second, passive plaquette with every link it updates).
//
// Kernels: one workgroup (256 threads) per configuration, plaquettes in LDS,
// the dense layers on FP32 MFMA (flow_mlp.h).
#include "flow_mlp.h"
#include "mesh_internal.h"
#include "rng_dev.h"
#include "u1_plaq.h"

using namespace mesh_dev;
namespace FM = flow_mlp;
using u1_plaq::block_sum;
using u1_plaq::PGeo;
using u1_plaq::plaqp;
using u1_plaq::wrapd;

namespace {

constexpr int NI = 40; // features: 20 loops, (cos, sin) each
constexpr int XS = FM::Layout<NI>::XS;
constexpr int NP_LAYER = FM::Layout<NI>::NP;
constexpr int MAX_LAYERS = 64;
constexpr int THREADS = 256;
constexpr bool FAST = true; // tanh_fast in every kernel of this flow
constexpr uint32_t TAG_PCFZ = 0x5043465au; // prior, holonomies, gauge
constexpr uint32_t TAG_PCFI = 0x50434649u; // initial weights
constexpr double LOG_2PI = 1.8378770664093454836;
using Weights = FM::WeightRegs<NI, THREADS>;

// offset of ring position k = 0..7 around a plaquette, in order around it
__device__ __forceinline__ void ring(int k, int &d0, int &d1)
{
	d0 = (k <= 2) ? -1 : ((k >= 4 && k <= 6) ? 1 : 0);
	d1 = (k == 0 || k == 6 || k == 7) ? -1 : ((k >= 2 && k <= 4) ? 1 : 0);
}

// site (x0, x1) with one periodic step either way
__device__ __forceinline__ int sitew(PGeo const &g, int x0, int x1)
{
	x0 = x0 < 0 ? x0 + g.L0 : (x0 >= g.L0 ? x0 - g.L0 : x0);
	x1 = x1 < 0 ? x1 + g.L1 : (x1 >= g.L1 ? x1 - g.L1 : x1);
	return x0 * g.L1 + x1;
}

// active plaquette r of mask class m
__device__ __forceinline__ void active_xy(PGeo const &g, int m, int r, int &x0, int &x1)
{
	const int h = g.L1 >> 1, i = r / h, j = r - i * h;
	x0 = (m & 1) + 2 * i;
	x1 = (m >> 1) + 2 * j;
}
__device__ __forceinline__ int n_active(PGeo const &g, int m) { return (g.V >> 2) - (m == 3 ? 1 : 0); }
__device__ __forceinline__ int class_of(PGeo const &g, int x) { return ((x / g.L1) & 1) | (((x % g.L1) & 1) << 1); }
__device__ __forceinline__ int active_index(PGeo const &g, int x)
{
	return ((x / g.L1) >> 1) * (g.L1 >> 1) + ((x % g.L1) >> 1);
}

// (cos, sin) of context loop k of the plaquette at (x0, x1): k < 8 ring
// plaquette k; 8 <= k < 16 the 1x2 loop of ring plaquettes k - 8 and k - 7
// (mod 8); 16 <= k < 20 the L-shaped loop of ring plaquettes 2j - 1, 2j and
// 2j + 1 (mod 8), j = k - 16 (the 2x2 square at corner 2j without its
// centre); by angle addition.  cs holds (cos, sin) of every plaquette, (1, 0)
// for the last, which enters the context as the angle 0.
__device__ __forceinline__ float2 ring_cs(PGeo const &g, const float2 *cs, int x0, int x1, int k)
{
	int d0, d1;
	ring(k & 7, d0, d1);
	return cs[sitew(g, x0 + d0, x1 + d1)];
}
__device__ __forceinline__ float2 cs_add(float2 a, float2 b)
{
	return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.y, b.x, a.x * b.y));
}
__device__ __forceinline__ float2 loop_cs(PGeo const &g, const float2 *cs, int x0, int x1, int k)
{
	if (k < 8)
		return ring_cs(g, cs, x0, x1, k);
	if (k < 16)
		return cs_add(ring_cs(g, cs, x0, x1, k - 8), ring_cs(g, cs, x0, x1, k - 7));
	const int j = k - 16;
	return cs_add(cs_add(ring_cs(g, cs, x0, x1, 2 * j + 7), ring_cs(g, cs, x0, x1, 2 * j)),
	              ring_cs(g, cs, x0, x1, 2 * j + 1));
}

// features of the 16 active plaquettes r0..r0+15 of class m (rows beyond NA
// zero): columns 2k, 2k+1 = (cos, sin) of context loop k, the row's columns
// from NI on zero.  First the 8 ring plaquettes of every row (the only
// gathers from the configuration), then the 1x2 and L-shaped loops from those
// columns, by the angle additions of loop_cs in its order (the same bits).
template <class TILE> __device__ void features(PGeo const &g, const float2 *cs, int m, int r0, int NA, TILE &T)
{
	const int l = threadIdx.x & 63;
	for (int e = l; e < 16 * 8; e += 64)
	{
		const int row = e >> 3, k = e & 7, r = r0 + row;
		float2 v = make_float2(0.f, 0.f);
		if (r < NA)
		{
			int x0, x1;
			active_xy(g, m, r, x0, x1);
			v = ring_cs(g, cs, x0, x1, k);
		}
		*reinterpret_cast<float2 *>(&T.x[row * XS + 2 * k]) = v;
	}
	FM::wave_sync();
	for (int e = l; e < 16 * 12; e += 64)
	{
		const int row = e / 12, q = e - row * 12;
		const float2 *rg = reinterpret_cast<const float2 *>(&T.x[row * XS]);
		float2 v;
		if (q < 8)
			v = cs_add(rg[q], rg[(q + 1) & 7]);
		else
		{
			const int j = q - 8;
			v = cs_add(cs_add(rg[(2 * j + 7) & 7], rg[2 * j]), rg[2 * j + 1]);
		}
		*reinterpret_cast<float2 *>(&T.x[row * XS + 16 + 2 * q]) = v;
	}
	for (int e = l; e < 16 * (XS - NI); e += 64)
		T.x[(e / (XS - NI)) * XS + NI + e % (XS - NI)] = 0.f;
}

__device__ __forceinline__ float2 sincos2(float a)
{
	float s, c;
	sincosf(a, &s, &c);
	return make_float2(c, s);
}

// LDS of the forward and inverse kernels (MLP outputs over the features) and
// of the backward kernel (outputs apart: they carry dL/dy)
template <bool SEP_Y> struct PCSharedT
{
	FM::LW<NI> lw;
	FM::Tile<NI, SEP_Y> tile[4];
	double red[4][4];
};
using PCShared = PCSharedT<false>;
using PCSharedB = PCSharedT<true>;

// One layer on the workgroup: wavefront w takes tiles w, w + 4, ... of the
// NA active plaquettes: features, MLP, then the circle map on quads (lane
// 4i + k: row i of the tile, mixture component k).  Features read only
// frozen plaquettes, so a wavefront may move its active ones while the others
// still read theirs.  INV: apply g^-1 (the inverse pass).  Adds log g' to ld.
template <bool INV>
__device__ void layer_apply(PGeo const &g, float *phi, float2 *cs, int m, PCShared &sh, double &ld)
{
	const int w = threadIdx.x >> 6, l = threadIdx.x & 63, NA = n_active(g, m), NT = (NA + 15) / 16;
	auto &T = sh.tile[w];
	for (int tix = w; tix < NT; tix += 4)
	{
		features(g, cs, m, tix * 16, NA, T);
		FM::wave_sync();
		FM::mlp_forward<NI, FAST>(sh.lw, T);
		const int row = l >> 2, r = tix * 16 + row;
		if (r < NA)
		{
			int x0, x1;
			active_xy(g, m, r, x0, x1);
			const int x = x0 * g.L1 + x1;
			const FM::NCPq p = FM::ncpq_params(&T.y()[row * 16]);
			const float a = FM::wrapf(phi[x] - p.t);
			float v, Gp;
			if (INV)
				v = FM::ncpq_inverse(p, a, Gp);
			else
				FM::ncpq_eval(p, a, v, Gp);
			if ((l & 3) == 0)
			{
				const float nx = FM::wrapf(v + p.t);
				phi[x] = nx;
				cs[x] = sincos2(nx);
				ld += (double)logf(Gp);
			}
		}
		FM::wave_sync();
	}
}

// z uniform -> free plaquettes -> links.  phi_layers (optional): the input of
// every layer; phi_final (optional): the output; theta_out (optional): the
// links.  out[2c + {0,1}] = {log q, S/beta}, S of the links when they are
// formed, else of the plaquettes (the last one = -sum of the free ones).
// LDS: phi[V], cs[V] (then the links' scratch), al[V].
__device__ void pc_forward(PGeo const &g, int nl, const float *__restrict__ params, uint64_t seed, uint32_t step,
                           float *theta_out, float *phi_layers, float *phi_final, double *out, int c, float *dyn,
                           PCShared &sh)
{
	float *phi = dyn;
	float2 *cs = reinterpret_cast<float2 *>(dyn + g.V);
	float *al = dyn + 3 * g.V;
	const uint32_t cc = (uint32_t)c << 1;
	Weights wr;
	wr.load(params);
	FM::zero_pad<NI>(sh.lw);
	for (int x = threadIdx.x; x < g.V; x += THREADS)
	{
		const float z = x < g.V - 1 ? u1_plaq::uniform_angle(seed, step, (uint32_t)x, cc, TAG_PCFZ) : 0.f;
		phi[x] = z;
		cs[x] = sincos2(z);
		al[x] = u1_plaq::uniform_angle(seed, step, (uint32_t)x, cc | 1u, TAG_PCFZ);
	}
	double ld = 0.0;
	for (int layer = 0; layer < nl; ++layer)
	{
		__syncthreads(); // the previous layer is done with the weights and the plaquettes
		wr.store(sh.lw);
		if (phi_layers)
			for (int x = threadIdx.x; x < g.V; x += THREADS)
				phi_layers[((long)c * nl + layer) * g.V + x] = phi[x];
		__syncthreads();
		if (layer + 1 < nl)
			wr.load(params + (long)(layer + 1) * NP_LAYER);
		layer_apply<false>(g, phi, cs, layer & 3, sh, ld);
	}
	__syncthreads();
	if (phi_final)
		for (int x = threadIdx.x; x < g.V; x += THREADS)
			phi_final[(long)c * g.V + x] = phi[x];
	double v[3] = {ld, 0.0, 0.0}; // log det, S of the free plaquettes, their sum
	for (int x = threadIdx.x; x < g.V - 1; x += THREADS)
	{
		v[1] += 1.0 - cos((double)phi[x]);
		v[2] += (double)phi[x];
	}
	block_sum<3>(v, sh.red);
	double S = v[1] + (1.0 - cos(wrapd(-v[2])));
	if (theta_out)
	{
		float *o = theta_out + (long)c * 2 * g.V;
		const float h0 = u1_plaq::uniform_angle(seed, step, (uint32_t)g.V, cc, TAG_PCFZ);
		const float h1 = u1_plaq::uniform_angle(seed, step, (uint32_t)g.V + 1u, cc, TAG_PCFZ);
		u1_plaq::links_from_plaquettes(g, phi, h0, h1, al, reinterpret_cast<float *>(cs), o, THREADS);
		__syncthreads();
		double s[1] = {0.0};
		for (int x = threadIdx.x; x < g.V; x += THREADS)
		{
			const int x0 = x / g.L1, x1 = x - x0 * g.L1;
			s[0] += 1.0 - cos((double)plaqp(g, o, x0, x1));
		}
		block_sum<1>(s, sh.red);
		S = s[0];
	}
	if (threadIdx.x == 0)
	{
		out[2 * c + 0] = -v[0] - 2.0 * g.V * LOG_2PI;
		out[2 * c + 1] = S;
	}
}

// log q of given links: plaquettes, then the layers in reverse with g^-1.
// out[2c + {0,1}] = {log q(U), S(U)/beta}; z_out (optional): the prior draws,
// V per configuration (the last 0).  LDS: the links (then cs[V]), phi[V].
__device__ void pc_inverse(PGeo const &g, int nl, const float *__restrict__ params, const float *__restrict__ theta,
                           float *z_out, double *out, int c, float *dyn, PCShared &sh)
{
	float *th = dyn;
	float2 *cs = reinterpret_cast<float2 *>(dyn);
	float *phi = dyn + 2 * g.V;
	Weights wr;
	wr.load(params + (long)(nl - 1) * NP_LAYER);
	FM::zero_pad<NI>(sh.lw);
	for (int i = threadIdx.x; i < 2 * g.V; i += THREADS)
		th[i] = theta[(long)c * 2 * g.V + i];
	__syncthreads();
	double v[2] = {0.0, 0.0}; // log det, S
	for (int x = threadIdx.x; x < g.V; x += THREADS)
	{
		const int x0 = x / g.L1, x1 = x - x0 * g.L1;
		const float tp = plaqp(g, th, x0, x1);
		v[1] += 1.0 - cos((double)tp);
		phi[x] = x < g.V - 1 ? FM::wrapf(tp) : 0.f;
	}
	__syncthreads(); // the links are read; cs takes their place
	for (int x = threadIdx.x; x < g.V; x += THREADS)
		cs[x] = sincos2(phi[x]);
	for (int layer = nl - 1; layer >= 0; --layer)
	{
		__syncthreads();
		wr.store(sh.lw);
		__syncthreads();
		if (layer > 0)
			wr.load(params + (long)(layer - 1) * NP_LAYER);
		layer_apply<true>(g, phi, cs, layer & 3, sh, v[0]);
	}
	__syncthreads();
	block_sum<2>(v, sh.red);
	if (threadIdx.x == 0)
	{
		out[2 * c + 0] = -v[0] - 2.0 * g.V * LOG_2PI;
		out[2 * c + 1] = v[1];
	}
	if (z_out)
		for (int x = threadIdx.x; x < g.V; x += THREADS)
			z_out[(long)c * g.V + x] = phi[x];
}

__global__ __launch_bounds__(THREADS) void k_pc_forward(PGeo g, int nl, const float *__restrict__ params,
                                                        uint64_t seed, uint32_t step, float *theta_out,
                                                        float *phi_layers, float *phi_final, double *out)
{
	extern __shared__ float dyn[];
	__shared__ PCShared sh;
	pc_forward(g, nl, params, seed, step, theta_out, phi_layers, phi_final, out, blockIdx.x, dyn, sh);
}

__global__ __launch_bounds__(THREADS) void k_pc_inverse(PGeo g, int nl, const float *__restrict__ params,
                                                        const float *__restrict__ theta, float *z_out, double *out)
{
	extern __shared__ float dyn[];
	__shared__ PCShared sh;
	pc_inverse(g, nl, params, theta, z_out, out, blockIdx.x, dyn, sh);
}

// one MH proposal round in one launch: workgroups 0..batch-1 take log q of
// the current configurations, batch..2 batch-1 draw the proposals, so that on
// small lattices (two workgroups' LDS per CU) the two passes share the CUs
__global__ __launch_bounds__(THREADS) void k_pc_propose(PGeo g, int nl, const float *__restrict__ params, int batch,
                                                        const float *__restrict__ theta, double *state,
                                                        uint64_t seed, uint32_t step, float *theta_prop,
                                                        double *prop_out)
{
	extern __shared__ float dyn[];
	__shared__ PCShared sh;
	if ((int)blockIdx.x < batch)
		pc_inverse(g, nl, params, theta, nullptr, state, blockIdx.x, dyn, sh);
	else
		pc_forward(g, nl, params, seed, step, theta_prop, nullptr, nullptr, prop_out, blockIdx.x - batch, dyn, sh);
}

// gradient of the per-configuration loss beta S - log det (S with the last
// plaquette = -sum of the free ones) with respect to every weight:
// partial[c][layer * NP_LAYER + i].  phi_layers / phi_final from k_pc_forward;
// dxs: scratch of (V/4) * NI floats per configuration.
// LDS: cs[V] (of the current layer's input), phi[V], a[V], gact[V/4].
__global__ __launch_bounds__(THREADS) void k_pc_backward(PGeo g, int nl, const float *__restrict__ params,
                                                         const float *__restrict__ phi_layers,
                                                         const float *__restrict__ phi_final, float beta,
                                                         float *__restrict__ dxs, float *__restrict__ partial)
{
	extern __shared__ float dyn[];
	float2 *cs = reinterpret_cast<float2 *>(dyn);
	float *phi = dyn + 2 * g.V;  // input of the current layer
	float *a = dyn + 3 * g.V;    // adjoint w.r.t. the current layer's output
	float *gact = dyn + 4 * g.V; // adjoint w.r.t. the active plaquettes' input
	__shared__ PCSharedB sh;
	const int c = blockIdx.x, w = threadIdx.x >> 6, l = threadIdx.x & 63;
	float *dx = dxs + (long)c * (g.V >> 2) * NI;
	float *part = partial + (long)c * nl * NP_LAYER;
	FM::zero_pad<NI>(sh.lw);
	for (int x = threadIdx.x; x < g.V; x += THREADS)
		phi[x] = phi_final[(long)c * g.V + x];
	double s[1] = {0.0};
	for (int x = threadIdx.x; x < g.V - 1; x += THREADS)
		s[0] += (double)phi[x];
	block_sum<1>(s, sh.red);
	const float sin_last = (float)sin(wrapd(-s[0]));
	for (int x = threadIdx.x; x < g.V; x += THREADS)
		a[x] = x < g.V - 1 ? beta * (sinf(phi[x]) - sin_last) : 0.f;
	for (int layer = nl - 1; layer >= 0; --layer)
	{
		const int m = layer & 3, NA = n_active(g, m), NT = (NA + 15) / 16;
		__syncthreads();
		for (int x = threadIdx.x; x < g.V; x += THREADS)
		{
			const float p = phi_layers[((long)c * nl + layer) * g.V + x];
			phi[x] = p;
			cs[x] = sincos2(p);
		}
		{
			// no prefetch here: the gradient accumulators need the registers
			Weights wr;
			wr.load(params + (long)layer * NP_LAYER);
			wr.store(sh.lw);
		}
		__syncthreads();
		FM::Grad<NI> gr;
		gr.zero();
		auto &T = sh.tile[w];
		for (int tix = w; tix < NT; tix += 4)
		{
			features(g, cs, m, tix * 16, NA, T);
			FM::wave_sync();
			FM::mlp_forward<NI, FAST>(sh.lw, T);
			if (l < 16)
			{
				float dy[16];
				for (int k = 0; k < 16; ++k)
					dy[k] = 0.f;
				const int r = tix * 16 + l;
				if (r < NA)
				{
					int x0, x1;
					active_xy(g, m, r, x0, x1);
					const int x = x0 * g.L1 + x1;
					const float *y = T.y() + l * 16;
					gact[r] = FM::ncp_adjoint(y, FM::wrapf(phi[x] - FM::ncp_params(y).t), a[x], dy);
				}
				for (int k = 0; k < 16; ++k)
					T.y()[l * 16 + k] = dy[k];
			}
			FM::wave_sync();
			FM::mlp_backward<NI>(sh.lw, T, gr, dx, tix * 16, NA);
		}
		__syncthreads();
		FM::store_grad<NI>(gr, part + (long)layer * NP_LAYER);
		// adjoints w.r.t. the layer's input: the active plaquettes from gact,
		// the frozen ones gain their features' adjoints
		for (int y = threadIdx.x; y < g.V - 1; y += THREADS)
		{
			if (class_of(g, y) == m)
			{
				a[y] = gact[active_index(g, y)];
				continue;
			}
			const int y0 = y / g.L1, y1 = y - y0 * g.L1;
			const float2 cy = cs[y];
			float acc = 0.f;
			for (int k = 0; k < 8; ++k)
			{
				int d0, d1;
				ring(k, d0, d1);
				const int x = sitew(g, y0 - d0, y1 - d1);
				if (x == g.V - 1 || class_of(g, x) != m)
					continue;
				const int x0 = x / g.L1, x1 = x - x0 * g.L1;
				const float *dr = dx + (long)active_index(g, x) * NI;
				acc = fmaf(-cy.y, dr[2 * k], acc);
				acc = fmaf(cy.x, dr[2 * k + 1], acc);
				// the two 1x2 loops that contain ring position k
				for (int j : {(k + 7) & 7, k})
				{
					const float2 ps = loop_cs(g, cs, x0, x1, 8 + j);
					acc = fmaf(-ps.y, dr[16 + 2 * j], acc);
					acc = fmaf(ps.x, dr[16 + 2 * j + 1], acc);
				}
				// the L-shaped loops (ring positions 2j - 1, 2j, 2j + 1) that
				// contain ring position k: one for a corner, two for an edge
				for (int j : {(k & 1) ? (k - 1) >> 1 : k >> 1, (k & 1) ? ((k + 1) >> 1) & 3 : -1})
				{
					if (j < 0)
						continue;
					const float2 ps = loop_cs(g, cs, x0, x1, 16 + j);
					acc = fmaf(-ps.y, dr[32 + 2 * j], acc);
					acc = fmaf(ps.x, dr[32 + 2 * j + 1], acc);
				}
			}
			a[y] += acc;
		}
	}
}

// grad[i] = scale * sum_c partial[c][i] (fixed order)
__global__ void k_pc_reduce(int n, const float *__restrict__ partial, int batch, float scale, float *__restrict__ grad)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	double s = 0.0;
	for (int c = 0; c < batch; ++c)
		s += partial[(long)c * n + i];
	grad[i] = (float)(s * scale);
}

__global__ void k_pc_init(int n, float *p, uint64_t seed)
{
	int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		FM::init_param<NI>(i, p, seed, TAG_PCFI);
}

int pc_setup(mesh_ctx *ctx, const mesh_grid *g, int batch, int layers, PGeo *pg, const float *params = nullptr)
{
	if (layers < 1 || layers > MAX_LAYERS)
		return mesh_fail(ctx, MESH_E_ARG, "the plaquette coupling flow has 1 to 64 layers");
	if (!g || g->nd != 2 || g->dims[0] % 2 || g->dims[1] % 2 || g->dims[0] < 4 || g->dims[1] < 4)
		return mesh_fail(ctx, MESH_E_SIZE, "the plaquette coupling flow needs a 2-d lattice with even extents >= 4");
	long V = (long)g->dims[0] * g->dims[1];
	if (V > 4096)
		return mesh_fail(ctx, MESH_E_SIZE, "the plaquette coupling flow holds a configuration in LDS: volume <= 4096");
	if (batch < 1)
		return mesh_fail(ctx, MESH_E_ARG, "batch must be >= 1");
	if ((uintptr_t)params % 8)
		return mesh_fail(ctx, MESH_E_ARG, "the flow's parameters must be 8-byte aligned");
	pg->L0 = g->dims[0];
	pg->L1 = g->dims[1];
	pg->V = (int)V;
	return MESH_OK;
}

template <class K> int set_lds(mesh_ctx *ctx, K kernel, size_t bytes)
{
	MESH_HIP(ctx, hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
	return MESH_OK;
}

} // namespace

extern "C" {

int mesh_pcflow_num_params(int layers) { return layers >= 1 && layers <= MAX_LAYERS ? layers * NP_LAYER : 0; }

int mesh_pcflow_init(mesh_ctx *ctx, int layers, float *params, uint64_t seed)
{
	MESH_CHECK_CTX(ctx);
	if (!params || layers < 1 || layers > MAX_LAYERS)
		return mesh_fail(ctx, MESH_E_ARG, "bad pcflow_init arguments");
	const int n = layers * NP_LAYER;
	k_pc_init<<<(n + 255) / 256, 256, 0, ctx->stream>>>(n, params, seed);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_pcflow_sample(mesh_ctx *ctx, const mesh_grid *g, int batch, int layers, const float *params, uint64_t seed,
                       uint32_t step, float *theta_out, double *out_dev)
{
	MESH_CHECK_CTX(ctx);
	PGeo pg;
	if (int rc = pc_setup(ctx, g, batch, layers, &pg, params))
		return rc;
	if (!params || !theta_out || !out_dev)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to pcflow_sample");
	const size_t lds = 4 * (size_t)pg.V * sizeof(float);
	if (int rc = set_lds(ctx, k_pc_forward, lds))
		return rc;
	k_pc_forward<<<batch, THREADS, lds, ctx->stream>>>(pg, layers, params, seed, step, theta_out, nullptr, nullptr,
	                                                   out_dev);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_pcflow_logq(mesh_ctx *ctx, const mesh_grid *g, int batch, int layers, const float *params,
                     const float *theta, float *z_out, double *out_dev)
{
	MESH_CHECK_CTX(ctx);
	PGeo pg;
	if (int rc = pc_setup(ctx, g, batch, layers, &pg, params))
		return rc;
	if (!params || !theta || !out_dev)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to pcflow_logq");
	const size_t lds = 3 * (size_t)pg.V * sizeof(float);
	if (int rc = set_lds(ctx, k_pc_inverse, lds))
		return rc;
	k_pc_inverse<<<batch, THREADS, lds, ctx->stream>>>(pg, layers, params, theta, z_out, out_dev);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_pcflow_propose(mesh_ctx *ctx, const mesh_grid *g, int batch, int layers, const float *params,
                        const float *theta, double *state_dev, uint64_t seed, uint32_t step, float *theta_prop,
                        double *prop_dev)
{
	MESH_CHECK_CTX(ctx);
	PGeo pg;
	if (int rc = pc_setup(ctx, g, batch, layers, &pg, params))
		return rc;
	if (!params || !theta || !state_dev || !theta_prop || !prop_dev)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to pcflow_propose");
	if (theta == theta_prop || state_dev == prop_dev)
		return mesh_fail(ctx, MESH_E_ARG, "pcflow_propose: the proposals and their log q need their own buffers");
	const size_t lds = 4 * (size_t)pg.V * sizeof(float);
	if (int rc = set_lds(ctx, k_pc_propose, lds))
		return rc;
	k_pc_propose<<<2 * batch, THREADS, lds, ctx->stream>>>(pg, layers, params, batch, theta, state_dev, seed, step,
	                                                       theta_prop, prop_dev);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

size_t mesh_pcflow_workspace_bytes(const mesh_grid *g, int batch, int layers)
{
	PGeo pg;
	if (pc_setup(nullptr, g, batch, layers, &pg))
		return 0;
	size_t n = (size_t)batch * layers * pg.V           // plaquettes per layer
	           + (size_t)batch * pg.V                  // final plaquettes
	           + (size_t)batch * (pg.V / 4) * NI       // feature adjoints
	           + (size_t)batch * layers * NP_LAYER;    // per-configuration gradients
	return n * sizeof(float);
}

int mesh_pcflow_gradient(mesh_ctx *ctx, const mesh_grid *g, int batch, int layers, const float *params, uint64_t seed,
                         uint32_t step, double beta, void *workspace, float *grad, double *out_dev)
{
	MESH_CHECK_CTX(ctx);
	PGeo pg;
	if (int rc = pc_setup(ctx, g, batch, layers, &pg, params))
		return rc;
	if (!params || !workspace || !grad || !out_dev)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to pcflow_gradient");
	float *pl = (float *)workspace;
	float *pf = pl + (size_t)batch * layers * pg.V;
	float *dxs = pf + (size_t)batch * pg.V;
	float *part = dxs + (size_t)batch * (pg.V / 4) * NI;
	const size_t lds_f = 4 * (size_t)pg.V * sizeof(float);
	if (int rc = set_lds(ctx, k_pc_forward, lds_f))
		return rc;
	k_pc_forward<<<batch, THREADS, lds_f, ctx->stream>>>(pg, layers, params, seed, step, nullptr, pl, pf, out_dev);
	MESH_LAUNCHED(ctx);
	const size_t lds_b = (4 * (size_t)pg.V + pg.V / 4) * sizeof(float);
	if (int rc = set_lds(ctx, k_pc_backward, lds_b))
		return rc;
	k_pc_backward<<<batch, THREADS, lds_b, ctx->stream>>>(pg, layers, params, pl, pf, (float)beta, dxs, part);
	MESH_LAUNCHED(ctx);
	const int n = layers * NP_LAYER;
	k_pc_reduce<<<(n + 255) / 256, 256, 0, ctx->stream>>>(n, part, batch, 1.0f / batch, grad);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_pcflow_adam(mesh_ctx *ctx, int layers, float *params, const float *grad, float *m, float *v, int t,
                     double lr)
{
	MESH_CHECK_CTX(ctx);
	if (!params || !grad || !m || !v || t < 1 || layers < 1 || layers > MAX_LAYERS)
		return mesh_fail(ctx, MESH_E_ARG, "bad pcflow_adam arguments");
	return mesh_launch_adam(ctx, layers * NP_LAYER, params, grad, m, v, t, lr);
}

} // extern "C"
