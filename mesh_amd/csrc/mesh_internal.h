// Internal declarations shared by the translation units of libmesh_hip.so.
#pragma once
#include "../../include/mesh_hip.h"
#include "su3_dev.h"
#include <hip/hip_runtime.h>
#include <string>
#include <vector>

// kernels timed by the optional per-launch event profiler
enum mesh_prof_kind
{
	PROF_FORCE_UPDATE = 0,
	PROF_LINK_UPDATE,
	PROF_DELTA_H,
	PROF_MOMENTA,
	PROF_PLAQ,
	PROF_LANDAU_OR,
	PROF_FORCE_LINK,
	PROF_MOMENTA_LINK,
	PROF_NKINDS
};

struct mesh_prof_rec
{
	int kind;
	hipEvent_t a, b;
};

struct mesh_ctx
{
	int magic;
	int device;
	hipStream_t stream;
	std::string err;
	// reduction scratch: per-block partial sums (device) and the final
	// values (device + pinned host)
	double *partials = nullptr;
	size_t partials_cap = 0; // doubles
	double *red_dev = nullptr;
	double *red_host = nullptr;
	double *aux = nullptr; // small device buffer (integrator step tables)
	size_t aux_cap = 0;    // doubles
	std::vector<double> aux_host; // host copy of what aux holds
	int xcd_remap = 3;     // block order of the stencil kernels (ordered_block)
	int rb_tile = 4;       // row-blocks per traversal tile of orders 3 and 4
	int force_variant = -1; // -1: the default kernel (variant 0)
	int fuse_md = 1;        // fuse each momentum step with the next link step
	int compact_p = 1;      // fused trajectory: momenta in 9 reals per link
	void *md_spare = nullptr; // third gauge field for the fused trajectory
	size_t md_spare_bytes = 0;
	int u1_general = 0;       // U(1): site-parallel kernels also for small 2-d lattices
	int float_sums = 0;       // FP32 trajectories: Delta H from the reference's FP32 energy sums
	void *u1_ws = nullptr;    // workspace of the site-parallel U(1) kernels
	size_t u1_ws_bytes = 0;
	// device address of the static scratch array of the no-allocation entry
	// point on this context's device (mesh_scratch_setup)
	double *scratch = nullptr;
	// profiler: events around launches, resolved after a stream sync
	int prof_on = 0;
	std::vector<hipEvent_t> prof_pool;
	std::vector<mesh_prof_rec> prof_pending;
	double prof_ms[PROF_NKINDS] = {};
	long long prof_cnt[PROF_NKINDS] = {};
};

// bracket a launch with events when profiling is on (returns a handle or -1)
int mesh_prof_begin(mesh_ctx *ctx, int kind);
void mesh_prof_end(mesh_ctx *ctx, int handle);
// resolve pending records (call after the stream has been synchronised)
void mesh_prof_collect(mesh_ctx *ctx);

constexpr int MESH_CTX_MAGIC = 0x4d455348;
constexpr int RED_MAX_Q = 18;

int mesh_fail(mesh_ctx *ctx, int code, std::string const &msg);
int mesh_hip_fail(mesh_ctx *ctx, hipError_t e, char const *what);
bool mesh_ctx_ok(mesh_ctx *ctx);

#define MESH_CHECK_CTX(ctx)                                                                        \
	do                                                                                             \
	{                                                                                              \
		if (!mesh_ctx_ok(ctx))                                                                     \
			return MESH_E_STATE;                                                                   \
	} while (0)

#define MESH_HIP(ctx, expr)                                                                        \
	do                                                                                             \
	{                                                                                              \
		hipError_t e_ = (expr);                                                                    \
		if (e_ != hipSuccess)                                                                      \
			return mesh_hip_fail(ctx, e_, #expr);                                                  \
	} while (0)

#define MESH_LAUNCHED(ctx)                                                                         \
	do                                                                                             \
	{                                                                                              \
		hipError_t e_ = hipGetLastError();                                                         \
		if (e_ != hipSuccess)                                                                      \
			return mesh_hip_fail(ctx, e_, "kernel launch");                                        \
	} while (0)

// validates a grid; fills the device geometry descriptor
int mesh_make_geo(mesh_ctx *ctx, const mesh_grid *g, mesh_dev::Geo *geo);
int mesh_check_prec(mesh_ctx *ctx, int prec);
// ensures ctx->partials holds at least n doubles
int mesh_reserve_partials(mesh_ctx *ctx, size_t n);
// copies n host doubles into ctx->aux (grown as needed), on the ctx stream
int mesh_stage_doubles(mesh_ctx *ctx, const double *host, size_t n);
// one Adam step (Kingma & Ba) over n parameters, t >= 1
int mesh_launch_adam(mesh_ctx *ctx, int n, float *params, const float *grad, float *m, float *v, int t, double lr);
// sums nq quantities of nb partials each into host out[nq] (synchronises)
int mesh_finish_reduction(mesh_ctx *ctx, int nb, int nq, double *out);
// the same fixed-order sum of nb partials (nq = 1) into device out_dev, on the
// context's stream, without copying or synchronising
int mesh_reduce_to_device(mesh_ctx *ctx, const double *partials, int nb, double *out_dev);
// resolves the device address of the no-allocation scratch array for the
// context's device (once per device; the device must be current)
int mesh_scratch_setup(mesh_ctx *ctx);

// launch shape of the site-blocked kernels: nb logical blocks; with XCD
// remapping the grid is padded to a multiple of 8
inline int mesh_grid_blocks(mesh_ctx *ctx, int nb) { return ctx->xcd_remap ? ((nb + 7) / 8) * 8 : nb; }
// block-order argument of the stencil kernels: order | (tile << 4)
inline int mesh_order(mesh_ctx *ctx) { return ctx->xcd_remap | (ctx->rb_tile << 4); }

namespace mesh_dev {

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X.  With
// remapping, XCD k processes the contiguous logical block range
// [k*chunk, (k+1)*chunk), so lattice neighbours a few blocks apart share
// that XCD's L2.
__device__ __forceinline__ int logical_block(int nb, int remap)
{
	int b = blockIdx.x;
	if (!remap)
		return b;
	int chunk = (nb + 7) >> 3;
	return (b & 7) * chunk + (b >> 3);
}

// Block order of the stencil kernels (sbk consecutive sites per workgroup,
// i.e. one "row-block" rb of the x2/x3 plane), argument order | (T << 4):
//   0: linear;  1: XCD-contiguous linear;
//   2: cache order for 4-d lattices with sbk | L2*L3 -- each XCD walks its own
//      range of row-blocks and, within it, x0 fastest then x1, so the
//      x0-neighbours of a workgroup were staged by the previous one and the
//      x1-neighbours L0 workgroups earlier, all in the same L2;
//   3: slab order for 4-d lattices with 8 | L0 and T | (L2*L3/sbk) -- XCD k
//      owns the slab x0 in [k L0/8, (k+1) L0/8) and walks it x0 fastest, then
//      T row-blocks, then x1, then the next tile of T row-blocks, so the x1-
//      and rb-neighbours of a workgroup are at most (L0/8) T workgroups back
//      and only the two x0 faces of the slab come from another XCD;
//   4: as 3 with XCD k owning x0 in a quarter (k % 4) and x1 in a half (k / 4)
//      of the lattice (4 | L0, 2 | L1); also order 3's fallback where 8 does
//      not divide L0.
// Returns the memory-order block index (-1: padding workgroup).  The order
// only permutes workgroups; every site's arithmetic is unchanged.  GEO: any
// geometry with nd and L[4] (the SU(3) and SU(2) ones).
template <class GEO> __device__ __forceinline__ int ordered_block_of(GEO const &g, int nb, int sbk, int arg, int b)
{
	const int order = arg & 15, T = arg >> 4;
	if (order == 0)
		return b < nb ? b : -1;
	const bool planes = g.nd == 4 && (g.L[2] * g.L[3]) % sbk == 0;
	const int nb2 = planes ? (g.L[2] * g.L[3]) / sbk : 1;
	if (order == 3 && planes && (g.L[0] & 7) == 0 && T > 0 && nb2 % T == 0)
	{
		const int X = g.L[0] >> 3;
		const int j = b >> 3;
		if (j >= nb >> 3)
			return -1;
		int x0 = (b & 7) * X + j % X;
		int t = j / X;
		int rb = t % T;
		t /= T;
		int x1 = t % g.L[1];
		rb += (t / g.L[1]) * T;
		return (x0 * g.L[1] + x1) * nb2 + rb;
	}
	if ((order == 4 || order == 3) && planes && (g.L[0] & 3) == 0 && (g.L[1] & 1) == 0 && T > 0 &&
	    nb2 % T == 0)
	{
		// XCD k owns x0 in a quarter (k % 4) and x1 in a half (k / 4) of the
		// lattice, walked as in order 3 (and order 3's fallback where 8 does
		// not divide L0)
		const int X0 = g.L[0] >> 2, X1 = g.L[1] >> 1, k = b & 7;
		const int j = b >> 3;
		if (j >= nb >> 3)
			return -1;
		int x0 = (k & 3) * X0 + j % X0;
		int t = j / X0;
		int rb = t % T;
		t /= T;
		int x1 = (k >> 2) * X1 + t % X1;
		rb += (t / X1) * T;
		return (x0 * g.L[1] + x1) * nb2 + rb;
	}
	int chunk = (nb + 7) >> 3;
	int o = (b & 7) * chunk + (b >> 3);
	if (o >= nb)
		return -1;
	if (order == 1 || !planes)
		return o;
	int x0 = o % g.L[0];
	int t = o / g.L[0];
	int x1 = t % g.L[1];
	int rb = t / g.L[1];
	return (x0 * g.L[1] + x1) * nb2 + rb;
}
__device__ __forceinline__ int ordered_block(Geo const &g, int nb, int sbk, int arg)
{
	return ordered_block_of(g, nb, sbk, arg, blockIdx.x);
}

// deterministic block sum of NQ doubles; thread 0 writes partials[q*nb + lb]
template <int NQ>
__device__ __forceinline__ void block_sum_store(double (&v)[NQ], double *partials, int nb, int lb)
{
	__shared__ double red[NQ][16];
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
	{
		int nw = (blockDim.x + 63) >> 6;
#pragma unroll
		for (int q = 0; q < NQ; ++q)
		{
			double s = red[q][0];
			for (int i = 1; i < nw; ++i)
				s += red[q][i];
			partials[(long)q * nb + lb] = s;
		}
	}
}

} // namespace mesh_dev
