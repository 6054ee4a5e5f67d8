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
	int xcd_remap = 2;     // block order of the stencil kernels (ordered_block)
	int force_variant = 0;
	int fuse_md = 1;        // fuse each momentum step with the next link step
	void *md_spare = nullptr; // third gauge field for the fused trajectory
	size_t md_spare_bytes = 0;
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
constexpr int RED_MAX_Q = 8;

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
// sums nq quantities of nb partials each into host out[nq] (synchronises)
int mesh_finish_reduction(mesh_ctx *ctx, int nb, int nq, double *out);

// launch shape of the site-blocked kernels: nb logical blocks; with XCD
// remapping the grid is padded to a multiple of 8
inline int mesh_grid_blocks(mesh_ctx *ctx, int nb) { return ctx->xcd_remap ? ((nb + 7) / 8) * 8 : nb; }

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

// Block order of the stencil kernels (sbk consecutive sites per workgroup):
//   0: linear;  1: XCD-contiguous linear;
//   2: cache order for 4-d lattices with sbk | L2*L3 -- each XCD walks its own
//      range of x2/x3 row-blocks and, within it, x0 fastest then x1, so the
//      x0-neighbours of a workgroup were staged by the previous one and the
//      x1-neighbours L0 workgroups earlier, all in the same L2.
// Returns the memory-order block index (-1: padding workgroup).  The order
// only permutes workgroups; every site's arithmetic is unchanged.
__device__ __forceinline__ int ordered_block(Geo const &g, int nb, int sbk, int order)
{
	int b = blockIdx.x;
	if (order == 0)
		return b < nb ? b : -1;
	int chunk = (nb + 7) >> 3;
	int o = (b & 7) * chunk + (b >> 3);
	if (o >= nb)
		return -1;
	if (order == 1 || g.nd != 4 || (g.L[2] * g.L[3]) % sbk != 0)
		return o;
	int nb2 = (g.L[2] * g.L[3]) / sbk;
	int x0 = o % g.L[0];
	int t = o / g.L[0];
	int x1 = t % g.L[1];
	int rb = t / g.L[1];
	return (x0 * g.L[1] + x1) * nb2 + rb;
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
