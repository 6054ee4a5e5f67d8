// Context, memory and field-transfer entry points of libmesh_hip.so.
//
// Replaces the untyped CUDA backend of the reference,
// src/lattice/device.h:14-24 / src/lattice/device.cpp:1-76, with one
// explicit context (stream + reduction scratch + error string) instead of
// global CUDA state and exceptions.
#include "mesh_internal.h"
#include <cstdlib>
#include <cstring>
#include <new>

using namespace mesh_dev;

int mesh_fail(mesh_ctx *ctx, int code, std::string const &msg)
{
	if (ctx && ctx->magic == MESH_CTX_MAGIC)
		ctx->err = msg;
	return code;
}

int mesh_hip_fail(mesh_ctx *ctx, hipError_t e, char const *what)
{
	return mesh_fail(ctx, MESH_E_HIP, std::string("hip error: ") + hipGetErrorString(e) + " (" + what + ")");
}

bool mesh_ctx_ok(mesh_ctx *ctx) { return ctx && ctx->magic == MESH_CTX_MAGIC; }

int mesh_check_prec(mesh_ctx *ctx, int prec)
{
	if (prec != MESH_F64 && prec != MESH_F32)
		return mesh_fail(ctx, MESH_E_ARG, "invalid precision " + std::to_string(prec));
	return MESH_OK;
}

int mesh_make_geo(mesh_ctx *ctx, const mesh_grid *g, Geo *geo)
{
	if (!g)
		return mesh_fail(ctx, MESH_E_ARG, "grid is NULL");
	if (g->nd < 2 || g->nd > 4)
		return mesh_fail(ctx, MESH_E_SIZE, "grid must have 2 to 4 dimensions");
	long V = 1;
	for (int i = 0; i < g->nd; ++i)
	{
		if (g->dims[i] < 2)
			return mesh_fail(ctx, MESH_E_SIZE, "every grid extent must be >= 2");
		V *= g->dims[i];
		if (V > (1L << 26))
			return mesh_fail(ctx, MESH_E_SIZE, "grid volume exceeds 2^26 sites");
	}
	geo->nd = g->nd;
	long s = 1;
	for (int i = 3; i >= 0; --i)
	{
		if (i >= g->nd)
		{
			geo->L[i] = 1;
			geo->st[i] = 0;
			continue;
		}
		geo->L[i] = g->dims[i];
		geo->st[i] = (int)s;
		s *= g->dims[i];
	}
	geo->V = V;
	geo->stride = (V + 255) / 256 * 256;
	return MESH_OK;
}

int mesh_reserve_partials(mesh_ctx *ctx, size_t n)
{
	if (n <= ctx->partials_cap)
		return MESH_OK;
	if (ctx->partials)
	{
		MESH_HIP(ctx, hipStreamSynchronize(ctx->stream));
		MESH_HIP(ctx, hipFree(ctx->partials));
		ctx->partials = nullptr;
		ctx->partials_cap = 0;
	}
	if (hipMalloc(&ctx->partials, n * sizeof(double)) != hipSuccess)
		return mesh_fail(ctx, MESH_E_ALLOC, "cannot allocate reduction scratch");
	ctx->partials_cap = n;
	return MESH_OK;
}

int mesh_stage_doubles(mesh_ctx *ctx, const double *host, size_t n)
{
	if (n > ctx->aux_cap)
	{
		if (ctx->aux)
		{
			MESH_HIP(ctx, hipStreamSynchronize(ctx->stream));
			MESH_HIP(ctx, hipFree(ctx->aux));
			ctx->aux = nullptr;
			ctx->aux_cap = 0;
		}
		if (hipMalloc(&ctx->aux, n * sizeof(double)) != hipSuccess)
			return mesh_fail(ctx, MESH_E_ALLOC, "cannot allocate the step-table buffer");
		ctx->aux_cap = n;
		ctx->aux_host.clear();
	}
	// a table equal to the one already staged is not copied again (a pageable
	// host-to-device copy can stall the host until the stream drains)
	if (ctx->aux_host.size() == n && std::memcmp(ctx->aux_host.data(), host, n * sizeof(double)) == 0)
		return MESH_OK;
	// the previous user of aux may still be in flight on the stream: the copy
	// is ordered after it on the same stream
	MESH_HIP(ctx, hipMemcpyAsync(ctx->aux, host, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
	ctx->aux_host.assign(host, host + n);
	return MESH_OK;
}

// second reduction pass: one workgroup, fixed summation order
template <int NQ> __global__ __launch_bounds__(1024) void k_reduce_partials(const double *p, int nb, double *out)
{
	double v[NQ];
#pragma unroll
	for (int q = 0; q < NQ; ++q)
	{
		double s = 0.0;
		for (int i = threadIdx.x; i < nb; i += blockDim.x)
			s += p[(long)q * nb + i];
		v[q] = s;
	}
	block_sum_store<NQ>(v, out, 1, 0);
}

static const char *const PROF_NAMES[PROF_NKINDS] = {"force_update", "link_update", "delta_h",   "momenta",
                                                     "plaquette",    "landau_or",   "force_link",
                                                     "momenta_link"};

int mesh_prof_begin(mesh_ctx *ctx, int kind)
{
	if (!ctx->prof_on)
		return -1;
	hipEvent_t ev[2];
	for (int i = 0; i < 2; ++i)
	{
		if (!ctx->prof_pool.empty())
		{
			ev[i] = ctx->prof_pool.back();
			ctx->prof_pool.pop_back();
		}
		else if (hipEventCreate(&ev[i]) != hipSuccess)
			return -1;
	}
	if (hipEventRecord(ev[0], ctx->stream) != hipSuccess)
		return -1;
	ctx->prof_pending.push_back({kind, ev[0], ev[1]});
	return (int)ctx->prof_pending.size() - 1;
}

void mesh_prof_end(mesh_ctx *ctx, int h)
{
	if (h < 0)
		return;
	(void)hipEventRecord(ctx->prof_pending[h].b, ctx->stream);
}

void mesh_prof_collect(mesh_ctx *ctx)
{
	for (auto &r : ctx->prof_pending)
	{
		float ms = 0.f;
		if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess)
		{
			ctx->prof_ms[r.kind] += ms;
			ctx->prof_cnt[r.kind] += 1;
		}
		ctx->prof_pool.push_back(r.a);
		ctx->prof_pool.push_back(r.b);
	}
	ctx->prof_pending.clear();
}

int mesh_finish_reduction(mesh_ctx *ctx, int nb, int nq, double *out)
{
	double *p = ctx->partials, *o = ctx->red_dev;
	hipStream_t st = ctx->stream;
	switch (nq)
	{
	case 1: k_reduce_partials<1><<<1, 1024, 0, st>>>(p, nb, o); break;
	case 2: k_reduce_partials<2><<<1, 1024, 0, st>>>(p, nb, o); break;
	case 3: k_reduce_partials<3><<<1, 1024, 0, st>>>(p, nb, o); break;
	case 4: k_reduce_partials<4><<<1, 1024, 0, st>>>(p, nb, o); break;
	case 5: k_reduce_partials<5><<<1, 1024, 0, st>>>(p, nb, o); break;
	case 6: k_reduce_partials<6><<<1, 1024, 0, st>>>(p, nb, o); break;
	case 7: k_reduce_partials<7><<<1, 1024, 0, st>>>(p, nb, o); break;
	case 8: k_reduce_partials<8><<<1, 1024, 0, st>>>(p, nb, o); break;
	case 18: k_reduce_partials<18><<<1, 1024, 0, st>>>(p, nb, o); break;
	default:
		return mesh_fail(ctx, MESH_E_ARG, "unsupported reduction width");
	}
	MESH_LAUNCHED(ctx);
	MESH_HIP(ctx, hipMemcpyAsync(ctx->red_host, ctx->red_dev, nq * sizeof(double), hipMemcpyDeviceToHost,
	                             ctx->stream));
	MESH_HIP(ctx, hipStreamSynchronize(ctx->stream));
	mesh_prof_collect(ctx);
	for (int q = 0; q < nq; ++q)
		out[q] = ctx->red_host[q];
	return MESH_OK;
}

int mesh_reduce_to_device(mesh_ctx *ctx, const double *partials, int nb, double *out_dev)
{
	k_reduce_partials<1><<<1, 1024, 0, ctx->stream>>>(partials, nb, out_dev);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

extern "C" {

int mesh_ctx_create(mesh_ctx **out, int device)
{
	if (!out)
		return MESH_E_ARG;
	*out = nullptr;
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
		return MESH_E_HIP;
	if (device < 0 || device >= n)
		return MESH_E_ARG;
	if (hipSetDevice(device) != hipSuccess)
		return MESH_E_HIP;
	mesh_ctx *ctx = new (std::nothrow) mesh_ctx();
	if (!ctx)
		return MESH_E_ALLOC;
	ctx->magic = MESH_CTX_MAGIC;
	ctx->device = device;
	if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess)
	{
		delete ctx;
		return MESH_E_HIP;
	}
	if (hipMalloc(&ctx->red_dev, RED_MAX_Q * sizeof(double)) != hipSuccess ||
	    hipHostMalloc(&ctx->red_host, RED_MAX_Q * sizeof(double), hipHostMallocDefault) != hipSuccess)
	{
		(void)hipStreamDestroy(ctx->stream);
		delete ctx;
		return MESH_E_ALLOC;
	}
	if (mesh_scratch_setup(ctx) != MESH_OK)
	{
		(void)hipFree(ctx->red_dev);
		(void)hipHostFree(ctx->red_host);
		(void)hipStreamDestroy(ctx->stream);
		delete ctx;
		return MESH_E_HIP;
	}
	*out = ctx;
	return MESH_OK;
}

int mesh_ctx_destroy(mesh_ctx *ctx)
{
	MESH_CHECK_CTX(ctx);
	(void)hipStreamSynchronize(ctx->stream);
	mesh_prof_collect(ctx);
	for (hipEvent_t e : ctx->prof_pool)
		(void)hipEventDestroy(e);
	if (ctx->partials)
		(void)hipFree(ctx->partials);
	if (ctx->aux)
		(void)hipFree(ctx->aux);
	if (ctx->md_spare)
		(void)hipFree(ctx->md_spare);
	if (ctx->u1_ws)
		(void)hipFree(ctx->u1_ws);
	(void)hipFree(ctx->red_dev);
	(void)hipHostFree(ctx->red_host);
	(void)hipStreamDestroy(ctx->stream);
	ctx->magic = 0;
	delete ctx;
	return MESH_OK;
}

const char *mesh_last_error(mesh_ctx *ctx)
{
	if (!mesh_ctx_ok(ctx))
		return "invalid mesh context";
	return ctx->err.c_str();
}

int mesh_synchronize(mesh_ctx *ctx)
{
	MESH_CHECK_CTX(ctx);
	MESH_HIP(ctx, hipStreamSynchronize(ctx->stream));
	mesh_prof_collect(ctx);
	return MESH_OK;
}

int mesh_profile_enable(mesh_ctx *ctx, int on)
{
	MESH_CHECK_CTX(ctx);
	ctx->prof_on = on != 0;
	return MESH_OK;
}

int mesh_profile_reset(mesh_ctx *ctx)
{
	MESH_CHECK_CTX(ctx);
	MESH_HIP(ctx, hipStreamSynchronize(ctx->stream));
	mesh_prof_collect(ctx);
	for (int k = 0; k < PROF_NKINDS; ++k)
	{
		ctx->prof_ms[k] = 0;
		ctx->prof_cnt[k] = 0;
	}
	return MESH_OK;
}

int mesh_profile_read(mesh_ctx *ctx, const char *kernel, double *total_ms, long long *count)
{
	MESH_CHECK_CTX(ctx);
	if (!kernel || !total_ms || !count)
		return mesh_fail(ctx, MESH_E_ARG, "NULL argument to mesh_profile_read");
	for (int k = 0; k < PROF_NKINDS; ++k)
		if (!std::strcmp(kernel, PROF_NAMES[k]))
		{
			*total_ms = ctx->prof_ms[k];
			*count = ctx->prof_cnt[k];
			return MESH_OK;
		}
	return mesh_fail(ctx, MESH_E_ARG, std::string("unknown profiled kernel ") + kernel);
}

void *mesh_stream(mesh_ctx *ctx) { return mesh_ctx_ok(ctx) ? (void *)ctx->stream : nullptr; }

int mesh_set_option(mesh_ctx *ctx, const char *name, long value)
{
	MESH_CHECK_CTX(ctx);
	if (!name)
		return mesh_fail(ctx, MESH_E_ARG, "option name is NULL");
	if (!std::strcmp(name, "xcd_remap") && value >= 0 && value <= 4)
		ctx->xcd_remap = (int)value;
	else if (!std::strcmp(name, "rb_tile") && value >= 1 && value <= 64)
		ctx->rb_tile = (int)value;
	else if (!std::strcmp(name, "force_variant") && value >= -1 && value <= 2)
		ctx->force_variant = (int)value;
	else if (!std::strcmp(name, "fuse_md"))
		ctx->fuse_md = value != 0;
	else if (!std::strcmp(name, "compact_momenta"))
		ctx->compact_p = value != 0;
	else if (!std::strcmp(name, "u1_general"))
		ctx->u1_general = value != 0;
	else if (!std::strcmp(name, "float_sums"))
		ctx->float_sums = value != 0;
	else
		return mesh_fail(ctx, MESH_E_ARG, std::string("unknown option ") + name);
	return MESH_OK;
}

int mesh_malloc(mesh_ctx *ctx, void **ptr, size_t bytes)
{
	MESH_CHECK_CTX(ctx);
	if (!ptr)
		return mesh_fail(ctx, MESH_E_ARG, "ptr is NULL");
	*ptr = nullptr;
	if (bytes == 0)
		return MESH_OK;
	if (hipMalloc(ptr, bytes) != hipSuccess)
	{
		*ptr = nullptr;
		(void)hipGetLastError();
		return mesh_fail(ctx, MESH_E_ALLOC, "device allocation of " + std::to_string(bytes) + " bytes failed");
	}
	return MESH_OK;
}

int mesh_free(mesh_ctx *ctx, void *ptr)
{
	MESH_CHECK_CTX(ctx);
	if (!ptr)
		return MESH_OK;
	MESH_HIP(ctx, hipStreamSynchronize(ctx->stream));
	MESH_HIP(ctx, hipFree(ptr));
	return MESH_OK;
}

int mesh_memcpy(mesh_ctx *ctx, void *dst, const void *src, size_t n)
{
	MESH_CHECK_CTX(ctx);
	if (n == 0)
		return MESH_OK;
	if (!dst || !src)
		return mesh_fail(ctx, MESH_E_ARG, "NULL pointer in mesh_memcpy");
	MESH_HIP(ctx, hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToDevice, ctx->stream));
	return MESH_OK;
}

int mesh_memcpy_2d(mesh_ctx *ctx, void *dst, size_t dpitch, const void *src, size_t spitch, size_t width,
                   size_t height)
{
	MESH_CHECK_CTX(ctx);
	if (width == 0 || height == 0)
		return MESH_OK;
	if (!dst || !src)
		return mesh_fail(ctx, MESH_E_ARG, "NULL pointer in mesh_memcpy_2d");
	MESH_HIP(ctx, hipMemcpy2DAsync(dst, dpitch, src, spitch, width, height, hipMemcpyDeviceToDevice, ctx->stream));
	return MESH_OK;
}

int mesh_memclear(mesh_ctx *ctx, void *ptr, size_t n)
{
	MESH_CHECK_CTX(ctx);
	if (n == 0)
		return MESH_OK;
	if (!ptr)
		return mesh_fail(ctx, MESH_E_ARG, "NULL pointer in mesh_memclear");
	MESH_HIP(ctx, hipMemsetAsync(ptr, 0, n, ctx->stream));
	return MESH_OK;
}

int mesh_memcpy_to_host(mesh_ctx *ctx, void *dst, const void *src, size_t n)
{
	MESH_CHECK_CTX(ctx);
	if (n == 0)
		return MESH_OK;
	if (!dst || !src)
		return mesh_fail(ctx, MESH_E_ARG, "NULL pointer in mesh_memcpy_to_host");
	MESH_HIP(ctx, hipMemcpyAsync(dst, src, n, hipMemcpyDeviceToHost, ctx->stream));
	MESH_HIP(ctx, hipStreamSynchronize(ctx->stream));
	return MESH_OK;
}

int mesh_memcpy_from_host(mesh_ctx *ctx, void *dst, const void *src, size_t n)
{
	MESH_CHECK_CTX(ctx);
	if (n == 0)
		return MESH_OK;
	if (!dst || !src)
		return mesh_fail(ctx, MESH_E_ARG, "NULL pointer in mesh_memcpy_from_host");
	MESH_HIP(ctx, hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, ctx->stream));
	MESH_HIP(ctx, hipStreamSynchronize(ctx->stream));
	return MESH_OK;
}

long long mesh_volume(const mesh_grid *g)
{
	Geo geo;
	if (mesh_make_geo(nullptr, g, &geo) != MESH_OK)
		return -1;
	return geo.V;
}

int mesh_site_block(void) { return MESH_SITE_LB ? (1 << MESH_SITE_LB) : 0; }

long long mesh_site_stride(const mesh_grid *g)
{
	Geo geo;
	if (mesh_make_geo(nullptr, g, &geo) != MESH_OK)
		return -1;
	return geo.stride;
}

size_t mesh_su3_field_bytes(const mesh_grid *g, int nf, int prec)
{
	Geo geo;
	if (mesh_make_geo(nullptr, g, &geo) != MESH_OK || nf < 1 || (prec != MESH_F64 && prec != MESH_F32))
		return 0;
	return (size_t)nf * 9 * geo.stride * 2 * (prec == MESH_F64 ? 8 : 4);
}

} // extern "C"

// ---- AoS <-> SoA --------------------------------------------------------------

template <class R> __global__ void k_aos_to_soa(const R *aos, R *soa, Geo g, int nf)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x; // matrix index f*V + x
	if (i >= g.V * nf)
		return;
	int f = (int)(i / g.V);
	long x = lane_site(g, i - (long)f * g.V);
	i = (long)f * g.V + x;
	m3<R> m;
#pragma unroll
	for (int k = 0; k < 9; ++k)
		m.e[k] = {aos[i * 18 + 2 * k], aos[i * 18 + 2 * k + 1]};
	store(soa, f, g, x, m);
}

template <class R> __global__ void k_soa_to_aos(const R *soa, R *aos, Geo g, int nf)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= g.V * nf)
		return;
	int f = (int)(i / g.V);
	long x = lane_site(g, i - (long)f * g.V);
	i = (long)f * g.V + x;
	m3<R> m = load(soa, f, g, x);
#pragma unroll
	for (int k = 0; k < 9; ++k)
	{
		aos[i * 18 + 2 * k] = m.e[k].re;
		aos[i * 18 + 2 * k + 1] = m.e[k].im;
	}
}

// 16 B per lane streaming copy (bandwidth reference)
__global__ __launch_bounds__(256) void k_stream_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, long n)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		dst[i] = src[i];
}

extern "C" int mesh_stream_copy(mesh_ctx *ctx, void *dst, const void *src, size_t bytes)
{
	MESH_CHECK_CTX(ctx);
	if (!dst || !src || bytes % 16)
		return mesh_fail(ctx, MESH_E_ARG, "stream_copy needs non-NULL pointers and a multiple of 16 bytes");
	long n = (long)(bytes / 16);
	if (n == 0)
		return MESH_OK;
	k_stream_copy<<<(int)((n + 255) / 256), 256, 0, ctx->stream>>>((const uint4 *)src, (uint4 *)dst, n);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

template <class D, class S> __global__ void k_convert(const S *src, D *dst, long n)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		dst[i] = (D)src[i];
}

template <class R> __global__ void k_unit(R *dst, Geo g, int nf)
{
	long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= g.V * nf)
		return;
	int f = (int)(i / g.V);
	store(dst, f, g, lane_site(g, i - (long)f * g.V), unit_m3<R>());
}

template <class R>
static int transfer(mesh_ctx *ctx, const mesh_grid *g, int nf, const void *host, void *dev, bool up)
{
	Geo geo;
	int rc = mesh_make_geo(ctx, g, &geo);
	if (rc)
		return rc;
	if (nf < 1 || !host || !dev)
		return mesh_fail(ctx, MESH_E_ARG, "bad field transfer arguments");
	size_t bytes = (size_t)nf * geo.V * 18 * sizeof(R);
	R *tmp = nullptr;
	if (hipMallocAsync((void **)&tmp, bytes, ctx->stream) != hipSuccess)
		return mesh_fail(ctx, MESH_E_ALLOC, "cannot allocate transfer staging buffer");
	long n = (long)nf * geo.V;
	int nb = (int)((n + 255) / 256);
	if (up)
	{
		MESH_HIP(ctx, hipMemcpyAsync(tmp, host, bytes, hipMemcpyHostToDevice, ctx->stream));
		k_aos_to_soa<R><<<nb, 256, 0, ctx->stream>>>(tmp, (R *)dev, geo, nf);
		MESH_LAUNCHED(ctx);
	}
	else
	{
		k_soa_to_aos<R><<<nb, 256, 0, ctx->stream>>>((const R *)dev, tmp, geo, nf);
		MESH_LAUNCHED(ctx);
		MESH_HIP(ctx, hipMemcpyAsync((void *)host, tmp, bytes, hipMemcpyDeviceToHost, ctx->stream));
	}
	MESH_HIP(ctx, hipFreeAsync(tmp, ctx->stream));
	MESH_HIP(ctx, hipStreamSynchronize(ctx->stream));
	return MESH_OK;
}

extern "C" {

int mesh_su3_upload(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, const void *host, void *dev)
{
	MESH_CHECK_CTX(ctx);
	if (int rc = mesh_check_prec(ctx, prec))
		return rc;
	return prec == MESH_F64 ? transfer<double>(ctx, g, nf, host, dev, true)
	                        : transfer<float>(ctx, g, nf, host, dev, true);
}

int mesh_su3_download(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, const void *dev, void *host)
{
	MESH_CHECK_CTX(ctx);
	if (int rc = mesh_check_prec(ctx, prec))
		return rc;
	return prec == MESH_F64 ? transfer<double>(ctx, g, nf, host, (void *)dev, false)
	                        : transfer<float>(ctx, g, nf, host, (void *)dev, false);
}

int mesh_su3_convert(mesh_ctx *ctx, const mesh_grid *g, int nf, int dst_prec, void *dst, int src_prec,
                     const void *src)
{
	MESH_CHECK_CTX(ctx);
	Geo geo;
	if (int rc = mesh_make_geo(ctx, g, &geo))
		return rc;
	if (int rc = mesh_check_prec(ctx, dst_prec))
		return rc;
	if (int rc = mesh_check_prec(ctx, src_prec))
		return rc;
	if (!dst || !src || nf < 1)
		return mesh_fail(ctx, MESH_E_ARG, "bad convert arguments");
	long n = (long)nf * 9 * geo.stride * 2;
	int nb = (int)((n + 255) / 256);
	if (dst_prec == src_prec)
		return mesh_memcpy(ctx, dst, src, n * (dst_prec == MESH_F64 ? 8 : 4));
	if (dst_prec == MESH_F64)
		k_convert<double, float><<<nb, 256, 0, ctx->stream>>>((const float *)src, (double *)dst, n);
	else
		k_convert<float, double><<<nb, 256, 0, ctx->stream>>>((const double *)src, (float *)dst, n);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

int mesh_su3_unit(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, void *dst)
{
	MESH_CHECK_CTX(ctx);
	Geo geo;
	if (int rc = mesh_make_geo(ctx, g, &geo))
		return rc;
	if (int rc = mesh_check_prec(ctx, prec))
		return rc;
	if (!dst || nf < 1)
		return mesh_fail(ctx, MESH_E_ARG, "bad unit-field arguments");
	long n = (long)nf * geo.V;
	int nb = (int)((n + 255) / 256);
	if (prec == MESH_F64)
		k_unit<double><<<nb, 256, 0, ctx->stream>>>((double *)dst, geo, nf);
	else
		k_unit<float><<<nb, 256, 0, ctx->stream>>>((float *)dst, geo, nf);
	MESH_LAUNCHED(ctx);
	return MESH_OK;
}

} // extern "C"
