/*
 * mesh_hip.h -- C ABI of libmesh_hip.so, the MI355X (gfx950) lattice library.
 *
 * Plain C: opaque context, plain pointers and sizes, int status codes.
 * Device buffers are allocated through mesh_malloc (or by any HIP allocator)
 * and hold fields in the structure-of-arrays layout described below.
 *
 * Each entry point names the reference (krox/mesh) interface it replaces;
 * paths are relative to the reference root.
 *
 * ---------------------------------------------------------------------------
 * Device layout of an SU(3) field with `nf` matrices per site (nf = Nd for a
 * gauge or momentum field, 1 for a single-direction lattice such as a staple
 * sum, a force or a gauge transformation):
 *
 *     matrix f < nf, complex entry k < 9 (row-major, k = 3 i + j), site x:
 *         elem(f, k, x) = base[f * 9 * S + (x / B) * 9 * B + k * B + slot(x % B)]
 *         slot(l) = (l % 2) * B / 2 + l / 2
 *     (complex: re, im), S = mesh_site_stride(grid) = V rounded up to a
 *     multiple of 256, B = mesh_site_block() (a power of two <= 256).
 *
 * i.e. each direction holds blocks of B consecutive sites; a block stores the
 * 9 complex entries as 9 planes of B sites (the block's even sites first,
 * then its odd sites, so each checkerboard parity of a row with an even
 * fastest extent is contiguous), and the lanes of a wavefront over B
 * consecutive sites read one contiguous B * 16 B (FP64) / B * 8 B (FP32)
 * segment per entry.  Sites
 * are row-major with dimension 0 slowest (src/lattice/grid.h:39-49).  Matrix
 * f of every site is the nf = 1 field at base + f * 9 * S complex elements.
 *
 * Host layout (mesh_*_upload / download): the reference's AoS order,
 * host[f][x][3][3] complex interleaved (src/lattice/lattice.h:159-192: one
 * buffer per Lorentz direction, one Matrix<complex,3> per site).
 *
 * Precision: MESH_F64 stores and computes in double.  MESH_F32 stores links
 * and momenta in float and computes the molecular dynamics in float; the
 * per-site action, kinetic energy, Delta H and every global sum are computed
 * in double in both modes.
 *
 * Every call is asynchronous on the context's stream unless it returns a
 * host scalar (reductions), in which case it synchronises that stream.
 * ---------------------------------------------------------------------------
 */
#ifndef MESH_HIP_H
#define MESH_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (src/lattice/device.cpp:10-19 throws instead) */
enum
{
	MESH_OK = 0,
	MESH_E_ARG = -1,   /* null pointer, bad direction, bad scheme ... */
	MESH_E_SIZE = -2,  /* unsupported geometry or size */
	MESH_E_ALLOC = -3, /* device allocation failed */
	MESH_E_HIP = -4,   /* HIP runtime error (message in mesh_last_error) */
	MESH_E_STATE = -5  /* invalid context */
};

enum
{
	MESH_F64 = 0,
	MESH_F32 = 1
};

typedef struct mesh_ctx mesh_ctx;

/* Lattice geometry: nd in [2, 4], dims[0..nd-1] >= 2 (src/lattice/grid.h:13-58). */
typedef struct
{
	int nd;
	int dims[4];
} mesh_grid;

/* ---- context ------------------------------------------------------------ */

/* Creates a context on HIP device `device` with its own stream and reduction
 * scratch.  Replaces the implicit global CUDA state of
 * src/lattice/device.cpp:1-76. */
int mesh_ctx_create(mesh_ctx **ctx, int device);
int mesh_ctx_destroy(mesh_ctx *ctx);
/* Message of the last failed call on this context ("" if none). */
const char *mesh_last_error(mesh_ctx *ctx);
/* Waits for all work queued on the context (device_synchronize,
 * src/lattice/device.cpp:21). */
int mesh_synchronize(mesh_ctx *ctx);
/* The context's hipStream_t, for callers that order their own work. */
void *mesh_stream(mesh_ctx *ctx);
/* Kernel launch tuning (no effect on results):
 *   "xcd_remap" = workgroup order of the stencil kernels: 0 linear, 1 each
 *       XCD takes a contiguous site range, 2 each XCD takes a range of x2/x3
 *       row-blocks and walks x0 fastest, then x1; 3 (default) each XCD takes
 *       a slab of L0/8 x0-slices and walks it x0 fastest, then "rb_tile"
 *       row-blocks, then x1, then the next tile (4-d lattices with 8 | L0;
 *       otherwise order 4, 2 or 1 applies); 4 as 3 with XCD k owning a
 *       quarter of x0 (k % 4) and a half of x1 (k / 4) (4 | L0, 2 | L1;
 *       otherwise order 2 or 1 applies);
 *   "rb_tile" = row-blocks per tile of order 3 (default 4; used when it
 *       divides the row-blocks of an x2/x3 plane);
 *   "force_variant" = -1 (default: 0); 0 wavefront per direction with the
 *       block's links staged in LDS, 1 thread per site, 2 LDS-staged x2/x3
 *       tile with halo rows;
 *   "fuse_md" = 1 (default): in mesh_su3_hmc_trajectory each momentum step
 *       and the link step after it run as one kernel (the thread that
 *       updates P_mu(x) writes exp(t P_mu(x)) U_mu(x) into a second field),
 *       and so do the momenta refresh and the first link step; same
 *       arithmetic, so the same bits;
 *   "u1_general" = 0 (default): U(1) lattices that fit use the LDS-resident
 *       kernels; 1 sends every U(1) lattice through the site-parallel ones;
 *   "compact_momenta" = 1 (default): the fused trajectory keeps its momenta
 *       in a compact antihermitian layout of 9 reals per link (upper
 *       off-diagonal entries and the imaginary diagonal); 0 keeps the
 *       18-real matrix layout.  The momenta are exactly antihermitian, so
 *       both give the same bits;
 *   "float_sums" = 0 (default): 1 makes FP32 trajectories form Delta H as
 *       the reference's float build (--precision 1) does: H_old and H_new
 *       each from FP32 sums over the lattice (mesh_su3_float_energy), then
 *       differenced.  A comparison baseline: its error grows with the volume
 *       (the default FP64 per-site differenced sums do not). */
int mesh_set_option(mesh_ctx *ctx, const char *name, long value);
/* Per-kernel timing (replaces the util::Stopwatch globals swStaples, swExp,
 * swPlaquette ... of src/gauge/utils.h:23, src/gauge/hmc.h:16, printed at
 * src/apps/hmc.cpp:138-155): when enabled, launches of the HMC and Landau
 * kernels are bracketed by HIP events on the context stream and resolved at
 * the next synchronisation.  Kernel names: "force_update", "link_update",
 * "delta_h", "momenta", "plaquette", "landau_or", "force_link" (fused
 * momentum + link step), "momenta_link" (momenta refresh fused with the
 * first link step of a trajectory). */
int mesh_profile_enable(mesh_ctx *ctx, int on);
int mesh_profile_reset(mesh_ctx *ctx);
int mesh_profile_read(mesh_ctx *ctx, const char *kernel, double *total_ms, long long *count);

/* ---- memory: the untyped backend of src/lattice/device.h:14-24 ---------- */

int mesh_malloc(mesh_ctx *ctx, void **ptr, size_t bytes);     /* device_malloc  device.cpp:23-31 */
int mesh_free(mesh_ctx *ctx, void *ptr);                      /* device_free    device.cpp:33-38 */
int mesh_memcpy(mesh_ctx *ctx, void *dst, const void *src, size_t bytes); /* device_memcpy device.cpp:40-45 */
int mesh_memcpy_2d(mesh_ctx *ctx, void *dst, size_t dpitch, const void *src, size_t spitch,
                   size_t width, size_t height);              /* device_memcpy_2d device.cpp:47-54 */
int mesh_memclear(mesh_ctx *ctx, void *ptr, size_t bytes);    /* device_memclear device.cpp:70-75 */
int mesh_memcpy_to_host(mesh_ctx *ctx, void *dst, const void *src, size_t bytes);   /* device.cpp:56-61 */
int mesh_memcpy_from_host(mesh_ctx *ctx, void *dst, const void *src, size_t bytes); /* device.cpp:63-68 */
/* Device-to-device copy by a plain 16 B-per-lane kernel (bytes % 16 == 0):
 * the streaming-bandwidth reference of the roofline figures (the reference
 * measures memory bandwidth with host loops, src/apps/ompbench.cpp:52-110). */
int mesh_stream_copy(mesh_ctx *ctx, void *dst, const void *src, size_t bytes);

/* ---- geometry and SU(3) field transfer ---------------------------------- */

long long mesh_volume(const mesh_grid *g);
long long mesh_site_stride(const mesh_grid *g);
/* site-block length B of the device layout, fixed when the library is
 * built; 0 means B = S, i.e. plain structure of arrays */
int mesh_site_block(void);
/* bytes of an SU(3) field with nf matrices per site in precision prec */
size_t mesh_su3_field_bytes(const mesh_grid *g, int nf, int prec);
/* host AoS (double for MESH_F64, float for MESH_F32) <-> device SoA;
 * replaces Lattice::copy_from_host / copy_to_host (src/lattice/lattice.h:102-115) */
int mesh_su3_upload(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, const void *host, void *dev);
int mesh_su3_download(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, const void *dev, void *host);
/* dst = src converted between precisions (same grid and nf) */
int mesh_su3_convert(mesh_ctx *ctx, const mesh_grid *g, int nf, int dst_prec, void *dst, int src_prec,
                     const void *src);
/* dst[x] = 1 for every matrix (cold start, Landau::reset src/gauge/landau.h:14) */
int mesh_su3_unit(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, void *dst);

/* ---- random fields (Philox4x32-10, counter = (site, dir, draw, stream)) --- */

/* Haar-random SU(3) field: randomizeGaugeField / random_gauge_field
 * (src/gauge/hmc.cpp:60-63, src/gauge/utils.h:43-52, groups.h:255-261).
 * nf matrices per site. */
int mesh_su3_random_gauge(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, void *U, uint64_t seed,
                          uint32_t stream);
/* Gaussian su(3) momenta P = sum_a p_a i T^a, p_a ~ N(0,1), for trajectory
 * `traj`; K_sites (device, V doubles, may be NULL) receives
 * k(x) = sum_mu ||P_mu(x)||_F^2.  randomizeMomenta (src/gauge/hmc.cpp:65-71,
 * utils.h:54-62, groups.h:245-253). */
int mesh_su3_momenta(mesh_ctx *ctx, const mesh_grid *g, int prec, void *P, double *K_sites, uint64_t seed,
                     uint32_t traj);

/* ---- observables --------------------------------------------------------- */

/* *out = normalised average plaquette in [0,1]: plaquette(U)
 * (src/gauge/utils.h:93-108). */
int mesh_su3_plaquette(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U, double *out);
/* The value of mesh_su3_plaquette, bit for bit, written to device memory
 * (out_dev: one double) by an entry point that allocates nothing.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The per-block partial sums live in a fixed
 * 512 KiB device array, one per GPU, held until the process ends: lattices
 * up to 2^24 sites (64^4); larger ones return MESH_E_SIZE.  Asynchronous on
 * the context's stream (no copy to the host, no synchronisation).  Every
 * call on one GPU shares that array: calls that overlap on one GPU (from
 * several contexts or streams) race. */
int mesh_su3_plaquette_noalloc(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U, double *out_dev);
/* *out = Wilson action beta * sum_P (1 - Re tr P / 3): wilsonAction
 * (src/gauge/utils.h:141-146). */
int mesh_su3_wilson_action(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U, double beta,
                           double *out);
/* *out = sum_{x,mu} ||P_mu(x)||_F^2: norm2(GaugeField) (src/gauge/utils.h:35-41) */
int mesh_su3_norm2(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, const void *P, double *out);
/* *out = sum_{x,mu} Re tr A_mu(x) B_mu(x): sum_real_trace(A*B) (utils.h:29) */
int mesh_su3_sum_real_trace_prod(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, const void *A,
                                 const void *B, double *out);

/* ---- molecular dynamics --------------------------------------------------- */

/* S = staple sum of direction mu (1 matrix per site): stapleSum(U, mu)
 * (src/gauge/utils.h:110-139). */
int mesh_su3_staple(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U, int mu, void *S);
/* F = (beta/6) TA(U_mu S_mu) (1 matrix per site): wilsonDeriv(U, mu, beta)
 * (src/gauge/utils.h:148-153). */
int mesh_su3_force(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U, int mu, double beta,
                   void *F);
/* Momentum step, all directions: P_mu -= delta (beta/6) TA(U_mu S_mu),
 * fused staple + force + update; the loop body of runHmd's odd steps
 * (src/gauge/hmc.h:54-58). */
int mesh_su3_force_update(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U, void *P, double beta,
                          double delta);
/* Link step, all directions: Uout_mu = exp(t TA(P_mu)) Uin_mu, then
 * project_on_group_fast if reunit != 0; evolve (src/gauge/hmc.h:26-36) fused
 * with reunitize (src/gauge/utils.h:80-91).  Uout may equal Uin. */
int mesh_su3_link_update(mesh_ctx *ctx, const mesh_grid *g, int prec, void *Uout, const void *Uin,
                         const void *P, double t, int reunit);
/* U = project_on_group_fast(U) for nf matrices per site: reunitize
 * (src/gauge/utils.h:80-91, groups.h:237-243). */
int mesh_su3_reunit(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, void *U);
/* U = exp(t TA(F)) U for nf = 1 fields (gauge transformation update,
 * src/gauge/landau.h:131). */
int mesh_su3_exp_mul(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, void *U, const void *F, double t);
/* Y += a X over nf matrices per site, the product rounded before the sum:
 * add_mul / `lhs += rhs * a` on lattices (src/lattice/lattice.h:302-309,
 * 349-363), e.g. the conjugate-gradient update of landau.h:128. */
int mesh_su3_axpy(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, void *Y, double a, const void *X);
/* dst[x] = src[x + offset e_dir], periodic, nf matrices per site; dst != src,
 * 0 <= dir < nd, 0 < |offset| < L[dir]: cshift (src/lattice/lattice.h:366-389,
 * device_cshift src/lattice/device.h:55-110). */
int mesh_su3_cshift(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, void *dst, const void *src, int dir,
                    int offset);
/* out[18] = sum over sites (and the nf matrices) of every complex entry,
 * (re, im) row-major, summed in FP64: lattice_sum (src/lattice/lattice.h:270-273,
 * device_sum src/lattice/device.h:238-265, which sums in the field's own type). */
int mesh_su3_sum(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, const void *src, double *out);

/* out[5] = {Delta H, sum_x ps_old(x), sum_x ps_new(x), K_old, K_new} with
 * per-site differencing d(x) = -(beta/3)(ps_new - ps_old)(x) + (k_new - k_old)(x),
 * ps(x) = sum_{mu<nu} Re tr of the plaquettes at x, computed in double.
 * Replaces H_new - H_old of src/gauge/hmc.cpp:85-90. */
int mesh_su3_delta_h(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U_old, const void *U_new,
                     const void *P, const double *K_old_sites, double beta, double *out);

/* Integrator step table: makeDeltas (src/gauge/hmc.cpp:7-51).  scheme is
 * "2lf", "2mn" or "4mn"; returns the length (> 0) or MESH_E_ARG; writes at
 * most cap values. */
int mesh_make_deltas(const char *scheme, double epsilon, int substeps, double *out, int cap);

/* One HMC trajectory: runHmcUpdate (src/gauge/hmc.cpp:75-107).
 *   *U      current field (device), *U_work a second field of the same size;
 *           on acceptance the two pointers are swapped (std::swap, hmc.cpp:96).
 *   P       momentum field, K_sites V doubles (device), both workspace
 *           (P needs mesh_su3_field_bytes(g, Nd, prec); the fused path keeps
 *           it in a compact antihermitian layout of 9 reals per link, so its
 *           contents on return are not a field in the layout above).
 *   metropolis == 0 accepts unconditionally (thermalisation).
 * out[6] = {Delta H, accepted, plaquette of the resulting U, H_old, H_new, u}.
 * The Metropolis uniform u is Philox(seed, traj); accept iff u < exp(-Delta H). */
int mesh_su3_hmc_trajectory(mesh_ctx *ctx, const mesh_grid *g, int prec, void **U, void **U_work, void *P,
                            double *K_sites, double beta, const double *deltas, int ndeltas, uint64_t seed,
                            uint32_t traj, int metropolis, double *out);

/* ---- Landau gauge (src/gauge/landau.h) ------------------------------------ */

/* *out = (1/(3 Nd V)) sum_{x,mu} Re tr h(x) U_mu(x) h(x+mu)^+; h == NULL
 * means h = 1.  Landau::operator() (src/gauge/landau.h:23-34). */
int mesh_landau_functional(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U, const void *h,
                           double *out);
/* *out = sum_x ||Delta(x)||_F^2, Delta(x) = sum_mu TA(U^h_mu(x)) - TA(U^h_mu(x-mu)).
 * Landau::condition() = *out / (Nd V) (src/gauge/landau.h:37-49); the
 * Landau condition theta = (1/V) sum_x |div A|^2 = *out / V. */
int mesh_landau_divsum(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U, const void *h,
                       double *out);
/* F(x) = -0.5/(3 Nd V) TA(h(x) d(x)), d(x) = sum_mu U_mu(x) h(x+mu)^+ + (h(x-mu) U_mu(x-mu))^+:
 * Landau::deriv (src/gauge/landau.h:52-65). */
int mesh_landau_deriv(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U, const void *h, void *F);
/* One overrelaxation half-sweep over sites of the given parity, in place on
 * the gauge-transformed links U (and on h if not NULL): replaces the update
 * step of Landau::run (src/gauge/landau.h:100-136); closest reference
 * ancestor is src/old/gauge_fixing.cpp:33-41.  Needs even dims. */
int mesh_landau_or_halfsweep(mesh_ctx *ctx, const mesh_grid *g, int prec, void *U, void *h, int parity,
                             double omega);
/* U_mu(x) <- h(x) U_mu(x) h(x+mu)^+ (out of place; Uout != U). */
int mesh_su3_gauge_transform(mesh_ctx *ctx, const mesh_grid *g, int prec, void *Uout, const void *U,
                             const void *h);

/* ---- U(1) (src/gauge/groups.h:29-108, hmc.cpp:109-110) ----------------------
 *
 * A batch of `batch` independent configurations ("chains") of a periodic
 * lattice of 2 to 4 dimensions, FP32 angles theta[c][mu][x] in [-pi, pi), sites
 * row-major (dimension 0 slowest; x = x0 * L1 + x1 in 2-d).  theta_{mu nu}(x) =
 * theta_mu(x) + theta_nu(x+mu) - theta_mu(x+nu) - theta_nu(x);
 * S = beta sum_x sum_{mu<nu} (1 - cos theta_{mu nu}); on 2-d lattices the
 * topological charge is Q = sum_x wrap(theta_01(x)) / 2 pi.  Device buffers:
 * theta Nd V floats per configuration; per-configuration outputs are device
 * doubles.  2-d lattices of up to 4096 sites (HMC) / 8192 sites (random start,
 * observables) run in LDS-resident kernels, one workgroup per configuration;
 * everything else in site-parallel kernels with the same arithmetic per link
 * (option "u1_general" = 1 selects these for every lattice). */

/* uniform (Haar) random angles: random_gauge_field for U1 (utils.h:43-52) */
int mesh_u1_random(mesh_ctx *ctx, const mesh_grid *g, int batch, float *theta, uint64_t seed);
/* out_dev[4c + {0,1,2,3}] = {sum (1 - cos theta_P), Q, sum cos theta_P, sum wrap(theta_P)} over
 * sites and planes; Q = 0 on lattices of more than 2 dimensions */
int mesh_u1_observables(mesh_ctx *ctx, const mesh_grid *g, int batch, const float *theta, double *out_dev);
/* One HMC trajectory per configuration (runHmcUpdate, hmc.cpp:75-107): momenta
 * q ~ N(0, 1/2) per link, K = sum q^2, link step theta += t q, momentum step
 * q -= t (beta/2) Im(U S), accept on u < exp(-Delta H) (metropolis != 0).
 * deltas: host step table (mesh_make_deltas), odd length.
 * out_dev[4c + {0,1,2,3}] = {Delta H, accepted, u, S_new / beta}. */
int mesh_u1_hmc(mesh_ctx *ctx, const mesh_grid *g, int batch, float *theta, const double *deltas, int ndeltas,
                double beta, uint64_t seed, uint32_t traj, int metropolis, double *out_dev);

/* ---- U(1) learned proposal (no counterpart in the reference) ---------------
 *
 * Gauge-equivariant coupling-layer flow of `layers` layers (8 in the
 * specified design, 1 to 64), each moving one direction's links on every
 * fourth row/column by h(theta_P) - theta_P, with h a mixture of non-compact
 * projections whose parameters an MLP (12 -> 64 -> 64 -> 10, tanh) computes
 * from the frozen plaquettes around the active one; layer l has direction
 * l % 2 and row offset (l / 2) % 4.  Lattice extents divisible by 4,
 * V <= 4096.  params: mesh_flow_num_params(layers) floats (device). */
int mesh_flow_num_params(int layers);
/* deterministic initial weights (near the identity transform: last layer
 * 1e-2 of the fan-in scale, bias of the identity) */
int mesh_flow_init(mesh_ctx *ctx, int layers, float *params, uint64_t seed);
/* batch samples U = f(z), z uniform from Philox(seed, step):
 * out_dev[2c + {0,1}] = {log q(U), S(U)/beta}; theta_layers may be NULL */
int mesh_flow_sample(mesh_ctx *ctx, const mesh_grid *g, int batch, int layers, const float *params, uint64_t seed,
                     uint32_t step, float *theta_out, float *theta_layers, double *out_dev);
/* log q of given configurations by the inverse flow (z_out may be NULL):
 * out_dev[2c + {0,1}] = {log q(U), S(U)/beta} */
int mesh_flow_logq(mesh_ctx *ctx, const mesh_grid *g, int batch, int layers, const float *params, const float *theta,
                   float *z_out, double *out_dev);
/* bytes of the workspace mesh_flow_gradient needs */
size_t mesh_flow_workspace_bytes(const mesh_grid *g, int batch, int layers);
/* reverse-KL gradient: grad = (1/batch) sum_c d/dparams [beta S(f(z_c)) - log det f'(z_c)]
 * over a fresh batch (Philox(seed, step)); out_dev as mesh_flow_sample */
int mesh_flow_gradient(mesh_ctx *ctx, const mesh_grid *g, int batch, int layers, const float *params, uint64_t seed,
                       uint32_t step, double beta, void *workspace, float *grad, double *out_dev);
/* Adam step t >= 1 (moments m, v zero-initialised by the caller) */
int mesh_flow_adam(mesh_ctx *ctx, int layers, float *params, const float *grad, float *m, float *v, int t, double lr);
/* independence Metropolis-Hastings: per chain accept the proposal with
 * probability min(1, exp[(-beta S' - log q') - (-beta S - log q)]); state and
 * prop hold {log q, S/beta} per chain; acc_dev[c] = 1 on acceptance.  Any 2-d
 * lattice (proposals of either flow). */
int mesh_flow_mh(mesh_ctx *ctx, const mesh_grid *g, int batch, float *theta, double *state, const float *prop_theta,
                 const double *prop, double beta, uint64_t seed, uint32_t step, double *acc_dev);

/* out[2] = {wilsonAction(U, beta), norm2(P)} of FP32 fields as the reference's
 * float build computes them (src/gauge/utils.h:94-108, 141-146, 27-41): each
 * plane's Re tr P and each direction's ||P_mu||^2 summed in FP32 by a 256-wide
 * pair tree per block and a sequential FP32 sum of the blocks
 * (device_sum_apply, src/lattice/device.h:184-210), accumulated in FP64 over
 * planes and directions.  P may be NULL (out[1] = 0); p_compact: P in the
 * compact layout of the fused trajectory. */
int mesh_su3_float_energy(mesh_ctx *ctx, const mesh_grid *g, const float *U, const float *P, int p_compact,
                          double beta, double *out);

/* ---- SU(2) pure-gauge HMC (Hmc<SU2<T>>, src/gauge/hmc.cpp:111-114) --------
 *
 * Links are unit quaternions (w, x, y, z), the 2x2 matrices w - i (x, y, z).sigma
 * -- the reference's SU2<T> =
 * util::quaternion<T> (src/gauge/groups.h:114-196) -- multiplied as Hamilton
 * products; adj = conjugate, Re tr = 2 w, project_on_algebra keeps (0, x, y, z).
 * Device fields hold each direction's quaternions in site order, component c
 * of direction f at base[4 (f S + x) + c] with S = V rounded up to 256
 * (mesh_su2_field_bytes); host arrays are [nf][V][4] in that component order.
 * A quaternion is read with vector loads, so device field pointers must be
 * aligned to 16 B (FP32) / 32 B (FP64) -- mesh_malloc's are; others are
 * refused with MESH_E_ARG.  Momenta are (0, p/2)
 * with p ~ N(0,1) (random_algebra_element, groups.h:171-181) and the kinetic
 * energy is 2 (x^2 + y^2 + z^2), the 2x2 matrix norm, which the force
 * (beta/4) project_on_algebra(U Sigma) conserves (the reference adds the
 * quaternion norm, half of it).  exp of an algebra element is
 * (cos a, v sin(a)/a) as in groups.h:159-169, with the limit at a = 0. */
size_t mesh_su2_field_bytes(const mesh_grid *g, int nf, int prec);
/* host <-> device copies of nf directions; upload != 0: host to device */
int mesh_su2_transfer(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, void *host, void *dev, int upload);
/* Haar-random links: four Philox normals normalised (random_group_element, groups.h:183-192) */
int mesh_su2_random(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, void *U, uint64_t seed);
/* the unit field */
int mesh_su2_unit(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, void *U);
/* momenta of trajectory `traj` (randomizeMomenta); K (device, V doubles, may be NULL) gets k(x) */
int mesh_su2_momenta(mesh_ctx *ctx, const mesh_grid *g, int prec, void *P, double *K, uint64_t seed, uint32_t traj);
/* P_mu -= delta (beta/4) project_on_algebra(U_mu Sigma_mu) for every mu (runHmd's momentum step) */
int mesh_su2_force_update(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U, void *P, double beta,
                          double delta);
/* Uout = exp(t project_on_algebra(P)) Uin (evolve, src/gauge/hmc.h:26-36), reunitarised if asked */
int mesh_su2_link_update(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, void *Uout, const void *Uin,
                         const void *P, double t, int reunit);
/* project_on_group_fast on every link (reunitize) */
int mesh_su2_reunit(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, void *U);
/* *out = mean Re tr P / 2 over all plaquettes (plaquette, src/gauge/utils.h:93-108) */
int mesh_su2_plaquette(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U, double *out);
/* out[5] = {Delta H, sum_x ps_old, sum_x ps_new, K_old, K_new}, ps(x) = sum of Re P / (Re tr / 2) at x */
int mesh_su2_delta_h(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U_old, const void *U_new,
                     const void *P, const double *K_old, double beta, double *out);
/* one trajectory as mesh_su3_hmc_trajectory: out[6] = {Delta H, accepted,
 * plaquette of the resulting U, H_old, H_new, u}; *U and *U_work swap on
 * acceptance */
int mesh_su2_hmc_trajectory(mesh_ctx *ctx, const mesh_grid *g, int prec, void **U, void **U_work, void *P, double *K,
                            double beta, const double *deltas, int ndeltas, uint64_t seed, uint32_t traj,
                            int metropolis, double *out);

/* ---- Landau gauge for SU(2) (Landau<SU2<T>>, src/gauge/landau.h; instantiated
 * by src/tests/lattice_gauge.cpp:46-53).  h is the gauge field, one quaternion
 * per site (nf = 1); h == NULL means h = 1 where allowed.  Sums in FP64. */
/* *out = sum_x,mu Re tr (h U_mu h(x+mu)^+) / (Nc Nd V) (operator(), landau.h:23-34) */
int mesh_su2_landau_functional(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U, const void *h,
                               double *out);
/* *out = sum_x |project_on_algebra(h d)|^2; condition() = *out / (Nd V) (landau.h:37-49) */
int mesh_su2_landau_divsum(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U, const void *h, double *out);
/* F = -0.5 / (Nc Nd V) project_on_algebra(h d) (deriv, landau.h:52-65) */
int mesh_su2_landau_deriv(mesh_ctx *ctx, const mesh_grid *g, int prec, const void *U, const void *h, void *F);
/* Uout_mu(x) = h(x) U_mu(x) h(x+mu)^+ */
int mesh_su2_gauge_transform(mesh_ctx *ctx, const mesh_grid *g, int prec, void *Uout, const void *U, const void *h);
/* overrelaxation of the sites of one parity on the transformed links U (in
 * place), h = G h when h != NULL; replaces the ascent loop of landau.h:100-136
 * as mesh_landau_or_halfsweep does for SU(3); needs even extents */
int mesh_su2_landau_or_halfsweep(mesh_ctx *ctx, const mesh_grid *g, int prec, void *U, void *h, int parity,
                                 double omega);
/* *out = sum Re tr (A B) over nf directions (sum_real_trace(A * B), src/gauge/utils.h:27-33) */
int mesh_su2_dot(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, const void *A, const void *B, double *out);
/* Y += a X (the F += last * beta of landau.h:128) */
int mesh_su2_axpy(mesh_ctx *ctx, const mesh_grid *g, int nf, int prec, void *Y, double a, const void *X);

/* ---- Landau gauge for U(1) (Landau<U1<T>>, src/gauge/landau.h) on angles:
 * links theta[mu][x] and the gauge field g(x) = exp(i phi(x)), phi[x], FP32;
 * values in FP64 from the angles, rounded once; phi == NULL means phi = 0. */
/* *out = sum_x,mu cos(phi(x) + theta_mu(x) - phi(x+mu)) / (Nd V) (landau.h:23-34) */
int mesh_u1_landau_functional(mesh_ctx *ctx, const mesh_grid *g, const float *theta, const float *phi, double *out);
/* *out = sum_x Im(g d)^2; condition() = *out / (Nd V) (landau.h:37-49) */
int mesh_u1_landau_divsum(mesh_ctx *ctx, const mesh_grid *g, const float *theta, const float *phi, double *out);
/* F[x] = -0.5 / (Nd V) Im(g d)(x), the algebra element i F (deriv, landau.h:52-65) */
int mesh_u1_landau_deriv(mesh_ctx *ctx, const mesh_grid *g, const float *theta, const float *phi, float *F);
/* theta_out_mu(x) = wrap(phi(x) + theta_mu(x) - phi(x+mu)) */
int mesh_u1_gauge_transform(mesh_ctx *ctx, const mesh_grid *g, float *theta_out, const float *theta, const float *phi);
/* overrelaxation of one parity on the transformed angles (in place): each
 * site turns by -omega arg K(x), phi += the same when phi != NULL; even extents */
int mesh_u1_landau_or_halfsweep(mesh_ctx *ctx, const mesh_grid *g, float *theta, float *phi, int parity,
                                double omega);
/* phi_out = wrap(phi_in + t F): g = exp(t i F) g (landau.h:85) */
int mesh_u1_gauge_step(mesh_ctx *ctx, const mesh_grid *g, float *phi_out, const float *phi_in, const float *F,
                       double t);
/* *out = sum Re tr (i A i B) = -sum A B over n values; Y += a X */
int mesh_u1_dot(mesh_ctx *ctx, long n, const float *A, const float *B, double *out);
int mesh_u1_axpy(mesh_ctx *ctx, long n, float *Y, double a, const float *X);

/* ---- U(1) plaquette flow (no counterpart in the reference) -----------------
 *
 * On a periodic 2-d lattice the plaquette angles are independent except for
 * one constraint (their sum is 0 mod 2 pi).  The proposal draws every
 * plaquette but the last (x0, x1) = (L0-1, L1-1) as theta_P = h(u), u uniform,
 * with a learned circle map h (mixture of 8 non-compact projections,
 * mesh_pflow_num_params() parameters), builds the links in a maximal-tree
 * gauge from them and two uniform holonomies, and applies a uniform random
 * gauge transformation.  The map is unimodular, so
 *   log q(U) = sum_{P != last} log q1(theta_P(U)) - (V + 1) log 2 pi,
 * q1 the density of h(u).  Any 2-d extents >= 2, V <= 4096. */
int mesh_pflow_num_params(void);
/* initial parameters: log s_k spread over [-3, 1], equal weights, t = 0 */
int mesh_pflow_init(mesh_ctx *ctx, float *params);
/* batch proposals (links) from Philox(seed, step) */
int mesh_pflow_sample(mesh_ctx *ctx, const mesh_grid *g, int batch, const float *params, uint64_t seed, uint32_t step,
                      float *theta_out);
/* out_dev[2c + {0,1}] = {log q(U), S(U)/beta} of configurations theta */
int mesh_pflow_logq(mesh_ctx *ctx, const mesh_grid *g, int batch, const float *params, const float *theta,
                    double *out_dev);
/* one MH proposal round in one launch: state_dev[2c + {0,1}] = {log q, S/beta}
 * of the current configurations theta (as mesh_pflow_logq) and proposals
 * theta_prop (as mesh_pflow_sample, the same links) with prop_dev = their
 * {log q, S/beta}, log q from the draws (the inverse map's value up to
 * rounding); theta_prop must not alias theta, nor prop_dev state_dev */
int mesh_pflow_propose(mesh_ctx *ctx, const mesh_grid *g, int batch, const float *params, const float *theta,
                       double *state_dev, uint64_t seed, uint32_t step, float *theta_prop, double *prop_dev);
/* bytes of the workspace mesh_pflow_gradient needs */
size_t mesh_pflow_workspace_bytes(const mesh_grid *g, int batch);
/* reverse-KL gradient grad = (1/batch) sum_c d/dparams [log q + beta S] over a
 * fresh batch of plaquette draws (Philox(seed, step)); out_dev as logq */
int mesh_pflow_gradient(mesh_ctx *ctx, const mesh_grid *g, int batch, const float *params, uint64_t seed,
                        uint32_t step, double beta, void *workspace, float *grad, double *out_dev);
/* Adam step t >= 1 (moments zero-initialised by the caller) */
int mesh_pflow_adam(mesh_ctx *ctx, float *params, const float *grad, float *m, float *v, int t, double lr);

/* ---- U(1) plaquette coupling flow (no counterpart in the reference) --------
 *
 * Coupling layers on the V - 1 free plaquettes of a 2-d torus (all but the
 * last, (L0-1, L1-1)), which start uniform; the links follow as for
 * mesh_pflow_* (maximal-tree gauge, two uniform holonomies, uniform gauge
 * transformation).  Layer l moves the plaquettes of parity class l % 4,
 * (x0 mod 2, x1 mod 2) = (m % 2, m / 2), by h(a) = g(wrap(a - t)) + t, g a
 * mixture of 4 non-compact projections whose parameters an MLP
 * 40 -> 64 -> 64 -> 10 (tanh) computes from (cos, sin) of the 8 surrounding
 * plaquettes, of the 8 1x2 loops formed by neighbouring pairs of them and of
 * the 4 L-shaped loops that the 2x2 squares of the 3x3 neighbourhood leave
 * without their (active) centre (the last plaquette enters as 0).  Exact
 * log-density:
 *   log q(U) = -sum_{layers, active} log g' - 2V log 2 pi.
 * Even extents >= 4, V <= 4096, 1 <= layers <= 64; weights as mesh_flow_*
 * with 40 input features (mesh_pcflow_num_params(layers) floats, device). */
int mesh_pcflow_num_params(int layers);
/* initial weights (Philox(seed)): the identity map plus 1e-2 noise */
int mesh_pcflow_init(mesh_ctx *ctx, int layers, float *params, uint64_t seed);
/* batch proposals theta_out (links, [batch][2][L0][L1]) from Philox(seed,
 * step); out_dev[2c + {0,1}] = {log q(U), S(U)/beta} */
int mesh_pcflow_sample(mesh_ctx *ctx, const mesh_grid *g, int batch, int layers, const float *params, uint64_t seed,
                       uint32_t step, float *theta_out, double *out_dev);
/* out_dev[2c + {0,1}] = {log q(U), S(U)/beta} of configurations theta;
 * z_out (optional, [batch][V]): the prior plaquettes the inverse recovers */
int mesh_pcflow_logq(mesh_ctx *ctx, const mesh_grid *g, int batch, int layers, const float *params,
                     const float *theta, float *z_out, double *out_dev);
/* one MH proposal round in one launch: state_dev[2c + {0,1}] = {log q, S/beta}
 * of the current configurations theta (as mesh_pcflow_logq) and proposals
 * theta_prop with prop_dev (as mesh_pcflow_sample); theta_prop must not alias
 * theta, nor prop_dev state_dev.  The two halves of the launch run side by side where two workgroups
 * fit a CU (V <= 256). */
int mesh_pcflow_propose(mesh_ctx *ctx, const mesh_grid *g, int batch, int layers, const float *params,
                        const float *theta, double *state_dev, uint64_t seed, uint32_t step, float *theta_prop,
                        double *prop_dev);
/* bytes of the workspace mesh_pcflow_gradient needs */
size_t mesh_pcflow_workspace_bytes(const mesh_grid *g, int batch, int layers);
/* reverse-KL gradient grad = (1/batch) sum_c d/dparams [log q + beta S] over a
 * fresh batch of draws (Philox(seed, step)); out_dev as mesh_pcflow_sample */
int mesh_pcflow_gradient(mesh_ctx *ctx, const mesh_grid *g, int batch, int layers, const float *params, uint64_t seed,
                         uint32_t step, double beta, void *workspace, float *grad, double *out_dev);
/* Adam step t >= 1 (moments zero-initialised by the caller) */
int mesh_pcflow_adam(mesh_ctx *ctx, int layers, float *params, const float *grad, float *m, float *v, int t,
                     double lr);

#ifdef __cplusplus
}
#endif

#endif /* MESH_HIP_H */
