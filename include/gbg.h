/*
 * gbg.h -- C ABI of libgbg, the MI355X (gfx950) spectral hot path of
 * GBGPeriodicFields.jl.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns a
 * gbg_status and never lets a C++ exception cross the boundary.  Field
 * buffers are device pointers owned by the caller (gbg_malloc or any HIP
 * allocation); plans own their tables and scratch.  Work is stream ordered
 * and asynchronous on the caller's stream (a hipStream_t passed as void*,
 * NULL = the default stream) unless the entry point returns a host value.
 *
 * Layout contract (the reference's Julia column-major storage,
 * src/Fields.jl:239, 369, 388):
 *   physical  real T      (Nx, Ny, Nz[, 3]) x fastest, component blocks contiguous
 *   spectral  complex T   (Nx/2+1, Ny, Nz[, 3]) interleaved re/im, x folded
 * Only fft_dims = (1,2,3) is supported (the default RealFourierLayout).
 *
 * Reference citations are relative to the GBGPeriodicFields.jl checkout.
 */
#ifndef GBG_H
#define GBG_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes map onto the reference's refusals (the Julia shim rethrows). */
typedef enum gbg_status {
  GBG_OK = 0,
  GBG_E_ARG = 1,          /* ArgumentError: grid/layout/eltype mismatch, bad value (src/Fields.jl:1199-1250, 1301-1305, 2057-2075) */
  GBG_E_DIM = 2,          /* DimensionMismatch: tensor order / component count (src/Fields.jl:1924-1929, src/Reductions.jl:893-896) */
  GBG_E_UNSUPPORTED = 3,  /* MethodError-class absence: size or layout this build does not implement */
  GBG_E_ALIAS = 4,        /* scratch === source refused (src/Fields.jl:1608-1610) */
  GBG_E_OOM = 5,          /* device allocation failed */
  GBG_E_DEVICE = 6        /* HIP runtime error (message in gbg_last_error) */
} gbg_status;

typedef struct gbg_plan gbg_plan;

/* Thread-local message for the last failing call on this thread. */
const char* gbg_last_error(void);
/* Library version string. */
const char* gbg_version(void);

/* ---- device memory and streams (the reference's adapt / adapt_structure
 *      family moves, src/Fields.jl:2810-2847) ---------------------------- */
gbg_status gbg_set_device(int device);
gbg_status gbg_device_count(int* count);
gbg_status gbg_malloc(void** ptr, size_t bytes);
gbg_status gbg_free(void* ptr);
gbg_status gbg_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream);
gbg_status gbg_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream);
gbg_status gbg_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream);
gbg_status gbg_memset(void* dst, int value, size_t bytes, void* stream);
gbg_status gbg_stream_create(void** stream);
gbg_status gbg_stream_destroy(void* stream);
gbg_status gbg_stream_sync(void* stream);

/* ---- plans ---------------------------------------------------------------
 * Replaces FieldTransformPlans (src/Fields.jl:1381-1405, 1463-1488) together
 * with the canonical FieldSpectralMultipliers (src/Fields.jl:1875-1943,
 * 1989-2041): transform radices, FP64 twiddle tables, FP64 wavenumber tables
 * (derivative family with the even-axis Nyquist zeroed, squared family with
 * the true k^2).  n = {Nx, Ny, Nz}, L = {Lx, Ly, Lz} (the box lengths),
 * elem_bytes = 8 (Float64) or 4 (Float32).  A plan serves any component
 * count (the reference's rank-free plans, decision 76). */
gbg_status gbg_plan_create(const int n[3], const double L[3], int elem_bytes, gbg_plan** out);
/* The same for a permuted RealFourierLayout(fft_dims) (src/Fields.jl:219-239):
 * fft_dims is a permutation of {1, 2, 3} (else GBG_E_ARG, as
 * _validate_fft_dims :122 refuses it) whose first entry is the folded axis,
 * stored with n/2 + 1 entries; the other two keep their lengths.  Spectral
 * buffers of such a plan are column-major over those extents.  Every
 * transform, operator, reduction and the composed u x omega / product
 * pipelines accept it (u x omega and the product through their composed
 * form); the batch and the server make plans of the default layout
 * (1, 2, 3). */
gbg_status gbg_plan_create_layout(const int n[3], const double L[3], int elem_bytes,
                                  const int fft_dims[3], gbg_plan** out);
gbg_status gbg_plan_destroy(gbg_plan* plan);
/* Sizes in elements per component (spectral complex elements, physical real
 * elements) and the plan's device bytes currently held. */
gbg_status gbg_plan_query(const gbg_plan* plan, size_t* spectral_elems, size_t* physical_elems,
                          size_t* device_bytes);

/* ---- transforms ----------------------------------------------------------- */
/* to_spectral! (src/Fields.jl:1649-1657 -> _forward_components! :1515-1527):
 * unnormalised forward rfft of ncomp components.  phys is not modified. */
gbg_status gbg_to_spectral(gbg_plan* plan, void* spec, const void* phys, int ncomp, void* stream);
/* to_physical! (3-arg src/Fields.jl:1589-1597, 4-arg :1600-1619 ->
 * _inverse_components! :1530-1542): inverse c2r carrying 1/N.  With scratch
 * == NULL the spectral source is consumed (overwritten), as the 3-arg form
 * may; with a scratch buffer of ncomp spectral components the source is
 * preserved and scratch == spec is refused with GBG_E_ALIAS. */
gbg_status gbg_to_physical(gbg_plan* plan, void* phys, void* spec, int ncomp, void* scratch,
                           void* stream);

/* ---- spectral operators (src/DifferentialOperators.jl) ------------------- */
/* gradient! scalar->vector (src_order 0, :812-824, kernel :357-368) and
 * vector->tensor (src_order 1, :827-837, kernel :414-441).  No aliasing. */
gbg_status gbg_gradient(gbg_plan* plan, void* dst, const void* src, int src_order, void* stream);
/* symmetric_gradient! vector->symmetric tensor (:1289-1300, kernel
 * :666-699): S_ii = (i a_i) u_i, S_ij = S_ji = (i/2)(a_j u_i + a_i u_j), the
 * nine blocks of dst (block i + 3 j) written, each off-diagonal value to both
 * of its blocks.  No aliasing. */
gbg_status gbg_symmetric_gradient(gbg_plan* plan, void* dst, const void* src, void* stream);
/* divergence! vector->scalar (src_order 1) and tensor->vector (src_order 2)
 * with the optional band kernel (has_band: spectral_band(k_lo, k_hi)) and
 * alpha/beta accumulate (:948-1037, _apply_spectral_divergence! :120-139). */
gbg_status gbg_divergence(gbg_plan* plan, void* dst, const void* src, int src_order, int has_band,
                          double k_lo, double k_hi, double alpha, double beta, void* stream);
/* curl! (:1135-1149, kernel :502-523).  dst may alias src. */
gbg_status gbg_curl(gbg_plan* plan, void* dst, const void* src, void* stream);
/* inverse_curl! (:1212-1224, kernel :580-605).  dst may alias src. */
gbg_status gbg_inverse_curl(gbg_plan* plan, void* dst, const void* src, void* stream);
/* laplacian! (:1369-1376, :147-150) over ncomp components.  Alias-safe. */
gbg_status gbg_laplacian(gbg_plan* plan, void* dst, const void* src, int ncomp, void* stream);
/* inverse_laplacian! (:1457-1464, :159-164).  Alias-safe. */
gbg_status gbg_inverse_laplacian(gbg_plan* plan, void* dst, const void* src, int ncomp, void* stream);
/* helmholtz_projection! (:1574-1593, kernel :191-230): the Leray projector,
 * optionally composed with spectral_band(k_lo, k_hi) and alpha/beta.  Alias-safe. */
gbg_status gbg_helmholtz(gbg_plan* plan, void* dst, const void* src, int has_band, double k_lo,
                         double k_hi, double alpha, double beta, void* stream);
/* convolve!(dst, src, spectral_band(m, k_lo, k_hi), alpha, beta) -- the
 * radial path (src/Convolution.jl:1035-1065, :735-787, :1832-1882).
 * spectral_cutoff(m, kc) is k_lo = 0, k_hi = kc.  Alias-safe. */
gbg_status gbg_convolve_band(gbg_plan* plan, void* dst, const void* src, int ncomp, double k_lo,
                             double k_hi, double alpha, double beta, void* stream);
/* mean_free_projection! (src/Fields.jl:3124-3128). */
gbg_status gbg_mean_free(gbg_plan* plan, void* spec, int ncomp, void* stream);

/* ---- physical products (src/FieldAlgebra.jl) ----------------------------- */
/* cross! (src/FieldAlgebra.jl:1088-1092, formula src/FiberSections.jl:641-652).  Alias-safe. */
gbg_status gbg_cross(gbg_plan* plan, void* dst, const void* u, const void* v, void* stream);
/* multiply! scalar x scalar (src/FieldAlgebra.jl:703-724, 778).  Alias-safe. */
gbg_status gbg_multiply(gbg_plan* plan, void* dst, const void* a, const void* b, void* stream);

/* ---- reductions (src/Reductions.jl) -------------------------------------- */
/* Spectral inner (Parseval, :883-925) over ncomp components, accumulated in
 * FP64 in a fixed order (deterministic).  Synchronises the stream and writes
 * the host double *out. */
gbg_status gbg_inner_spectral(gbg_plan* plan, const void* f, const void* g, int ncomp, double* out,
                              void* stream);

/* Physical inner (:838-853): sum over every grid point and component of
 * f g, divided by the grid point count; FP64 accumulation in a fixed order.
 * Synchronises the stream and writes the host double *out. */
gbg_status gbg_inner_physical(gbg_plan* plan, const void* f, const void* g, int ncomp, double* out,
                              void* stream);

/* Radial (shell-binned) spectrum, radial_spectrum(f[, g], m)
 * (src/Reductions.jl:1309-1455): shell m collects the modes with
 * round(|k|/dk) = m (ties to even), dk = 2 pi / max L, Parseval-weighted
 * sum_c |f_c|^2 (or Re f_c conj g_c when g is not NULL) over (Nx Ny Nz)^2.
 * gbg_radial_shells gives the shell count and width (:1334-1338);
 * gbg_radial_spectrum writes nshells FP64 values to the host array
 * `spectrum` (summed in a fixed order: deterministic) and synchronises. */
gbg_status gbg_radial_shells(gbg_plan* plan, int* nshells, double* dk);
gbg_status gbg_radial_spectrum(gbg_plan* plan, const void* f, const void* g, int ncomp,
                               double* spectrum, int nshells, void* stream);

/* ---- exponential kernel (src/Convolution.jl:1197-1795) --------------------- */
/* exponential_convolve!(dst, f, m, alpha, beta; coefficient, offset,
 * offset_scale) (:1750; _exponential_convolution_kernel! :1326-1346):
 *   dst = alpha exp(c |k|^2 + s w) f  (+ beta dst when beta != 0),
 * |k|^2 the true squared wavenumber, w = offset (NULL, or one spectral
 * component of the plan's precision; a complex w gives a complex factor).
 * The factor is evaluated per mode in FP64 and never stored.  dst may be f or
 * offset. */
gbg_status gbg_exponential_convolve(gbg_plan* plan, void* dst, const void* f, int ncomp,
                                    double coefficient, const void* offset, double offset_scale,
                                    double alpha, double beta, void* stream);
/* The two-operand form exponential_convolve!(dst, alpha, a, beta, b, m; ...)
 * (_exponential_combination_kernel! :1349-1362): dst = exp(c |k|^2 + s w)
 * (alpha a + beta b) in one pass.  dst may be a, b or offset. */
gbg_status gbg_exponential_combine(gbg_plan* plan, void* dst, double alpha, const void* a,
                                   double beta, const void* b, int ncomp, double coefficient,
                                   const void* offset, double offset_scale, void* stream);

/* ---- reality condition (src/Fields.jl:2858-3094) ---------------------------- */
/* reality_residual (:2918-2928): the largest |f(k) - conj f(-k)| over the
 * folded axis's k = 0 plane and, for an even folded length, its Nyquist
 * plane (x for the default layout), every component.  Host FP64 result;
 * synchronises. */
gbg_status gbg_reality_residual(gbg_plan* plan, const void* spec, int ncomp, double* residual,
                                void* stream);
/* reality_projection!(dst, src) / reality_projection!(f) (:2940-3094): each
 * stored conjugate pair on those planes becomes its mean (x + conj y)/2 and
 * the conjugate of that; a self-paired corner its real part; every other
 * coefficient untouched.  dst == src projects in place (one or two planes);
 * otherwise src is copied first. */
gbg_status gbg_reality_projection(gbg_plan* plan, void* dst, const void* src, int ncomp,
                                  void* stream);

/* ---- fused pipelines ------------------------------------------------------ */
/* The dealiased rotational nonlinear term of examples/HIST_CPU.jl:229-247:
 *   N^ = Q M P F( F^-1(M u^) x F^-1(curl M u^) )
 * with M = spectral_cutoff(kc) (applied to the input and the output),
 * P = Leray projection, Q = mean-free.  u_hat (3 components) is read only.
 * energy / enstrophy (host pointers, may be NULL) receive 1/2<u,u> and
 * 1/2<w,w> of the truncated fields in FP64; when either is non-NULL the call
 * synchronises the stream.  flags: GBG_UXW_UNFUSED runs the composition of
 * the standalone operators instead of the fused pass schedule;
 * GBG_UXW_SLABS runs the fused schedule's three middle passes z-slab by
 * z-slab over two small slots (on two streams, forked from and joined back
 * into `stream`), which drops the work area's six full-y fields: the same
 * results bit for bit, 4.4 GB less work area at 512^3 FP64 and 35 GB less at
 * 1024^3 FP64, and within 1 % of the default time on the MI355X. */
#define GBG_UXW_UNFUSED 1u
#define GBG_UXW_SLABS 2u
gbg_status gbg_uxw(gbg_plan* plan, const void* u_hat, void* n_hat, double kc, double* energy,
                   double* enstrophy, unsigned flags, void* stream);
/* gbg_uxw for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Same arguments, kernels and results as
 * gbg_uxw, but its work area is a fixed 128 MiB device array of the library
 * (one per device, held until the process ends) and the call creates nothing
 * on the device: no hipMalloc, no tile list, no stream.  Grids and cutoffs
 * whose work area exceeds 128 MiB are refused with GBG_E_UNSUPPORTED (it
 * holds 128^3 FP64 at the 2/3 cutoff, 128^3 FP32 at any cutoff).  The plan
 * must belong to the current device.  Every call on a device shares the
 * array: calls that overlap in time on one device race. */
gbg_status gbg_uxw_noalloc(gbg_plan* plan, const void* u_hat, void* n_hat, double kc,
                           double* energy, double* enstrophy, unsigned flags, void* stream);
/* u x omega with the integrating-factor step folded into its last pass: the
 * update on either side of N^ in an integrating-factor RK2 step
 * (examples/HIST_CPU.jl:267-282: u* = E(u_n + dt N_1), then
 * u_{n+1} = E u_n + dt/2 (E N_1 + N_2)) as
 *   dst = exp(c |k|^2) (alpha a + beta b + delta N^) + gamma N^
 * with exp(c |k|^2) the exponential kernel of exponential_convolve!
 * (src/Convolution.jl:1289-1294, 1326-1346), |k|^2 = (k1^2 + k2^2) + k3^2.
 * Stage 1: n_hat = N_1, a = u_n, alpha = 1, delta = dt, dst = u*.
 * Stage 2: n_hat = NULL, a = u_n, b = N_1, alpha = 1, beta = gamma = dt/2,
 * dst = u_n (in place).  a, b may be NULL; n_hat NULL skips writing N^.
 * dst must not overlap u_hat or n_hat; it may be a or b itself but no other
 * overlap of them (GBG_E_ALIAS).  energy / enstrophy as gbg_uxw.  No
 * counterpart entry point in the reference (it composes rhs! and
 * exponential_convolve! calls). */
typedef struct gbg_uxw_combination {
  void* dst;
  const void* a;
  double alpha;
  const void* b;
  double beta;
  double delta;       /* weight of N^ inside the exponential factor */
  double gamma;       /* weight of N^ outside it */
  double coefficient; /* c */
} gbg_uxw_combination;
gbg_status gbg_uxw_combine(gbg_plan* plan, const void* u_hat, void* n_hat, double kc,
                           const gbg_uxw_combination* comb, double* energy, double* enstrophy,
                           unsigned flags, void* stream);
/* Dealiased scalar product p^ = M F(F^-1(M a^) * F^-1(M b^)).  Inputs read only. */
gbg_status gbg_dealiased_product(gbg_plan* plan, void* p_hat, const void* a_hat, const void* b_hat,
                                 double kc, void* stream);

/* ---- autotuning ------------------------------------------------------------
 * Times, per axis, the candidate pass configurations (register-resident pass
 * for power-of-two lengths; mixed-radix LDS pass under several radix
 * orderings and pencils-per-workgroup widths) on a scratch vector field and
 * keeps the fastest in the plan; on power-of-two grids it then times the
 * fused pipelines' x pass in each tile shape (row-pair units per workgroup)
 * over one u x omega term and keeps the faster.  Allocates and frees
 * scratch, synchronises the stream.  *best_ms (may be NULL) = forward +
 * inverse time of the transform choice.
 * No counterpart in the reference (FFTW's planner flags play this role,
 * src/Fields.jl:1352-1362). */
gbg_status gbg_plan_autotune(gbg_plan* plan, void* stream, double* best_ms);

/* ---- batched form -----------------------------------------------------------
 * Many independent items, each a field on its own grid and precision, run as
 * one hipGraph launch.  Item kinds:
 *   GBG_ITEM_SCALAR_GRAD_PRODUCT: in = a^ (1 spectral component);
 *     out0 = grad a^ (3 components), out1 = M F(F^-1(M a^)^2) (1 component)
 *   GBG_ITEM_VECTOR_GRAD_UXW: in = u^ (3 components);
 *     out0 = grad u^ (9 components, block i + 3 j = (i a_j) u_i), out1 = the
 *     dealiased u x w term (3 components, as gbg_uxw)
 * gbg_batch_instantiate assigns items to nstreams streams, creates (and with
 * autotune_plans != 0, tunes) one plan per grid, precision and stream, runs
 * every item once, then captures all of them into one graph.  Buffers must
 * stay valid and in place until gbg_batch_destroy.  No counterpart in the
 * reference (it launches per call, src/DifferentialOperators.jl:535-543). */
typedef struct gbg_batch gbg_batch;
#define GBG_ITEM_SCALAR_GRAD_PRODUCT 0
#define GBG_ITEM_VECTOR_GRAD_UXW 1
gbg_status gbg_batch_create(int nstreams, int autotune_plans, gbg_batch** out);
gbg_status gbg_batch_add(gbg_batch* batch, int kind, const int n[3], const double L[3],
                         int elem_bytes, const void* in, void* out0, void* out1, double kc);
gbg_status gbg_batch_instantiate(gbg_batch* batch, void* stream);
gbg_status gbg_batch_launch(gbg_batch* batch, void* stream);
gbg_status gbg_batch_destroy(gbg_batch* batch);

/* ---- serving loop -----------------------------------------------------------
 * Continuous batching of independent jobs.  gbg_server_submit queues a job
 * and returns at once; a dispatcher thread admits queued jobs in arrival
 * order as soon as one of `slots` streams is free and the job's plan (per
 * grid, precision and slot; created on first use, then resident) fits
 * `hbm_budget` bytes, evicting idle plans least recently used first, so jobs
 * that arrive while others run overlap on the device.  Job kinds:
 *   GBG_JOB_TO_SPECTRAL        in0 physical (ncomp) -> out spectral
 *   GBG_JOB_TO_PHYSICAL        in0 spectral (ncomp, preserved) -> out physical
 *   GBG_JOB_CURL               in0 spectral vector -> out
 *   GBG_JOB_INVERSE_LAPLACIAN  in0 spectral (ncomp) -> out
 *   GBG_JOB_DEALIASED_PRODUCT  in0 a^, in1 b^ (scalars) -> out p^ (cutoff kc)
 *   GBG_JOB_UXW                in0 u^ -> out N^ (cutoff kc)
 * Field buffers are the caller's and must stay valid until the job is
 * waited for.  Each job runs on a plan of its own slot, so its result does
 * not depend on the other jobs in flight.  No counterpart in the reference
 * (a consumer-side request loop over the calls above). */
typedef struct gbg_server gbg_server;
#define GBG_JOB_TO_SPECTRAL 0
#define GBG_JOB_TO_PHYSICAL 1
#define GBG_JOB_CURL 2
#define GBG_JOB_INVERSE_LAPLACIAN 3
#define GBG_JOB_DEALIASED_PRODUCT 4
#define GBG_JOB_UXW 5
typedef struct gbg_job {
  int kind;
  int n[3];
  double L[3];
  int elem_bytes;
  int ncomp;
  const void* in0;
  const void* in1;
  void* out;
  double kc;
} gbg_job;
gbg_status gbg_server_create(size_t hbm_budget, int slots, gbg_server** out);
gbg_status gbg_server_submit(gbg_server* server, const gbg_job* job, unsigned long long* ticket);
/* Blocks until the job is done; *latency_ms = completion - submission. */
gbg_status gbg_server_wait(gbg_server* server, unsigned long long ticket, double* latency_ms);
/* Returns at once: *done = 0 while the job runs (the ticket stays valid);
 * *done = 1 once it has finished, with the job's status and *latency_ms as
 * gbg_server_wait gives them (the ticket is then retired). */
gbg_status gbg_server_poll(gbg_server* server, unsigned long long ticket, int* done,
                           double* latency_ms);
/* Latency percentiles over completed jobs and their throughput since the
 * first submission. */
gbg_status gbg_server_stats(gbg_server* server, double* p50_ms, double* p99_ms, double* jobs_per_s,
                            unsigned long long* completed);
/* Device bytes held by the resident plans and their count. */
gbg_status gbg_server_resident(gbg_server* server, size_t* bytes, int* plans);
gbg_status gbg_server_destroy(gbg_server* server);

/* ---- profiling aid --------------------------------------------------------
 * When enabled, fused pipelines record HIP events between their passes on
 * the caller's stream; gbg_plan_stage_times synchronises and returns the
 * per-pass milliseconds of the last call (up to *n entries; *n is set to the
 * count). */
gbg_status gbg_plan_set_profiling(gbg_plan* plan, int enabled);
gbg_status gbg_plan_stage_times(gbg_plan* plan, double* ms, int* n);

#ifdef __cplusplus
}
#endif
#endif /* GBG_H */
