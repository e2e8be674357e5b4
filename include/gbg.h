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

/* ---- fused pipelines ------------------------------------------------------ */
/* The dealiased rotational nonlinear term of examples/HIST_CPU.jl:229-247:
 *   N^ = Q M P F( F^-1(M u^) x F^-1(curl M u^) )
 * with M = spectral_cutoff(kc) (applied to the input and the output),
 * P = Leray projection, Q = mean-free.  u_hat (3 components) is read only.
 * energy / enstrophy (host pointers, may be NULL) receive 1/2<u,u> and
 * 1/2<w,w> of the truncated fields in FP64; when either is non-NULL the call
 * synchronises the stream.  flags: GBG_UXW_UNFUSED runs the composition of
 * the standalone operators instead of the fused pass schedule. */
#define GBG_UXW_UNFUSED 1u
gbg_status gbg_uxw(gbg_plan* plan, const void* u_hat, void* n_hat, double kc, double* energy,
                   double* enstrophy, unsigned flags, void* stream);
/* Dealiased scalar product p^ = M F(F^-1(M a^) * F^-1(M b^)).  Inputs read only. */
gbg_status gbg_dealiased_product(gbg_plan* plan, void* p_hat, const void* a_hat, const void* b_hat,
                                 double kc, void* stream);

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
