// Internal declarations shared by the libgbg translation units: the complex
// type, the plan object, status plumbing and the kernel-launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/gbg.h"

namespace gbg {

// ---------------------------------------------------------------------------
// complex arithmetic (interleaved re/im, the layout of Julia's Complex{T})
// ---------------------------------------------------------------------------
template <typename T>
struct alignas(2 * sizeof(T)) cpx {
  T x, y;
};

template <typename T>
__host__ __device__ __forceinline__ cpx<T> mk(T x, T y) {
  cpx<T> r;
  r.x = x;
  r.y = y;
  return r;
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) {
  return mk<T>(a.x + b.x, a.y + b.y);
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) {
  return mk<T>(a.x - b.x, a.y - b.y);
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> cscale(cpx<T> a, T s) {
  return mk<T>(a.x * s, a.y * s);
}
// i * a
template <typename T>
__host__ __device__ __forceinline__ cpx<T> cmuli(cpx<T> a) {
  return mk<T>(-a.y, a.x);
}

// Streaming accesses of the fused pencil passes.  GBG_NT (a build option) sets
// which carry the nontemporal hint: bit 0 the loads, bit 1 the stores, bit 2
// the x pass's LDS staging copies.  Measured on the MI355X (512^3 u x omega,
// one term): stores only 5.26 -> 5.15 ms in FP64 and 2.87 -> 2.73 ms in FP32;
// nontemporal loads made every pass that reads a component slower (FP64
// 5.44 ms with loads alone), nontemporal staging copies the x pass too.  A
// streaming copy kernel shows the same split: 1 GiB copy 5.76 TB/s plain,
// 5.91 with nontemporal stores.
#ifndef GBG_NT
#define GBG_NT 2
#endif
template <typename T>
__device__ __forceinline__ cpx<T> ld_s(const cpx<T>* p) {
  if constexpr ((GBG_NT & 1) != 0) {
    using V = T __attribute__((ext_vector_type(2)));
    const V v = __builtin_nontemporal_load(reinterpret_cast<const V*>(p));
    return mk<T>(v.x, v.y);
  } else {
    return *p;
  }
}
// The hint only pays where a tile's contiguous piece of a line fills it: a
// 64-byte piece shares its 128-byte line with the next tile's, and with the
// hint the two halves can reach memory as partial lines instead of meeting in
// L2 (a probe of pencil-tile writes: 64-byte pieces 3.3-3.8 TB/s with the hint,
// 128-byte pieces 5.0-6.0 TB/s).
constexpr bool piece_store_nt(int piece_bytes) { return (GBG_NT & 2) != 0 && piece_bytes >= 128; }
template <bool NTS = (GBG_NT & 2) != 0, typename T>
__device__ __forceinline__ void st_s(cpx<T>* p, cpx<T> v) {
  if constexpr (NTS) {
    using V = T __attribute__((ext_vector_type(2)));
    V w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, reinterpret_cast<V*>(p));
  } else {
    *p = v;
  }
}

// ---------------------------------------------------------------------------
// 1-D transform plan: a mixed-radix factorisation of the axis length
// ---------------------------------------------------------------------------
constexpr int kMaxStages = 16;
struct AxisPlan {
  int n;
  int nst;
  int radix[kMaxStages];
  // 0: radix[] factors n (a direct mixed-radix transform).  Otherwise n has a
  // prime factor above 13 and the axis is transformed as a chirp-z
  // convolution of length m (Bluestein): radix[] factors m, and the axis's
  // twiddle table holds the m-entry table, then the n chirps e^{-i pi j^2/n},
  // then the m-point spectra of the forward and inverse chirp filters.
  int m;
};

// Factor n into the radices the butterflies implement (8, 4, 2, 3, 5, 7, 11,
// 13); when n has a larger prime factor, set up the chirp-z form over the
// smallest m >= 2n - 1 of the form 2^a 3^b 5^c.  Returns false only when no
// factorisation fits kMaxStages.
bool factor_axis(int n, AxisPlan* ap);
// LDS entries one pencil of the axis occupies: n, or m for a chirp-z axis
inline int pencil_len(const AxisPlan& ap) { return ap.m ? ap.m : ap.n; }

// FP64 per-axis wavenumber tables (device pointers): derivative family kd*,
// squared family ks*.
struct Tables {
  const double* kd1;
  const double* kd2;
  const double* kd3;
  const double* ks1;
  const double* ks2;
  const double* ks3;
};

// ---------------------------------------------------------------------------
// The plan object behind gbg_plan*
// ---------------------------------------------------------------------------
constexpr int kMaxEvents = 16;
constexpr int kRedBlocks = 1024;   // fixed reduction grid: deterministic sums
constexpr int kSlabStreams = 4;    // at most this many slabs of the z-slab chain in flight

// Output tiles (B consecutive (ky, kx) pencils) of the fused pipeline's last
// pass for one cutoff: dev[0, nd) lists the tiles that hold a kept (kx, ky)
// pencil, dev[nd, nd + nt) flags every tile 1 if it is in that list, else 0.
// Built once per (plan, cutoff) outside any stream capture and kept, so a
// captured graph may reference it for the plan's lifetime.
constexpr int kTileSets = 8;
struct TileSet {
  double kc;
  int b, nd, nt;
  int* dev;
};

}  // namespace gbg

struct gbg_plan {
  int n[3];          // Nx, Ny, Nz
  double L[3];       // box lengths
  int elem;          // 4 or 8
  // the folded (real-FFT, halved) axis, 0 = x unless the plan was made for a
  // permuted RealFourierLayout, and the layout's transform order fft_dims
  int fold;
  int fft_dims[3];
  int s[3];          // spectral extents: n, with n[fold] / 2 + 1 on the folded axis
  int nh;            // s[0], the spectral x extent (Nx/2 + 1 when x is folded)
  size_t M;          // stored modes per component
  size_t P;          // grid points per component
  int device;
  int ncu;           // compute units of that device
  int generic_only;  // GBG_GENERIC_FFT=1 at creation: mixed-radix LDS passes only
  int wide_offsets;  // GBG_WIDE_OFFSETS=1 at creation: 64-bit offsets in u x omega K1
  gbg::AxisPlan ax[3];
  // per-axis transform choice (set by gbg_plan_autotune): use_fast = take the
  // register-resident pass where the length allows it; bsel = pencils per
  // workgroup of the mixed-radix pass (0 = the default choice)
  int use_fast[3];
  int bsel[3];
  // tile shape of the fused pipelines' x pass (0: the default row-pair units
  // per workgroup, 1: half as many); chosen at creation (x_default_shape),
  // changed by gbg_plan_autotune or GBG_X_SHAPE=0|1 at creation
  int x_var;
  // FP64 per-axis tables (device): derivative family and squared family
  double* kd[3];
  double* ks[3];
  // twiddle tables, n entries e^{-2 pi i m / n} in the plan's precision (a
  // chirp-z axis: its m-entry table and chirp tables, AxisPlan::m)
  void* tw[3];
  // pipeline work area (allocated on first use, kept)
  void* work;
  size_t work_bytes;
  // reduction partials and results (FP64)
  double* red_part;
  double* red_out;
  // tile tables of the fused u x omega pipeline's last pass, one per cutoff
  gbg::TileSet tset[gbg::kTileSets];
  int ntset;
  size_t device_bytes;
  // second stream of the fused pipelines and its fork / join events (created
  // on first use outside a graph capture)
  hipStream_t aux;
  hipEvent_t aux_fork, aux_join;
  // further streams of the fused pipelines' z-slab chain (gbg_uxw.hip), its
  // fork event and one join event per stream (created on first use outside a
  // graph capture)
  hipStream_t slab_st[gbg::kSlabStreams];
  hipEvent_t slab_fork;
  hipEvent_t slab_join[gbg::kSlabStreams];
  // GBG_UXW_SLAB=Z,R (or 0) at creation: slab_set = 1 and the chain's planes
  // per slab and slabs in flight (slab_z = 0: off)
  int slab_set, slab_z, slab_r;
  // profiling events
  int profiling;
  int nev;
  hipEvent_t ev[gbg::kMaxEvents];
  hipStream_t ev_stream;
};

namespace gbg {

// the x-pass tile shape a plan starts with (gbg_uxw.hip, XCfg): half-size
// row-pair groups only for FP64 rows of 1024 and longer, which span two waves
inline int x_default_shape(int nx, int elem) { return (nx >= 1024 && elem == 8) ? 1 : 0; }

// status plumbing --------------------------------------------------------------
void set_error(const std::string& msg);
gbg_status fail(gbg_status st, const std::string& msg);
gbg_status check_hip(hipError_t e, const char* what);
gbg_status check_launch(const char* what);

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// ensure the plan's work area holds at least `bytes`
gbg_status ensure_work(gbg_plan* p, size_t bytes);

inline Tables tables_of(gbg_plan* p) {
  return Tables{p->kd[0], p->kd[1], p->kd[2], p->ks[0], p->ks[1], p->ks[2]};
}

// event marks for the profiling aid
void prof_begin(gbg_plan* p, hipStream_t s);
void prof_mark(gbg_plan* p, hipStream_t s);

// ---- launchers implemented in the .hip translation units ---------------------
// transforms (gbg_fft.hip)
gbg_status fwd_transform(gbg_plan* p, void* spec, const void* phys, int ncomp, hipStream_t s);
gbg_status inv_transform(gbg_plan* p, void* phys, void* spec_in, void* spec_work, int ncomp,
                         hipStream_t s);

// operators (gbg_ops.hip)
gbg_status op_gradient(gbg_plan* p, void* dst, const void* src, int src_order, hipStream_t s);
gbg_status op_symmetric_gradient(gbg_plan* p, void* dst, const void* src, hipStream_t s);
gbg_status op_divergence(gbg_plan* p, void* dst, const void* src, int src_order, int has_band,
                         double lo, double hi, double alpha, double beta, hipStream_t s);
gbg_status op_curl(gbg_plan* p, void* dst, const void* src, int inverse, hipStream_t s);
gbg_status op_laplacian(gbg_plan* p, void* dst, const void* src, int ncomp, int inverse,
                        hipStream_t s);
gbg_status op_helmholtz(gbg_plan* p, void* dst, const void* src, int has_band, double lo,
                        double hi, double alpha, double beta, hipStream_t s);
gbg_status op_convolve_band(gbg_plan* p, void* dst, const void* src, int ncomp, double lo,
                            double hi, double alpha, double beta, hipStream_t s);
gbg_status op_mean_free(gbg_plan* p, void* spec, int ncomp, hipStream_t s);
gbg_status op_cross(gbg_plan* p, void* dst, const void* u, const void* v, hipStream_t s);
gbg_status op_multiply(gbg_plan* p, void* dst, const void* a, const void* b, hipStream_t s);
// Parseval inner into p->red_out[slot] (device); no synchronisation
gbg_status op_inner(gbg_plan* p, const void* f, const void* g, int ncomp, int slot, hipStream_t s);
// physical inner (sum over grid points / P) into p->red_out[slot]; no synchronisation
gbg_status op_inner_physical(gbg_plan* p, const void* f, const void* g, int ncomp, int slot,
                             hipStream_t s);

// exponential kernel, radial spectrum, reality condition (gbg_next.hip)
gbg_status op_exponential(gbg_plan* p, void* dst, const void* src, int ncomp, double coefficient,
                          const void* offset, double offset_scale, double alpha, double beta,
                          hipStream_t s);
gbg_status op_exponential_combine(gbg_plan* p, void* dst, double alpha, const void* a, double beta,
                                  const void* b, int ncomp, double coefficient, const void* offset,
                                  double offset_scale, hipStream_t s);
void radial_shells(const gbg_plan* p, int* nshells, double* dk);
size_t radial_work_bytes(const gbg_plan* p);
// part: radial_work_bytes of device scratch; out_dev: nsh doubles on the device
gbg_status op_radial_spectrum(gbg_plan* p, const void* f, const void* g, int ncomp, int nsh,
                              double* out_dev, double* part, hipStream_t s);
gbg_status op_reality_project(gbg_plan* p, void* spec, int ncomp, hipStream_t s);
// the residual into *out_dev (device); uses the plan's reduction partials
gbg_status op_reality_residual(gbg_plan* p, const void* spec, int ncomp, double* out_dev,
                               hipStream_t s);

// pipelines (gbg_uxw.hip)
// the argument checks of gbg_uxw (null buffers, cutoff, u_hat / n_hat overlap)
gbg_status uxw_check_args(gbg_plan* p, const void* u_hat, const void* n_hat, double kc);
// energy and enstrophy of the last u x omega call on p into the host doubles
// (either may be NULL; when both are, nothing happens): synchronises s
gbg_status uxw_reductions(gbg_plan* p, double* energy, double* enstrophy, hipStream_t s);
gbg_status pipe_uxw(gbg_plan* p, const void* u_hat, void* n_hat, double kc, unsigned flags,
                    hipStream_t s);
gbg_status pipe_dealiased_product(gbg_plan* p, void* p_hat, const void* a_hat, const void* b_hat,
                                  double kc, hipStream_t s);
// pipe_uxw with the caller's work area of work_bytes (refused with
// GBG_E_UNSUPPORTED when smaller than uxw_work_bytes): creates nothing on
// the device (no work area, tile set or stream)
gbg_status pipe_uxw_in(gbg_plan* p, const void* u_hat, void* n_hat, double kc, unsigned flags,
                       void* work, size_t work_bytes, hipStream_t s);
// u x omega followed by the integrating-factor combination
//   dst = exp(coef |k|^2) (alpha a + beta b + delta N^) + gamma N^
// applied as N^ is formed (in the last fused pass, or one streaming pass after
// the composed form).  n_hat may be null (N^ then is not written); a, b may be
// null.  The caller has checked the buffers (dst off u_hat, n_hat and the
// inputs apart; dst may be a or b itself).
struct UxwComb {
  void* dst;
  const void* a;
  const void* b;
  double alpha, beta, delta, gamma, coef;
};
gbg_status pipe_uxw_comb(gbg_plan* p, const void* u_hat, void* n_hat, double kc, const UxwComb& cb,
                         unsigned flags, hipStream_t s);
// work-area bytes pipe_uxw needs on this plan's grid for cutoff kc (reads
// only the plan's n, L, nh, M, P, elem and generic_only)
size_t uxw_work_bytes(gbg_plan* p, double kc, unsigned flags);
// work-area bytes pipe_dealiased_product needs (same fields read as above)
size_t dealiased_product_work_bytes(gbg_plan* p, double kc);

}  // namespace gbg
