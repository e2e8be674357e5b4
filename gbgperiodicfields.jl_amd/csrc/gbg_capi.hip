// extern "C" entry points of libgbg: argument checks (mirroring the
// reference's refusals), plan construction, device memory, and dispatch to
// the kernel launchers.  No C++ exception crosses this boundary.
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "gbg_internal.hpp"

namespace gbg {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

gbg_status fail(gbg_status st, const std::string& msg) {
  set_error(msg);
  return st;
}

gbg_status check_hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return GBG_OK;
  if (e == hipErrorOutOfMemory) return fail(GBG_E_OOM, std::string(what) + ": out of device memory");
  return fail(GBG_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

gbg_status check_launch(const char* what) { return check_hip(hipGetLastError(), what); }

static bool factor_into(int n, AxisPlan* ap) {
  ap->nst = 0;
  int m = n;
  const int order[] = {8, 4, 2, 3, 5, 7, 11, 13};
  for (int r : order) {
    while (m % r == 0) {
      if (ap->nst >= kMaxStages) return false;
      ap->radix[ap->nst++] = r;
      m /= r;
    }
  }
  return m == 1;
}

bool factor_axis(int n, AxisPlan* ap) {
  ap->n = n;
  ap->m = 0;
  if (factor_into(n, ap)) return true;
  // chirp-z: the smallest 2^a 3^b 5^c >= 2n - 1
  long long best = 0;
  for (long long a = 1; a < (1LL << 40); a *= 2)
    for (long long b = a; b < (1LL << 40); b *= 3)
      for (long long c = b; c < (1LL << 40); c *= 5)
        if (c >= 2LL * n - 1 && (!best || c < best)) best = c;
  if (!best || best > (1 << 30)) return false;
  ap->m = (int)best;
  return factor_into(ap->m, ap);
}

gbg_status ensure_work(gbg_plan* p, size_t bytes) {
  if (p->work_bytes >= bytes) return GBG_OK;
  if (p->work) {
    hipError_t e = hipFree(p->work);
    p->device_bytes -= p->work_bytes;
    p->work = nullptr;
    p->work_bytes = 0;
    if (e != hipSuccess) return check_hip(e, "free work area");
  }
  gbg_status st = check_hip(hipMalloc(&p->work, bytes), "plan work area");
  if (st) {
    p->work = nullptr;
    return st;
  }
  p->work_bytes = bytes;
  p->device_bytes += bytes;
  return GBG_OK;
}

void prof_begin(gbg_plan* p, hipStream_t s) {
  p->nev = 0;
  if (!p->profiling) return;
  p->ev_stream = s;
  prof_mark(p, s);
}

void prof_mark(gbg_plan* p, hipStream_t s) {
  if (!p->profiling || p->nev >= kMaxEvents) return;
  if (hipEventRecord(p->ev[p->nev], s) == hipSuccess) ++p->nev;
}

// FP64-accurate twiddles e^{-2 pi i m / n}: angles reduced exactly to the
// first octant in integer arithmetic, evaluated in long double.
static void make_twiddles(int n, std::vector<long double>& re, std::vector<long double>& im) {
  re.resize(n);
  im.resize(n);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (int m = 0; m < n; ++m) {
    // angle = -2 pi m / n; use symmetry so the argument stays in [0, pi/4]
    long long a = 8LL * m;  // angle in units of (2 pi / (8 n)) : a / n octants
    long double c, s;
    // reduce: theta = 2 pi m / n = (pi/4) * (8 m / n)
    long long q = a / n;       // octant index 0..7
    long long r = a - q * n;   // remainder in [0, n)
    long double t = (pi / 4.0L) * ((long double)r / (long double)n);
    long double ct = cosl(t), st = sinl(t);
    long double tc = cosl((pi / 4.0L) - t), ts = sinl((pi / 4.0L) - t);
    // cos/sin of theta = q*pi/4 + t, expressed with exact octant symmetries
    switch (q & 7) {
      case 0: c = ct; s = st; break;
      case 1: c = ts; s = tc; break;          // pi/4 + t = pi/2 - (pi/4 - t)
      case 2: c = -st; s = ct; break;         // pi/2 + t
      case 3: c = -tc; s = ts; break;         // pi - (pi/4 - t)
      case 4: c = -ct; s = -st; break;        // pi + t
      case 5: c = -ts; s = -tc; break;        // 3pi/2 - (pi/4 - t)
      case 6: c = st; s = -ct; break;         // 3pi/2 + t
      default: c = tc; s = -ts; break;        // 2pi - (pi/4 - t)
    }
    re[m] = c;
    im[m] = -s;  // e^{-i theta}
  }
}

// The chirp-z tables of an axis of length n over convolution length m
// (Bluestein's algorithm): with w_j = e^{-i pi j^2 / n},
//   X_k = w_k sum_j (x_j w_j) conj(w_{k-j}),
// a cyclic convolution of length m >= 2n - 1 of a_j = x_j w_j (zero past n)
// with b_j = conj(w_j) for |j| < n (b at -j stored at m - j).  Appended to the
// m-entry twiddle table: the n chirps w_j, then B = (1/m) DFT_m(b) for the
// forward transform and the same for the inverse, whose chirps are conj(w_j).
// w_j = e^{-2 pi i (j^2 mod 2n) / (2n)} from the exact-octant table of 2n.
static void chirp_tables(int n, int m, std::vector<long double>& re, std::vector<long double>& im) {
  std::vector<long double> cr, ci, mr, mi;
  make_twiddles(2 * n, cr, ci);
  make_twiddles(m, mr, mi);
  std::vector<long double> wr(n), wi(n);
  for (int j = 0; j < n; ++j) {
    const long long q = ((long long)j * j) % (2LL * n);
    wr[j] = cr[q];
    wi[j] = ci[q];
  }
  re.insert(re.end(), wr.begin(), wr.end());
  im.insert(im.end(), wi.begin(), wi.end());
  for (int dir = 0; dir < 2; ++dir) {
    // b_j = conj(w_j) forward, w_j inverse, on j in (-n, n), wrapped mod m
    const long double sg = dir == 0 ? -1.0L : 1.0L;
    for (int k = 0; k < m; ++k) {
      long double sr = 0.0L, si = 0.0L;
      for (int j = -(n - 1); j <= n - 1; ++j) {
        const int a = j < 0 ? -j : j;
        const long double br = wr[a], bi = sg * wi[a];
        const long long e = (((long long)(j + m) % m) * k) % m;   // e^{-2 pi i j k / m}
        sr += br * mr[e] - bi * mi[e];
        si += br * mi[e] + bi * mr[e];
      }
      re.push_back(sr / m);
      im.push_back(si / m);
    }
  }
}

template <typename T>
static gbg_status upload_twiddles(gbg_plan* p, int axis) {
  std::vector<long double> re, im;
  const AxisPlan& ap = p->ax[axis];
  make_twiddles(pencil_len(ap), re, im);
  if (ap.m) chirp_tables(ap.n, ap.m, re, im);
  const size_t cnt = re.size();
  std::vector<cpx<T>> h(cnt);
  for (size_t m = 0; m < cnt; ++m) h[m] = mk<T>((T)(double)re[m], (T)(double)im[m]);
  const size_t bytes = sizeof(cpx<T>) * cnt;
  gbg_status st = check_hip(hipMalloc(&p->tw[axis], bytes), "twiddle table");
  if (st) return st;
  p->device_bytes += bytes;
  return check_hip(hipMemcpy(p->tw[axis], h.data(), bytes, hipMemcpyHostToDevice), "twiddle upload");
}

// _axis_wavenumber_values (src/Fields.jl:2006-2026) in FP64
static void axis_tables(int n, double L, bool half, std::vector<double>& kd, std::vector<double>& ks) {
  const double unit = (2.0 * M_PI) / L;
  const int stored = half ? n / 2 + 1 : n;
  kd.resize(stored);
  ks.resize(stored);
  for (int i = 0; i < stored; ++i) {
    const int f = (half || i < (n + 1) / 2) ? i : i - n;
    const double k = unit * (double)f;
    ks[i] = k * k;
    kd[i] = k;
  }
  if (n % 2 == 0) kd[half ? stored - 1 : n / 2] = 0.0;
}

static void destroy_plan(gbg_plan* p) {
  if (!p) return;
  for (int a = 0; a < 3; ++a) {
    if (p->kd[a]) (void)hipFree(p->kd[a]);
    if (p->ks[a]) (void)hipFree(p->ks[a]);
    if (p->tw[a]) (void)hipFree(p->tw[a]);
  }
  if (p->work) (void)hipFree(p->work);
  if (p->red_part) (void)hipFree(p->red_part);
  if (p->red_out) (void)hipFree(p->red_out);
  for (int i = 0; i < p->ntset; ++i)
    if (p->tset[i].dev) (void)hipFree(p->tset[i].dev);
  for (int i = 0; i < kMaxEvents; ++i)
    if (p->ev[i]) (void)hipEventDestroy(p->ev[i]);
  if (p->aux_fork) (void)hipEventDestroy(p->aux_fork);
  if (p->aux_join) (void)hipEventDestroy(p->aux_join);
  if (p->aux) (void)hipStreamDestroy(p->aux);
  for (int i = 0; i < kSlabStreams; ++i) {
    if (p->slab_join[i]) (void)hipEventDestroy(p->slab_join[i]);
    if (p->slab_st[i]) (void)hipStreamDestroy(p->slab_st[i]);
  }
  if (p->slab_fork) (void)hipEventDestroy(p->slab_fork);
  delete p;
}

static gbg_status create_plan(const int n[3], const double L[3], int elem, const int fft_dims[3],
                              gbg_plan** out) {
  if (!n || !L || !out || !fft_dims) return fail(GBG_E_ARG, "null argument");
  *out = nullptr;
  if (elem != 4 && elem != 8) return fail(GBG_E_ARG, "elem_bytes must be 4 (Float32) or 8 (Float64)");
  for (int a = 0; a < 3; ++a) {
    if (n[a] < 1) return fail(GBG_E_ARG, "grid sizes must be positive");
    if (!(L[a] > 0.0) || !std::isfinite(L[a])) return fail(GBG_E_ARG, "domain lengths must be positive and finite");
  }
  // _validate_fft_dims (src/Fields.jl:122)
  {
    bool seen[4] = {false, false, false, false};
    for (int a = 0; a < 3; ++a) {
      const int d = fft_dims[a];
      if (d < 1 || d > 3 || seen[d]) return fail(GBG_E_ARG, "fft_dims must be a permutation of (1, 2, 3)");
      seen[d] = true;
    }
  }
  const int fold = fft_dims[0] - 1;
  int ext[3];
  for (int a = 0; a < 3; ++a) ext[a] = a == fold ? n[a] / 2 + 1 : n[a];
  const size_t M = (size_t)ext[0] * ext[1] * ext[2];
  if (3 * M >= (size_t)1 << 32) return fail(GBG_E_UNSUPPORTED, "grid too large for 32-bit mode indexing");
  gbg_plan* p = new (std::nothrow) gbg_plan();
  if (!p) return fail(GBG_E_OOM, "host allocation");
  std::memset(p, 0, sizeof(*p));
  p->fold = fold;
  for (int a = 0; a < 3; ++a) {
    p->fft_dims[a] = fft_dims[a];
    p->s[a] = ext[a];
  }
  for (int a = 0; a < 3; ++a) {
    p->n[a] = n[a];
    p->L[a] = L[a];
    if (!factor_axis(n[a], &p->ax[a])) {
      destroy_plan(p);
      return fail(GBG_E_UNSUPPORTED, "axis length " + std::to_string(n[a]) + " cannot be factored");
    }
    // the LDS pencil kernels hold at least one whole pencil per workgroup (a
    // chirp-z axis: one pencil of its convolution length)
    const int cap = 65536 / (2 * elem);
    if (pencil_len(p->ax[a]) + 1 > cap) {
      const std::string msg =
          "axis length " + std::to_string(n[a]) +
          (p->ax[a].m ? " (a prime factor above 13: chirp-z length " + std::to_string(p->ax[a].m) + ")"
                      : std::string()) +
          " exceeds one LDS pencil";
      destroy_plan(p);
      return fail(GBG_E_UNSUPPORTED, msg);
    }
  }
  p->elem = elem;
  {
    const char* env = std::getenv("GBG_GENERIC_FFT");
    p->generic_only = (env && env[0] == '1') ? 1 : 0;
    const char* wide = std::getenv("GBG_WIDE_OFFSETS");
    p->wide_offsets = (wide && wide[0] == '1') ? 1 : 0;
    p->x_var = x_default_shape(n[0], elem);
    const char* xs = std::getenv("GBG_X_SHAPE");
    if (xs && (xs[0] == '0' || xs[0] == '1') && xs[1] == 0) p->x_var = xs[0] - '0';
    // the z-slab chain of the fused pipelines (gbg_uxw.hip): Z,R or 0 (off)
    const char* sl = std::getenv("GBG_UXW_SLAB");
    if (sl) {
      p->slab_set = 1;
      char* end = nullptr;
      const long z = std::strtol(sl, &end, 10);
      if (z > 0 && end && *end == ',') {
        const long r = std::strtol(end + 1, &end, 10);
        if (r >= 1 && r <= kSlabStreams && end && *end == 0) {
          p->slab_z = (int)(z < n[2] ? z : n[2]);
          p->slab_r = (int)r;
        }
      }
    }
  }
  for (int a = 0; a < 3; ++a) {
    p->use_fast[a] = 1;
    p->bsel[a] = 0;
  }
  p->nh = ext[0];
  p->M = M;
  p->P = (size_t)n[0] * n[1] * n[2];
  gbg_status st = check_hip(hipGetDevice(&p->device), "hipGetDevice");
  if (!st)
    st = check_hip(hipDeviceGetAttribute(&p->ncu, hipDeviceAttributeMultiprocessorCount, p->device),
                   "hipDeviceGetAttribute");
  if (st) {
    destroy_plan(p);
    return st;
  }
  for (int a = 0; a < 3; ++a) {
    std::vector<double> kd, ks;
    axis_tables(n[a], L[a], a == fold, kd, ks);
    const size_t bytes = sizeof(double) * kd.size();
    if ((st = check_hip(hipMalloc((void**)&p->kd[a], bytes), "wavenumber table")) ||
        (st = check_hip(hipMalloc((void**)&p->ks[a], bytes), "wavenumber table")) ||
        (st = check_hip(hipMemcpy(p->kd[a], kd.data(), bytes, hipMemcpyHostToDevice), "upload")) ||
        (st = check_hip(hipMemcpy(p->ks[a], ks.data(), bytes, hipMemcpyHostToDevice), "upload"))) {
      destroy_plan(p);
      return st;
    }
    p->device_bytes += 2 * bytes;
    st = elem == 8 ? upload_twiddles<double>(p, a) : upload_twiddles<float>(p, a);
    if (st) {
      destroy_plan(p);
      return st;
    }
  }
  const size_t rb = sizeof(double) * kRedBlocks * 4;
  if ((st = check_hip(hipMalloc((void**)&p->red_part, rb), "reduction partials")) ||
      (st = check_hip(hipMalloc((void**)&p->red_out, sizeof(double) * 4), "reduction results")) ||
      (st = check_hip(hipMemset(p->red_out, 0, sizeof(double) * 4), "reduction init"))) {
    destroy_plan(p);
    return st;
  }
  p->device_bytes += rb + sizeof(double) * 4;
  *out = p;
  return GBG_OK;
}

}  // namespace gbg

using namespace gbg;

#define GBG_GUARD_BEGIN try {
#define GBG_GUARD_END                                              \
  }                                                                \
  catch (const std::bad_alloc&) {                                  \
    return fail(GBG_E_OOM, "host allocation failed");              \
  }                                                                \
  catch (const std::exception& e) {                                \
    return fail(GBG_E_DEVICE, e.what());                           \
  }                                                                \
  catch (...) {                                                    \
    return fail(GBG_E_DEVICE, "unknown exception");                \
  }

static gbg_status need_plan(gbg_plan* p) {
  if (!p) return fail(GBG_E_ARG, "null plan");
  return GBG_OK;
}

static bool overlaps(const void* a, size_t na, const void* b, size_t nb) {
  const char* x = (const char*)a;
  const char* y = (const char*)b;
  return x < y + nb && y < x + na;
}

namespace gbg {

gbg_status uxw_check_args(gbg_plan* p, const void* u_hat, const void* n_hat, double kc) {
  gbg_status st = need_plan(p);
  if (st) return st;
  if (!u_hat || !n_hat) return fail(GBG_E_ARG, "null field buffer");
  if (!std::isfinite(kc) || kc < 0) return fail(GBG_E_ARG, "cutoff must be finite and >= 0");
  const size_t b = 3 * p->M * 2 * p->elem;
  if (overlaps(u_hat, b, n_hat, b)) return fail(GBG_E_ALIAS, "u_hat and n_hat must not overlap");
  return GBG_OK;
}

gbg_status uxw_reductions(gbg_plan* p, double* energy, double* enstrophy, hipStream_t s) {
  if (!energy && !enstrophy) return GBG_OK;
  double h[2] = {0, 0};
  gbg_status st =
      check_hip(hipMemcpyAsync(h, p->red_out, 2 * sizeof(double), hipMemcpyDeviceToHost, s),
                "reductions d2h");
  if (st) return st;
  if ((st = check_hip(hipStreamSynchronize(s), "reductions sync"))) return st;
  if (energy) *energy = 0.5 * h[0];
  if (enstrophy) *enstrophy = 0.5 * h[1];
  return GBG_OK;
}

}  // namespace gbg

extern "C" {

const char* gbg_last_error(void) { return g_err.c_str(); }
const char* gbg_version(void) { return "gbg-mi355x 0.1.0 (gfx950)"; }

gbg_status gbg_set_device(int device) {
  GBG_GUARD_BEGIN
  return check_hip(hipSetDevice(device), "hipSetDevice");
  GBG_GUARD_END
}

gbg_status gbg_device_count(int* count) {
  GBG_GUARD_BEGIN
  if (!count) return fail(GBG_E_ARG, "null count");
  *count = 0;
  hipError_t e = hipGetDeviceCount(count);
  if (e != hipSuccess) {
    *count = 0;
    return check_hip(e, "hipGetDeviceCount");
  }
  return GBG_OK;
  GBG_GUARD_END
}

gbg_status gbg_malloc(void** ptr, size_t bytes) {
  GBG_GUARD_BEGIN
  if (!ptr) return fail(GBG_E_ARG, "null pointer");
  *ptr = nullptr;
  if (bytes == 0) return GBG_OK;
  return check_hip(hipMalloc(ptr, bytes), "gbg_malloc");
  GBG_GUARD_END
}

gbg_status gbg_free(void* ptr) {
  GBG_GUARD_BEGIN
  if (!ptr) return GBG_OK;
  return check_hip(hipFree(ptr), "gbg_free");
  GBG_GUARD_END
}

gbg_status gbg_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
  GBG_GUARD_BEGIN
  if (bytes == 0) return GBG_OK;
  if (!dst || !src) return fail(GBG_E_ARG, "null pointer");
  gbg_status st = check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)), "h2d");
  if (st) return st;
  return check_hip(hipStreamSynchronize(as_stream(stream)), "h2d sync");
  GBG_GUARD_END
}

gbg_status gbg_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
  GBG_GUARD_BEGIN
  if (bytes == 0) return GBG_OK;
  if (!dst || !src) return fail(GBG_E_ARG, "null pointer");
  gbg_status st = check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)), "d2h");
  if (st) return st;
  return check_hip(hipStreamSynchronize(as_stream(stream)), "d2h sync");
  GBG_GUARD_END
}

gbg_status gbg_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream) {
  GBG_GUARD_BEGIN
  if (bytes == 0) return GBG_OK;
  if (!dst || !src) return fail(GBG_E_ARG, "null pointer");
  return check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)), "d2d");
  GBG_GUARD_END
}

gbg_status gbg_memset(void* dst, int value, size_t bytes, void* stream) {
  GBG_GUARD_BEGIN
  if (bytes == 0) return GBG_OK;
  if (!dst) return fail(GBG_E_ARG, "null pointer");
  return check_hip(hipMemsetAsync(dst, value, bytes, as_stream(stream)), "memset");
  GBG_GUARD_END
}

gbg_status gbg_stream_create(void** stream) {
  GBG_GUARD_BEGIN
  if (!stream) return fail(GBG_E_ARG, "null pointer");
  hipStream_t s = nullptr;
  gbg_status st = check_hip(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream create");
  *stream = (void*)s;
  return st;
  GBG_GUARD_END
}

gbg_status gbg_stream_destroy(void* stream) {
  GBG_GUARD_BEGIN
  if (!stream) return GBG_OK;
  return check_hip(hipStreamDestroy(as_stream(stream)), "stream destroy");
  GBG_GUARD_END
}

gbg_status gbg_stream_sync(void* stream) {
  GBG_GUARD_BEGIN
  return check_hip(hipStreamSynchronize(as_stream(stream)), "stream sync");
  GBG_GUARD_END
}

gbg_status gbg_plan_create(const int n[3], const double L[3], int elem_bytes, gbg_plan** out) {
  GBG_GUARD_BEGIN
  const int dims[3] = {1, 2, 3};
  return create_plan(n, L, elem_bytes, dims, out);
  GBG_GUARD_END
}

gbg_status gbg_plan_create_layout(const int n[3], const double L[3], int elem_bytes,
                                  const int fft_dims[3], gbg_plan** out) {
  GBG_GUARD_BEGIN
  return create_plan(n, L, elem_bytes, fft_dims, out);
  GBG_GUARD_END
}

gbg_status gbg_plan_destroy(gbg_plan* plan) {
  GBG_GUARD_BEGIN
  destroy_plan(plan);
  return GBG_OK;
  GBG_GUARD_END
}

gbg_status gbg_plan_query(const gbg_plan* p, size_t* spectral_elems, size_t* physical_elems,
                          size_t* device_bytes) {
  GBG_GUARD_BEGIN
  if (!p) return fail(GBG_E_ARG, "null plan");
  if (spectral_elems) *spectral_elems = p->M;
  if (physical_elems) *physical_elems = p->P;
  if (device_bytes) *device_bytes = p->device_bytes;
  return GBG_OK;
  GBG_GUARD_END
}

gbg_status gbg_to_spectral(gbg_plan* p, void* spec, const void* phys, int ncomp, void* stream) {
  GBG_GUARD_BEGIN
  gbg_status st = need_plan(p);
  if (st) return st;
  if (!spec || !phys) return fail(GBG_E_ARG, "null field buffer");
  if (ncomp != 1 && ncomp != 3 && ncomp != 9) return fail(GBG_E_DIM, "ncomp must be 1, 3 or 9");
  if (overlaps(spec, ncomp * p->M * 2 * p->elem, phys, ncomp * p->P * p->elem))
    return fail(GBG_E_ALIAS, "spectral destination overlaps the physical source");
  return fwd_transform(p, spec, phys, ncomp, as_stream(stream));
  GBG_GUARD_END
}

gbg_status gbg_to_physical(gbg_plan* p, void* phys, void* spec, int ncomp, void* scratch,
                           void* stream) {
  GBG_GUARD_BEGIN
  gbg_status st = need_plan(p);
  if (st) return st;
  if (!spec || !phys) return fail(GBG_E_ARG, "null field buffer");
  if (ncomp != 1 && ncomp != 3 && ncomp != 9) return fail(GBG_E_DIM, "ncomp must be 1, 3 or 9");
  const size_t sb = ncomp * p->M * 2 * p->elem;
  if (scratch == spec) return fail(GBG_E_ALIAS, "scratch must not be the spectral source");
  if (overlaps(phys, ncomp * p->P * p->elem, spec, sb))
    return fail(GBG_E_ALIAS, "physical destination overlaps the spectral source");
  void* work = scratch ? scratch : spec;
  if (scratch && overlaps(scratch, sb, spec, sb))
    return fail(GBG_E_ALIAS, "scratch overlaps the spectral source");
  if (scratch && overlaps(scratch, sb, phys, ncomp * p->P * p->elem))
    return fail(GBG_E_ALIAS, "scratch overlaps the physical destination");
  return inv_transform(p, phys, spec, work, ncomp, as_stream(stream));
  GBG_GUARD_END
}

gbg_status gbg_gradient(gbg_plan* p, void* dst, const void* src, int src_order, void* stream) {
  GBG_GUARD_BEGIN
  gbg_status st = need_plan(p);
  if (st) return st;
  if (!dst || !src) return fail(GBG_E_ARG, "null field buffer");
  if (src_order != 0 && src_order != 1) return fail(GBG_E_DIM, "gradient source order must be 0 or 1");
  const size_t cb = p->M * 2 * p->elem;
  const int nsrc = src_order == 0 ? 1 : 3;
  if (overlaps(dst, 3 * nsrc * cb, src, nsrc * cb))
    return fail(GBG_E_ALIAS, "gradient raises the order and cannot alias");
  return op_gradient(p, dst, src, src_order, as_stream(stream));
  GBG_GUARD_END
}

gbg_status gbg_symmetric_gradient(gbg_plan* p, void* dst, const void* src, void* stream) {
  GBG_GUARD_BEGIN
  gbg_status st = need_plan(p);
  if (st) return st;
  if (!dst || !src) return fail(GBG_E_ARG, "null field buffer");
  const size_t cb = p->M * 2 * p->elem;
  if (overlaps(dst, 9 * cb, src, 3 * cb))
    return fail(GBG_E_ALIAS, "symmetric gradient raises the order and cannot alias");
  return op_symmetric_gradient(p, dst, src, as_stream(stream));
  GBG_GUARD_END
}

gbg_status gbg_divergence(gbg_plan* p, void* dst, const void* src, int src_order, int has_band,
                          double k_lo, double k_hi, double alpha, double beta, void* stream) {
  GBG_GUARD_BEGIN
  gbg_status st = need_plan(p);
  if (st) return st;
  if (!dst || !src) return fail(GBG_E_ARG, "null field buffer");
  if (src_order != 1 && src_order != 2) return fail(GBG_E_DIM, "divergence source order must be 1 or 2");
  if (has_band && (!std::isfinite(k_lo) || !std::isfinite(k_hi) || !(0 <= k_lo && k_lo <= k_hi)))
    return fail(GBG_E_ARG, "band edges must satisfy 0 <= low <= high");
  const size_t cb = p->M * 2 * p->elem;
  const int nout = src_order == 1 ? 1 : 3;
  if (overlaps(dst, nout * cb, src, 3 * nout * cb))
    return fail(GBG_E_ALIAS, "divergence lowers the order and cannot alias");
  return op_divergence(p, dst, src, src_order, has_band, k_lo, k_hi, alpha, beta, as_stream(stream));
  GBG_GUARD_END
}

static gbg_status same_or_disjoint(const void* dst, const void* src, size_t bytes) {
  if (dst == src || !overlaps(dst, bytes, src, bytes)) return GBG_OK;
  return fail(GBG_E_ALIAS, "destination partially overlaps the source");
}

gbg_status gbg_curl(gbg_plan* p, void* dst, const void* src, void* stream) {
  GBG_GUARD_BEGIN
  gbg_status st = need_plan(p);
  if (st) return st;
  if (!dst || !src) return fail(GBG_E_ARG, "null field buffer");
  if ((st = same_or_disjoint(dst, src, 3 * p->M * 2 * p->elem))) return st;
  return op_curl(p, dst, src, 0, as_stream(stream));
  GBG_GUARD_END
}

gbg_status gbg_inverse_curl(gbg_plan* p, void* dst, const void* src, void* stream) {
  GBG_GUARD_BEGIN
  gbg_status st = need_plan(p);
  if (st) return st;
  if (!dst || !src) return fail(GBG_E_ARG, "null field buffer");
  if ((st = same_or_disjoint(dst, src, 3 * p->M * 2 * p->elem))) return st;
  return op_curl(p, dst, src, 1, as_stream(stream));
  GBG_GUARD_END
}

gbg_status gbg_laplacian(gbg_plan* p, void* dst, const void* src, int ncomp, void* stream) {
  GBG_GUARD_BEGIN
  gbg_status st = need_plan(p);
  if (st) return st;
  if (!dst || !src) return fail(GBG_E_ARG, "null field buffer");
  if (ncomp != 1 && ncomp != 3 && ncomp != 9) return fail(GBG_E_DIM, "ncomp must be 1, 3 or 9");
  if ((st = same_or_disjoint(dst, src, ncomp * p->M * 2 * p->elem))) return st;
  return op_laplacian(p, dst, src, ncomp, 0, as_stream(stream));
  GBG_GUARD_END
}

gbg_status gbg_inverse_laplacian(gbg_plan* p, void* dst, const void* src, int ncomp, void* stream) {
  GBG_GUARD_BEGIN
  gbg_status st = need_plan(p);
  if (st) return st;
  if (!dst || !src) return fail(GBG_E_ARG, "null field buffer");
  if (ncomp != 1 && ncomp != 3 && ncomp != 9) return fail(GBG_E_DIM, "ncomp must be 1, 3 or 9");
  if ((st = same_or_disjoint(dst, src, ncomp * p->M * 2 * p->elem))) return st;
  return op_laplacian(p, dst, src, ncomp, 1, as_stream(stream));
  GBG_GUARD_END
}

gbg_status gbg_helmholtz(gbg_plan* p, void* dst, const void* src, int has_band, double k_lo,
                         double k_hi, double alpha, double beta, void* stream) {
  GBG_GUARD_BEGIN
  gbg_status st = need_plan(p);
  if (st) return st;
  if (!dst || !src) return fail(GBG_E_ARG, "null field buffer");
  if (has_band && (!std::isfinite(k_lo) || !std::isfinite(k_hi) || !(0 <= k_lo && k_lo <= k_hi)))
    return fail(GBG_E_ARG, "band edges must satisfy 0 <= low <= high");
  if ((st = same_or_disjoint(dst, src, 3 * p->M * 2 * p->elem))) return st;
  return op_helmholtz(p, dst, src, has_band, k_lo, k_hi, alpha, beta, as_stream(stream));
  GBG_GUARD_END
}

gbg_status gbg_convolve_band(gbg_plan* p, void* dst, const void* src, int ncomp, double k_lo,
                             double k_hi, double alpha, double beta, void* stream) {
  GBG_GUARD_BEGIN
  gbg_status st = need_plan(p);
  if (st) return st;
  if (!dst || !src) return fail(GBG_E_ARG, "null field buffer");
  if (ncomp != 1 && ncomp != 3 && ncomp != 9) return fail(GBG_E_DIM, "ncomp must be 1, 3 or 9");
  if (!std::isfinite(k_lo) || !std::isfinite(k_hi) || !(0 <= k_lo && k_lo <= k_hi))
    return fail(GBG_E_ARG, "band edges must satisfy 0 <= low <= high");
  if ((st = same_or_disjoint(dst, src, ncomp * p->M * 2 * p->elem))) return st;
  return op_convolve_band(p, dst, src, ncomp, k_lo, k_hi, alpha, beta, as_stream(stream));
  GBG_GUARD_END
}

gbg_status gbg_mean_free(gbg_plan* p, void* spec, int ncomp, void* stream) {
  GBG_GUARD_BEGIN
  gbg_status st = need_plan(p);
  if (st) return st;
  if (!spec) return fail(GBG_E_ARG, "null field buffer");
  if (ncomp < 1 || ncomp > 64) return fail(GBG_E_DIM, "ncomp out of range");
  return op_mean_free(p, spec, ncomp, as_stream(stream));
  GBG_GUARD_END
}

gbg_status gbg_cross(gbg_plan* p, void* dst, const void* u, const void* v, void* stream) {
  GBG_GUARD_BEGIN
  gbg_status st = need_plan(p);
  if (st) return st;
  if (!dst || !u || !v) return fail(GBG_E_ARG, "null field buffer");
  const size_t b = 3 * p->P * p->elem;
  if ((st = same_or_disjoint(dst, u, b)) || (st = same_or_disjoint(dst, v, b))) return st;
  return op_cross(p, dst, u, v, as_stream(stream));
  GBG_GUARD_END
}

gbg_status gbg_multiply(gbg_plan* p, void* dst, const void* a, const void* b, void* stream) {
  GBG_GUARD_BEGIN
  gbg_status st = need_plan(p);
  if (st) return st;
  if (!dst || !a || !b) return fail(GBG_E_ARG, "null field buffer");
  const size_t by = p->P * p->elem;
  if ((st = same_or_disjoint(dst, a, by)) || (st = same_or_disjoint(dst, b, by))) return st;
  return op_multiply(p, dst, a, b, as_stream(stream));
  GBG_GUARD_END
}

gbg_status gbg_inner_spectral(gbg_plan* p, const void* f, const void* g, int ncomp, double* out,
                              void* stream) {
  GBG_GUARD_BEGIN
  gbg_status st = need_plan(p);
  if (st) return st;
  if (!f || !g || !out) return fail(GBG_E_ARG, "null argument");
  if (ncomp < 1 || ncomp > 9) return fail(GBG_E_DIM, "ncomp out of range");
  hipStream_t s = as_stream(stream);
  if ((st = op_inner(p, f, g, ncomp, 2, s))) return st;
  if ((st = check_hip(hipMemcpyAsync(out, p->red_out + 2, sizeof(double), hipMemcpyDeviceToHost, s), "inner d2h")))
    return st;
  return check_hip(hipStreamSynchronize(s), "inner sync");
  GBG_GUARD_END
}

gbg_status gbg_inner_physical(gbg_plan* p, const void* f, const void* g, int ncomp, double* out,
                              void* stream) {
  GBG_GUARD_BEGIN
  gbg_status st = need_plan(p);
  if (st) return st;
  if (!f || !g || !out) return fail(GBG_E_ARG, "null argument");
  if (ncomp != 1 && ncomp != 3 && ncomp != 9) return fail(GBG_E_DIM, "ncomp must be 1, 3 or 9");
  hipStream_t s = as_stream(stream);
  if ((st = op_inner_physical(p, f, g, ncomp, 2, s))) return st;
  if ((st = check_hip(hipMemcpyAsync(out, p->red_out + 2, sizeof(double), hipMemcpyDeviceToHost, s),
                      "inner d2h")))
    return st;
  return check_hip(hipStreamSynchronize(s), "inner sync");
  GBG_GUARD_END
}

gbg_status gbg_radial_shells(gbg_plan* p, int* nshells, double* dk) {
  GBG_GUARD_BEGIN
  gbg_status st = need_plan(p);
  if (st) return st;
  if (!nshells || !dk) return fail(GBG_E_ARG, "null argument");
  radial_shells(p, nshells, dk);
  return GBG_OK;
  GBG_GUARD_END
}

gbg_status gbg_radial_spectrum(gbg_plan* p, const void* f, const void* g, int ncomp,
                               double* spectrum, int nshells, void* stream) {
  GBG_GUARD_BEGIN
  gbg_status st = need_plan(p);
  if (st) return st;
  if (!f || !spectrum) return fail(GBG_E_ARG, "null argument");
  if (ncomp != 1 && ncomp != 3 && ncomp != 9) return fail(GBG_E_DIM, "ncomp must be 1, 3 or 9");
  int n = 0;
  double dk = 0.0;
  radial_shells(p, &n, &dk);
  if (nshells != n) return fail(GBG_E_DIM, "nshells does not match the plan's shell count");
  hipStream_t s = as_stream(stream);
  if ((st = ensure_work(p, radial_work_bytes(p)))) return st;
  double* part = (double*)p->work;
  double* out = part + (radial_work_bytes(p) / sizeof(double) - (size_t)n);
  if ((st = op_radial_spectrum(p, f, g, ncomp, n, out, part, s))) return st;
  if ((st = check_hip(hipMemcpyAsync(spectrum, out, sizeof(double) * n, hipMemcpyDeviceToHost, s),
                      "spectrum d2h")))
    return st;
  return check_hip(hipStreamSynchronize(s), "spectrum sync");
  GBG_GUARD_END
}

gbg_status gbg_exponential_convolve(gbg_plan* p, void* dst, const void* f, int ncomp,
                                    double coefficient, const void* offset, double offset_scale,
                                    double alpha, double beta, void* stream) {
  GBG_GUARD_BEGIN
  gbg_status st = need_plan(p);
  if (st) return st;
  if (!dst || !f) return fail(GBG_E_ARG, "null field buffer");
  if (ncomp != 1 && ncomp != 3 && ncomp != 9) return fail(GBG_E_DIM, "ncomp must be 1, 3 or 9");
  if (!std::isfinite(coefficient) || !std::isfinite(offset_scale) || !std::isfinite(alpha) ||
      !std::isfinite(beta))
    return fail(GBG_E_ARG, "coefficient, offset_scale, alpha and beta must be finite");
  const size_t b = ncomp * p->M * 2 * p->elem, b1 = p->M * 2 * p->elem;
  if ((st = same_or_disjoint(dst, f, b))) return st;
  // the offset may be a component of dst (read per mode before any write there)
  if (offset && offset != dst && overlaps(dst, b, offset, b1))
    return fail(GBG_E_ALIAS, "destination partially overlaps the offset");
  return op_exponential(p, dst, f, ncomp, coefficient, offset, offset_scale, alpha, beta,
                        as_stream(stream));
  GBG_GUARD_END
}

gbg_status gbg_exponential_combine(gbg_plan* p, void* dst, double alpha, const void* a,
                                   double beta, const void* b, int ncomp, double coefficient,
                                   const void* offset, double offset_scale, void* stream) {
  GBG_GUARD_BEGIN
  gbg_status st = need_plan(p);
  if (st) return st;
  if (!dst || !a || !b) return fail(GBG_E_ARG, "null field buffer");
  if (ncomp != 1 && ncomp != 3 && ncomp != 9) return fail(GBG_E_DIM, "ncomp must be 1, 3 or 9");
  if (!std::isfinite(coefficient) || !std::isfinite(offset_scale) || !std::isfinite(alpha) ||
      !std::isfinite(beta))
    return fail(GBG_E_ARG, "coefficient, offset_scale, alpha and beta must be finite");
  const size_t n = ncomp * p->M * 2 * p->elem, b1 = p->M * 2 * p->elem;
  if ((st = same_or_disjoint(dst, a, n)) || (st = same_or_disjoint(dst, b, n))) return st;
  if (offset && offset != dst && overlaps(dst, n, offset, b1))
    return fail(GBG_E_ALIAS, "destination partially overlaps the offset");
  return op_exponential_combine(p, dst, alpha, a, beta, b, ncomp, coefficient, offset,
                                offset_scale, as_stream(stream));
  GBG_GUARD_END
}

gbg_status gbg_reality_residual(gbg_plan* p, const void* spec, int ncomp, double* residual,
                                void* stream) {
  GBG_GUARD_BEGIN
  gbg_status st = need_plan(p);
  if (st) return st;
  if (!spec || !residual) return fail(GBG_E_ARG, "null argument");
  if (ncomp != 1 && ncomp != 3 && ncomp != 9) return fail(GBG_E_DIM, "ncomp must be 1, 3 or 9");
  hipStream_t s = as_stream(stream);
  if ((st = op_reality_residual(p, spec, ncomp, p->red_out + 3, s))) return st;
  if ((st = check_hip(hipMemcpyAsync(residual, p->red_out + 3, sizeof(double),
                                     hipMemcpyDeviceToHost, s), "residual d2h")))
    return st;
  return check_hip(hipStreamSynchronize(s), "residual sync");
  GBG_GUARD_END
}

gbg_status gbg_reality_projection(gbg_plan* p, void* dst, const void* src, int ncomp,
                                  void* stream) {
  GBG_GUARD_BEGIN
  gbg_status st = need_plan(p);
  if (st) return st;
  if (!dst || !src) return fail(GBG_E_ARG, "null field buffer");
  if (ncomp != 1 && ncomp != 3 && ncomp != 9) return fail(GBG_E_DIM, "ncomp must be 1, 3 or 9");
  const size_t b = ncomp * p->M * 2 * p->elem;
  if ((st = same_or_disjoint(dst, src, b))) return st;
  hipStream_t s = as_stream(stream);
  if (dst != src &&
      (st = check_hip(hipMemcpyAsync(dst, src, b, hipMemcpyDeviceToDevice, s), "projection copy")))
    return st;
  return op_reality_project(p, dst, ncomp, s);
  GBG_GUARD_END
}

gbg_status gbg_uxw(gbg_plan* p, const void* u_hat, void* n_hat, double kc, double* energy,
                   double* enstrophy, unsigned flags, void* stream) {
  GBG_GUARD_BEGIN
  gbg_status st = uxw_check_args(p, u_hat, n_hat, kc);
  if (st) return st;
  hipStream_t s = as_stream(stream);
  if ((st = pipe_uxw(p, u_hat, n_hat, kc, flags, s))) return st;
  return uxw_reductions(p, energy, enstrophy, s);
  GBG_GUARD_END
}

gbg_status gbg_uxw_combine(gbg_plan* p, const void* u_hat, void* n_hat, double kc,
                           const gbg_uxw_combination* comb, double* energy, double* enstrophy,
                           unsigned flags, void* stream) {
  GBG_GUARD_BEGIN
  gbg_status st = need_plan(p);
  if (st) return st;
  if (!u_hat || !comb || !comb->dst) return fail(GBG_E_ARG, "null field buffer");
  if (!std::isfinite(kc) || kc < 0) return fail(GBG_E_ARG, "cutoff must be finite and >= 0");
  const double w[5] = {comb->alpha, comb->beta, comb->delta, comb->gamma, comb->coefficient};
  for (double v : w)
    if (!std::isfinite(v)) return fail(GBG_E_ARG, "combination weights must be finite");
  const size_t b = 3 * p->M * 2 * p->elem;
  const void* dst = comb->dst;
  if (overlaps(dst, b, u_hat, b)) return fail(GBG_E_ALIAS, "dst must not overlap u_hat");
  if (n_hat) {
    if (overlaps(n_hat, b, u_hat, b)) return fail(GBG_E_ALIAS, "u_hat and n_hat must not overlap");
    if (overlaps(n_hat, b, dst, b)) return fail(GBG_E_ALIAS, "n_hat must not overlap dst");
    if ((comb->a && overlaps(n_hat, b, comb->a, b)) || (comb->b && overlaps(n_hat, b, comb->b, b)))
      return fail(GBG_E_ALIAS, "n_hat must not overlap a or b");
  }
  // dst may be a or b itself (each mode is read before it is written), not a
  // shifted view of either
  for (const void* x : {comb->a, comb->b})
    if (x && x != dst && overlaps(dst, b, x, b))
      return fail(GBG_E_ALIAS, "dst overlaps a or b without being it");
  UxwComb cb;
  cb.dst = comb->dst;
  cb.a = comb->a;
  cb.b = comb->b;
  cb.alpha = comb->alpha;
  cb.beta = comb->beta;
  cb.delta = comb->delta;
  cb.gamma = comb->gamma;
  cb.coef = comb->coefficient;
  hipStream_t s = as_stream(stream);
  if ((st = pipe_uxw_comb(p, u_hat, n_hat, kc, cb, flags, s))) return st;
  return uxw_reductions(p, energy, enstrophy, s);
  GBG_GUARD_END
}

gbg_status gbg_dealiased_product(gbg_plan* p, void* p_hat, const void* a_hat, const void* b_hat,
                                 double kc, void* stream) {
  GBG_GUARD_BEGIN
  gbg_status st = need_plan(p);
  if (st) return st;
  if (!p_hat || !a_hat || !b_hat) return fail(GBG_E_ARG, "null field buffer");
  if (!std::isfinite(kc) || kc < 0) return fail(GBG_E_ARG, "cutoff must be finite and >= 0");
  const size_t b = p->M * 2 * p->elem;
  if (overlaps(p_hat, b, a_hat, b) || overlaps(p_hat, b, b_hat, b))
    return fail(GBG_E_ALIAS, "output must not overlap the inputs");
  return pipe_dealiased_product(p, p_hat, a_hat, b_hat, kc, as_stream(stream));
  GBG_GUARD_END
}

gbg_status gbg_plan_set_profiling(gbg_plan* p, int enabled) {
  GBG_GUARD_BEGIN
  gbg_status st = need_plan(p);
  if (st) return st;
  if (enabled && !p->ev[0]) {
    for (int i = 0; i < kMaxEvents; ++i)
      if ((st = check_hip(hipEventCreate(&p->ev[i]), "event create"))) return st;
  }
  p->profiling = enabled ? 1 : 0;
  p->nev = 0;
  return GBG_OK;
  GBG_GUARD_END
}

gbg_status gbg_plan_stage_times(gbg_plan* p, double* ms, int* n) {
  GBG_GUARD_BEGIN
  gbg_status st = need_plan(p);
  if (st) return st;
  if (!n) return fail(GBG_E_ARG, "null count");
  const int cap = *n;
  const int cnt = p->nev > 1 ? p->nev - 1 : 0;
  *n = cnt;
  if (cnt == 0) return GBG_OK;
  if ((st = check_hip(hipEventSynchronize(p->ev[p->nev - 1]), "event sync"))) return st;
  for (int i = 0; i < cnt && i < cap && ms; ++i) {
    float t = 0.f;
    if ((st = check_hip(hipEventElapsedTime(&t, p->ev[i], p->ev[i + 1]), "event elapsed"))) return st;
    ms[i] = (double)t;
  }
  return GBG_OK;
  GBG_GUARD_END
}

}  // extern "C"
