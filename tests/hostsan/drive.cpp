// Drives libgbg's C ABI on the CPU stand-in runtime (hip_stub.cpp) under
// AddressSanitizer: every entry point on fused, composed, permuted-layout and
// chirp-z grids, the batch and the server, then each scenario again with the
// k-th resource call made to fail for every k the scenario reaches.  A failed
// call must return an error status (or finish correctly where the library
// declines an optional resource), release every device allocation it made,
// and leave memory intact.  Test infrastructure only.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "gbg.h"

extern "C" void stub_fail_at(long k);
extern "C" long stub_calls();
extern "C" long stub_fired();
extern "C" long stub_live();

static int failures = 0;
#define EXPECT(c)                                                                   \
  do {                                                                              \
    if (!(c)) {                                                                     \
      fprintf(stderr, "FAIL %s:%d: %s (%s)\n", __FILE__, __LINE__, #c, gbg_last_error()); \
      ++failures;                                                                   \
    }                                                                               \
  } while (0)

#define TRY(x)                      \
  do {                              \
    gbg_status s_ = (x);            \
    if (s_ != GBG_OK) return s_;    \
  } while (0)

// device buffers released when the holder goes out of scope
struct Bufs {
  std::vector<void*> p;
  ~Bufs() {
    for (void* x : p) gbg_free(x);
  }
  gbg_status get(void** out, size_t bytes) {
    *out = nullptr;
    gbg_status s = gbg_malloc(out, bytes);
    if (s == GBG_OK) p.push_back(*out);
    return s;
  }
};
struct Plan {
  gbg_plan* h = nullptr;
  ~Plan() {
    if (h) gbg_plan_destroy(h);
  }
};

// every operator, transform, reduction and pipeline on one plan
static gbg_status ops_on(const int n[3], int elem, const int dims[3]) {
  const double L[3] = {6.283185307179586, 6.283185307179586, 6.283185307179586};
  Plan pl;
  TRY(dims ? gbg_plan_create_layout(n, L, elem, dims, &pl.h) : gbg_plan_create(n, L, elem, &pl.h));
  size_t M = 0, P = 0, dev = 0;
  TRY(gbg_plan_query(pl.h, &M, &P, &dev));
  const size_t sc = M * 2 * elem, ph = P * elem;
  Bufs b;
  void *phys3, *spec3, *spec3b, *spec9, *spec1, *scratch, *out3;
  TRY(b.get(&phys3, 3 * ph));
  TRY(b.get(&spec3, 3 * sc));
  TRY(b.get(&spec3b, 3 * sc));
  TRY(b.get(&spec9, 9 * sc));
  TRY(b.get(&spec1, sc));
  TRY(b.get(&scratch, 3 * sc));
  TRY(b.get(&out3, 3 * sc));
  TRY(gbg_memset(phys3, 0, 3 * ph, nullptr));
  TRY(gbg_memset(spec3, 0, 3 * sc, nullptr));
  TRY(gbg_to_spectral(pl.h, spec3, phys3, 3, nullptr));
  TRY(gbg_to_physical(pl.h, phys3, spec3, 3, scratch, nullptr));
  TRY(gbg_to_physical(pl.h, phys3, spec3b, 3, nullptr, nullptr));
  TRY(gbg_gradient(pl.h, spec3, spec1, 0, nullptr));
  TRY(gbg_gradient(pl.h, spec9, spec3, 1, nullptr));
  TRY(gbg_symmetric_gradient(pl.h, spec9, spec3, nullptr));
  TRY(gbg_divergence(pl.h, spec1, spec3, 1, 1, 1.0, 3.0, 0.5, 2.0, nullptr));
  TRY(gbg_curl(pl.h, spec3b, spec3, nullptr));
  TRY(gbg_curl(pl.h, spec3b, spec3b, nullptr));
  TRY(gbg_inverse_curl(pl.h, spec3b, spec3, nullptr));
  TRY(gbg_laplacian(pl.h, spec3b, spec3, 3, nullptr));
  TRY(gbg_inverse_laplacian(pl.h, spec1, spec1, 1, nullptr));
  TRY(gbg_helmholtz(pl.h, spec3b, spec3, 1, 0.0, 4.0, 1.0, 0.0, nullptr));
  TRY(gbg_convolve_band(pl.h, spec3b, spec3, 3, 0.0, 4.0, 1.0, 0.0, nullptr));
  TRY(gbg_mean_free(pl.h, spec3b, 3, nullptr));
  TRY(gbg_cross(pl.h, phys3, phys3, phys3, nullptr));
  TRY(gbg_multiply(pl.h, phys3, phys3, phys3, nullptr));
  double v = 0;
  TRY(gbg_inner_spectral(pl.h, spec3, spec3b, 3, &v, nullptr));
  TRY(gbg_inner_physical(pl.h, phys3, phys3, 3, &v, nullptr));
  int nsh = 0;
  double dk = 0;
  TRY(gbg_radial_shells(pl.h, &nsh, &dk));
  std::vector<double> sp((size_t)nsh);
  TRY(gbg_radial_spectrum(pl.h, spec3, nullptr, 3, sp.data(), nsh, nullptr));
  TRY(gbg_radial_spectrum(pl.h, spec3, spec3b, 3, sp.data(), nsh, nullptr));
  TRY(gbg_exponential_convolve(pl.h, spec3b, spec3, 3, -0.1, spec1, 1.0, 1.0, 0.0, nullptr));
  TRY(gbg_exponential_combine(pl.h, spec3b, 1.0, spec3, 0.5, spec3b, 3, -0.1, nullptr, 1.0, nullptr));
  TRY(gbg_reality_residual(pl.h, spec3, 3, &v, nullptr));
  TRY(gbg_reality_projection(pl.h, spec3b, spec3, 3, nullptr));
  TRY(gbg_reality_projection(pl.h, spec3b, spec3b, 3, nullptr));
  double E = 0, Z = 0;
  const double kc = 2.0;
  for (unsigned flags : {0u, (unsigned)GBG_UXW_UNFUSED, (unsigned)GBG_UXW_SLABS}) {
    TRY(gbg_uxw(pl.h, spec3, out3, kc, &E, &Z, flags, nullptr));
    gbg_uxw_combination cb;
    cb.dst = spec3b;
    cb.a = spec3b;
    cb.alpha = 1.0;
    cb.b = nullptr;
    cb.beta = 0.0;
    cb.delta = 0.01;
    cb.gamma = 0.0;
    cb.coefficient = -0.001;
    TRY(gbg_uxw_combine(pl.h, spec3, out3, kc, &cb, &E, &Z, flags, nullptr));
    cb.b = out3;
    cb.beta = 0.005;
    cb.gamma = 0.005;
    TRY(gbg_uxw_combine(pl.h, spec3, nullptr, kc, &cb, nullptr, nullptr, flags, nullptr));
  }
  TRY(gbg_dealiased_product(pl.h, spec1, spec3, spec3b, kc, nullptr));
  const gbg_status na = gbg_uxw_noalloc(pl.h, spec3, out3, kc, &E, &Z, 0, nullptr);
  if (na != GBG_OK && na != GBG_E_UNSUPPORTED) return na;
  TRY(gbg_plan_set_profiling(pl.h, 1));
  TRY(gbg_uxw(pl.h, spec3, out3, kc, nullptr, nullptr, 0, nullptr));
  double ms[16];
  int nms = 16;
  TRY(gbg_plan_stage_times(pl.h, ms, &nms));
  return GBG_OK;
}

static gbg_status sc_ops_fused64() {
  const int n[3] = {32, 32, 32};
  return ops_on(n, 8, nullptr);
}
static gbg_status sc_ops_fused32() {
  const int n[3] = {64, 32, 32};
  return ops_on(n, 4, nullptr);
}
// the z-slab chain on several slabs and two streams: GBG_UXW_SLAB at plan
// creation sets 5 planes per slab (a last slab of 4) and two slots
static gbg_status sc_uxw_slabs(int elem) {
  const int n[3] = {32, 32, 64};
  const double L[3] = {6.283185307179586, 6.283185307179586, 6.283185307179586};
  Plan pl;
  setenv("GBG_UXW_SLAB", "5,2", 1);
  const gbg_status cs = gbg_plan_create(n, L, elem, &pl.h);
  unsetenv("GBG_UXW_SLAB");
  TRY(cs);
  size_t M = 0, P = 0, dev = 0;
  TRY(gbg_plan_query(pl.h, &M, &P, &dev));
  const size_t sc = M * 2 * elem;
  Bufs b;
  void *spec3, *out3, *dst3;
  TRY(b.get(&spec3, 3 * sc));
  TRY(b.get(&out3, 3 * sc));
  TRY(b.get(&dst3, 3 * sc));
  TRY(gbg_memset(spec3, 0, 3 * sc, nullptr));
  double E = 0, Z = 0;
  TRY(gbg_uxw(pl.h, spec3, out3, 2.0, &E, &Z, GBG_UXW_SLABS, nullptr));
  TRY(gbg_uxw(pl.h, spec3, out3, 2.0, nullptr, nullptr, GBG_UXW_SLABS, nullptr));
  gbg_uxw_combination cb;
  cb.dst = dst3;
  cb.a = spec3;
  cb.alpha = 1.0;
  cb.b = nullptr;
  cb.beta = 0.0;
  cb.delta = 0.01;
  cb.gamma = 0.0;
  cb.coefficient = -0.001;
  TRY(gbg_uxw_combine(pl.h, spec3, out3, 2.0, &cb, &E, &Z, GBG_UXW_SLABS, nullptr));
  return GBG_OK;
}
static gbg_status sc_uxw_slabs64() { return sc_uxw_slabs(8); }
static gbg_status sc_uxw_slabs32() { return sc_uxw_slabs(4); }
// the fused u x omega at Ny = 1024, where the z pass writes all six fields
// (the y pass forms the kx and ky terms of w only up to Ny = 512)
static gbg_status sc_uxw_ny1024(int elem) {
  const int n[3] = {32, 1024, 32};
  const double L[3] = {6.283185307179586, 6.283185307179586, 6.283185307179586};
  Plan pl;
  TRY(gbg_plan_create(n, L, elem, &pl.h));
  size_t M = 0, P = 0, dev = 0;
  TRY(gbg_plan_query(pl.h, &M, &P, &dev));
  const size_t sc = M * 2 * elem;
  Bufs b;
  void *spec3, *out3;
  TRY(b.get(&spec3, 3 * sc));
  TRY(b.get(&out3, 3 * sc));
  TRY(gbg_memset(spec3, 0, 3 * sc, nullptr));
  double E = 0, Z = 0;
  TRY(gbg_uxw(pl.h, spec3, out3, 12.0, &E, &Z, 0, nullptr));
  return GBG_OK;
}
static gbg_status sc_uxw_ny1024_64() { return sc_uxw_ny1024(8); }
static gbg_status sc_uxw_ny1024_32() { return sc_uxw_ny1024(4); }
static gbg_status sc_ops_composed() {
  const int n[3] = {16, 12, 10};
  return ops_on(n, 8, nullptr);
}
static gbg_status sc_ops_layout() {
  const int n[3] = {8, 7, 6}, d[3] = {3, 2, 1};
  return ops_on(n, 8, d);
}
static gbg_status sc_ops_chirp() {
  const int n[3] = {17, 19, 6};
  return ops_on(n, 4, nullptr);
}

static gbg_status sc_autotune() {
  const int n[3] = {32, 48, 17};
  const double L[3] = {1.0, 2.0, 3.0};
  Plan pl;
  TRY(gbg_plan_create(n, L, 8, &pl.h));
  double best = 0;
  return gbg_plan_autotune(pl.h, nullptr, &best);
}

static gbg_status sc_batch() {
  gbg_batch* bt = nullptr;
  TRY(gbg_batch_create(3, 1, &bt));
  struct Guard {
    gbg_batch* b;
    ~Guard() { gbg_batch_destroy(b); }
  } guard{bt};
  Bufs b;
  const int grids[4][3] = {{32, 32, 32}, {16, 12, 10}, {64, 32, 32}, {17, 8, 6}};
  const double L[3] = {6.283185307179586, 6.283185307179586, 6.283185307179586};
  for (int i = 0; i < 4; ++i) {
    const int* n = grids[i];
    const size_t M = (size_t)(n[0] / 2 + 1) * n[1] * n[2];
    const int elem = i % 2 ? 4 : 8, kind = i % 2;
    const int nin = kind ? 3 : 1;
    void *in, *o0, *o1;
    TRY(b.get(&in, nin * M * 2 * elem));
    TRY(b.get(&o0, 3 * nin * M * 2 * elem));
    TRY(b.get(&o1, nin * M * 2 * elem));
    TRY(gbg_batch_add(bt, kind, n, L, elem, in, o0, o1, 3.0));
  }
  TRY(gbg_batch_instantiate(bt, nullptr));
  TRY(gbg_batch_launch(bt, nullptr));
  TRY(gbg_batch_launch(bt, nullptr));
  return GBG_OK;
}

static gbg_status sc_server() {
  gbg_server* sv = nullptr;
  TRY(gbg_server_create((size_t)1 << 30, 2, &sv));
  struct Guard {
    gbg_server* s;
    ~Guard() { gbg_server_destroy(s); }
  } guard{sv};
  Bufs b;
  const int n[3] = {16, 16, 16};
  const size_t M = 9 * 16 * 16, P = 16 * 16 * 16;
  void *phys, *s0, *s1;
  TRY(b.get(&phys, 3 * P * 8));
  TRY(b.get(&s0, 3 * M * 16));
  TRY(b.get(&s1, 3 * M * 16));
  std::vector<unsigned long long> tickets;
  const int kinds[6] = {GBG_JOB_TO_SPECTRAL, GBG_JOB_TO_PHYSICAL, GBG_JOB_CURL,
                        GBG_JOB_INVERSE_LAPLACIAN, GBG_JOB_DEALIASED_PRODUCT, GBG_JOB_UXW};
  for (int k : kinds) {
    gbg_job j;
    memset(&j, 0, sizeof(j));
    j.kind = k;
    for (int a = 0; a < 3; ++a) {
      j.n[a] = n[a];
      j.L[a] = 6.283185307179586;
    }
    j.elem_bytes = 8;
    j.ncomp = k == GBG_JOB_DEALIASED_PRODUCT ? 1 : 3;
    j.kc = 5.0;
    j.in0 = k == GBG_JOB_TO_SPECTRAL ? phys : s0;
    j.in1 = k == GBG_JOB_DEALIASED_PRODUCT ? s0 : nullptr;
    j.out = k == GBG_JOB_TO_PHYSICAL ? phys : s1;
    unsigned long long t = 0;
    TRY(gbg_server_submit(sv, &j, &t));
    tickets.push_back(t);
  }
  gbg_status first = GBG_OK;
  for (unsigned long long t : tickets) {
    double ms = 0;
    const gbg_status s = gbg_server_wait(sv, t, &ms);
    if (s != GBG_OK && first == GBG_OK) first = s;
  }
  double p50, p99, jps;
  unsigned long long done = 0;
  TRY(gbg_server_stats(sv, &p50, &p99, &jps, &done));
  size_t bytes = 0;
  int plans = 0;
  TRY(gbg_server_resident(sv, &bytes, &plans));
  return first;
}

// argument refusals: each must be refused without touching the device
static void refusals() {
  const int n[3] = {8, 8, 8}, bad_dims[3] = {1, 1, 3}, zero[3] = {0, 8, 8};
  const double L[3] = {1, 1, 1}, Lbad[3] = {1, -1, 1};
  gbg_plan* p = nullptr;
  EXPECT(gbg_plan_create(zero, L, 8, &p) == GBG_E_ARG);
  EXPECT(gbg_plan_create(n, Lbad, 8, &p) == GBG_E_ARG);
  EXPECT(gbg_plan_create(n, L, 2, &p) == GBG_E_ARG);
  EXPECT(gbg_plan_create_layout(n, L, 8, bad_dims, &p) == GBG_E_ARG);
  const int big[3] = {2053, 2, 2};
  EXPECT(gbg_plan_create(big, L, 8, &p) == GBG_E_UNSUPPORTED);
  EXPECT(gbg_plan_create(n, L, 8, &p) == GBG_OK);
  Bufs b;
  void* x;
  EXPECT(b.get(&x, 9 * 5 * 8 * 8 * 16) == GBG_OK);
  char* c = (char*)x;
  EXPECT(gbg_curl(p, nullptr, c, nullptr) == GBG_E_ARG);
  EXPECT(gbg_gradient(p, c, c + 16, 1, nullptr) == GBG_E_ALIAS);
  EXPECT(gbg_symmetric_gradient(p, c, c + 16, nullptr) == GBG_E_ALIAS);
  EXPECT(gbg_gradient(p, c, c, 3, nullptr) == GBG_E_DIM);
  EXPECT(gbg_uxw(p, c, c + 32, 2.0, nullptr, nullptr, 0, nullptr) == GBG_E_ALIAS);
  EXPECT(gbg_uxw(p, c, c + 3 * 5 * 8 * 8 * 16, -1.0, nullptr, nullptr, 0, nullptr) == GBG_E_ARG);
  gbg_uxw_combination cb;
  memset(&cb, 0, sizeof(cb));
  cb.dst = c;
  EXPECT(gbg_uxw_combine(p, c + 64, nullptr, 2.0, &cb, nullptr, nullptr, 0, nullptr) == GBG_E_ALIAS);
  cb.delta = NAN;
  EXPECT(gbg_uxw_combine(p, c + 3 * 5 * 8 * 8 * 16, nullptr, 2.0, &cb, nullptr, nullptr, 0, nullptr) ==
         GBG_E_ARG);
  int nsh = 0;
  double dk = 0;
  EXPECT(gbg_radial_shells(p, &nsh, &dk) == GBG_OK);
  std::vector<double> sp((size_t)nsh + 1);
  EXPECT(gbg_radial_spectrum(p, c, nullptr, 3, sp.data(), nsh + 1, nullptr) == GBG_E_DIM);
  EXPECT(gbg_plan_destroy(p) == GBG_OK);
  gbg_batch* bt = nullptr;
  EXPECT(gbg_batch_create(0, 0, &bt) == GBG_E_ARG);
  gbg_server* sv = nullptr;
  EXPECT(gbg_server_create(0, 1, &sv) == GBG_E_ARG || sv == nullptr);
  if (sv) gbg_server_destroy(sv);
}

int main() {
  refusals();
  struct Sc {
    const char* name;
    std::function<gbg_status()> run;
  } scs[] = {{"ops fused FP64 32^3", sc_ops_fused64},
             {"ops fused FP32 64x32x32", sc_ops_fused32},
             {"uxw slabs FP64 32x32x64", sc_uxw_slabs64},
             {"uxw slabs FP32 32x32x64", sc_uxw_slabs32},
             {"uxw six fields FP64 32x1024x32", sc_uxw_ny1024_64},
             {"uxw six fields FP32 32x1024x32", sc_uxw_ny1024_32},
             {"ops composed 16x12x10", sc_ops_composed},
             {"ops layout (3,2,1) 8x7x6", sc_ops_layout},
             {"ops chirp-z FP32 17x19x6", sc_ops_chirp},
             {"autotune 32x48x17", sc_autotune},
             {"batch", sc_batch},
             {"server", sc_server}};
  const long base = stub_live();
  for (auto& sc : scs) {
    stub_fail_at(0);
    const gbg_status s = sc.run();
    EXPECT(s == GBG_OK);
    EXPECT(stub_live() == base);
    const long reach = stub_calls();
    long refused = 0, absorbed = 0;
    for (long k = 1; k <= reach; ++k) {
      stub_fail_at(k);
      const gbg_status f = sc.run();
      const bool hit = stub_fired() != 0;
      stub_fail_at(0);
      if (!hit) break;   // the run made fewer resource calls than counted
      if (f == GBG_OK)
        ++absorbed;      // an optional resource declined (tile list, second stream)
      else
        ++refused;
      if (f != GBG_OK && f != GBG_E_OOM && f != GBG_E_DEVICE) {
        fprintf(stderr, "FAIL %s: injected failure %ld returned %d (%s)\n", sc.name, k, (int)f,
                gbg_last_error());
        ++failures;
      }
      if (stub_live() != base) {
        fprintf(stderr, "FAIL %s: injected failure %ld left %ld device allocations\n", sc.name, k,
                stub_live() - base);
        ++failures;
      }
    }
    printf("%-28s %3ld resource calls: %3ld failures refused, %3ld absorbed\n", sc.name, reach,
           refused, absorbed);
  }
  printf(failures ? "hostsan: %d failures\n" : "hostsan: ok\n", failures);
  return failures ? 1 : 0;
}
