// A CPU stand-in for the HIP runtime, for running libgbg's host code under
// AddressSanitizer without a GPU.  Kernel launches check their shape and do
// nothing; device memory is a host mapping with no access rights (a host
// read or write through a device pointer faults), backed by a shadow copy
// that the copy and fill calls use.  Resource calls (device allocations,
// streams, events, graph instantiation) can be made to fail at the k-th call
// (stub_fail_at), and the live device allocations are counted (stub_live).
// Test infrastructure only.
#include <hip/hip_runtime_api.h>
#include <sys/mman.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <set>

namespace {
struct Region {
  size_t n;
  char* shadow;
};
std::mutex mu;
std::map<char*, Region> regions;   // device allocations
std::set<void*> recorded;          // events recorded since the last capture began
long calls = 0, fail_at = 0, fired = 0;
int capturing = 0;

// true when this resource call is the one to fail
bool inject() {
  std::lock_guard<std::mutex> g(mu);
  ++calls;
  if (fail_at && calls == fail_at) {
    fired = 1;
    return true;
  }
  return false;
}

char* shadow_of(const void* p) {
  std::lock_guard<std::mutex> g(mu);
  char* c = (char*)p;
  auto it = regions.upper_bound(c);
  if (it == regions.begin()) return c;
  --it;
  if (c >= it->first && c < it->first + it->second.n) return it->second.shadow + (c - it->first);
  return c;
}
}  // namespace

extern "C" {
// fail the k-th resource call from now on (0: none); resets the call count
void stub_fail_at(long k) {
  std::lock_guard<std::mutex> g(mu);
  fail_at = k;
  calls = 0;
  fired = 0;
}
long stub_calls() { return calls; }
long stub_fired() { return fired; }
static char* symbol_area_ptr();
long stub_live() {
  std::lock_guard<std::mutex> g(mu);
  return (long)regions.size() - (symbol_area_ptr() ? 1 : 0);
}

hipError_t hipMalloc(void** p, size_t n) {
  if (inject()) return hipErrorOutOfMemory;
  if (!n) n = 1;
  void* m = mmap(nullptr, n, PROT_NONE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
  if (m == MAP_FAILED) return hipErrorOutOfMemory;
  std::lock_guard<std::mutex> g(mu);
  regions[(char*)m] = Region{n, (char*)calloc(1, n)};
  *p = m;
  return hipSuccess;
}
hipError_t hipFree(void* p) {
  if (!p) return hipSuccess;
  std::lock_guard<std::mutex> g(mu);
  auto it = regions.find((char*)p);
  if (it == regions.end()) {
    fprintf(stderr, "stub: hipFree of a pointer that is no allocation\n");
    abort();
  }
  munmap(p, it->second.n);
  free(it->second.shadow);
  regions.erase(it);
  return hipSuccess;
}
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) {
  memmove(shadow_of(d), shadow_of(s), n);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t) {
  return hipMemcpy(d, s, n, k);
}
hipError_t hipMemset(void* d, int v, size_t n) {
  memset(shadow_of(d), v, n);
  return hipSuccess;
}
hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) { return hipMemset(d, v, n); }
hipError_t hipGetDevice(int* d) {
  *d = 0;
  return hipSuccess;
}
hipError_t hipSetDevice(int d) { return d == 0 ? hipSuccess : hipErrorInvalidDevice; }
hipError_t hipGetDeviceCount(int* c) {
  *c = 1;
  return hipSuccess;
}
hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t, int) {
  *v = 256;
  return hipSuccess;
}
hipError_t hipGetLastError() { return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "stub runtime error"; }
hipError_t hipEventCreate(hipEvent_t* e) {
  if (inject()) return hipErrorOutOfMemory;
  *e = (hipEvent_t)malloc(8);
  return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) {
  {
    std::lock_guard<std::mutex> g(mu);
    recorded.erase((void*)e);
  }
  free(e);
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) {
  std::lock_guard<std::mutex> g(mu);
  recorded.insert((void*)e);
  return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) {
  *ms = 1.0f;
  return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
  if (inject()) return hipErrorOutOfMemory;
  *s = (hipStream_t)malloc(8);
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
  free(s);
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
// a wait on an event that was never recorded (since the capture began) is a
// dependency the real runtime would not order
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t e, unsigned) {
  std::lock_guard<std::mutex> g(mu);
  if (!e || !recorded.count((void*)e)) {
    fprintf(stderr, "stub: wait on an event that was not recorded\n");
    abort();
  }
  return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
  std::lock_guard<std::mutex> g(mu);
  capturing = 1;
  recorded.clear();
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* gr) {
  capturing = 0;
  *gr = (hipGraph_t)malloc(8);
  return hipSuccess;
}
hipError_t hipStreamIsCapturing(hipStream_t, hipStreamCaptureStatus* s) {
  *s = capturing ? hipStreamCaptureStatusActive : hipStreamCaptureStatusNone;
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t, hipGraphNode_t*, char*, size_t) {
  if (inject()) return hipErrorOutOfMemory;
  *e = (hipGraphExec_t)malloc(8);
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t gr) {
  free(gr);
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
  free(e);
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipSuccess; }
hipError_t hipLaunchHostFunc(hipStream_t, hipHostFn_t fn, void* a) {
  fn(a);
  return hipSuccess;
}
// kernel launches: the shape must be one the device accepts
static thread_local dim3 cfg_grid, cfg_block;
hipError_t __hipPushCallConfiguration(dim3 gr, dim3 bl, size_t, hipStream_t) {
  cfg_grid = gr;
  cfg_block = bl;
  return hipSuccess;
}
hipError_t __hipPopCallConfiguration(dim3* gr, dim3* bl, size_t* sh, hipStream_t* st) {
  *gr = cfg_grid;
  *bl = cfg_block;
  *sh = 0;
  *st = nullptr;
  return hipSuccess;
}
hipError_t hipLaunchKernel(const void*, dim3 gr, dim3 bl, void**, size_t, hipStream_t) {
  const unsigned long long threads = (unsigned long long)bl.x * bl.y * bl.z;
  if (gr.x == 0 || gr.y == 0 || gr.z == 0 || threads == 0 || threads > 1024 || gr.y > 65535 ||
      gr.z > 65535) {
    fprintf(stderr, "stub: launch of grid (%u, %u, %u) block (%u, %u, %u)\n", gr.x, gr.y, gr.z, bl.x,
            bl.y, bl.z);
    abort();
  }
  return hipSuccess;
}
// every __device__ symbol resolves to one 256 MiB device region, outside the
// count of live allocations
static char* symbol_area = nullptr;
hipError_t hipGetSymbolAddress(void** p, const void*) {
  std::lock_guard<std::mutex> g(mu);
  if (!symbol_area) {
    const size_t n = (size_t)256 << 20;
    void* m = mmap(nullptr, n, PROT_NONE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
    if (m == MAP_FAILED) return hipErrorOutOfMemory;
    symbol_area = (char*)m;
    regions[symbol_area] = Region{n, (char*)calloc(1, n)};
  }
  *p = symbol_area;
  return hipSuccess;
}
static char* symbol_area_ptr() { return symbol_area; }
void** __hipRegisterFatBinary(const void*) {
  static void* h[4];
  return h;
}
void __hipUnregisterFatBinary(void**) {}
void __hipRegisterFunction(void**, const void*, char*, const char*, unsigned, void*, void*, void*,
                           void*, int*) {}
void __hipRegisterVar(void**, char*, char*, const char*, int, size_t, int, int) {}
}
