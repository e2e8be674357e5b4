// gbg_uxw_noalloc: the dealiased u x omega term for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  It runs the same pass schedule as
// gbg_uxw, with the compact intermediates and the K1 partial sums in a
// file-scope device array of kScratchBytes instead of the plan's work area,
// and it creates nothing on the device: no work area, no tile list of the
// last pass (one the plan already caches for this cutoff is used; otherwise
// the last pass classifies its tiles itself) and no second stream.
//
// Each device the library runs on holds its own copy of the array, at its
// own address, until the process ends.  The host-side name `scratch` is only
// a stand-in for that memory: the device address is resolved once per device
// with hipGetSymbolAddress (with that device current) on the entry point's
// first call there, and only that address is ever passed to a kernel.  Every
// call on a device shares the array: calls that overlap in time on one
// device (on different streams or host threads) race on it.  Nothing in it
// needs to start at zero (each pass writes everything the next one reads),
// so a call does not clear it.
#include <exception>
#include <mutex>
#include <new>
#include <string>

#include "gbg_internal.hpp"

namespace gbg {

// 128 MiB: the fused u x omega of a 128^3 FP64 field at the 2/3 cutoff
// (119.9 MiB), of a 128^3 FP32 field at any cutoff (120.2 MiB), of
// 128x128x256 FP32 or 256x64x64 FP64 at the 2/3 cutoff, and the composed
// path of a 96^3 FP64 field (81.8 MiB).
constexpr size_t kScratchBytes = (size_t)128 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices];

// the device address of `scratch` on the current device (resolved once)
static gbg_status scratch_address(int device, void** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    gbg_status st = check_hip(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)),
                              "hipGetSymbolAddress of the no-allocation scratch");
    if (st) return st;
    if (!ptr) return fail(GBG_E_DEVICE, "hipGetSymbolAddress gave a null scratch address");
    g_scratch_ptr[device] = ptr;
  }
  *out = g_scratch_ptr[device];
  return GBG_OK;
}

}  // namespace gbg

using namespace gbg;

extern "C" gbg_status gbg_uxw_noalloc(gbg_plan* p, const void* u_hat, void* n_hat, double kc,
                                      double* energy, double* enstrophy, unsigned flags,
                                      void* stream) {
  try {
    gbg_status st = uxw_check_args(p, u_hat, n_hat, kc);
    if (st) return st;
    int dev = -1;
    if ((st = check_hip(hipGetDevice(&dev), "hipGetDevice"))) return st;
    if (dev < 0 || dev >= kMaxDevices)
      return fail(GBG_E_UNSUPPORTED, "device index " + std::to_string(dev) + " beyond " +
                                         std::to_string(kMaxDevices));
    if (dev != p->device)
      return fail(GBG_E_ARG, "the plan was created on device " + std::to_string(p->device) +
                                 " but device " + std::to_string(dev) + " is current");
    const size_t need = uxw_work_bytes(p, kc, flags);
    if (need > kScratchBytes)
      return fail(GBG_E_UNSUPPORTED,
                  "u x omega on this grid and cutoff needs " + std::to_string(need) +
                      " bytes of scratch; gbg_uxw_noalloc holds " + std::to_string(kScratchBytes) +
                      " (use gbg_uxw)");
    void* ptr = nullptr;
    if ((st = scratch_address(dev, &ptr))) return st;
    hipStream_t s = as_stream(stream);
    if ((st = pipe_uxw_in(p, u_hat, n_hat, kc, flags, ptr, kScratchBytes, s))) return st;
    return uxw_reductions(p, energy, enstrophy, s);
  } catch (const std::bad_alloc&) {
    return fail(GBG_E_OOM, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(GBG_E_DEVICE, e.what());
  } catch (...) {
    return fail(GBG_E_DEVICE, "unknown exception");
  }
}
