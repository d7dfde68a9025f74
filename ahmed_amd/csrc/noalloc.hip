// noalloc.hip -- the H-matvec for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  Its workspace (V^T x, chunk partial sums, the interleaved
// right-hand sides) lives in one static device array per GPU instead of the
// buffers the matrix handle grows on demand.
#include <mutex>

#include "ahm_internal.h"

namespace {

// 256 MiB: one right-hand side of an FP64 matrix of about three million
// unknowns at the density of the N = 1M cube benchmark, or fewer unknowns
// with several right-hand sides.  Every GPU this library runs on holds the
// array from the first call on that GPU until the process ends.
constexpr size_t kScratchBytes = (size_t)256 << 20;
constexpr int kMaxDevices = 64;

// Host code must not use this name as a device address (it is the host-side
// shadow); scratch_address() gives the device's own copy.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// device address of `scratch` on `device` (current), looked up once per device
int scratch_address(int device, void** out) {
  if (device < 0 || device >= kMaxDevices)
    return ahm::fail(AHM_ERR_INVALID, "ahm_hmat_matvec_noalloc: device %d out of range", device);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return ahm::fail(AHM_ERR_HIP, "ahm_hmat_matvec_noalloc: hipGetSymbolAddress(scratch): %s",
                       hipGetErrorString(e));
    g_ptr[device] = ptr;
  }
  *out = g_ptr[device];
  return AHM_OK;
}

}  // namespace

extern "C" int ahm_hmat_matvec_noalloc(ahm_hmat h, double alpha, const void* X, uint32_t ldx, void* Y,
                                       uint32_t ldy, uint32_t p, ahm_dtype vec) {
  if (!h) return ahm::fail(AHM_ERR_INVALID, "ahm_hmat_matvec_noalloc: NULL handle");
  const hipError_t e = hipSetDevice(h->ctx->device);
  if (e != hipSuccess)
    return ahm::fail(AHM_ERR_HIP, "ahm_hmat_matvec_noalloc: hipSetDevice: %s", hipGetErrorString(e));
  void* base = nullptr;
  const int st = scratch_address(h->ctx->device, &base);
  if (st != AHM_OK) return st;
  return ahm::matvec_on_workspace(h, base, kScratchBytes, alpha, X, ldx, Y, ldy, p, vec,
                                  "ahm_hmat_matvec_noalloc");
}
