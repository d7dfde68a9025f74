// ahm_core.hip -- error state, contexts and device buffers of libahmed_amd.
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "ahm_internal.h"

namespace ahm {

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

namespace {
constexpr size_t kCacheMin = (size_t)1 << 30;     // blocks at least this large are cached
// per device: three quarters of its memory, so that the large buffers of one
// assembly (slab scratch / leaf buffer, arena, result parts) all stay cached
// for the next one
size_t cache_max(int dev) {
  static size_t cap[64] = {};
  if (dev < 0 || dev >= 64) return (size_t)128 << 30;
  if (!cap[dev]) {
    size_t fr = 0, tot = 0;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(dev);
    cap[dev] = hipMemGetInfo(&fr, &tot) == hipSuccess && tot ? tot / 4 * 3 : (size_t)128 << 30;
    (void)hipSetDevice(cur);
  }
  return cap[dev];
}
struct CachedBlock {
  void* p;
  size_t cap;
  int dev;
};
std::mutex g_cache_mu;
std::vector<CachedBlock> g_cache;

size_t cache_bytes_locked(int dev) {
  size_t s = 0;
  for (const CachedBlock& c : g_cache)
    if (c.dev == dev) s += c.cap;
  return s;
}
}  // namespace

size_t cache_bytes(int device) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  return cache_bytes_locked(device);
}

void cache_flush(int device) {
  std::lock_guard<std::mutex> lk(g_cache_mu);
  int cur = 0;
  (void)hipGetDevice(&cur);
  for (size_t i = 0; i < g_cache.size();) {
    if (g_cache[i].dev == device) {
      (void)hipSetDevice(device);
      (void)hipFree(g_cache[i].p);
      g_cache.erase(g_cache.begin() + (ptrdiff_t)i);
    } else {
      ++i;
    }
  }
  (void)hipSetDevice(cur);
}

int DeviceBuf::alloc(size_t b) {
  release();
  if (b == 0) return AHM_OK;
  (void)hipGetDevice(&dev);
  if (b >= kCacheMin) {
    // best fit among the cached blocks of at most twice the size: a small
    // request must not take a block that a later large one would reuse
    std::lock_guard<std::mutex> lk(g_cache_mu);
    size_t best = g_cache.size();
    for (size_t i = 0; i < g_cache.size(); ++i) {
      const CachedBlock& c = g_cache[i];
      if (c.dev == dev && c.cap >= b && c.cap / 2 <= b && (best == g_cache.size() || c.cap < g_cache[best].cap))
        best = i;
    }
    if (best < g_cache.size()) {
      p = g_cache[best].p;
      cap = g_cache[best].cap;
      bytes = b;
      g_cache.erase(g_cache.begin() + (ptrdiff_t)best);
      return AHM_OK;
    }
  }
  const bool timing = b >= kCacheMin && std::getenv("AHM_TIMING") != nullptr;
  const auto t0 = std::chrono::steady_clock::now();
  hipError_t e = hipMalloc(&p, b);
  if (timing)
    std::fprintf(stderr, "[ahm timing] hipMalloc %.2f GB: %.1f ms (cache holds %.2f GB)\n", b / 1e9,
                 std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(),
                 cache_bytes(dev) / 1e9);
  if (e != hipSuccess && cache_bytes(dev) > 0) {
    (void)hipGetLastError();
    cache_flush(dev);
    e = hipMalloc(&p, b);
  }
  if (e != hipSuccess) {
    p = nullptr;
    return fail(AHM_ERR_NOMEM, "hipMalloc(%zu bytes): %s", b, hipGetErrorString(e));
  }
  bytes = cap = b;
  return AHM_OK;
}

void DeviceBuf::release() {
  if (p) {
    if (cap >= kCacheMin) {
      // the same wait as hipFree's: work still using the block finishes first
      int cur = 0;
      (void)hipGetDevice(&cur);
      (void)hipSetDevice(dev);
      (void)hipDeviceSynchronize();
      std::lock_guard<std::mutex> lk(g_cache_mu);
      g_cache.push_back(CachedBlock{p, cap, dev});
      // over the limit: free the oldest blocks of the device
      const size_t cmax = cache_max(dev);
      while (cache_bytes_locked(dev) > cmax) {
        for (size_t i = 0; i < g_cache.size(); ++i)
          if (g_cache[i].dev == dev) {
            (void)hipFree(g_cache[i].p);
            g_cache.erase(g_cache.begin() + (ptrdiff_t)i);
            break;
          }
      }
      (void)hipSetDevice(cur);
    } else {
      (void)hipFree(p);
    }
  }
  p = nullptr;
  bytes = cap = 0;
}

static double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

PhaseTimer::PhaseTimer(const char* w) : on(std::getenv("AHM_TIMING") != nullptr), what(w), t_last(now_ms()) {}

void PhaseTimer::mark(const char* phase) {
  if (!on) return;
  const double t = now_ms();
  std::fprintf(stderr, "[ahm timing] %s: %s %.1f ms\n", what, phase, t - t_last);
  t_last = t;
}

size_t dtype_size(ahm_dtype t) {
  switch (t) {
    case AHM_F64: return 8;
    case AHM_F32: return 4;
    case AHM_BF16: return 2;
  }
  return 0;
}

}  // namespace ahm

extern "C" {

const char* ahm_last_error(void) { return ahm::g_last_error.c_str(); }

int ahm_ctx_create(int device, ahm_ctx* out) {
  if (!out) return ahm::fail(AHM_ERR_INVALID, "ahm_ctx_create: out is NULL");
  int n = 0;
  AHM_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n)
    return ahm::fail(AHM_ERR_INVALID, "ahm_ctx_create: device %d of %d", device, n);
  AHM_HIP(hipSetDevice(device));
  ahm_ctx c = new ahm_ctx_s;
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return ahm::fail(AHM_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
  }
  c->stream = c->own;
  *out = c;
  return AHM_OK;
}

int ahm_ctx_destroy(ahm_ctx c) {
  if (!c) return AHM_OK;
  (void)hipSetDevice(c->device);
  if (c->own) {
    (void)hipStreamSynchronize(c->own);
    (void)hipStreamDestroy(c->own);
  }
  ahm::cache_flush(c->device);
  delete c;
  return AHM_OK;
}

int ahm_ctx_release_cached(ahm_ctx c) {
  if (!c) return ahm::fail(AHM_ERR_INVALID, "ahm_ctx_release_cached: NULL context");
  ahm::cache_flush(c->device);
  return AHM_OK;
}

int ahm_ctx_cached_bytes(ahm_ctx c, uint64_t* bytes) {
  if (!c || !bytes) return ahm::fail(AHM_ERR_INVALID, "ahm_ctx_cached_bytes: NULL argument");
  *bytes = ahm::cache_bytes(c->device);
  return AHM_OK;
}

int ahm_ctx_set_stream(ahm_ctx c, void* s) {
  if (!c) return ahm::fail(AHM_ERR_INVALID, "ahm_ctx_set_stream: NULL context");
  c->stream = (hipStream_t)s;  // NULL: HIP's default (null) stream
  return AHM_OK;
}

int ahm_ctx_sync(ahm_ctx c) {
  if (!c) return ahm::fail(AHM_ERR_INVALID, "ahm_ctx_sync: NULL context");
  AHM_HIP(hipSetDevice(c->device));
  AHM_HIP(hipStreamSynchronize(c->stream));
  return AHM_OK;
}

}  // extern "C"
