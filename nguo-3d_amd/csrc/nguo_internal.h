// Shared host-side plumbing of libnguo_hip: error reporting (thread-local last error, the
// reference's GpuError / Result<(), String> split, engine-gpu/src/lib.rs:58-66), HIP call
// checking and an owning device buffer.
#pragma once
#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "nguo_hip.h"

namespace nguo {

// Records `msg` as the calling thread's last error and returns `code`.
int fail(int code, const std::string &msg);
void clear_error();

}  // namespace nguo

#ifdef __HIPCC__
#include <hip/hip_runtime.h>

namespace nguo {

struct HipError {
  int code;
  std::string msg;
};

#define NGUO_HIP(call)                                                                      \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      throw ::nguo::HipError{NGUO_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)}; \
  } while (0)

#define NGUO_REQUIRE(cond, code, msg)                      \
  do {                                                     \
    if (!(cond)) throw ::nguo::HipError{(code), (msg)};    \
  } while (0)

// Runs `body`, converting thrown HipError into a status code + last error message.
template <class F> int guarded(F &&body) {
  try {
    clear_error();
    return body();
  } catch (const HipError &e) {
    return fail(e.code, e.msg);
  } catch (const std::bad_alloc &) {
    return fail(NGUO_E_NOMEM, "host allocation failed");
  }
}

// Owning device allocation of T[n] (hipMalloc / hipFree). Grows, never shrinks.
// Device allocations and frees made by DevBuf on the calling thread (the per-step
// device_alloc / device_free metrics).
inline thread_local uint64_t g_dev_allocs = 0, g_dev_frees = 0;

template <class T> struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) {
      (void)hipFree(p);
      ++g_dev_frees;
    }
    p = nullptr;
    n = 0;
  }
  // Grows to at least `count` elements (contents are not kept). A buffer that grows again
  // takes 1.5x headroom, so sizes that creep up step by step (contact lists, sort keys)
  // stop reallocating after a few steps.
  void reserve(size_t count) {
    if (count <= n) return;
    if (p) count = std::max(count, n + n / 2);
    release();
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc((void **)&p, bytes);
    ++g_dev_allocs;
    if (e != hipSuccess) {
      p = nullptr;
      throw HipError{NGUO_E_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e)};
    }
    n = count;
  }
  void upload(const T *src, size_t count, hipStream_t s) {
    reserve(count);
    if (count) NGUO_HIP(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T> &v, hipStream_t s) { upload(v.data(), v.size(), s); }
  void download(T *dst, size_t count, hipStream_t s) const {
    if (count) NGUO_HIP(hipMemcpyAsync(dst, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
  }
  void zero(size_t count, hipStream_t s) {
    reserve(count);
    if (count) NGUO_HIP(hipMemsetAsync(p, 0, count * sizeof(T), s));
  }
};

inline unsigned div_up(size_t a, unsigned b) { return (unsigned)((a + b - 1) / b); }

// Looks up this device's copy of the allocation-free sort's scratch array (once per device,
// with that device current; nguo_ctx_create calls it).
void noalloc_scratch_init(int device);

}  // namespace nguo

// Device context: one HIP device and the (default-priority) stream all work is queued on.
struct nguo_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipDeviceProp_t prop{};
};
#endif
