// Internal helpers shared by the libdust_gpu translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/dust_gpu.h"

namespace dust {

void set_error(const char *fmt, ...);

// Evaluate a HIP call; on failure record the error text and return DUST_E_HIP
// from the enclosing function.
#define DUST_HIP(call)                                                            \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) {                                                       \
      ::dust::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,                 \
                        hipGetErrorString(e_));                                   \
      return DUST_E_HIP;                                                          \
    }                                                                             \
  } while (0)

#define DUST_CHECK(cond, code, ...)                                               \
  do {                                                                            \
    if (!(cond)) {                                                                \
      ::dust::set_error(__VA_ARGS__);                                             \
      return (code);                                                              \
    }                                                                             \
  } while (0)

// Propagate a non-zero status from a helper.
#define DUST_TRY(expr)                                                            \
  do {                                                                            \
    int s_ = (expr);                                                              \
    if (s_ != 0) return s_;                                                       \
  } while (0)

// Release of the library-global workspaces of each translation unit
// (buffers, page-locked records, streams, events), for dust_shutdown.  The
// caller has synchronised the device.
void aic_release();
void asm_release();
void elem_release();
void fmm_release();
void lu_release();
void mixed_release();
int mixed_reserved_cus();
int mixed_set_reserved_cus(int v);
void pres_release();
void wake_release();

constexpr double kPi = 3.1415926535897932384626433832795028841971693993751;

inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

// Device buffer that grows on demand (no shrinking).
template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  int reserve(size_t count) {
    if (count <= n) return 0;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    size_t bytes = (count ? count : 1) * sizeof(T);
    if (hipMalloc(&p, bytes) != hipSuccess) {
      set_error("hipMalloc of %zu bytes failed", bytes);
      return DUST_E_NOMEM;
    }
    n = count;
    return 0;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

// Device allocation owned by one scope (the staging buffers of the
// host-pointer entry points): freed on every return path.
template <typename T>
struct TmpDev {
  T *p = nullptr;
  TmpDev() = default;
  TmpDev(const TmpDev &) = delete;
  TmpDev &operator=(const TmpDev &) = delete;
  ~TmpDev() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t count) {
    size_t bytes = (count ? count : 1) * sizeof(T);
    if (hipMalloc(&p, bytes) != hipSuccess) {
      p = nullptr;
      set_error("hipMalloc of %zu bytes failed", bytes);
      return DUST_E_NOMEM;
    }
    return 0;
  }
};

// Page-locked host buffer that grows on demand, the source or target of
// asynchronous copies.  A copy between the device and pageable memory is
// staged by the runtime and returns only once the stream has reached it,
// which would tie the host to the GPU at every read-back.  Zero-filled.
template <typename T>
struct HostBuf {
  T *p = nullptr;
  size_t n = 0;
  int reserve(size_t count) {
    if (count <= n) return 0;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
    size_t bytes = (count ? count : 1) * sizeof(T);
    void *q = nullptr;
    if (hipHostMalloc(&q, bytes, hipHostMallocDefault) != hipSuccess) {
      set_error("hipHostMalloc of %zu bytes failed", bytes);
      return DUST_E_NOMEM;
    }
    std::memset(q, 0, bytes);
    p = (T *)q;
    n = count;
    return 0;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
  }
};

}  // namespace dust
