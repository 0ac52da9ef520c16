// Host-side utilities shared by the synthesis objects and the serving pool: error reporting,
// device buffers and small integer helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/stormm_amd.h"

namespace stormm_amd {

// Records `msg` as the calling thread's last error (stormm_last_error) and returns `code`.
int fail(int code, const std::string &msg);

#define HIP_TRY(expr)                                                                      \
  do {                                                                                     \
    hipError_t _e = (expr);                                                                \
    if (_e != hipSuccess) {                                                                \
      return ::stormm_amd::fail(STORMM_ERR_HIP, std::string(#expr) + ": " +              \
                                                    hipGetErrorString(_e));                \
    }                                                                                      \
  } while (0)

// One device allocation, freed with its owner.
struct DevMem {
  void *ptr = nullptr;
  size_t bytes = 0;
  DevMem() = default;
  DevMem(const DevMem &) = delete;
  DevMem &operator=(const DevMem &) = delete;
  ~DevMem() {
    if (ptr != nullptr) {
      (void)hipFree(ptr);
    }
  }
  hipError_t alloc(size_t n) {
    if (ptr != nullptr) {
      (void)hipFree(ptr);
      ptr = nullptr;
    }
    bytes = n;
    if (n == 0) {
      return hipSuccess;
    }
    return hipMalloc(&ptr, n);
  }
  template <typename T> hipError_t upload(const std::vector<T> &v) {
    hipError_t e = alloc(v.size() * sizeof(T));
    if (e != hipSuccess || v.empty()) {
      return e;
    }
    return hipMemcpy(ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
  }
  template <typename T> T *as() const { return static_cast<T *>(ptr); }
};

inline int round_up(int v, int m) { return ((v + m - 1) / m) * m; }

// Compute units of the current device (256 when no device is visible, for host-only planning).
inline int cu_count_or_default() {
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
      prop.multiProcessorCount > 0) {
    return prop.multiProcessorCount;
  }
  return 256;
}

}  // namespace stormm_amd
