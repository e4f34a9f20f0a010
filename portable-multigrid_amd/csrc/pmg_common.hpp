// pmg_common.hpp -- shared host/device helpers of the MI355X multigrid library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "pmg.h"

namespace pmg {

constexpr uint32_t kInvalid = 0xFFFFFFFFu;

// Exception type carried up to the C-ABI, where it becomes a status code.
struct Error : std::runtime_error {
  int status;
  Error(int s, const std::string &m) : std::runtime_error(m), status(s) {}
};

void set_last_error(const std::string &msg);

#define PMG_HIP_CHECK(expr)                                                              \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      throw ::pmg::Error(PMG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define PMG_REQUIRE(cond, status, msg)                  \
  do {                                                  \
    if (!(cond)) throw ::pmg::Error((status), (msg));   \
  } while (0)

// Owning device buffer (hipMalloc'ed at setup time, never inside apply calls).
template <typename T>
struct DeviceArray {
  T *ptr = nullptr;
  size_t n = 0;
  DeviceArray() = default;
  explicit DeviceArray(size_t count) { resize(count); }
  DeviceArray(const DeviceArray &) = delete;
  DeviceArray &operator=(const DeviceArray &) = delete;
  DeviceArray(DeviceArray &&o) noexcept : ptr(o.ptr), n(o.n) { o.ptr = nullptr; o.n = 0; }
  DeviceArray &operator=(DeviceArray &&o) noexcept {
    if (this != &o) { release(); ptr = o.ptr; n = o.n; o.ptr = nullptr; o.n = 0; }
    return *this;
  }
  ~DeviceArray() { release(); }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    n = 0;
  }
  void resize(size_t count) {
    release();
    if (count == 0) return;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&ptr), count * sizeof(T));
    if (e != hipSuccess)
      throw Error(PMG_ERR_ALLOC, "hipMalloc of " + std::to_string(count * sizeof(T)) +
                                     " bytes failed: " + hipGetErrorString(e));
    n = count;
  }
  void upload(const T *host, size_t count) {
    resize(count);
    if (count) PMG_HIP_CHECK(hipMemcpy(ptr, host, count * sizeof(T), hipMemcpyHostToDevice));
  }
  void upload(const std::vector<T> &v) { upload(v.data(), v.size()); }
  std::vector<T> download() const {
    std::vector<T> v(n);
    if (n) PMG_HIP_CHECK(hipMemcpy(v.data(), ptr, n * sizeof(T), hipMemcpyDeviceToHost));
    return v;
  }
  size_t bytes() const { return n * sizeof(T); }
};

// Pinned, host-coherent doubles: kernels store into them directly and the
// host reads them after synchronising, without a device-to-host copy.
struct PinnedDoubles {
  double *ptr = nullptr;
  PinnedDoubles() = default;
  PinnedDoubles(const PinnedDoubles &) = delete;
  PinnedDoubles &operator=(const PinnedDoubles &) = delete;
  ~PinnedDoubles() {
    if (ptr) (void)hipHostFree(ptr);
  }
  void allocate(size_t count) {
    if (ptr) return;
    hipError_t e = hipHostMalloc(reinterpret_cast<void **>(&ptr), count * sizeof(double),
                                 hipHostMallocCoherent);
    if (e != hipSuccess) {
      ptr = nullptr;
      throw Error(PMG_ERR_ALLOC, std::string("hipHostMalloc failed: ") + hipGetErrorString(e));
    }
    for (size_t i = 0; i < count; ++i) ptr[i] = 0.0;
  }
};

inline size_t element_size(int precision) { return precision == PMG_FP32 ? 4 : 8; }

inline void check_precision(int precision) {
  PMG_REQUIRE(precision == PMG_FP64 || precision == PMG_FP32, PMG_ERR_INVALID_ARG,
              "precision must be PMG_FP64 or PMG_FP32");
}

// Number of XCDs of an MI355X; workgroups are dispatched round-robin over them.
constexpr int kNumXcd = 8;

// Remap a hardware block index so that each XCD (and its private L2) works on a
// contiguous range of logical blocks.  The grid must be padded to a multiple of
// kNumXcd; logical indices >= n_logical are idle.
__device__ inline unsigned xcd_logical_block(unsigned bid, unsigned n_padded) {
  const unsigned chunk = n_padded / kNumXcd;
  return (bid % kNumXcd) * chunk + bid / kNumXcd;
}

inline unsigned pad_to_xcd(unsigned n) { return (n + kNumXcd - 1) / kNumXcd * kNumXcd; }

// Hardware floating-point atomics (global_atomic_add_f64 / _f32), never CAS loops.
__device__ inline void atomic_add(double *a, double v) { unsafeAtomicAdd(a, v); }
__device__ inline void atomic_add(float *a, float v) { unsafeAtomicAdd(a, v); }

}  // namespace pmg
