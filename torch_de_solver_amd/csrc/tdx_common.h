// Internal helpers shared by the tdx translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "tdx.h"

namespace tdx {

void set_error(const char *fmt, ...);

#define TDX_TRY_HIP(expr)                                                    \
  do {                                                                       \
    hipError_t _e = (expr);                                                  \
    if (_e != hipSuccess) {                                                  \
      ::tdx::set_error("%s failed: %s", #expr, hipGetErrorString(_e));       \
      return TDX_ELAUNCH;                                                    \
    }                                                                        \
  } while (0)

#define TDX_REQUIRE(cond, ...)                                               \
  do {                                                                       \
    if (!(cond)) {                                                           \
      ::tdx::set_error(__VA_ARGS__);                                         \
      return TDX_EINVAL;                                                     \
    }                                                                        \
  } while (0)

// Device copy of an immutable host table, owned by a plan.
struct DevTable {
  void *ptr = nullptr;
  size_t bytes = 0;
  int upload(const void *host, size_t n) {
    bytes = n;
    if (n == 0) return TDX_OK;
    if (hipMalloc(&ptr, n) != hipSuccess) {
      set_error("hipMalloc of %zu bytes for a plan table failed", n);
      return TDX_ENOMEM;
    }
    if (hipMemcpy(ptr, host, n, hipMemcpyHostToDevice) != hipSuccess) {
      set_error("hipMemcpy of a plan table failed");
      return TDX_ELAUNCH;
    }
    return TDX_OK;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
  }
};

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// ---- device helpers -------------------------------------------------------

typedef float f4_t __attribute__((ext_vector_type(4)));
typedef double d4_t __attribute__((ext_vector_type(4)));

template <typename T> struct V4;
template <> struct V4<float> { typedef f4_t type; };
template <> struct V4<double> { typedef d4_t type; };

__device__ __forceinline__ f4_t mma16(float a, float b, f4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ d4_t mma16(double a, double b, d4_t c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Row of the 16x16 MFMA accumulator held in register i of a lane in
// quarter q (= lane / 16); the column is lane % 16.  The f32 and f64
// 16x16x4 instructions place rows differently.
template <typename T> __device__ __forceinline__ int crow(int q, int i);
template <> __device__ __forceinline__ int crow<float>(int q, int i) { return 4 * q + i; }
template <> __device__ __forceinline__ int crow<double>(int q, int i) { return q + 4 * i; }
template <typename T> inline int crow_host(int q, int i) {
  return sizeof(T) == 4 ? 4 * q + i : q + 4 * i;
}

__device__ __forceinline__ float dtanh(float x) { return tanhf(x); }
__device__ __forceinline__ double dtanh(double x) { return tanh(x); }
__device__ __forceinline__ float dpow(float x, float p) { return powf(x, p); }
__device__ __forceinline__ double dpow(double x, double p) { return pow(x, p); }

// x ** p for integer p with the special cases torch uses for a scalar
// exponent (copy for 1, x*x for 2, x*x*x for 3, reciprocals for -1/-2).
template <typename T>
__device__ __forceinline__ T ipow(T x, int p) {
  switch (p) {
    case 0: return T(1);
    case 1: return x;
    case 2: return x * x;
    case 3: return (x * x) * x;
    case -1: return T(1) / x;
    case -2: return T(1) / (x * x);
    default: return dpow(x, T(p));
  }
}

}  // namespace tdx
