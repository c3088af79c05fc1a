// Internal helpers shared by the tdx translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <mutex>
#include <set>
#include <string>
#include <utility>
#include <vector>

#include "tdx.h"

namespace tdx {

void set_error(const char *fmt, ...);
void clear_error();

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

// Raises the dynamic LDS limit of each given kernel to 160 KB, once per
// (kernel, device) and process; safe to call from several threads.
inline void raise_lds_limit(std::initializer_list<const void *> kernels) {
  static std::mutex mu;
  static std::set<std::pair<const void *, int>> done;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lock(mu);
  for (const void *k : kernels)
    if (done.insert({k, dev}).second)
      (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
}

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

// tanh in float32 without branches: for |x| < 0.625 an odd polynomial
// x + x^3 q(x^2) (q fitted by weighted least squares to (tanh x - x)/x^3,
// max error 0.8 ulp on that interval in float32 Horner form); otherwise
// 1 - 2/(1 + 2^(2|x| log2 e)) with the sign restored (about 2 ulp; exp2
// overflows to inf for large |x| and gives exactly 1).
__device__ __forceinline__ float dtanh(float x) {
  const float ax = fabsf(x);
  const float x2 = x * x;
  float p = 0.002142803743481636f;
  p = fmaf(p, x2, -0.008177093230187893f);
  p = fmaf(p, x2, 0.021700501441955566f);
  p = fmaf(p, x2, -0.05394670367240906f);
  p = fmaf(p, x2, 0.1333320587873459f);
  p = fmaf(p, x2, -0.3333333134651184f);
  const float small = fmaf(x * x2, p, x);
  const float e = __builtin_amdgcn_exp2f(ax * 2.8853900817779268f);
  const float r = __builtin_amdgcn_rcpf(1.0f + e);
  const float large = copysignf(fmaf(-2.0f, r, 1.0f), x);
  return ax < 0.625f ? small : large;
}
__device__ __forceinline__ double dtanh(double x) { return tanh(x); }

// Sum over the 16 lanes of a row of the wavefront (lanes 16k .. 16k+15);
// every lane of the row receives the sum.  Float: DPP lane swizzles
// (quad xor 1, quad xor 2, half-row mirror, row mirror) fused into adds.
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
__device__ __forceinline__ float rowsum16(float v) {
  v += dpp_f<0xB1>(v);
  v += dpp_f<0x4E>(v);
  v += dpp_f<0x141>(v);
  v += dpp_f<0x140>(v);
  return v;
}
__device__ __forceinline__ double rowsum16(double v) {
  v += __shfl_xor(v, 1);
  v += __shfl_xor(v, 2);
  v += __shfl_xor(v, 4);
  v += __shfl_xor(v, 8);
  return v;
}
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

struct AdamHyper {
  double lr, b1, b2, eps;
};

// torch.optim.Adam (single-tensor form): m = b1 m + (1-b1) g;
// v = b2 v + (1-b2) g^2; p -= lr/(1-b1^t) * m / (sqrt(v)/sqrt(1-b2^t) + eps)
// The per-step scalars (bias corrections in double, as torch computes them
// in Python floats) depend only on the step t.
template <typename T>
struct AdamScalars {
  T w1, w2, b2, step_size, bc2s, eps;
};
template <typename T>
__device__ __forceinline__ AdamScalars<T> adam_scalars(int64_t t, const AdamHyper &h) {
  AdamScalars<T> s;
  s.b2 = T(h.b2);
  s.w1 = T(1.0 - h.b1);
  s.w2 = T(1.0 - h.b2);
  const double bc1 = 1.0 - pow(h.b1, (double)t);
  const double bc2 = 1.0 - pow(h.b2, (double)t);
  s.step_size = T(h.lr / bc1);
  s.bc2s = T(sqrt(bc2));
  s.eps = T(h.eps);
  return s;
}
template <typename T>
__device__ __forceinline__ void adam_apply(T &p, T g, T &m, T &v, const AdamScalars<T> &s) {
  m = m + s.w1 * (g - m);
  v = v * s.b2 + s.w2 * g * g;
  const T denom = sqrt(v) / s.bc2s + s.eps;
  p = p - s.step_size * (m / denom);
}

}  // namespace tdx
