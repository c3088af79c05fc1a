// Device helpers shared by the HIP translation units and by the kernels the
// library compiles at run time (hipRTC): MFMA tiles, accumulator layout,
// float32 tanh, row sums, integer powers.  Self-contained (no includes), so
// that its text can be compiled by hipRTC as it is.
#ifndef TDX_DEV_H
#define TDX_DEV_H

namespace tdx {

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


__device__ __forceinline__ float dsin(float x) { return sinf(x); }
__device__ __forceinline__ double dsin(double x) { return sin(x); }
__device__ __forceinline__ float dcos(float x) { return cosf(x); }
__device__ __forceinline__ double dcos(double x) { return cos(x); }
__device__ __forceinline__ float dexp(float x) { return expf(x); }
__device__ __forceinline__ double dexp(double x) { return exp(x); }
__device__ __forceinline__ float dlog(float x) { return logf(x); }
__device__ __forceinline__ double dlog(double x) { return log(x); }
__device__ __forceinline__ float dsqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double dsqrt(double x) { return sqrt(x); }

// Table access by word index: records (i) and real values (r).
template <typename T>
struct PtrTab {
  const int32_t *I;
  const T *R;
  __device__ __forceinline__ int i(int k) const { return I[k]; }
  __device__ __forceinline__ T r(int k) const { return R[k]; }
};

// A callable power of a residual program (the reference applies it to the
// term's running product, tedeous/derivative.py:54-55, 128-129, 317-318):
// the chain of one-argument operations at word ref of g's records -- n,
// then n (op, arg) pairs, op TDX_AD_OP_*, arg a real's index or the integer
// power -- applied to x in order; v its value, d its derivative (dual
// numbers).
template <typename T, typename G>
__device__ __forceinline__ void chain_apply(const G &g, int ref, T x, T &v, T &d) {
  const int n = g.i(ref);
  v = x;
  d = T(1);
  for (int k = 0; k < n; ++k) {
    const int op = g.i(ref + 1 + 2 * k), arg = g.i(ref + 2 + 2 * k);
    switch (op) {
      case TDX_AD_OP_ADD: v = v + g.r(arg); break;
      case TDX_AD_OP_SUB: v = v - g.r(arg); break;
      case TDX_AD_OP_RSUB: v = g.r(arg) - v; d = -d; break;
      case TDX_AD_OP_MUL: v = g.r(arg) * v; d = g.r(arg) * d; break;
      case TDX_AD_OP_DIV: v = v / g.r(arg); d = d / g.r(arg); break;
      case TDX_AD_OP_RDIV: {
        const T nv = g.r(arg) / v;
        d = -(nv / v) * d;
        v = nv;
      } break;
      case TDX_AD_OP_POWI:
        d = (arg == 0 ? T(0) : T(arg) * ipow(v, arg - 1)) * d;
        v = ipow(v, arg);
        break;
      case TDX_AD_OP_POWR: {
        const T p = g.r(arg);
        d = (p * dpow(v, p - T(1))) * d;
        v = dpow(v, p);
      } break;
      case TDX_AD_OP_NEG: v = -v; d = -d; break;
      case TDX_AD_OP_SIN: d = dcos(v) * d; v = dsin(v); break;
      case TDX_AD_OP_COS: d = -dsin(v) * d; v = dcos(v); break;
      case TDX_AD_OP_EXP: v = dexp(v); d = v * d; break;
      case TDX_AD_OP_LOG: d = d / v; v = dlog(v); break;
      case TDX_AD_OP_TANH: v = dtanh(v); d = (T(1) - v * v) * d; break;
      case TDX_AD_OP_SQRT: v = dsqrt(v); d = d / (T(2) * v); break;
      default: break;
    }
  }
}
}  // namespace tdx

#endif  // TDX_DEV_H
