// Plan-specialized 'mat' kernel: source generation and runtime compilation.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace tdx {

struct JitFactor {
  int var;
  int nax;
  int ax[3];
  int pow_kind;     // 0 integer, 1 real
  int pow_int;
  double pow_real;
};

struct JitTerm {
  int coeff_kind;   // 0 scalar, 1 grid array in the data table
  double coeff;     // scalar value (working precision, exactly representable)
  int64_t data_off; // data-table offset of the array
  std::vector<JitFactor> f;
};

struct JitEq {
  double lam;       // lambda / N of the equation (working precision)
  std::vector<JitTerm> terms;
};

struct JitSpec {
  int n_eq;
  int dp;
  int div_mode;     // 0 reciprocal product, 1 IEEE division, 2 corrected reciprocal product
  int64_t N;
  float ab[8], af[8];
  float h2[2], ih2[2];     // 2h and RN(1/(2h)) per axis
  float tb2[2][8], tf2[2][8];  // transposed products ab/(2h), af/(2h) per axis
  int R0, R1;              // reach per axis (halo 2R per side)
  std::vector<JitEq> eqs;
};

// Arguments of the generated kernel (same layout in the generated source).
struct JitArgs {
  const float *u;
  float *grad;
  float *op;
  float *lp;
  const float *data;
  const float *dloss;
  int64_t N;
  int n0, n1;
  int r_lo, r_hi, c_lo, c_hi;  // output region (rows, columns)
  int hb;                      // rows per band
  int nstrips, nbands;
  int lp_base;                 // first loss-partial pair of this kernel
  int adj;
  // fused Adam step: updated grid, moments, and the step's scalars
  // {1-b1, 1-b2, b2, lr/(1-b1^t), sqrt(1-b2^t), eps}
  float *u_out;
  float *m;
  float *v;
  const float *adam;
};

struct JitKernel {
  hipModule_t module = nullptr;
  hipFunction_t fn_adj = nullptr;  // residual + loss + gradient
  hipFunction_t fn_fwd = nullptr;  // residual + loss
  hipFunction_t fn_step = nullptr; // residual + loss + Adam update into u_out, m, v
};

// Generates the source for `spec` (returns "" when the program is outside
// what the generator serves; reason in *why).
std::string mat_jit_source(const JitSpec &spec, std::string *why);

// Compiles `src` to a code object for `arch` (hipRTC; needs no device);
// returns 0, or a negative code with the compiler log in the error message.
int mat_jit_build(const std::string &src, std::vector<char> *code,
                  const std::string &arch = "gfx950");

// Compiles and loads `src` on the current device once per process (cached by
// device and source); returns 0 and fills *k, or a negative code with the
// message set.
int mat_jit_compile(const std::string &src, JitKernel *k);

}  // namespace tdx
