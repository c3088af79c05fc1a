// Plan-specialized residual programs of the fused 'NN' training step.
//
// The library's fused kernel (k_fused) interprets each group's residual
// program: in every 16-entry round each wavefront walks the program's
// records, a chain of dependent table reads per stencil point, factor and
// term.  For large plans the library instead generates, at plan creation,
// one straight-line function per distinct program structure -- slots,
// stencil points, powers and coefficient kinds are constants of the code;
// real values and data offsets are read from the group's own tables at fixed
// positions, loads the compiler issues together -- and compiles the fused
// kernel body (tdx_nn_dev.h, the same text the library is built from) with
// them by hipRTC.  The generated functions perform entry_program's
// operations in the same order, so the results are the same bits as the
// interpreted kernel's.

#include "tdx_nn_jit.h"

#include <hip/hiprtc.h>

#include <map>
#include <mutex>
#include <sstream>
#include <utility>

#include "tdx_common.h"
#include "tdx_nn_dev.h"
// kTdxDevSrc, kTdxNnDevSrc: the texts of tdx_dev.h and tdx_nn_dev.h,
// written by the build (torch_de_solver_amd/build.py)
#include "tdx_embed.inc"

namespace tdx {

std::vector<int32_t> nn_prog_key(const FusedProg &p) {
  std::vector<int32_t> k = p.ints;
  for (int q = 0; q < p.n_eq; ++q) {
    const int E = q * TDX_EQ_INTS;
    k[E + 2] = k[E + 2] >= 0 ? -2 : -1;  // a target array or none
    k[E + 3] = 0;                         // residual output slot (not read)
    const int nT = k[E], ti = k[E + 1];
    for (int t = 0; t < nT; ++t) {
      const int Tm = ti + t * TDX_TERM_INTS;
      if (k[Tm + 2] == 1) k[Tm + 3] = 0;  // coefficient array offset
    }
  }
  return k;
}

namespace {

std::string num(int v) { return std::to_string(v); }

// gs = sum over the factor's stencil points of u[slot * n_out + var] * w,
// in point order
void emit_stencil_sum(std::ostringstream &o, const std::vector<int32_t> &k, int F,
                      const std::string &dst, int n_out) {
  const int ns = k[F + 3], si = k[F + 4], var = k[F];
  o << dst << " = 0.0f;\n";
  for (int j = 0; j < ns; ++j) {
    const int S = si + j * TDX_STENCIL_INTS;
    o << dst << " = " << dst << " + u[" << k[S] * n_out + var << "] * R[" << k[S + 1]
      << "];\n";
  }
}

std::string pow_expr(const std::vector<int32_t> &k, int F, const std::string &x) {
  if (k[F + 1] == 0) return "ipow(" + x + ", " + num(k[F + 2]) + ")";
  return "dpow(" + x + ", R[" + num(k[F + 5]) + "])";
}

// d(x**p)/dx as one value (parenthesized: the caller multiplies by it)
std::string dpow_expr(const std::vector<int32_t> &k, int F, const std::string &x) {
  if (k[F + 1] == 0) {
    const int p = k[F + 2];
    if (p == 0) return "0.0f";
    return "((float)(" + num(p) + ") * ipow(" + x + ", " + num(p - 1) + "))";
  }
  const std::string pr = "R[" + num(k[F + 5]) + "]";
  return "(" + pr + " * dpow(" + x + ", " + pr + " - 1.0f))";
}

std::string coeff_expr(const std::vector<int32_t> &k, int Tm) {
  if (k[Tm + 2] == 0) return "R[" + num(k[Tm + 3]) + "]";
  return "data[(int64_t)I[" + num(Tm + 3) + "] + e]";
}

// entry_program of one structure as straight-line code (network values u,
// seed accumulators sdv, both indexed slot * n_out + var)
void emit_program(std::ostringstream &o, const std::vector<int32_t> &k, int n_eq, int n_out) {
  for (int q = 0; q < n_eq; ++q) {
    const int E = q * TDX_EQ_INTS;
    const int nT = k[E], ti = k[E + 1];
    o << "{\nfloat r = 0.0f;\n";
    for (int t = 0; t < nT; ++t) {
      const int Tm = ti + t * TDX_TERM_INTS;
      const int nf = k[Tm], fi = k[Tm + 1];
      o << "{\nfloat der = 1.0f;\n";
      for (int f = 0; f < nf; ++f) {
        const int F = fi + f * TDX_FACTOR_INTS;
        o << "{\nfloat gs;\n";
        emit_stencil_sum(o, k, F, "gs", n_out);
        o << "der = der * " << pow_expr(k, F, "gs") << ";\n}\n";
      }
      o << "const float tv = " << coeff_expr(k, Tm) << " * der;\n"
        << (t == 0 ? "r = tv;\n" : "r = r + tv;\n") << "}\n";
    }
    o << "const float target = "
      << (k[E + 2] == -2 ? "data[(int64_t)I[" + num(E + 2) + "] + e]" : std::string("0.0f"))
      << ";\n"
         "const float diff = r - target;\n"
         "const float sq = diff * diff;\n"
         "const float lam_d = R["
      << k[E + 4] << "];\nls += lam_d * sq;\nlns += R[" << k[E + 5] << "] * sq;\n"
      << "const float dr = 2.0f * R[" << k[E + 6] << "] * diff;\n";
    for (int t = 0; t < nT; ++t) {
      const int Tm = ti + t * TDX_TERM_INTS;
      const int nf = k[Tm], fi = k[Tm + 1];
      o << "{\nfloat gsv[" << MAXF << "], pw[" << MAXF << "], pre[" << MAXF + 1 << "], suf["
        << MAXF + 1 << "];\n";
      for (int f = 0; f < MAXF; ++f) {
        if (f < nf) {
          const int F = fi + f * TDX_FACTOR_INTS;
          const std::string g = "gsv[" + num(f) + "]";
          emit_stencil_sum(o, k, F, g, n_out);
          o << "pw[" << f << "] = " << pow_expr(k, F, g) << ";\n";
        } else {
          o << "gsv[" << f << "] = 0.0f;\npw[" << f << "] = 1.0f;\n";
        }
      }
      o << "pre[0] = 1.0f;\n";
      for (int f = 0; f < MAXF; ++f)
        o << "pre[" << f + 1 << "] = pre[" << f << "] * pw[" << f << "];\n";
      o << "suf[" << MAXF << "] = 1.0f;\n";
      for (int f = MAXF - 1; f >= 0; --f)
        o << "suf[" << f << "] = suf[" << f + 1 << "] * pw[" << f << "];\n";
      o << "const float cdr = " << coeff_expr(k, Tm) << " * dr;\n";
      for (int f = 0; f < nf; ++f) {
        const int F = fi + f * TDX_FACTOR_INTS;
        o << "{\nconst float gf = cdr * " << dpow_expr(k, F, "gsv[" + num(f) + "]") << " * pre["
          << f << "] * suf[" << f + 1 << "];\n";
        const int ns = k[F + 3], si = k[F + 4], var = k[F];
        for (int j = 0; j < ns; ++j) {
          const int S = si + j * TDX_STENCIL_INTS;
          o << "sdv[" << k[S] * n_out + var << "] += gf * R[" << k[S + 1] << "];\n";
        }
        o << "}\n";
      }
      o << "}\n";
    }
    o << "}\n";
  }
}

}  // namespace

namespace {

// Shared prelude: integer types, the table record sizes, the device headers.
void emit_prelude(std::ostringstream &o) {
  o << "typedef __hip_internal::int32_t int32_t;\n"
       "typedef __hip_internal::int64_t int64_t;\n"
    << "#define TDX_GROUP_INTS " << TDX_GROUP_INTS << "\n"
    << "#define TDX_EQ_INTS " << TDX_EQ_INTS << "\n"
    << "#define TDX_TERM_INTS " << TDX_TERM_INTS << "\n"
    << "#define TDX_FACTOR_INTS " << TDX_FACTOR_INTS << "\n"
    << "#define TDX_STENCIL_INTS " << TDX_STENCIL_INTS << "\n";
  const std::pair<const char *, int> ops[] = {
      {"ADD", TDX_AD_OP_ADD},   {"SUB", TDX_AD_OP_SUB},   {"RSUB", TDX_AD_OP_RSUB},
      {"MUL", TDX_AD_OP_MUL},   {"DIV", TDX_AD_OP_DIV},   {"RDIV", TDX_AD_OP_RDIV},
      {"POWI", TDX_AD_OP_POWI}, {"POWR", TDX_AD_OP_POWR}, {"NEG", TDX_AD_OP_NEG},
      {"SIN", TDX_AD_OP_SIN},   {"COS", TDX_AD_OP_COS},   {"EXP", TDX_AD_OP_EXP},
      {"LOG", TDX_AD_OP_LOG},   {"TANH", TDX_AD_OP_TANH}, {"SQRT", TDX_AD_OP_SQRT}};
  for (const auto &op : ops) o << "#define TDX_AD_OP_" << op.first << " " << op.second << "\n";
  o << kTdxDevSrc << "\n"
    << kTdxNnDevSrc << "\n";
}

}  // namespace

std::string nn_fused_jit_source(const std::vector<std::vector<int32_t>> &keys,
                                const std::vector<int> &n_eq, int NB, int NH) {
  std::ostringstream o;
  emit_prelude(o);
  o << "namespace tdx {\n"
       "// generated residual programs, one per structure (kind)\n"
       "struct GenResid {\n"
       "  const int32_t *I;\n"
       "  const float *R;\n"
       "  const float *data;\n"
       "  int kind;\n"
       "  __device__ GenResid(const NNArgs<float> &a, int g, int, float *, int) {\n"
       "    const int4 pi = a.prog_idx[g];\n"
       "    I = a.prog_ints + pi.x;\n"
       "    R = a.prog_reals + pi.z;\n"
       "    data = a.data;\n"
       "    kind = a.prog_kind[g];\n"
       "  }\n"
       "  __device__ __forceinline__ void run(int e, bool valid, int row, const float *u,\n"
       "                                      float (&sdv)[FUSE_MAX_SLOTS], float &ls,\n"
       "                                      float &lns) {\n"
       "#pragma unroll\n"
       "    for (int s = 0; s < FUSE_MAX_SLOTS; ++s) sdv[s] = 0.0f;\n"
       "    if (!valid) return;\n"
       "    switch (kind) {\n";
  for (size_t i = 0; i < keys.size(); ++i) {
    o << "    case " << i << ": {\n";
    emit_program(o, keys[i], n_eq[i], 1);
    o << "    } break;\n";
  }
  o << "    default: break;\n"
       "    }\n"
       "  }\n"
       "};\n"
       "}  // namespace tdx\n"
       "extern \"C\" __global__ __launch_bounds__(64 * tdx::FUSE_WAVES)\n"
       "__attribute__((amdgpu_waves_per_eu(2))) void tdx_fused_jit(tdx::NNArgs<float> a) {\n"
       "  extern __shared__ __align__(16) unsigned char smem_raw[];\n"
       "  tdx::fused_body<"
    << NB << ", " << NH
    << ", tdx::GenResid>(a, reinterpret_cast<float *>(smem_raw));\n"
       "}\n";
  return o.str();
}

std::string nn_block_jit_source(const std::vector<std::vector<int32_t>> &keys,
                                const std::vector<int> &n_eq, int NB, int NH, int n_out) {
  std::ostringstream o;
  emit_prelude(o);
  o << "namespace tdx {\n"
       "// generated residual programs, one per structure (kind); seeds added\n"
       "// into the entry's LDS row sdv (zeroed by the caller)\n"
       "struct GenBlockResid {\n"
       "  const int32_t *I;\n"
       "  const float *R;\n"
       "  const float *data;\n"
       "  int kind;\n"
       "  __device__ GenBlockResid(const NNArgs<float> &a, int g, int32_t *, float *) {\n"
       "    data = a.data;\n"
       "    if (g < 0) return;\n"
       "    const int4 pi = a.prog_idx[g];\n"
       "    I = a.prog_ints + pi.x;\n"
       "    R = a.prog_reals + pi.z;\n"
       "    kind = a.prog_kind[g];\n"
       "  }\n"
       "  __device__ __forceinline__ void run(int, int e, int, const float *u, float *sdv,\n"
       "                                      float &ls, float &lns) {\n"
       "    switch (kind) {\n";
  for (size_t i = 0; i < keys.size(); ++i) {
    o << "    case " << i << ": {\n";
    emit_program(o, keys[i], n_eq[i], n_out);
    o << "    } break;\n";
  }
  o << "    default: break;\n"
       "    }\n"
       "  }\n"
       "};\n"
       "}  // namespace tdx\n"
       "extern \"C\" __global__ __launch_bounds__(256) void tdx_block_jit(tdx::NNArgs<float> a) {\n"
       "  extern __shared__ __align__(16) unsigned char smem_raw[];\n"
       "  tdx::block_body<float, "
    << NB << ", " << NH
    << ", true, tdx::GenBlockResid>(a, reinterpret_cast<float *>(smem_raw));\n"
       "}\n";
  return o.str();
}

int nn_fused_jit_build(const std::string &src, std::vector<char> *code, const std::string &arch) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "tdx_nn_fused_jit.hip", 0, nullptr, nullptr) !=
      HIPRTC_SUCCESS) {
    set_error("hiprtcCreateProgram failed");
    return TDX_ELAUNCH;
  }
  const std::string arch_opt = "--offload-arch=" + arch;
  const char *opts[] = {arch_opt.c_str(), "-O3",     "-std=c++17",      "-ffp-contract=off",
                        "-Xclang",        "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult rc = hiprtcCompileProgram(prog, 8, opts);
  if (rc != HIPRTC_SUCCESS) {
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string log(ls, '\0');
    if (ls) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    set_error("run-time compiled kernel failed to compile: %.400s", log.c_str());
    return TDX_ELAUNCH;
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  code->assign(cs, 0);
  hiprtcGetCode(prog, code->data());
  hiprtcDestroyProgram(&prog);
  return TDX_OK;
}

int nn_fused_jit_load(const std::string &src, hipFunction_t *fn, const char *name) {
  static std::mutex mu;
  static std::map<std::pair<int, std::string>, hipFunction_t> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    set_error("hipGetDevice failed");
    return TDX_ELAUNCH;
  }
  std::lock_guard<std::mutex> g(mu);
  auto it = cache.find({dev, src});
  if (it != cache.end()) {
    *fn = it->second;
    return TDX_OK;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
    set_error("hipGetDeviceProperties failed");
    return TDX_ELAUNCH;
  }
  std::string arch(prop.gcnArchName);
  arch = arch.substr(0, arch.find(':'));
  std::vector<char> code;
  const int rc = nn_fused_jit_build(src, &code, arch);
  if (rc) return rc;
  hipModule_t mod = nullptr;
  if (hipModuleLoadData(&mod, code.data()) != hipSuccess) {
    set_error("hipModuleLoadData of a run-time compiled kernel failed");
    return TDX_ELAUNCH;
  }
  hipFunction_t f = nullptr;
  if (hipModuleGetFunction(&f, mod, name) != hipSuccess) {
    (void)hipModuleUnload(mod);
    set_error("run-time compiled kernel: entry point missing");
    return TDX_ELAUNCH;
  }
  cache[{dev, src}] = f;
  *fn = f;
  return TDX_OK;
}

}  // namespace tdx
