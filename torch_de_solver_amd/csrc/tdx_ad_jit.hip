// Plan-specialized jet maps and residual programs of the fused 'autograd'
// training step.
//
// The library's fused kernel (k_ad_step) walks the partition tables for
// every (point, unit) item and the residual records for every entry: chains
// of dependent table reads.  For large plans the library instead generates,
// at plan creation, the jet maps of the plan's jets as straight-line code
// (every pre-activation jet of an item read at once, the Faa di Bruno sums
// and their adjoint written out in the interpreter's order) and one
// straight-line residual function per group (stencil points, powers and
// coefficient kinds constants of the code; real values read from the
// plan's table at fixed positions), and compiles the step body
// (tdx_ad_dev.h, the text the library is built from) with them by hipRTC.

#include "tdx_ad_jit.h"

#include <sstream>
#include <utility>

#include "tdx_common.h"
// kTdxDevSrc, kTdxNnDevSrc, kTdxAdDevSrc: the texts of tdx_dev.h,
// tdx_nn_dev.h and tdx_ad_dev.h, written by the build
// (torch_de_solver_amd/build.py)
#include "tdx_embed.inc"

namespace tdx {

namespace {

std::string num(int64_t v) { return std::to_string(v); }

// product of tanh^(m) and the partition's blocks, left to right
std::string part_product(const int32_t *I, int pr, int deriv, int skip) {
  const int m = I[pr];
  std::string e = "s.s" + num(deriv);
  for (int q = 0; q < m; ++q)
    if (q != skip) e = "(" + e + " * z" + num(I[pr + 1 + q]) + ")";
  return e;
}

void emit_jets(std::ostringstream &o, const AdJitSpec &sp) {
  const int32_t *I = sp.ints;
  const int J = sp.J;
  auto loads = [&](const char *arr, const char *src) {
    for (int b = 0; b < J; ++b) o << "      const float " << arr << b << " = " << src << "[off[r] + " << b << "];\n";
  };
  o << "struct GenJets {\n"
       "  static __device__ __forceinline__ void fwd(const float *zb, float *hb, int r0, int NP,\n"
       "                                             float rNP, int J, int H, const Tables &,\n"
       "                                             int lane) {\n"
       "    const int n = NP * H;\n"
       "    for (int kb = lane; kb < n; kb += 64 * ADF_R) {\n"
       "      int off[ADF_R];\n"
       "      bool ok[ADF_R];\n"
       "      ad_items(kb, n, NP, rNP, r0, J, off, ok);\n"
       "#pragma unroll\n"
       "      for (int r = 0; r < ADF_R; ++r) {\n";
  loads("z", "zb");
  o << "      const TanhDerivs<float> s = tanh_derivs(z0);\n"
       "      if (ok[r]) hb[off[r]] = s.s0;\n";
  for (int al = 1; al < J; ++al) {
    const int k0 = I[sp.pptr_i + al], k1 = I[sp.pptr_i + al + 1];
    o << "      {\n";
    for (int k = k0; k < k1; ++k) {
      const int pr = sp.part_i + TDX_AD_PART_INTS * k;
      o << "        " << (k == k0 ? "float v = " : "v = v + ") << part_product(I, pr, I[pr], -1)
        << ";\n";
    }
    o << "        if (ok[r]) hb[off[r] + " << al << "] = v;\n      }\n";
  }
  o << "      }\n    }\n  }\n";
  o << "  static __device__ __forceinline__ void adj(const float *zb, const float *dab, float *dzb,\n"
       "                                             int r0, int NP, float rNP, int J, int H,\n"
       "                                             const Tables &, int lane) {\n"
       "    const int n = NP * H;\n"
       "    for (int kb = lane; kb < n; kb += 64 * ADF_R) {\n"
       "      int off[ADF_R];\n"
       "      bool ok[ADF_R];\n"
       "      ad_items(kb, n, NP, rNP, r0, J, off, ok);\n"
       "#pragma unroll\n"
       "      for (int r = 0; r < ADF_R; ++r) {\n";
  loads("z", "zb");
  loads("d", "dab");
  o << "      const TanhDerivs<float> s = tanh_derivs(z0);\n"
       "      float dz0 = d0 * s.s1;\n";
  for (int al = 1; al < J; ++al) {
    const int k0 = I[sp.pptr_i + al], k1 = I[sp.pptr_i + al + 1];
    for (int k = k0; k < k1; ++k) {
      const int pr = sp.part_i + TDX_AD_PART_INTS * k;
      o << "      dz0 = dz0 + d" << al << " * " << part_product(I, pr, I[pr] + 1, -1) << ";\n";
    }
  }
  o << "      if (ok[r]) dzb[off[r]] = dz0;\n";
  const int32_t *tptr = sp.occ.data();
  const int32_t *tent = sp.occ.data() + J + 1;
  for (int b = 1; b < J; ++b) {
    o << "      {\n        float acc = 0.0f;\n";
    for (int t = tptr[b]; t < tptr[b + 1]; ++t) {
      const int al = tent[3 * t], pk = tent[3 * t + 1], qo = tent[3 * t + 2];
      const int pr = sp.part_i + TDX_AD_PART_INTS * pk;
      // d_al * tanh^(m), then the other blocks left to right
      std::string e = "(d" + num(al) + " * s.s" + num(I[pr]) + ")";
      for (int q = 0; q < I[pr]; ++q)
        if (q != qo) e = "(" + e + " * z" + num(I[pr + 1 + q]) + ")";
      o << "        acc = acc + " << e << ";\n";
    }
    o << "        if (ok[r]) dzb[off[r] + " << b << "] = acc;\n      }\n";
  }
  o << "      }\n    }\n  }\n};\n";
}

std::string pow_expr(int kind, int ref, const std::string &x) {
  if (kind == 0) return "ipow(" + x + ", " + num(ref) + ")";
  return "dpow(" + x + ", R[" + num(ref) + "])";
}
std::string dpow_expr(int kind, int ref, const std::string &x) {
  if (kind == 0)
    return ref == 0 ? std::string("0.0f")
                    : "((float)(" + num(ref) + ") * ipow(" + x + ", " + num(ref - 1) + "))";
  const std::string p = "R[" + num(ref) + "]";
  return "(" + p + " * dpow(" + x + ", " + p + " - 1.0f))";
}

// the residual program of one group (ad_entry's operations, in its order)
void emit_group(std::ostringstream &o, const AdJitSpec &sp, int g) {
  const int32_t *I = sp.ints;
  const int J = sp.J, no = sp.n_out;
  const int32_t *G = I + sp.groups_i + TDX_AD_GROUP_INTS * g;
  const int neq = G[3], eq_i = G[4];
  auto coeff = [&](int Tr) {
    return I[Tr + 2] == 0 ? "R[" + num(I[Tr + 3]) + "]" : "D[" + num(I[Tr + 3]) + " + e]";
  };
  auto stencil = [&](int F, const std::string &dst) {
    const int nst = I[F], sti = I[F + 1], var = I[F + 2];
    for (int k = 0; k < nst; ++k) {
      const int S = sti + TDX_AD_STENCIL_INTS * k;
      const std::string t = "R[" + num(I[S + 2]) + "] * U[" +
                            num(((int64_t)I[S] * J + I[S + 1]) * no + var) + "]";
      o << "        " << dst << (k == 0 ? " = " + t : " = " + dst + " + " + t) << ";\n";
    }
  };
  auto has_chain = [&](int fi, int nf) {
    for (int f = 0; f < nf; ++f)
      if (I[fi + TDX_AD_FACTOR_INTS * f + 3] == 2) return true;
    return false;
  };
  // a term with a callable power: the factor values into fvv, then the
  // running product and its derivatives from the tables (chain_term)
  auto chain_call = [&](int fi, int nf) {
    o << "        float fvv[AD_MAXF], dfv[AD_MAXF];\n";
    for (int f = 0; f < 4; ++f) {
      if (f < nf) {
        stencil(fi + TDX_AD_FACTOR_INTS * f, "fvv[" + num(f) + "]");
      } else {
        o << "        fvv[" << f << "] = 0.0f;\n";
      }
    }
    o << "        const float pch = chain_term(a.ints, " << fi << ", " << nf
      << ", a.reals, fvv, dfv);\n";
  };
  for (int q = 0; q < neq; ++q) {
    const int E = eq_i + TDX_AD_EQ_INTS * q;
    const int nt = I[E], ti = I[E + 1];
    o << "    {\n      float r;\n";
    for (int t = 0; t < nt; ++t) {
      const int Tr = ti + TDX_AD_TERM_INTS * t;
      const int nf = I[Tr], fi = I[Tr + 1];
      o << "      {\n        const float c = " << coeff(Tr) << ";\n        float prod = 1.0f;\n";
      if (has_chain(fi, nf)) {
        chain_call(fi, nf);
        o << "        prod = pch;\n";
      } else {
        for (int f = 0; f < nf; ++f) {
          const int F = fi + TDX_AD_FACTOR_INTS * f;
          o << "        {\n        float fv;\n";
          stencil(F, "fv");
          o << "        prod = prod * " << pow_expr(I[F + 3], I[F + 4], "fv") << ";\n        }\n";
        }
      }
      o << "        const float tv = c * prod;\n        " << (t == 0 ? "r = tv;" : "r = r + tv;")
        << "\n      }\n";
    }
    const int tgt = I[E + 4];
    o << "      const float diff = "
      << (tgt >= 0 ? "r - D[" + num(tgt) + " + e]" : std::string("r")) << ";\n"
      << "      const float d2 = diff * diff;\n"
      << "      ls += R[" << I[E + 2] << "] * d2;\n"
      << "      lns += R[" << I[E + 3] << "] * d2;\n"
      << "      const float gq = 2.0f * R[" << I[E + 6] << "] * diff;\n";
    for (int t = 0; t < nt; ++t) {
      const int Tr = ti + TDX_AD_TERM_INTS * t;
      const int nf = I[Tr], fi = I[Tr + 1];
      o << "      {\n        const float c = " << coeff(Tr) << ";\n";
      if (has_chain(fi, nf)) {
        chain_call(fi, nf);
        for (int f = 0; f < nf; ++f) {
          const int F = fi + TDX_AD_FACTOR_INTS * f;
          o << "        {\n          const float dv = gq * c * dfv[" << f << "];\n";
          const int nst = I[F], sti = I[F + 1], var = I[F + 2];
          for (int k = 0; k < nst; ++k) {
            const int S = sti + TDX_AD_STENCIL_INTS * k;
            o << "          SD[" << ((int64_t)I[S] * J + I[S + 1]) * no + var << "] += dv * R["
              << I[S + 2] << "];\n";
          }
          o << "        }\n";
        }
        o << "      }\n";
        continue;
      }
      for (int f = 0; f < nf; ++f) {
        const int F = fi + TDX_AD_FACTOR_INTS * f;
        o << "        float fv" << f << ";\n";
        stencil(F, "fv" + num(f));
        o << "        const float pw" << f << " = " << pow_expr(I[F + 3], I[F + 4], "fv" + num(f))
          << ";\n";
      }
      for (int f = 0; f < nf; ++f) {
        const int F = fi + TDX_AD_FACTOR_INTS * f;
        o << "        {\n          float dv = gq * c * "
          << dpow_expr(I[F + 3], I[F + 4], "fv" + num(f)) << ";\n";
        for (int f2 = 0; f2 < nf; ++f2)
          if (f2 != f) o << "          dv = dv * pw" << f2 << ";\n";
        const int nst = I[F], sti = I[F + 1], var = I[F + 2];
        for (int k = 0; k < nst; ++k) {
          const int S = sti + TDX_AD_STENCIL_INTS * k;
          o << "          SD[" << ((int64_t)I[S] * J + I[S + 1]) * no + var << "] += dv * R["
            << I[S + 2] << "];\n";
        }
        o << "        }\n";
      }
      o << "      }\n";
    }
    o << "    }\n";
  }
}

}  // namespace

std::string ad_step_jit_source(const AdJitSpec &sp) {
  std::ostringstream o;
  o << "typedef __hip_internal::int32_t int32_t;\n"
       "typedef __hip_internal::int64_t int64_t;\n"
    << "#define TDX_GROUP_INTS " << TDX_GROUP_INTS << "\n"
    << "#define TDX_EQ_INTS " << TDX_EQ_INTS << "\n"
    << "#define TDX_TERM_INTS " << TDX_TERM_INTS << "\n"
    << "#define TDX_FACTOR_INTS " << TDX_FACTOR_INTS << "\n"
    << "#define TDX_STENCIL_INTS " << TDX_STENCIL_INTS << "\n"
    << "#define TDX_AD_MAX_ORDER " << TDX_AD_MAX_ORDER << "\n"
    << "#define TDX_AD_PART_INTS " << TDX_AD_PART_INTS << "\n"
    << "#define TDX_AD_GROUP_INTS " << TDX_AD_GROUP_INTS << "\n"
    << "#define TDX_AD_EQ_INTS " << TDX_AD_EQ_INTS << "\n"
    << "#define TDX_AD_TERM_INTS " << TDX_AD_TERM_INTS << "\n"
    << "#define TDX_AD_FACTOR_INTS " << TDX_AD_FACTOR_INTS << "\n"
    << "#define TDX_AD_STENCIL_INTS " << TDX_AD_STENCIL_INTS << "\n"
    << "#define TDX_AD_MAX_CHAIN " << TDX_AD_MAX_CHAIN << "\n";
  const std::pair<const char *, int> ops[] = {
      {"ADD", TDX_AD_OP_ADD},   {"SUB", TDX_AD_OP_SUB},   {"RSUB", TDX_AD_OP_RSUB},
      {"MUL", TDX_AD_OP_MUL},   {"DIV", TDX_AD_OP_DIV},   {"RDIV", TDX_AD_OP_RDIV},
      {"POWI", TDX_AD_OP_POWI}, {"POWR", TDX_AD_OP_POWR}, {"NEG", TDX_AD_OP_NEG},
      {"SIN", TDX_AD_OP_SIN},   {"COS", TDX_AD_OP_COS},   {"EXP", TDX_AD_OP_EXP},
      {"LOG", TDX_AD_OP_LOG},   {"TANH", TDX_AD_OP_TANH}, {"SQRT", TDX_AD_OP_SQRT}};
  for (const auto &op : ops) o << "#define TDX_AD_OP_" << op.first << " " << op.second << "\n";
  o
    << kTdxDevSrc << "\n"
    << kTdxNnDevSrc << "\n"
    << kTdxAdDevSrc << "\n";
  o << "namespace tdx {\n// jet maps of the plan's jets\n";
  emit_jets(o, sp);
  o << "// residual programs, one per group\n"
       "struct GenADResid {\n"
       "  static __device__ __forceinline__ void run(const ADStepArgs &a, int g, int e, int n_out,\n"
       "                                             const float *U, float *SD, float &ls,\n"
       "                                             float &lns) {\n"
       "    const float *R = a.reals;\n"
       "    const float *D = a.data;\n"
       "    switch (g) {\n";
  for (int g = 0; g < sp.n_groups; ++g) {
    o << "    case " << g << ": {\n";
    emit_group(o, sp, g);
    o << "    } break;\n";
  }
  o << "    default: break;\n    }\n  }\n};\n}  // namespace tdx\n"
       "extern \"C\" __global__ __launch_bounds__(256) void tdx_ad_step_jit(tdx::ADStepArgs a) {\n"
       "  extern __shared__ __align__(16) unsigned char smem_raw[];\n"
       "  tdx::ad_step_body<"
    << sp.NB << ", " << sp.NHM
    << ", tdx::GenJets, tdx::GenADResid>(a, reinterpret_cast<float *>(smem_raw));\n"
       "}\n";
  return o.str();
}

}  // namespace tdx
