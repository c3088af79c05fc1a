// Plan-specialized 'mat' kernel.
//
// The general tile kernel (k_mat_tile) interprets the residual program and
// passes every derivative level through LDS; it spends hundreds of
// instructions per grid point.  For the interior of the grid -- every point
// whose stencils and halo lie inside the grid, where all stencils take the
// central form -- the plan generates a kernel for its own program and
// compiles it at plan creation (hipRTC):
//
// * one wavefront owns a strip of 64 columns (lanes; the outer 2R1 on each
//   side are halo) and marches down a band of rows, loading one row of u per
//   step;
// * every derived field -- each level of each derivative chain, the seeds
//   dL/d(factor), each level of the transposed chains -- lives in a small
//   ring of registers indexed by row; a derivative along axis 0 reads the
//   ring of its input, one along axis 1 reads the neighbouring lanes through
//   DPP wavefront shifts;
// * each field lags the load front by the rows its axis-0 stencils reach, so
//   at step t the residual is formed for row t - LAGF and the gradient for
//   row t - LAGG; the march is unrolled so every ring index is a constant;
// * no LDS, no barriers; coefficients, powers, lambdas and the stencil
//   weights are constants of the generated code.
//
// The forward arithmetic is the interior form of the general kernel's (the
// reference's) operation for operation, including the division mode; the
// transposed chains use the same products ab/(2h), af/(2h).

#include "tdx_mat_jit.h"

#include <hip/hiprtc.h>

#include <cstdio>
#include <map>
#include <mutex>
#include <sstream>

#include "tdx_common.h"

namespace tdx {

namespace {

std::string fl(float v) {
  char b[64];
  std::snprintf(b, sizeof b, "%af", (double)v);
  return b;
}

int pow2_at_least(int n) {
  int w = 1;
  while (w < n) w <<= 1;
  return w;
}

struct Field {
  int kind;   // 0 load(var), 1 forward D(src, axis), 2 seed(fs), 3 transposed D(src, axis)
  int src = -1, axis = 0, var = 0, fs = -1;
  int lag = 0, back = 0, W = 1;
  std::string name;
};

}  // namespace

std::string mat_jit_source(const JitSpec &s, std::string *why) {
  const int dp = s.dp;
  if (dp < 2 || dp > 8) {
    *why = "derivative points";
    return "";
  }
  std::vector<Field> F;
  std::map<std::vector<int>, int> chain_id;  // (var, axes prefix...) -> field
  auto add = [&](Field f) {
    f.name = "f" + std::to_string(F.size());
    F.push_back(f);
    return (int)F.size() - 1;
  };
  // factor slots in (equation, term, factor) order
  struct Slot {
    const JitFactor *f;
    int fwd;   // field holding the factor value
    int seed;  // seed field
    int out;   // last transposed level (adds into the gradient of var)
  };
  std::vector<Slot> slots;
  for (const JitEq &e : s.eqs)
    for (const JitTerm &t : e.terms)
      for (const JitFactor &f : t.f) {
        std::vector<int> key{f.var};
        auto it = chain_id.find(key);
        int cur;
        if (it == chain_id.end()) {
          Field ld;
          ld.kind = 0;
          ld.var = f.var;
          cur = chain_id[key] = add(ld);
        } else {
          cur = it->second;
        }
        for (int k = 0; k < f.nax; ++k) {
          key.push_back(f.ax[k]);
          auto jt = chain_id.find(key);
          if (jt != chain_id.end()) {
            cur = jt->second;
            continue;
          }
          Field d;
          d.kind = 1;
          d.src = cur;
          d.axis = f.ax[k];
          d.lag = F[cur].lag + (d.axis == 0 ? dp - 1 : 0);
          cur = chain_id[key] = add(d);
        }
        slots.push_back(Slot{&f, cur, -1, -1});
      }
  int LAGF = 0;
  for (const Slot &sl : slots) LAGF = std::max(LAGF, F[sl.fwd].lag);
  for (size_t q = 0; q < slots.size(); ++q) {
    Field sd;
    sd.kind = 2;
    sd.fs = (int)q;
    sd.lag = LAGF;
    slots[q].seed = add(sd);
  }
  int LAGG = LAGF;
  for (size_t q = 0; q < slots.size(); ++q) {
    const JitFactor &f = *slots[q].f;
    int cur = slots[q].seed;
    for (int k = f.nax - 1; k >= 0; --k) {
      Field d;
      d.kind = 3;
      d.src = cur;
      d.axis = f.ax[k];
      d.lag = F[cur].lag + (d.axis == 0 ? dp - 1 : 0);
      cur = add(d);
    }
    slots[q].out = cur;
    LAGG = std::max(LAGG, F[cur].lag);
  }
  // how far back each field is read
  for (const Field &f : F)
    if ((f.kind == 1 || f.kind == 3) && f.axis == 0) F[f.src].back = std::max(F[f.src].back, 2 * (dp - 1));
  for (const Slot &sl : slots) {
    F[sl.fwd].back = std::max(F[sl.fwd].back, LAGF - F[sl.fwd].lag);
    F[sl.out].back = std::max(F[sl.out].back, LAGG - F[sl.out].lag);
  }
  std::vector<int> load_of(s.n_eq, -1);
  for (size_t k = 0; k < F.size(); ++k)
    if (F[k].kind == 0) load_of[F[k].var] = (int)k;
  // The march is unrolled by UNR rows (a power of two at least every ring
  // window, so ring slots are constants) and loads run one unrolled block
  // ahead, so at least 8 rows are in flight per wavefront.  The fused Adam
  // step also keeps u in the load ring until its gradient row when that
  // does not deepen the unrolling (else it reads u again).
  std::vector<int> back0(F.size());
  for (size_t k = 0; k < F.size(); ++k) back0[k] = F[k].back;
  int UNR = 8;
  bool u_in_ring = false;
  auto windows = [&](bool step) {
    int unr = 8;
    for (size_t k = 0; k < F.size(); ++k) {
      F[k].back = back0[k];
      F[k].W = pow2_at_least(F[k].back + 1);
      unr = std::max(unr, F[k].W);
    }
    u_in_ring = false;
    if (step) {
      int unr2 = unr;
      for (const Field &f : F)
        if (f.kind == 0) unr2 = std::max(unr2, pow2_at_least(std::max(f.back, LAGG) + 1));
      if (unr2 == unr) {
        u_in_ring = true;
        for (Field &f : F)
          if (f.kind == 0) {
            f.back = std::max(f.back, LAGG);
            f.W = pow2_at_least(f.back + 1);
          }
      }
    }
    UNR = unr;
  };
  windows(false);
  if (UNR > 16) {
    *why = "stencil windows too deep";
    return "";
  }
  const int R0 = s.R0, R1 = s.R1;
  const int OW = 64 - 4 * R1;
  if (OW < 16) {
    *why = "axis-1 reach too large for a 64-lane strip";
    return "";
  }

  std::ostringstream o;
  o << "// generated by tdx_mat_jit for one 'mat' program\n"
       "typedef long long i64;\n"
       "struct JitArgs { const float *u; float *grad; float *op; float *lp; const float *data;\n"
       "  const float *dloss; i64 N; int n0, n1; int r_lo, r_hi, c_lo, c_hi; int hb;\n"
       "  int nstrips, nbands; int lp_base; int adj;\n"
       "  float *u_out; float *m; float *v; const float *adam; };\n"
       "__device__ __forceinline__ float shl1(float v) {  // lane l gets lane l+1\n"
       "  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xF, 0xF, false)); }\n"
       "__device__ __forceinline__ float shr1(float v) {  // lane l gets lane l-1\n"
       "  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xF, 0xF, false)); }\n";
  // division
  o << "__device__ __forceinline__ float qdiv(float x, float h, float inv) {\n";
  if (s.div_mode == 0) {
    o << "  return x * inv;\n";
  } else if (s.div_mode == 2) {
    o << "  const float q = x * inv;\n"
         "  const float r = __builtin_fmaf(-q, h, x);\n"
         "  float res = __builtin_copysignf(__builtin_fmaf(r, inv, q), x);\n"
         "  const unsigned ab = __float_as_uint(x) & 0x7fffffffu;\n"
         "  const bool slow = ab != 0u && (ab - 0x12800000u) > (0x6C800000u - 0x12800000u);\n"
         "  if (__builtin_amdgcn_ballot_w64(slow) != 0) res = slow ? x / h : res;\n"
         "  return res;\n";
  } else {
    o << "  return x / h;\n";
  }
  o << "}\n";

  auto ring = [&](int fid, int j, int rows_back) {
    // slot of the row that is rows_back rows before field fid's newest row at
    // unrolled step j
    const Field &f = F[fid];
    const int idx = (((j - f.lag - rows_back) % f.W) + f.W) % f.W;
    return f.name + "[" + std::to_string(idx) + "]";
  };
  auto powexpr = [&](const JitFactor &f, const std::string &x) -> std::string {
    if (f.pow_kind != 0) return "powf(" + x + ", " + fl((float)f.pow_real) + ")";
    switch (f.pow_int) {
      case 0: return "1.0f";
      case 1: return x;
      case 2: return "(" + x + " * " + x + ")";
      case 3: return "((" + x + " * " + x + ") * " + x + ")";
      case -1: return "(1.0f / " + x + ")";
      case -2: return "(1.0f / (" + x + " * " + x + "))";
      default: return "powf(" + x + ", " + fl((float)f.pow_int) + ")";
    }
  };
  auto dpowexpr = [&](const JitFactor &f, const std::string &x) -> std::string {
    if (f.pow_kind != 0)
      return "(" + fl((float)f.pow_real) + " * powf(" + x + ", " +
             fl((float)f.pow_real) + " - 1.0f))";
    const int p = f.pow_int;
    if (p == 0) return "0.0f";
    JitFactor g = f;
    g.pow_int = p - 1;
    return "(" + fl((float)p) + " * " + powexpr(g, x) + ")";
  };

  // mode 2: residual + loss + adjoint + Adam update into u_out, m, v;
  // 1: residual + loss + gradient; 0: residual + loss
  for (int mode = 2; mode >= 0; --mode) {
    const bool adj = mode >= 1, step = mode == 2;
    windows(step);
    o << "extern \"C\" __global__ __launch_bounds__(256) void tdx_mat_jit_"
      << (step ? "step" : adj ? "adj" : "fwd") << "(JitArgs a) {\n"
         "  // blocks go to the 8 XCDs round-robin: block b takes logical block\n"
         "  // (b % 8) * per + b / 8, so each XCD walks one contiguous range of\n"
         "  // strips and the halo columns neighbouring strips share stay in its L2\n"
         "  int blk = blockIdx.x;\n"
         "  { const int full = gridDim.x & ~7; if (blk < full) blk = (blk & 7) * (full >> 3) + (blk >> 3); }\n"
         "  const int wv = blk * 4 + (threadIdx.x >> 6);\n"
         "  const int lane = threadIdx.x & 63;\n"
         "  if (wv >= a.nstrips * a.nbands) return;\n"
         "  const int strip = wv % a.nstrips, band = wv / a.nstrips;\n"
         "  const int c = a.c_lo + strip * " << OW << " - " << 2 * R1 << " + lane;\n"
         "  const bool own = lane >= " << 2 * R1 << " && lane < " << 2 * R1 + OW << " && c < a.c_hi;\n"
         "  const bool cin = c >= 0 && c < a.n1;\n"
         "  const int b_lo = a.r_lo + band * a.hb;\n"
         "  const int b_hi = min(b_lo + a.hb, a.r_hi);\n"
         "  int t0 = b_lo - " << 2 * R0 << ";\n"
         "  t0 -= ((t0 % " << UNR << ") + " << UNR << ") % " << UNR << ";\n"
         "  // last row step, rounded up to whole unrolled blocks (the extra\n"
         "  // steps' rows lie outside the band and are not stored)\n"
         "  const int tl = b_hi - 1 + " << (adj ? LAGG : LAGF) << ";\n"
         "  const int t1 = t0 + ((tl - t0) / " << UNR << " + 1) * " << UNR << " - 1;\n"
         "  // sum of r^2 per equation; lambda and 1/N are applied once at the end\n"
         "  float esum[" << s.n_eq << "];\n"
         "  for (int k = 0; k < " << s.n_eq << "; ++k) esum[k] = 0.0f;\n";
    if (adj) o << "  const float dl = *a.dloss;\n";
    if (step)
      o << "  // torch.optim.Adam scalars of this step: 1-b1, 1-b2, b2, lr/(1-b1^t),\n"
           "  // sqrt(1-b2^t), eps\n"
           "  const float w1 = a.adam[0], w2 = a.adam[1], b2 = a.adam[2], ss = a.adam[3],\n"
           "              bc2s = a.adam[4], eps = a.adam[5];\n";
    // m and v of the gradient rows, fetched one unrolled block ahead
    auto mvload = [&](int v, const std::string &row, int slot) {
      const std::string sl = std::to_string(slot), vs = std::to_string(v);
      return "  { const int r = " + row + "; float xm = 0.0f, xv = 0.0f;\n"
             "    if (own && r >= b_lo && r < b_hi) { const i64 gi = (i64)" + vs +
             " * a.N + (i64)r * a.n1 + c; xm = a.m[gi]; xv = a.v[gi]; }\n"
             "    pm" + vs + "[" + sl + "] = xm; pv" + vs + "[" + sl + "] = xv; }\n";
    };
    if (step)
      for (int v = 0; v < s.n_eq; ++v) {
        o << "  float pm" << v << "[" << UNR << "], pv" << v << "[" << UNR << "];\n";
        for (int j = 0; j < UNR; ++j)
          o << mvload(v, "t0 - " + std::to_string(LAGG) + " + " + std::to_string(j), j);
      }
    auto load = [&](const Field &f, const std::string &row) {
      return "{ const int r = " + row + "; float x = 0.0f; if (r >= 0 && r < a.n0 && cin) x = a.u[(i64)" +
             std::to_string(f.var) + " * a.N + (i64)r * a.n1 + c]; ";
    };
    for (const Field &f : F) {
      if (!adj && (f.kind == 2 || f.kind == 3)) continue;
      o << "  float " << f.name << "[" << f.W << "];\n"
        << "  for (int k = 0; k < " << f.W << "; ++k) " << f.name << "[k] = 0.0f;\n";
      if (f.kind == 0) {
        // rows of the first unrolled block
        o << "  float pf_" << f.name << "[" << UNR << "];\n";
        for (int j = 0; j < UNR; ++j)
          o << "  " << load(f, "t0 + " + std::to_string(j)) << "pf_" << f.name << "[" << j << "] = x; }\n";
      }
    }
    o << "  for (int tb = t0; tb <= t1; tb += " << UNR << ") {\n";
    for (int j = 0; j < UNR; ++j) {
      o << "  {\n"
        << "    const int t = tb + " << j << ";\n";
      // loads, then forward levels in creation order
      for (size_t fid = 0; fid < F.size(); ++fid) {
        const Field &f = F[fid];
        if (f.kind == 0) {
          // take row t from the prefetch registers, then fetch row t + UNR
          o << "    " << ring((int)fid, j, 0) << " = pf_" << f.name << "[" << j << "];\n"
            << "    " << load(f, "t + " + std::to_string(UNR)) << "pf_" << f.name << "[" << j << "] = x; }\n";
        } else if (f.kind == 1) {
          // central stencil of src for the row (t - lag): b = sum ab[i] src[r + i - (dp-1)],
          // f = sum af[i] src[r + i], (b + f) / (2h)
          const int ax = f.axis;
          o << "    {\n";
          auto tap = [&](int k) {  // src at offset k (rows or columns) from the output point
            if (ax == 0) return ring(f.src, j, (dp - 1) - k);
            std::string v = ring(f.src, j, 0);
            std::string e = v;
            for (int m = 0; m < (k > 0 ? k : -k); ++m) e = std::string(k > 0 ? "shl1(" : "shr1(") + e + ")";
            return e;
          };
          if (ax == 1) {
            // lane shifts of the source row, computed once
            o << "      const float s0 = " << ring(f.src, j, 0) << ";\n";
            for (int k = 1; k <= dp - 1; ++k) {
              o << "      const float sp" << k << " = shl1(" << (k == 1 ? std::string("s0") : "sp" + std::to_string(k - 1)) << ");\n";
              o << "      const float sm" << k << " = shr1(" << (k == 1 ? std::string("s0") : "sm" + std::to_string(k - 1)) << ");\n";
            }
          }
          auto val = [&](int k) -> std::string {
            if (ax == 0) return tap(k);
            if (k == 0) return "s0";
            return (k > 0 ? "sp" : "sm") + std::to_string(k > 0 ? k : -k);
          };
          o << "      float b = " << val(-(dp - 1)) << " * " << fl(s.ab[0]) << ";\n";
          for (int i = 1; i < dp; ++i) o << "      b = b + " << val(i - (dp - 1)) << " * " << fl(s.ab[i]) << ";\n";
          o << "      float fw = " << val(0) << " * " << fl(s.af[0]) << ";\n";
          for (int i = 1; i < dp; ++i) o << "      fw = fw + " << val(i) << " * " << fl(s.af[i]) << ";\n";
          o << "      " << ring((int)fid, j, 0) << " = qdiv(b + fw, " << fl(s.h2[ax]) << ", " << fl(s.ih2[ax]) << ");\n"
            << "    }\n";
        }
      }
      // residual, loss, seeds for row i = t - LAGF
      o << "    {\n"
        << "      const int i = t - " << LAGF << ";\n"
        << "      const bool rin = own && i >= b_lo && i < b_hi;\n"
        << "      const i64 gp = (i64)i * a.n1 + c;\n";
      size_t q = 0;
      for (size_t e = 0; e < s.eqs.size(); ++e) {
        const JitEq &eq = s.eqs[e];
        o << "      {\n        float r;\n";
        std::vector<std::string> cexpr;
        const size_t q0 = q;
        for (size_t ti = 0; ti < eq.terms.size(); ++ti) {
          const JitTerm &tm = eq.terms[ti];
          const std::string cn = "c" + std::to_string(ti);
          if (tm.coeff_kind == 0)
            o << "        const float " << cn << " = " << fl((float)tm.coeff) << ";\n";
          else
            o << "        const float " << cn << " = (i >= 0 && i < a.n0 && cin) ? a.data["
              << tm.data_off << "LL + gp] : 0.0f;\n";
          o << "        float der" << ti << " = 1.0f;\n";
          for (size_t fi = 0; fi < tm.f.size(); ++fi, ++q)
            o << "        der" << ti << " = der" << ti << " * " << powexpr(tm.f[fi], ring(slots[q].fwd, j, LAGF - F[slots[q].fwd].lag)) << ";\n";
          o << "        { const float tv = " << cn << " * der" << ti << "; r = " << (ti == 0 ? "tv" : "r + tv") << "; }\n";
        }
        o << "        if (rin) {\n"
          << "          esum[" << e << "] += r * r;\n"
          << "          if (a.op) a.op[(i64)" << e << " * a.N + gp] = r;\n"
          << "        }\n";
        if (adj) {
          o << "        const float g = 2.0f * " << fl((float)eq.lam) << " * r * dl;\n";
          size_t qq = q0;
          for (size_t ti = 0; ti < eq.terms.size(); ++ti) {
            const JitTerm &tm = eq.terms[ti];
            const size_t nf = tm.f.size();
            std::vector<std::string> fv(nf), pw(nf);
            for (size_t fi = 0; fi < nf; ++fi) {
              fv[fi] = ring(slots[qq + fi].fwd, j, LAGF - F[slots[qq + fi].fwd].lag);
              pw[fi] = "pw" + std::to_string(ti) + "_" + std::to_string(fi);
              o << "        const float " << pw[fi] << " = " << powexpr(tm.f[fi], fv[fi]) << ";\n";
            }
            for (size_t fi = 0; fi < nf; ++fi) {
              o << "        { float d = c" << ti << " * g * " << dpowexpr(tm.f[fi], fv[fi]) << ";\n";
              for (size_t f2 = 0; f2 < nf; ++f2)
                if (f2 != fi) o << "          d *= " << pw[f2] << ";\n";
              o << "          " << ring(slots[qq + fi].seed, j, 0) << " = d; }\n";
            }
            qq += nf;
          }
        }
        o << "      }\n";
      }
      o << "    }\n";
      if (adj) {
        // transposed levels
        for (size_t fid = 0; fid < F.size(); ++fid) {
          const Field &f = F[fid];
          if (f.kind != 3) continue;
          const int ax = f.axis;
          o << "    {\n";
          if (ax == 1) {
            o << "      const float s0 = " << ring(f.src, j, 0) << ";\n";
            for (int k = 1; k <= dp - 1; ++k) {
              o << "      const float sp" << k << " = shl1(" << (k == 1 ? std::string("s0") : "sp" + std::to_string(k - 1)) << ");\n";
              o << "      const float sm" << k << " = shr1(" << (k == 1 ? std::string("s0") : "sm" + std::to_string(k - 1)) << ");\n";
            }
          }
          auto val = [&](int k) -> std::string {
            if (ax == 0) return ring(f.src, j, (dp - 1) - k);
            if (k == 0) return "s0";
            return (k > 0 ? "sp" : "sm") + std::to_string(k > 0 ? k : -k);
          };
          // acc = sum_i src[g + (dp-1) - i] ab_i/(2h) + sum_i src[g - i] af_i/(2h)
          o << "      float acc = " << val(dp - 1) << " * " << fl(s.tb2[ax][0]) << ";\n";
          for (int i = 1; i < dp; ++i) o << "      acc += " << val((dp - 1) - i) << " * " << fl(s.tb2[ax][i]) << ";\n";
          for (int i = 0; i < dp; ++i) o << "      acc += " << val(-i) << " * " << fl(s.tf2[ax][i]) << ";\n";
          o << "      " << ring((int)fid, j, 0) << " = acc;\n"
            << "    }\n";
        }
        // gradient for row t - LAGG: per variable, slots in order
        o << "    {\n"
          << "      const int gr = t - " << LAGG << ";\n"
          << "      if (" << (step ? "" : "a.adj && ") << "own && gr >= b_lo && gr < b_hi) {\n";
        for (int v = 0; v < s.n_eq; ++v) {
          // the general kernel adds the slots into a zeroed accumulator in
          // slot order: ((0 + s0) + s1) + ...; a variable no factor reads
          // has a zero gradient
          std::string acc = "0.0f";
          {
            bool first = true;
            for (size_t qi = 0; qi < slots.size(); ++qi) {
              if (slots[qi].f->var != v) continue;
              const std::string term = ring(slots[qi].out, j, LAGG - F[slots[qi].out].lag);
              acc = first ? "(0.0f + " + term + ")" : "(" + acc + " + " + term + ")";
              first = false;
            }
          }
          if (!step) {
            o << "        a.grad[(i64)" << v << " * a.N + (i64)gr * a.n1 + c] = " << acc << ";\n";
            continue;
          }
          // Adam, in the operation order of the library's update kernel
          const std::string vs = std::to_string(v), js = std::to_string(j);
          o << "        {\n"
            << "          const i64 gi = (i64)" << vs << " * a.N + (i64)gr * a.n1 + c;\n"
            << "          const float g = " << acc << ";\n"
            << "          float p = " << (u_in_ring && load_of[v] >= 0 ? ring(load_of[v], j, LAGG) : "a.u[gi]") << ";\n"
            << "          float mm = pm" << vs << "[" << js << "], vv = pv" << vs << "[" << js << "];\n"
            << "          mm = mm + w1 * (g - mm);\n"
            << "          vv = vv * b2 + w2 * g * g;\n"
            << "          const float den = sqrtf(vv) / bc2s + eps;\n"
            << "          p = p - ss * (mm / den);\n"
            << "          a.u_out[gi] = p; a.m[gi] = mm; a.v[gi] = vv;\n"
            << "        }\n";
        }
        o << "      }\n    }\n";
        if (step)
          for (int v = 0; v < s.n_eq; ++v)
            o << mvload(v, "t + " + std::to_string(UNR) + " - " + std::to_string(LAGG), j);
      }
      o << "  }\n";
    }
    o << "  }\n"
         "  float lsum = 0.0f, lnsum = 0.0f;\n";
    for (size_t e = 0; e < s.eqs.size(); ++e)
      o << "  lsum += " << fl((float)s.eqs[e].lam) << " * esum[" << e << "];\n"
        << "  lnsum += esum[" << e << "];\n";
    o << "  lnsum = lnsum / " << fl((float)s.N) << ";\n"
         "  for (int m = 32; m >= 1; m >>= 1) {\n"
         "    lsum += __shfl_xor(lsum, m);\n"
         "    lnsum += __shfl_xor(lnsum, m);\n"
         "  }\n"
         "  if (lane == 0) {\n"
         "    a.lp[2 * (a.lp_base + wv)] = lsum;\n"
         "    a.lp[2 * (a.lp_base + wv) + 1] = lnsum;\n"
         "  }\n"
         "}\n";
  }
  return o.str();
}

int mat_jit_build(const std::string &src, std::vector<char> *code, const std::string &arch) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "tdx_mat_jit.hip", 0, nullptr, nullptr) !=
      HIPRTC_SUCCESS) {
    set_error("hiprtcCreateProgram failed");
    return TDX_ELAUNCH;
  }
  const std::string arch_opt = "--offload-arch=" + arch;
  const char *opts[] = {arch_opt.c_str(), "-O3", "-ffp-contract=off", "-Xclang",
                        "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult rc = hiprtcCompileProgram(prog, 7, opts);
  if (rc != HIPRTC_SUCCESS) {
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string log(ls, '\0');
    if (ls) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    set_error("specialized 'mat' kernel failed to compile: %.400s", log.c_str());
    return TDX_ELAUNCH;
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  code->assign(cs, 0);
  hiprtcGetCode(prog, code->data());
  hiprtcDestroyProgram(&prog);
  return TDX_OK;
}

int mat_jit_compile(const std::string &src, JitKernel *k) {
  static std::mutex mu;
  static std::map<std::pair<int, std::string>, JitKernel> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    set_error("hipGetDevice failed");
    return TDX_ELAUNCH;
  }
  std::lock_guard<std::mutex> g(mu);
  auto it = cache.find({dev, src});
  if (it != cache.end()) {
    *k = it->second;
    return TDX_OK;
  }
  // code for the current device's architecture (the name before any
  // feature suffix, e.g. "gfx950" of "gfx950:sramecc+:xnack-")
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
    set_error("hipGetDeviceProperties failed");
    return TDX_ELAUNCH;
  }
  std::string arch(prop.gcnArchName);
  arch = arch.substr(0, arch.find(':'));
  std::vector<char> code;
  const int rc = mat_jit_build(src, &code, arch);
  if (rc) return rc;
  JitKernel out;
  if (hipModuleLoadData(&out.module, code.data()) != hipSuccess) {
    set_error("hipModuleLoadData of the specialized 'mat' kernel failed");
    return TDX_ELAUNCH;
  }
  if (hipModuleGetFunction(&out.fn_adj, out.module, "tdx_mat_jit_adj") != hipSuccess ||
      hipModuleGetFunction(&out.fn_fwd, out.module, "tdx_mat_jit_fwd") != hipSuccess ||
      hipModuleGetFunction(&out.fn_step, out.module, "tdx_mat_jit_step") != hipSuccess) {
    (void)hipModuleUnload(out.module);
    set_error("specialized 'mat' kernel: entry point missing");
    return TDX_ELAUNCH;
  }
  cache[{dev, src}] = out;
  *k = out;
  return TDX_OK;
}

}  // namespace tdx
