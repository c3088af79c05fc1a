// 'mat' mode: the unknown is the grid tensor itself and derivatives are
// per-axis stencils (tedeous/derivative.py:135-323).
//
// The reference applies D along an axis with two flattened torch.roll sums,
// divides by 2h and overwrites the first/last dp-1 entries with one-sided
// quotients by h; a k-th derivative is k such applications, every one of
// them a round trip through HBM plus a device sync in torch.unique for h.
//
// Here one workgroup owns a T0 x T1 tile of output points.  It loads u on the
// tile plus a halo of 2R per side (R = reach of the deepest derivative) into
// LDS, applies the derivative chains of every factor there (the halo shrinks
// by dp-1 per application), evaluates the residual program and its squared
// loss on the tile plus R, forms the seeds 2*lambda/N*op*d(term)/d(factor),
// runs the transposed chains back down to the tile and writes the gradient.
// One HBM read of u and one write of the gradient per point.
//
// D at global index g of an axis of length n, in the reference's arithmetic:
//   back = sum_i u[g+i-(dp-1)]*ab[i], forw = sum_i u[g+i]*af[i] (in order)
//   g <  dp-1   : forw / h
//   g >= n-dp+1 : back / h
//   otherwise   : (back + forw) / (2h)
// With exact_div the quotients are IEEE divisions and the fields are
// bit-identical to the reference's on the same input; otherwise they are
// products with the rounded reciprocals.

#include "tdx_common.h"
#include "tdx_mat_jit.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>

namespace tdx {

constexpr int MMAX_DP = 8;
constexpr int MMAX_EQ = 4;
constexpr int MMAX_FS = 16;   // factor slots over all terms of all equations
constexpr int MMAX_TF = 4;    // factors of one term
constexpr int MAT_THREADS = 256;
constexpr int MFS_INTS = 2 + TDX_MAT_MAX_CHAIN;  // factor-slot record: var, n_axes, axes
// Stencil coefficient table (T values): ab[MMAX_DP] af[MMAX_DP], then per
// axis (three) 4*MMAX_DP products for the transposed stencils:
//   ab/(2h) | af/(2h) | ab/h | af/h
constexpr int MCOEF_AX = 2 * MMAX_DP;
constexpr int MCOEF_N = 2 * MMAX_DP + 3 * 4 * MMAX_DP;

template <typename T>
struct MatArgs {
  int n_eq, n0, n1, dp;
  int64_t N;  // n0 * n1
  T h0, h1, h20, h21, ih0, ih1, ih20, ih21;  // h, 2h and their reciprocals per axis
  const T *coef;      // stencil coefficient table
  const int32_t *fs;  // factor-slot records, n_fs x MFS_INTS
  // residual program, flattened (copied to LDS by the tile kernel):
  //   pint: per equation {n_terms, first term}; per term {n_factors,
  //         coefficient kind, coefficient ref, first factor slot}; per
  //         factor slot {pow kind, pow int, pow real index}
  //   preal: lambda per equation, scalar coefficients, real powers
  const int32_t *pint;
  const T *preal;
  int n_pint, n_preal, pterm0, pfac0;
  const int32_t *ints;
  const T *reals;
  const T *data;
  int eq_i;
  int n_fs;
  int R0, R1;         // reach per axis
  int T0, T1;         // output tile
  int E0, E1;         // tile with its halo of 2R per side; E1 is 16, 32 or 64 * 2^k
  int tpr_log;        // log2 of the threads per tile row, min(E1, 256)
  int tiles1;         // tiles along axis 1
  const int32_t *tile_list;  // tiles of this launch; null: every tile
  // bc
  int n_bc, bc_i;
  int n_touch, bct_ptr_i, bct_pos_i, bct_seed_i, bct_w_i;
  // buffers
  const T *u;
  T *grad;     // may be null (forward only)
  T *op;       // may be null
  T *bval;     // may be null
  T *lp;       // 2 per tile workgroup then 2 per bc workgroup
  T *lp_mid;   // 2 x 256: first stage of the loss sum
  T *bc_seeds;
  const T *dloss;
  int n_tile_wg;
  // full-grid plans: chain fields [n_fs][N], and per bc entry the first
  // field of its operator's factors (-1: plain value)
  T *fields;
  const int32_t *bc_slot;
  // fused train step (u_in -> u_out): moments and the step's Adam scalars
  int fused;
  T *u_out;
  T *m;
  T *v;
  const T *adam;
};

// x / h by division mode DIV: 0 multiplies by the rounded reciprocal inv
// (within an ulp); 1 is the IEEE division; 2 (float only) is the
// correctly rounded quotient from inv = RN(1/h) by one correction step,
// q = RN(x inv), r = x - q h (exact by fma), RN(q + r inv), which the plan
// has checked against the correctly rounded quotient for every significand
// of x in [1, 2) (h > 0).  It holds for every x whose steps scale exactly,
// 2^-90 <= |x| <= 2^90; the sign of x is copied onto the result, which
// makes it exact for +-0 too.  Any other x (subnormal, huge, inf, NaN)
// takes the IEEE division: a wavefront-uniform test keeps the common case
// free of branches.
template <typename T, int DIV>
__device__ __forceinline__ T qdiv(T x, T h, T inv) {
  if constexpr (DIV == 0) {
    return x * inv;
  } else if constexpr (DIV == 2 && sizeof(T) == 4) {
    const float q = x * inv;
    const float r = fmaf(-q, h, x);
    float res = copysignf(fmaf(r, inv, q), x);
    const uint32_t ab = __float_as_uint(x) & 0x7fffffffu;
    const bool slow = ab != 0u && (ab - 0x12800000u) > (0x6C800000u - 0x12800000u);
    if (__builtin_amdgcn_ballot_w64(slow) != 0) res = slow ? x / h : res;
    return res;
  } else {
    return x / h;
  }
}

// One application of D at local point p of an LDS field (stride between
// neighbours along the axis); g is the global index along the axis, n its
// length, h/h2 the step and twice the step.  DP > 0 fixes the number of
// derivative points at compile time (the loops unroll and ab/af index
// registers); DP = 0 takes it from dp.
template <typename T, int DP, int DIV>
__device__ __forceinline__ T dstencil(const T *src, int p, int stride, int g, int n, int dp,
                                      const T *ab, const T *af, T h, T h2, T ih, T ih2) {
  const int np = DP > 0 ? DP : dp;
  if (g < np - 1) {
    T f = src[p] * af[0];
#pragma unroll
    for (int i = 1; i < np; ++i) f = f + src[p + i * stride] * af[i];
    return qdiv<T, DIV>(f, h, ih);
  }
  T b = src[p - (np - 1) * stride] * ab[0];
#pragma unroll
  for (int i = 1; i < np; ++i) b = b + src[p + (i - (np - 1)) * stride] * ab[i];
  if (g >= n - np + 1) return qdiv<T, DIV>(b, h, ih);
  T f = src[p] * af[0];
#pragma unroll
  for (int i = 1; i < np; ++i) f = f + src[p + i * stride] * af[i];
  return qdiv<T, DIV>(b + f, h2, ih2);
}

// D at a point whose stencil rows are all interior: the third case of
// dstencil, the same operations in the same order.
template <typename T, int DP, int DIV>
__device__ __forceinline__ T dstencil_in(const T *src, int p, int stride, int dp, const T *ab,
                                         const T *af, T h2, T ih2) {
  const int np = DP > 0 ? DP : dp;
  T b = src[p - (np - 1) * stride] * ab[0];
#pragma unroll
  for (int i = 1; i < np; ++i) b = b + src[p + (i - (np - 1)) * stride] * ab[i];
  T f = src[p] * af[0];
#pragma unroll
  for (int i = 1; i < np; ++i) f = f + src[p + i * stride] * af[i];
  return qdiv<T, DIV>(b + f, h2, ih2);
}

// Transpose of D at global index g: sum over the rows k whose stencil reads
// g, each weighted by its coefficient over h (one-sided rows) or 2h; c holds
// ab/(2h), af/(2h), ab/h, af/h at a distance of cs values.
template <typename T, int DP>
__device__ __forceinline__ T dtstencil(const T *src, int p, int stride, int g, int n, int dp,
                                       const T *c, int cs) {
  const int np = DP > 0 ? DP : dp;
  const T *cb2 = c, *cf2 = c + cs, *cb1 = c + 2 * cs, *cf1 = c + 3 * cs;
  T acc = T(0);
  // backward labels: row k = g + (np-1) - i reads g with ab[i]; front rows
  // (k < np-1) use forward labels only
#pragma unroll
  for (int i = 0; i < np; ++i) {
    const int k = g + (np - 1) - i;
    if (k < np - 1 || k >= n) continue;
    acc += src[p + ((np - 1) - i) * stride] * ((k >= n - np + 1) ? cb1[i] : cb2[i]);
  }
  // forward labels: row k = g - i reads g with af[i]; back rows use
  // backward labels only
#pragma unroll
  for (int i = 0; i < np; ++i) {
    const int k = g - i;
    if (k < 0 || (k >= np - 1 && k >= n - np + 1)) continue;
    acc += src[p - i * stride] * ((k < np - 1) ? cf1[i] : cf2[i]);
  }
  return acc;
}

// dtstencil where every row k that reads g is interior.
template <typename T, int DP>
__device__ __forceinline__ T dtstencil_in(const T *src, int p, int stride, int dp, const T *c,
                                          int cs) {
  const int np = DP > 0 ? DP : dp;
  const T *cb2 = c, *cf2 = c + cs;
  T acc = src[p + (np - 1) * stride] * cb2[0];
#pragma unroll
  for (int i = 1; i < np; ++i) acc += src[p + ((np - 1) - i) * stride] * cb2[i];
#pragma unroll
  for (int i = 0; i < np; ++i) acc += src[p - i * stride] * cf2[i];
  return acc;
}

template <typename T>
__device__ __forceinline__ T mfpow(const MatArgs<T> &a, const int32_t *F, T x) {
  return F[1] == 0 ? ipow(x, F[2]) : dpow(x, a.reals[F[3]]);
}
template <typename T>
__device__ __forceinline__ T mdfpow(const MatArgs<T> &a, const int32_t *F, T x) {
  if (F[1] == 0) {
    const int p = F[2];
    return p == 0 ? T(0) : T(p) * ipow(x, p - 1);
  }
  const T p = a.reals[F[3]];
  return p * dpow(x, p - T(1));
}

// x ** p for a factor record {pow kind, pow int, pow real index} of the
// flattened program, and its derivative
template <typename T>
__device__ __forceinline__ T ppow(const int32_t *Fr, const T *pr, T x) {
  const int kind = __builtin_amdgcn_readfirstlane(Fr[0]);
  const int pi = __builtin_amdgcn_readfirstlane(Fr[1]);
  return kind == 0 ? ipow(x, pi) : dpow(x, pr[__builtin_amdgcn_readfirstlane(Fr[2])]);
}
template <typename T>
__device__ __forceinline__ T pdpow(const int32_t *Fr, const T *pr, T x) {
  const int kind = __builtin_amdgcn_readfirstlane(Fr[0]);
  if (kind == 0) {
    const int p = __builtin_amdgcn_readfirstlane(Fr[1]);
    return p == 0 ? T(0) : T(p) * ipow(x, p - 1);
  }
  const T p = pr[__builtin_amdgcn_readfirstlane(Fr[2])];
  return p * dpow(x, p - T(1));
}

// Coefficient of a term record of the flattened program: a scalar of the
// program reals (LDS) or the grid array in global memory at point gp.  The
// kind is uniform; reading it through readfirstlane keeps the two loads on
// separate scalar branches instead of one flat load.
template <typename T>
__device__ __forceinline__ T term_coeff(const int32_t *Tr, const T *pr, const T *data, int64_t gp) {
  const int kind = __builtin_amdgcn_readfirstlane(Tr[1]);
  const int ref = __builtin_amdgcn_readfirstlane(Tr[2]);
  if (kind == 0) return pr[ref];
  return data[ref + gp];
}

// Calls fn(i0, i1) for every point of the box [lo0, hi0) x [lo1, hi1) of an
// E0 x E1 tile.  A tile row is covered by 2^tpr_log consecutive threads
// (coalesced along axis 1), 256 >> tpr_log rows per sweep.
template <typename Fn>
__device__ __forceinline__ void for_box(int E1, int tpr_log, int lo0, int hi0, int lo1, int hi1,
                                        Fn fn) {
  const int tpr = 1 << tpr_log;
  const int c0 = threadIdx.x & (tpr - 1);
  const int r0 = lo0 + (int)(threadIdx.x >> tpr_log), rstep = MAT_THREADS >> tpr_log;
  if (E1 == tpr && tpr >= 64) {
    // one column per thread: the column test is made once, and since a row
    // spans whole wavefronts every lane of a wavefront has the same rows,
    // which readfirstlane tells the compiler (a scalar loop)
    const int r0u = __builtin_amdgcn_readfirstlane(r0);
    if (c0 >= lo1 && c0 < hi1)
      for (int i0 = r0u; i0 < hi0; i0 += rstep) fn(i0, c0);
  } else if (E1 == tpr) {
    // rows narrower than a wavefront: one column per thread, several rows
    // per wavefront
    if (c0 >= lo1 && c0 < hi1)
      for (int i0 = r0; i0 < hi0; i0 += rstep) fn(i0, c0);
  } else {
    for (int i0 = r0; i0 < hi0; i0 += rstep)
      for (int i1 = c0; i1 < E1; i1 += tpr)
        if (i1 >= lo1 && i1 < hi1) fn(i0, i1);
  }
}

// ---------------------------------------------------------------------------
// tiled residual + loss + adjoint
// ---------------------------------------------------------------------------
// LDS: U[n_eq][S] | F[n_fs][S] (factor values, then seeds) | A[S] | B[S] |
// 2 x 256 reduction words | coefficient table | program reals | program ints,
// S = E0 * E1.  The adjoint G[n_eq][S] takes U's place: u is last read by the
// forward chains.  DP: number of derivative points, 0 = any (read from dp).
// INT: the tile with its halo lies inside the grid, so every point exists and
// every stencil row the tile evaluates or transposes is an interior row.
template <typename T, int DP, int DIV, bool ADJ, bool INT>
__device__ __forceinline__ void mat_tile_body(const MatArgs<T> &a, T *smem, int t0, int t1) {
  const int E0 = a.E0, E1 = a.E1, tl = a.tpr_log;
  const int S = E0 * E1;
  const int R0 = a.R0, R1 = a.R1, T0 = a.T0, T1 = a.T1;
  const int n0 = a.n0, n1 = a.n1, dp = a.dp;
  const int64_t N = a.N;
  const int base0 = t0 * T0 - 2 * R0, base1 = t1 * T1 - 2 * R1;
  T *U = smem;
  T *F = U + a.n_eq * S;
  T *A = F + a.n_fs * S;
  T *B = A + S;
  T *G = U;
  T *red = B + S;
  const int tid = threadIdx.x;
  auto in_grid = [&](int g0, int g1) {
    return INT || (g0 >= 0 && g0 < n0 && g1 >= 0 && g1 < n1);
  };
  // stencil coefficients: registers when DP is fixed, else an LDS copy
  constexpr int CN = DP > 0 ? DP : 1;
  T abr[CN], afr[CN], t0r[4 * CN], t1r[4 * CN];
  const T *ab, *af, *tc0, *tc1;
  int cs;
  if constexpr (DP > 0) {
#pragma unroll
    for (int i = 0; i < DP; ++i) {
      abr[i] = a.coef[i];
      afr[i] = a.coef[MMAX_DP + i];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        t0r[k * DP + i] = a.coef[MCOEF_AX + k * MMAX_DP + i];
        t1r[k * DP + i] = a.coef[MCOEF_AX + 4 * MMAX_DP + k * MMAX_DP + i];
      }
    }
    ab = abr;
    af = afr;
    tc0 = t0r;
    tc1 = t1r;
    cs = DP;
  } else {
    T *cl = red + 2 * MAT_THREADS;
    ab = cl;
    af = cl + MMAX_DP;
    tc0 = cl + MCOEF_AX;
    tc1 = cl + MCOEF_AX + 4 * MMAX_DP;
    cs = MMAX_DP;
  }
  // coefficient table (used when DP = 0) and residual program in LDS
  T *cl = red + 2 * MAT_THREADS;
  T *PR = cl + MCOEF_N;
  int32_t *PI = reinterpret_cast<int32_t *>(PR + a.n_preal);
  for (int k = tid; k < MCOEF_N; k += MAT_THREADS) cl[k] = a.coef[k];
  for (int k = tid; k < a.n_preal; k += MAT_THREADS) PR[k] = a.preal[k];
  for (int k = tid; k < a.n_pint; k += MAT_THREADS) PI[k] = a.pint[k];

  // u with halo (zero outside the grid)
  for (int v = 0; v < a.n_eq; ++v)
    for_box(E1, tl, 0, E0, 0, E1, [&](int i0, int i1) {
      const int g0 = base0 + i0, g1 = base1 + i1;
      T val = T(0);
      if (in_grid(g0, g1)) val = a.u[v * N + (int64_t)g0 * n1 + g1];
      U[v * S + i0 * E1 + i1] = val;
    });
  __syncthreads();

  // forward chains: the valid box shrinks by dp-1 per application
  for (int fs = 0; fs < a.n_fs; ++fs) {
    const int32_t *Rf = a.fs + fs * MFS_INTS;
    const int nax = Rf[1];
    const T *src = U + Rf[0] * S;
    int lo0 = 0, hi0 = E0, lo1 = 0, hi1 = E1;
    if (nax == 0)
      for_box(E1, tl, 0, E0, 0, E1, [&](int i0, int i1) {
        F[fs * S + i0 * E1 + i1] = src[i0 * E1 + i1];
      });
    for (int j = 0; j < nax; ++j) {
      const int ax = Rf[2 + j];
      T *dst = (j == nax - 1) ? F + fs * S : ((j & 1) ? B : A);
      if (ax == 0) {
        lo0 += dp - 1;
        hi0 -= dp - 1;
      } else {
        lo1 += dp - 1;
        hi1 -= dp - 1;
      }
      if (INT) {
        const int st = ax == 0 ? E1 : 1;
        const T h2 = ax == 0 ? a.h20 : a.h21, ih2 = ax == 0 ? a.ih20 : a.ih21;
        for_box(E1, tl, lo0, hi0, lo1, hi1, [&](int i0, int i1) {
          const int p = i0 * E1 + i1;
          dst[p] = dstencil_in<T, DP, DIV>(src, p, st, dp, ab, af, h2, ih2);
        });
      } else if (ax == 0) {
        for_box(E1, tl, lo0, hi0, lo1, hi1, [&](int i0, int i1) {
          const int g0 = base0 + i0, g1 = base1 + i1, p = i0 * E1 + i1;
          dst[p] = in_grid(g0, g1)
                       ? dstencil<T, DP, DIV>(src, p, E1, g0, n0, dp, ab, af, a.h0, a.h20, a.ih0, a.ih20)
                       : T(0);
        });
      } else {
        for_box(E1, tl, lo0, hi0, lo1, hi1, [&](int i0, int i1) {
          const int g0 = base0 + i0, g1 = base1 + i1, p = i0 * E1 + i1;
          dst[p] = in_grid(g0, g1)
                       ? dstencil<T, DP, DIV>(src, p, 1, g1, n1, dp, ab, af, a.h1, a.h21, a.ih1, a.ih21)
                       : T(0);
        });
      }
      __syncthreads();
      src = dst;
    }
  }
  __syncthreads();
  // the adjoint overwrites u (read by the forward chains only); the barrier
  // before the transposed chains orders these stores before their sums
  if (ADJ)
    for (int v = 0; v < a.n_eq; ++v)
      for_box(E1, tl, 0, E0, 0, E1, [&](int i0, int i1) { G[v * S + i0 * E1 + i1] = T(0); });

  // residual, loss and seeds on the tile plus R; each point is read and
  // written by one thread, so a term's seeds overwrite its factor values
  T lsum = T(0), lnsum = T(0);
  const T dl = ADJ ? *a.dloss : T(0);
  for_box(E1, tl, R0, 3 * R0 + T0, R1, 3 * R1 + T1, [&](int i0, int i1) {
    const int g0 = base0 + i0, g1 = base1 + i1, p = i0 * E1 + i1;
    const bool ing = in_grid(g0, g1);
    const bool in_tile = i0 >= 2 * R0 && i0 < 2 * R0 + T0 && i1 >= 2 * R1 && i1 < 2 * R1 + T1;
    const int64_t gp = ing ? (int64_t)g0 * n1 + g1 : 0;
    for (int eq = 0; eq < a.n_eq; ++eq) {
      // program words are the same for every lane: readfirstlane makes the
      // loops scalar
      const int n_terms = __builtin_amdgcn_readfirstlane(PI[2 * eq]);
      const int term0 = __builtin_amdgcn_readfirstlane(PI[2 * eq + 1]);
      const T lam_n = PR[eq];
      T r = T(0);
      for (int t = 0; t < n_terms; ++t) {
        const int32_t *Tr = PI + a.pterm0 + 4 * (term0 + t);
        const int nf = __builtin_amdgcn_readfirstlane(Tr[0]);
        const int fs = __builtin_amdgcn_readfirstlane(Tr[3]);
        T der = T(1);
        for (int f = 0; f < nf; ++f)
          der = der * ppow(PI + a.pfac0 + 3 * (fs + f), PR, F[(fs + f) * S + p]);
        const T c = term_coeff(Tr, PR, a.data, gp);
        const T tv = c * der;
        r = (t == 0) ? tv : r + tv;
      }
      if (ing && in_tile) {
        lsum += lam_n * (r * r);
        lnsum += r * r;
        if (a.op) a.op[eq * N + gp] = r;
      }
      if (ADJ) {
        const T g = T(2) * lam_n * r * dl;
        for (int t = 0; t < n_terms; ++t) {
          const int32_t *Tr = PI + a.pterm0 + 4 * (term0 + t);
          const int nf = __builtin_amdgcn_readfirstlane(Tr[0]);
          const int fs = __builtin_amdgcn_readfirstlane(Tr[3]);
          const T c = term_coeff(Tr, PR, a.data, gp);
          T fv[MMAX_TF], pw[MMAX_TF];
#pragma unroll
          for (int f = 0; f < MMAX_TF; ++f) {
            fv[f] = f < nf ? F[(fs + f) * S + p] : T(0);
            pw[f] = f < nf ? ppow(PI + a.pfac0 + 3 * (fs + f), PR, fv[f]) : T(1);
          }
#pragma unroll
          for (int f = 0; f < MMAX_TF; ++f) {
            if (f < nf) {
              T d = c * g * pdpow(PI + a.pfac0 + 3 * (fs + f), PR, fv[f]);
#pragma unroll
              for (int f2 = 0; f2 < MMAX_TF; ++f2)
                if (f2 != f && f2 < nf) d *= pw[f2];
              F[(fs + f) * S + p] = ing ? d : T(0);
            }
          }
        }
      }
    }
  });

  if (ADJ) {
    __syncthreads();
    // transposed chains from the tile plus R back to the tile; the last
    // application accumulates into G
    for (int fs = 0; fs < a.n_fs; ++fs) {
      const int32_t *Rf = a.fs + fs * MFS_INTS;
      const int nax = Rf[1];
      T *Gv = G + Rf[0] * S;
      const T *src = F + fs * S;
      int lo0 = R0, hi0 = 3 * R0 + T0, lo1 = R1, hi1 = 3 * R1 + T1;
      if (nax == 0)
        for_box(E1, tl, lo0, hi0, lo1, hi1, [&](int i0, int i1) {
          Gv[i0 * E1 + i1] += src[i0 * E1 + i1];
        });
      for (int j = nax - 1; j >= 0; --j) {
        const int ax = Rf[2 + j];
        T *dst = ((nax - 1 - j) & 1) ? B : A;
        const bool last = j == 0;
        if (ax == 0) {
          lo0 += dp - 1;
          hi0 -= dp - 1;
        } else {
          lo1 += dp - 1;
          hi1 -= dp - 1;
        }
        if (INT && ax == 0) {
          for_box(E1, tl, lo0, hi0, lo1, hi1, [&](int i0, int i1) {
            const int p = i0 * E1 + i1;
            const T val = dtstencil_in<T, DP>(src, p, E1, dp, tc0, cs);
            if (last)
              Gv[p] += val;
            else
              dst[p] = val;
          });
        } else if (INT) {
          for_box(E1, tl, lo0, hi0, lo1, hi1, [&](int i0, int i1) {
            const int p = i0 * E1 + i1;
            const T val = dtstencil_in<T, DP>(src, p, 1, dp, tc1, cs);
            if (last)
              Gv[p] += val;
            else
              dst[p] = val;
          });
        } else if (ax == 0)
          for_box(E1, tl, lo0, hi0, lo1, hi1, [&](int i0, int i1) {
            const int g0 = base0 + i0, g1 = base1 + i1, p = i0 * E1 + i1;
            const T val = in_grid(g0, g1) ? dtstencil<T, DP>(src, p, E1, g0, n0, dp, tc0, cs) : T(0);
            if (last)
              Gv[p] += val;
            else
              dst[p] = val;
          });
        else
          for_box(E1, tl, lo0, hi0, lo1, hi1, [&](int i0, int i1) {
            const int g0 = base0 + i0, g1 = base1 + i1, p = i0 * E1 + i1;
            const T val = in_grid(g0, g1) ? dtstencil<T, DP>(src, p, 1, g1, n1, dp, tc1, cs) : T(0);
            if (last)
              Gv[p] += val;
            else
              dst[p] = val;
          });
        __syncthreads();
        src = dst;
      }
      if (nax == 0) __syncthreads();
    }
    for (int v = 0; v < a.n_eq; ++v)
      for_box(E1, tl, 2 * R0, 2 * R0 + T0, 2 * R1, 2 * R1 + T1, [&](int i0, int i1) {
        const int g0 = base0 + i0, g1 = base1 + i1;
        if (INT || (g0 < n0 && g1 < n1))
          a.grad[v * N + (int64_t)g0 * n1 + g1] = G[v * S + i0 * E1 + i1];
      });
  }

  // deterministic workgroup sum of the loss partials (the normalized sum
  // divided by N once per thread)
  red[tid] = lsum;
  red[MAT_THREADS + tid] = lnsum / T(N);
  __syncthreads();
  for (int st = MAT_THREADS / 2; st > 0; st >>= 1) {
    if (tid < st) {
      red[tid] += red[tid + st];
      red[MAT_THREADS + tid] += red[MAT_THREADS + tid + st];
    }
    __syncthreads();
  }
  if (tid == 0) {
    a.lp[2 * blockIdx.x] = red[0];
    a.lp[2 * blockIdx.x + 1] = red[MAT_THREADS];
  }
}

// Workgroups are handed to the 8 XCDs round-robin (block b to XCD b % 8),
// each XCD with its own L2.  Block b takes tile (b % 8) * per + b / 8 for
// the first 8 * per blocks, so every XCD works through one contiguous band
// of tiles and the halo rows its neighbouring tiles share stay in its L2.
__device__ __forceinline__ int xcd_band_tile(int b, int n) {
  const int full = n & ~7;
  if (b >= full) return b;
  const int per = full >> 3;
  return (b & 7) * per + (b >> 3);
}

template <typename T, int DP, int DIV, bool ADJ>
__global__ __launch_bounds__(MAT_THREADS) void k_mat_tile(MatArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *smem = reinterpret_cast<T *>(smem_raw);
  const int tile = a.tile_list ? a.tile_list[blockIdx.x] : xcd_band_tile((int)blockIdx.x, (int)gridDim.x);
  const int t0 = tile / a.tiles1, t1 = tile - (tile / a.tiles1) * a.tiles1;
  const int base0 = t0 * a.T0 - 2 * a.R0, base1 = t1 * a.T1 - 2 * a.R1;
  const bool interior = base0 >= 0 && base0 + a.E0 <= a.n0 && base1 >= 0 && base1 + a.E1 <= a.n1;
  if (interior)
    mat_tile_body<T, DP, DIV, ADJ, true>(a, smem, t0, t1);
  else
    mat_tile_body<T, DP, DIV, ADJ, false>(a, smem, t0, t1);
}

// ---------------------------------------------------------------------------
// boundary entries: pointwise derivative chains straight from HBM
// ---------------------------------------------------------------------------
template <typename T, int DIV>
__device__ T point_chain(const MatArgs<T> &a, const T *plane, int g0, int g1, const int *axes,
                         int depth) {
  // D_{axes[depth-1]} ... D_{axes[0]} u at (g0, g1), evaluated recursively
  if (depth == 0) return plane[(int64_t)g0 * a.n1 + g1];
  const int ax = axes[depth - 1];
  const int g = ax == 0 ? g0 : g1;
  const int n = ax == 0 ? a.n0 : a.n1;
  const int dp = a.dp;
  auto at = [&](int off) {
    const int h0 = ax == 0 ? g0 + off : g0, h1 = ax == 0 ? g1 : g1 + off;
    return point_chain<T, DIV>(a, plane, h0, h1, axes, depth - 1);
  };
  const T *ab = a.coef, *af = a.coef + MMAX_DP;
  const T h = ax == 0 ? a.h0 : a.h1, ih = ax == 0 ? a.ih0 : a.ih1;
  const T h2 = ax == 0 ? a.h20 : a.h21, ih2 = ax == 0 ? a.ih20 : a.ih21;
  if (g < dp - 1) {
    T f = at(0) * af[0];
    for (int i = 1; i < dp; ++i) f = f + at(i) * af[i];
    return qdiv<T, DIV>(f, h, ih);
  }
  T b = at(-(dp - 1)) * ab[0];
  for (int i = 1; i < dp; ++i) b = b + at(i - (dp - 1)) * ab[i];
  if (g >= n - dp + 1) return qdiv<T, DIV>(b, h, ih);
  T f = at(0) * af[0];
  for (int i = 1; i < dp; ++i) f = f + at(i) * af[i];
  return qdiv<T, DIV>(b + f, h2, ih2);
}

// Boundary entries: value, loss partials and (adj) seeds.  ND (full-grid
// plans): the factor values are the chain fields at the entry's positions
// (factor k of the entry's operator in field bc_slot[e] + k); otherwise the
// chains are evaluated at the point (2-D, at most 3 applications).
template <typename T, int DIV, bool ND>
__global__ __launch_bounds__(MAT_THREADS) void k_mat_bc(MatArgs<T> a, int adj) {
  __shared__ T red[2 * MAT_THREADS];
  const int tid = threadIdx.x;
  const int e = blockIdx.x * MAT_THREADS + tid;
  T lsum = T(0), lnsum = T(0);
  if (e < a.n_bc) {
    const int32_t *Bc = a.ints + a.bc_i + e * TDX_MBC_INTS;
    const int n_sets = Bc[0], pos_i = Bc[1], n_terms = Bc[2], term_i = Bc[3], var = Bc[4];
    const T target = a.data[Bc[5]];
    const T lam_d = a.reals[Bc[6]], invden = a.reals[Bc[7]];
    const int seed_i = Bc[8];
    const int slot0 = ND ? a.bc_slot[e] : 0;
    // value of factor record Fc (the k-th factor of the operator) at pos
    auto fval = [&](const int32_t *Fc, int k, int pos) -> T {
      if constexpr (ND) {
        return a.fields[(int64_t)(slot0 + k) * a.N + pos];
      } else {
        const int g0 = pos / a.n1, g1 = pos - (pos / a.n1) * a.n1;
        return point_chain<T, DIV>(a, a.u + Fc[0] * a.N, g0, g1, Fc + 5, Fc[4]);
      }
    };
    T val = T(0);
    for (int s = 0; s < n_sets; ++s) {
      const int pos = a.ints[pos_i + s];
      T v;
      if (n_terms == 0) {
        v = a.u[var * a.N + pos];
      } else {
        T r = T(0);
        int k = 0;
        for (int t = 0; t < n_terms; ++t) {
          const int32_t *Tm = a.ints + term_i + t * TDX_MTERM_INTS;
          T der = T(1);
          for (int f = 0; f < Tm[0]; ++f, ++k) {
            const int32_t *Fc = a.ints + Tm[1] + f * TDX_MFACTOR_INTS;
            der = der * mfpow(a, Fc, fval(Fc, k, pos));
          }
          const T c = Tm[2] == 0 ? a.reals[Tm[3]] : a.data[Tm[3] + pos];
          const T tv = c * der;
          r = (t == 0) ? tv : r + tv;
        }
        v = r;
        if (Bc[9] > 0) {
          // Robin: alpha u[var] + (0 + beta_1 r) + beta_2 r ... in the
          // reference's order (tedeous/eval.py:372-387)
          const int32_t *Rb = a.ints + Bc[9] - 1;
          const T *R = a.reals + Rb[1];
          T dt = T(0);
          for (int q = 0; q < Rb[0]; ++q) dt = dt + R[1 + q] * r;
          v = R[0] * a.u[var * a.N + pos] + dt;
        }
      }
      val = (s == 0) ? v : val - v;
    }
    if (a.bval) a.bval[e] = val;
    const T diff = val - target;
    lsum = lam_d * (diff * diff);
    lnsum = invden * (diff * diff);
    if (adj) {
      const T dr = T(2) * lam_d * diff * (*a.dloss);
      int sidx = seed_i;
      T robin_w = T(1);
      if (Bc[9] > 0) {
        // d value / d u[var] = alpha (its own seed, first); the operator's
        // seeds carry the sum of the betas
        const int32_t *Rb = a.ints + Bc[9] - 1;
        const T *R = a.reals + Rb[1];
        a.bc_seeds[sidx++] = dr * R[0];
        robin_w = T(0);
        for (int q = 0; q < Rb[0]; ++q) robin_w = robin_w + R[1 + q];
      }
      for (int s = 0; s < n_sets; ++s) {
        const T sg = (s == 0 ? dr : -dr) * robin_w;
        if (n_terms == 0) {
          a.bc_seeds[sidx++] = sg;
          continue;
        }
        const int pos = a.ints[pos_i + s];
        int k0 = 0;
        for (int t = 0; t < n_terms; ++t) {
          const int32_t *Tm = a.ints + term_i + t * TDX_MTERM_INTS;
          const int nf = Tm[0];
          const T c = Tm[2] == 0 ? a.reals[Tm[3]] : a.data[Tm[3] + pos];
          for (int f = 0; f < nf; ++f) {
            const int32_t *Fc = a.ints + Tm[1] + f * TDX_MFACTOR_INTS;
            T d = c * sg * mdfpow(a, Fc, fval(Fc, k0 + f, pos));
            for (int f2 = 0; f2 < nf; ++f2) {
              if (f2 == f) continue;
              const int32_t *F2 = a.ints + Tm[1] + f2 * TDX_MFACTOR_INTS;
              d *= mfpow(a, F2, fval(F2, k0 + f2, pos));
            }
            a.bc_seeds[sidx++] = d;
          }
          k0 += nf;
        }
      }
    }
  }
  red[tid] = lsum;
  red[MAT_THREADS + tid] = lnsum;
  __syncthreads();
  for (int st = MAT_THREADS / 2; st > 0; st >>= 1) {
    if (tid < st) {
      red[tid] += red[tid + st];
      red[MAT_THREADS + tid] += red[MAT_THREADS + tid + st];
    }
    __syncthreads();
  }
  if (tid == 0) {
    a.lp[2 * (a.n_tile_wg + blockIdx.x)] = red[0];
    a.lp[2 * (a.n_tile_wg + blockIdx.x) + 1] = red[MAT_THREADS];
  }
}

// gradient of the boundary terms: one thread per touched grid point, fixed
// order over its contributions (no atomics)
template <typename T>
__global__ void k_mat_bc_adj(MatArgs<T> a) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= a.n_touch) return;
  const int32_t *ptr = a.ints + a.bct_ptr_i;
  const int64_t pos = a.ints[a.bct_pos_i + j];
  T s = T(0);
  for (int k = ptr[j]; k < ptr[j + 1]; ++k)
    s += a.bc_seeds[a.ints[a.bct_seed_i + k]] * a.reals[a.ints[a.bct_w_i + k]];
  a.grad[pos] += s;
}

// Sum of n (loss, normalized loss) pairs: workgroup b of the grid adds the
// pairs b, b + G, b + 2G, ... (G = grid size) in a fixed order and writes one
// pair to out[b]; a second launch with one workgroup adds those.
template <typename T>
__device__ __forceinline__ void mat_loss_block(const T *lp, int n, T *loss, int blk, int nblk) {
  __shared__ T red[2 * 256];
  T s0 = T(0), s1 = T(0);
  for (int k = blk * 256 + threadIdx.x; k < n; k += 256 * nblk) {
    s0 += lp[2 * k];
    s1 += lp[2 * k + 1];
  }
  red[threadIdx.x] = s0;
  red[256 + threadIdx.x] = s1;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      red[threadIdx.x] += red[threadIdx.x + st];
      red[256 + threadIdx.x] += red[256 + threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    loss[2 * blk] = red[0];
    loss[2 * blk + 1] = red[256];
  }
}

template <typename T>
__global__ void k_mat_loss(const T *lp, int n, T *loss) {
  mat_loss_block(lp, n, loss, blockIdx.x, gridDim.x);
}

// The fused step's set-up in one launch: the adjoint's unit seed scale, the
// Adam scalars of step t = *step + 1, computed once for every kernel of the
// step ({1-b1, 1-b2, b2, lr/(1-b1^t), sqrt(1-b2^t), eps}), and the step
// counter advanced to t (no later kernel of the fused step reads it).
template <typename T>
__global__ void k_mat_step_prep(T *one, int64_t *step, AdamHyper h, T *out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    *one = T(1);
    const int64_t t = *step + 1;
    const AdamScalars<T> sc = adam_scalars<T>(t, h);
    out[0] = sc.w1;
    out[1] = sc.w2;
    out[2] = sc.b2;
    out[3] = sc.step_size;
    out[4] = sc.bc2s;
    out[5] = sc.eps;
    *step = t;
  }
}

// Adam update u_out = step(u, grad, m, v) of every grid point outside the
// rectangle [r_lo, r_hi) x [c_lo, c_hi) of each plane (the points the
// generated kernel does not update; an empty rectangle covers the grid).
template <typename T>
__global__ __launch_bounds__(256) void k_mat_frame_adam(MatArgs<T> a, int r_lo, int r_hi,
                                                        int c_lo, int c_hi, const T *lp,
                                                        int n_lp, T *loss) {
  // with loss: the last workgroup forms the loss pair from the n_lp
  // partial pairs (k_mat_loss in one workgroup) instead of updating points
  if (loss && blockIdx.x == gridDim.x - 1) {
    mat_loss_block(lp, n_lp, loss, 0, 1);
    return;
  }
  const int n_upd = gridDim.x - (loss ? 1 : 0);
  const int n0 = a.n0, n1 = a.n1;
  const int64_t top = (int64_t)r_lo * n1, bot = (int64_t)(n0 - r_hi) * n1;
  const int wl = c_lo, wm = c_lo + (n1 - c_hi);
  const int64_t per = top + bot + (int64_t)(r_hi - r_lo) * wm;
  const int64_t tot = per * a.n_eq;
  AdamScalars<T> sc;
  sc.w1 = a.adam[0];
  sc.w2 = a.adam[1];
  sc.b2 = a.adam[2];
  sc.step_size = a.adam[3];
  sc.bc2s = a.adam[4];
  sc.eps = a.adam[5];
  const int64_t stride = (int64_t)n_upd * blockDim.x;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < tot; idx += stride) {
    const int64_t e = idx / per, k = idx - e * per;
    int64_t row, col;
    if (k < top) {
      row = k / n1;
      col = k - row * n1;
    } else if (k < top + bot) {
      const int64_t k2 = k - top;
      row = r_hi + k2 / n1;
      col = k2 - (k2 / n1) * n1;
    } else {
      const int64_t k3 = k - top - bot;
      const int64_t j = k3 % wm;
      row = r_lo + k3 / wm;
      col = j < wl ? j : c_hi + (j - wl);
    }
    const int64_t gi = e * a.N + row * n1 + col;
    T p = a.u[gi], m = a.m[gi], v = a.v[gi];
    adam_apply(p, a.grad[gi], m, v, sc);
    a.u_out[gi] = p;
    a.m[gi] = m;
    a.v[gi] = v;
  }
}

// ---------------------------------------------------------------------------
// full-grid plans: 3-D grids, chains of more than 3 applications, reaches
// beyond an LDS tile.  Every factor's chain is a sequence of passes over the
// whole plane (one thread per point, coalesced along the last axis), the
// residual program runs per point on the chain fields, and the adjoint runs
// the transposed passes back, accumulating into the gradient.  Bounded by
// HBM: each pass reads and writes the plane once.
// ---------------------------------------------------------------------------
template <typename T, int DIV>
__global__ __launch_bounds__(256) void k_mat_nd_d(const T *src, T *dst, int N, int stride,
                                                  int n, int dp, const T *coef, T h, T h2, T ih,
                                                  T ih2) {
  for (int p = blockIdx.x * 256 + threadIdx.x; p < N; p += gridDim.x * 256) {
    const int g = (p / stride) % n;
    dst[p] = dstencil<T, 0, DIV>(src, p, stride, g, n, dp, coef, coef + MMAX_DP, h, h2, ih, ih2);
  }
}

// dst = D^T src along one axis (acc: dst += D^T src); c is the axis' block of
// the coefficient table (ab/(2h), af/(2h), ab/h, af/h)
template <typename T>
__global__ __launch_bounds__(256) void k_mat_nd_dt(const T *src, T *dst, int N, int stride,
                                                   int n, int dp, const T *c, int acc) {
  for (int p = blockIdx.x * 256 + threadIdx.x; p < N; p += gridDim.x * 256) {
    const int g = (p / stride) % n;
    const T v = dtstencil<T, 0>(src, p, stride, g, n, dp, c, MMAX_DP);
    dst[p] = acc ? dst[p] + v : v;
  }
}

// dst = src (acc: dst += src)
template <typename T>
__global__ __launch_bounds__(256) void k_mat_nd_copy(const T *src, T *dst, int64_t n, int acc) {
  for (int64_t p = blockIdx.x * 256 + threadIdx.x; p < n; p += (int64_t)gridDim.x * 256)
    dst[p] = acc ? dst[p] + src[p] : src[p];
}

template <typename T>
__global__ __launch_bounds__(256) void k_mat_nd_zero(T *dst, int64_t n) {
  for (int64_t p = blockIdx.x * 256 + threadIdx.x; p < n; p += (int64_t)gridDim.x * 256)
    dst[p] = T(0);
}

// Residual program per point on the chain fields F (the n_fs factor slots of
// the equations, [n_fs][N]): residual field, loss partials per workgroup
// (fixed order) and, with adj, each factor's seed written over its value
// (each slot is read by its own term only).
template <typename T>
__global__ __launch_bounds__(MAT_THREADS) void k_mat_nd_resid(MatArgs<T> a, int adj) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *PR = reinterpret_cast<T *>(smem_raw);
  int32_t *PI = reinterpret_cast<int32_t *>(PR + a.n_preal);
  __shared__ T red[2 * MAT_THREADS];
  const int tid = threadIdx.x;
  for (int k = tid; k < a.n_preal; k += MAT_THREADS) PR[k] = a.preal[k];
  for (int k = tid; k < a.n_pint; k += MAT_THREADS) PI[k] = a.pint[k];
  __syncthreads();
  T *F = a.fields;
  const int64_t N = a.N;
  T lsum = T(0), lnsum = T(0);
  const T dl = adj ? *a.dloss : T(0);
  for (int64_t p = blockIdx.x * (int64_t)MAT_THREADS + tid; p < N;
       p += (int64_t)gridDim.x * MAT_THREADS) {
    for (int eq = 0; eq < a.n_eq; ++eq) {
      const int n_terms = __builtin_amdgcn_readfirstlane(PI[2 * eq]);
      const int term0 = __builtin_amdgcn_readfirstlane(PI[2 * eq + 1]);
      const T lam_n = PR[eq];
      T r = T(0);
      for (int t = 0; t < n_terms; ++t) {
        const int32_t *Tr = PI + a.pterm0 + 4 * (term0 + t);
        const int nf = __builtin_amdgcn_readfirstlane(Tr[0]);
        const int fs = __builtin_amdgcn_readfirstlane(Tr[3]);
        T der = T(1);
        for (int f = 0; f < nf; ++f)
          der = der * ppow(PI + a.pfac0 + 3 * (fs + f), PR, F[(fs + f) * N + p]);
        const T tv = term_coeff(Tr, PR, a.data, p) * der;
        r = (t == 0) ? tv : r + tv;
      }
      lsum += lam_n * (r * r);
      lnsum += r * r;
      if (a.op) a.op[eq * N + p] = r;
      if (adj) {
        const T g = T(2) * lam_n * r * dl;
        for (int t = 0; t < n_terms; ++t) {
          const int32_t *Tr = PI + a.pterm0 + 4 * (term0 + t);
          const int nf = __builtin_amdgcn_readfirstlane(Tr[0]);
          const int fs = __builtin_amdgcn_readfirstlane(Tr[3]);
          const T c = term_coeff(Tr, PR, a.data, p);
          T fv[MMAX_TF], pw[MMAX_TF];
#pragma unroll
          for (int f = 0; f < MMAX_TF; ++f) {
            fv[f] = f < nf ? F[(fs + f) * N + p] : T(0);
            pw[f] = f < nf ? ppow(PI + a.pfac0 + 3 * (fs + f), PR, fv[f]) : T(1);
          }
#pragma unroll
          for (int f = 0; f < MMAX_TF; ++f) {
            if (f < nf) {
              T d = c * g * pdpow(PI + a.pfac0 + 3 * (fs + f), PR, fv[f]);
#pragma unroll
              for (int f2 = 0; f2 < MMAX_TF; ++f2)
                if (f2 != f && f2 < nf) d *= pw[f2];
              F[(fs + f) * N + p] = d;
            }
          }
        }
      }
    }
  }
  red[tid] = lsum;
  red[MAT_THREADS + tid] = lnsum / T(N);
  __syncthreads();
  for (int st = MAT_THREADS / 2; st > 0; st >>= 1) {
    if (tid < st) {
      red[tid] += red[tid + st];
      red[MAT_THREADS + tid] += red[MAT_THREADS + tid + st];
    }
    __syncthreads();
  }
  if (tid == 0) {
    a.lp[2 * blockIdx.x] = red[0];
    a.lp[2 * blockIdx.x + 1] = red[MAT_THREADS];
  }
}

// Boundary seeds back onto the fields: row j of the CSR (key, seed list)
// adds its seeds, in order, to fields[key] (key < n_field_words) or to
// grad[key - n_field_words].  Keys are distinct: no two threads write one
// word.
template <typename T>
__global__ void k_mat_nd_scatter(const int32_t *ptr, const int64_t *key, const int32_t *sidx,
                                 int rows, const T *seeds, T *fields, int64_t n_field_words,
                                 T *grad) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= rows) return;
  T v = T(0);
  for (int k = ptr[j]; k < ptr[j + 1]; ++k) v += seeds[sidx[k]];
  const int64_t kk = key[j];
  if (kk < n_field_words)
    fields[kk] += v;
  else
    grad[kk - n_field_words] += v;
}

template <typename T>
__global__ void k_mat_one(T *x) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *x = T(1);
}

}  // namespace tdx

using namespace tdx;

struct tdx_mat_plan {
  tdx_mat_desc d;
  int n_fs;
  int R0, R1, T0, T1, E0, E1, tpr_log, tiles0, tiles1, n_tile_wg, n_bc_wg;
  size_t tile_lds, tile_lds_fwd;
  DevTable ints, reals, data, fs_tab, coef_tab, pint_tab, preal_tab;
  int n_pint, n_preal, pterm0, pfac0;
  int div_mode;     // division mode of the launches (0 product, 1 IEEE, 2 corrected product)
  // Specialized interior kernel: output rows [jr_lo, jr_hi) x columns
  // [jc_lo, jc_hi), the union of the interior tiles; the general kernel
  // takes the n_gen_wg tiles of gen_list (every tile when jit_on is 0).
  int jit_on;
  JitKernel jit;
  int jr_lo, jr_hi, jc_lo, jc_hi, j_nstrips;
  // per kernel variant (0 forward, 1 adjoint, 2 fused step): rows per band,
  // bands, wavefronts
  int j_hb[3], j_nbands[3], j_waves[3], j_waves_max;
  int jit_fuse;     // no boundary adjoint touches the rectangle: the fused
                    // step updates it inside the generated kernel
  DevTable gen_list;
  int n_gen_wg;
  size_t off_lp, off_mid, off_seeds, off_one, off_grad, off_adam, ws_bytes;
  double ab[MMAX_DP], af[MMAX_DP];
  // full-grid plans (nd = 1): chain fields per factor slot -- the equations'
  // n_fs slots first, then the boundary operators' -- as {var, axes...};
  // per bc entry its first slot; the boundary seed scatter CSR (rows keyed
  // by field word, then by gradient word)
  int nd = 0;
  int64_t N = 0;
  int n_slots = 0, n_nd_wg = 0, sc_rows_f = 0, sc_rows = 0;
  std::vector<std::vector<int>> chains;
  DevTable bc_slot_tab, sc_ptr, sc_key, sc_idx;
  size_t off_fields = 0, off_tmp = 0;
};


namespace {

// True when, for the float step d and y = RN(1/d), RN(RN(x y) + (x - RN(x y) d) y)
// equals the correctly rounded x / d for every x in [1, 2) (checked one by
// one).  Scaling x or d by a power of two scales every step exactly, so the
// result then holds for all x in the range qdiv guards, given 2^-30 <= d <=
// 2^30.  The reference quotient is computed in double and rounded once to
// float, which is the correctly rounded float quotient (53 >= 2*24 + 2).
int64_t quotient_mismatches(float d, int mode, bool stop_at_first) {
  const float y = (float)(1.0 / (double)d);
  int64_t bad = 0;
  for (uint32_t m = 0; m < (1u << 23); ++m) {
    const uint32_t bits = 0x3f800000u | m;
    float x;
    std::memcpy(&x, &bits, 4);
    const float q = x * y;
    float res = q;
    if (mode == 0) {
      const float r = std::fma(-q, d, x);
      res = std::fma(r, y, q);
    }
    if (res != (float)((double)x / (double)d)) {
      ++bad;
      if (stop_at_first) break;
    }
  }
  return bad;
}

bool quotient_step_exact(float d) {
  static std::mutex mu;
  static std::map<uint32_t, bool> cache;
  uint32_t key;
  std::memcpy(&key, &d, 4);
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  const bool ok = std::isfinite(d) && d >= 0x1p-30f && d <= 0x1p30f &&
                  quotient_mismatches(d, 0, true) == 0;
  std::lock_guard<std::mutex> g(mu);
  cache[key] = ok;
  return ok;
}

template <typename T>
MatArgs<T> mat_args(const tdx_mat_plan *p, const void *u, void *ws) {
  MatArgs<T> a;
  const tdx_mat_desc &d = p->d;
  a.n_eq = d.n_eq;
  a.n0 = d.n0;
  a.n1 = d.n1;
  a.dp = d.dp;
  a.N = p->N;
  a.h0 = T(d.h[0]);
  a.h1 = T(d.h[1]);
  a.h20 = T(2.0 * d.h[0]);
  a.h21 = T(2.0 * d.h[1]);
  a.ih0 = T(1.0 / d.h[0]);
  a.ih1 = T(1.0 / d.h[1]);
  a.ih20 = T(1.0 / (2.0 * d.h[0]));
  a.ih21 = T(1.0 / (2.0 * d.h[1]));
  a.coef = (const T *)p->coef_tab.ptr;
  a.fs = (const int32_t *)p->fs_tab.ptr;
  a.pint = (const int32_t *)p->pint_tab.ptr;
  a.preal = (const T *)p->preal_tab.ptr;
  a.n_pint = p->n_pint;
  a.n_preal = p->n_preal;
  a.pterm0 = p->pterm0;
  a.pfac0 = p->pfac0;
  a.ints = (const int32_t *)p->ints.ptr;
  a.reals = (const T *)p->reals.ptr;
  a.data = (const T *)p->data.ptr;
  a.eq_i = d.eq_i;
  a.n_fs = p->n_fs;
  a.R0 = p->R0;
  a.R1 = p->R1;
  a.T0 = p->T0;
  a.T1 = p->T1;
  a.E0 = p->E0;
  a.E1 = p->E1;
  a.tpr_log = p->tpr_log;
  a.tiles1 = p->tiles1;
  a.tile_list = p->jit_on ? (const int32_t *)p->gen_list.ptr : nullptr;
  a.n_bc = d.n_bc;
  a.bc_i = d.bc_i;
  a.n_touch = d.n_touch;
  a.bct_ptr_i = d.bct_ptr_i;
  a.bct_pos_i = d.bct_pos_i;
  a.bct_seed_i = d.bct_seed_i;
  a.bct_w_i = d.bct_w_i;
  a.u = (const T *)u;
  a.grad = nullptr;
  a.op = nullptr;
  a.bval = nullptr;
  char *w = (char *)ws;
  a.lp = (T *)(w + p->off_lp);
  a.lp_mid = (T *)(w + p->off_mid);
  a.bc_seeds = (T *)(w + p->off_seeds);
  a.dloss = (const T *)(w + p->off_one);
  a.n_tile_wg = p->n_gen_wg;
  a.fields = p->nd ? (T *)(w + p->off_fields) : nullptr;
  a.bc_slot = (const int32_t *)p->bc_slot_tab.ptr;
  a.fused = 0;
  a.u_out = nullptr;
  a.m = nullptr;
  a.v = nullptr;
  a.adam = nullptr;
  return a;
}

template <typename T, int DP, int DIV>
int mat_tile_launch(const tdx_mat_plan *p, const MatArgs<T> &a, bool adj, hipStream_t s) {
  raise_lds_limit({(const void *)k_mat_tile<T, DP, DIV, true>,
                   (const void *)k_mat_tile<T, DP, DIV, false>});
  if (p->n_gen_wg == 0) return TDX_OK;
  if (adj)
    hipLaunchKernelGGL((k_mat_tile<T, DP, DIV, true>), dim3(p->n_gen_wg), dim3(MAT_THREADS),
                       p->tile_lds, s, a);
  else
    hipLaunchKernelGGL((k_mat_tile<T, DP, DIV, false>), dim3(p->n_gen_wg), dim3(MAT_THREADS),
                       p->tile_lds_fwd, s, a);
  TDX_TRY_HIP(hipGetLastError());
  return TDX_OK;
}

// The specialized interior kernel: 4 wavefronts per workgroup, one per
// (strip, band); its loss pairs follow the general tiles' and the boundary
// workgroups'.
template <typename T>
int mat_jit_launch(const tdx_mat_plan *p, const MatArgs<T> &a, bool adj, hipStream_t s) {
  JitArgs j;
  j.u = (const float *)a.u;
  j.grad = (float *)a.grad;
  j.op = (float *)a.op;
  j.lp = (float *)a.lp;
  j.data = (const float *)a.data;
  j.dloss = (const float *)a.dloss;
  j.N = a.N;
  j.n0 = a.n0;
  j.n1 = a.n1;
  j.r_lo = p->jr_lo;
  j.r_hi = p->jr_hi;
  j.c_lo = p->jc_lo;
  j.c_hi = p->jc_hi;
  const bool step = adj && a.fused && p->jit_fuse;
  const int var = step ? 2 : adj ? 1 : 0;
  j.hb = p->j_hb[var];
  j.nstrips = p->j_nstrips;
  j.nbands = p->j_nbands[var];
  j.lp_base = p->n_gen_wg + p->n_bc_wg;
  j.adj = adj ? 1 : 0;
  j.u_out = step ? (float *)a.u_out : nullptr;
  j.m = step ? (float *)a.m : nullptr;
  j.v = step ? (float *)a.v : nullptr;
  j.adam = step ? (const float *)a.adam : nullptr;
  void *params[] = {&j};
  const unsigned blocks = (unsigned)((p->j_waves[var] + 3) / 4);
  hipFunction_t fn = step ? p->jit.fn_step : adj ? p->jit.fn_adj : p->jit.fn_fwd;
  TDX_TRY_HIP(hipModuleLaunchKernel(fn, blocks, 1, 1, 256, 1, 1, 0, s, params, nullptr));
  return TDX_OK;
}

template <typename T, int DIV>
int mat_run(const tdx_mat_plan *p, MatArgs<T> a, bool adj, T *loss, hipStream_t s) {
  int rc = p->d.dp == 2   ? mat_tile_launch<T, 2, DIV>(p, a, adj, s)
           : p->d.dp == 3 ? mat_tile_launch<T, 3, DIV>(p, a, adj, s)
                          : mat_tile_launch<T, 0, DIV>(p, a, adj, s);
  if (rc) return rc;
  if (p->jit_on) {
    rc = mat_jit_launch<T>(p, a, adj, s);
    if (rc) return rc;
  }
  if (p->n_bc_wg > 0) {
    hipLaunchKernelGGL((k_mat_bc<T, DIV, false>), dim3(p->n_bc_wg), dim3(MAT_THREADS), 0, s, a,
                       adj ? 1 : 0);
    TDX_TRY_HIP(hipGetLastError());
    if (adj && p->d.n_touch > 0) {
      hipLaunchKernelGGL((k_mat_bc_adj<T>), dim3((p->d.n_touch + 255) / 256), dim3(256), 0, s,
                         a);
      TDX_TRY_HIP(hipGetLastError());
    }
  }
  // loss partial pairs of the step's kernels
  const int var = adj && a.fused && p->jit_fuse ? 2 : adj ? 1 : 0;
  const int n = p->n_gen_wg + p->n_bc_wg + (p->jit_on ? p->j_waves[var] : 0);
  const int g = std::min(256, (n + 4095) / 4096);
  if (adj && a.fused) {
    // Adam on the points the generated kernel did not update; its last
    // workgroup also forms the loss pair when one workgroup sums them
    const bool in_jit = p->jit_on && p->jit_fuse;
    const int r_lo = in_jit ? p->jr_lo : 0, r_hi = in_jit ? p->jr_hi : 0;
    const int c_lo = in_jit ? p->jc_lo : 0, c_hi = in_jit ? p->jc_hi : 0;
    const int64_t per = (int64_t)p->d.n0 * p->d.n1 - (int64_t)(r_hi - r_lo) * (c_hi - c_lo);
    const int64_t tot = per * p->d.n_eq;
    if (tot > 0) {
      T *lw = (loss && g == 1) ? loss : nullptr;
      const unsigned blocks =
          (unsigned)std::max<int64_t>(1, std::min<int64_t>((tot + 255) / 256, 16384));
      hipLaunchKernelGGL((k_mat_frame_adam<T>), dim3(blocks + (lw ? 1 : 0)), dim3(256), 0, s, a,
                         r_lo, r_hi, c_lo, c_hi, (const T *)a.lp, n, lw);
      TDX_TRY_HIP(hipGetLastError());
      if (lw) loss = nullptr;  // formed
    }
  }
  if (loss) {
    T *mid = a.lp_mid;
    if (g > 1) {
      hipLaunchKernelGGL((k_mat_loss<T>), dim3(g), dim3(256), 0, s, a.lp, n, mid);
      TDX_TRY_HIP(hipGetLastError());
      hipLaunchKernelGGL((k_mat_loss<T>), dim3(1), dim3(256), 0, s, (const T *)mid, g, loss);
    } else {
      hipLaunchKernelGGL((k_mat_loss<T>), dim3(1), dim3(256), 0, s, a.lp, n, loss);
    }
    TDX_TRY_HIP(hipGetLastError());
  }
  return TDX_OK;
}

// Axis geometry of a full-grid plan: stride, extent, step.
struct NdAxis {
  int stride, n;
  double h;
};
inline NdAxis nd_axis(const tdx_mat_plan *p, int ax) {
  const tdx_mat_desc &d = p->d;
  if (ax == 0) return {d.n1 * d.n2, d.n0, d.h[0]};
  if (ax == 1) return {d.n2, d.n1, d.h[1]};
  return {1, d.n2, d.h[2]};
}

inline unsigned nd_blocks(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192));
}

// A full-grid plan's forward (and adjoint): chain passes for every factor
// slot, the residual program, the boundary entries; with adj the scatter of
// the boundary seeds, the transposed passes into the gradient and the
// fused Adam update.
template <typename T, int DIV>
int mat_run_nd(const tdx_mat_plan *p, MatArgs<T> a, bool adj, T *loss, hipStream_t s) {
  const int N = (int)p->N;
  const int dp = p->d.dp;
  T *F = a.fields;
  T *tmp[2] = {(T *)((char *)F + (p->off_tmp - p->off_fields)),
               (T *)((char *)F + (p->off_tmp - p->off_fields)) + N};
  const unsigned nb = nd_blocks(N);
  for (int sl = 0; sl < p->n_slots; ++sl) {
    const std::vector<int> &c = p->chains[sl];
    const T *src = a.u + (int64_t)c[0] * N;
    T *out = F + (int64_t)sl * N;
    const int nax = (int)c.size() - 1;
    if (nax == 0)
      hipLaunchKernelGGL((k_mat_nd_copy<T>), dim3(nb), dim3(256), 0, s, src, out, (int64_t)N, 0);
    for (int j = 0; j < nax; ++j) {
      const NdAxis ax = nd_axis(p, c[1 + j]);
      T *dst = j == nax - 1 ? out : tmp[j & 1];
      hipLaunchKernelGGL((k_mat_nd_d<T, DIV>), dim3(nb), dim3(256), 0, s, src, dst, N, ax.stride,
                         ax.n, dp, a.coef, T(ax.h), T(2.0 * ax.h), T(1.0 / ax.h),
                         T(1.0 / (2.0 * ax.h)));
      src = dst;
    }
  }
  TDX_TRY_HIP(hipGetLastError());
  const size_t lds = (size_t)p->n_preal * sizeof(T) + (size_t)p->n_pint * 4;
  hipLaunchKernelGGL((k_mat_nd_resid<T>), dim3(p->n_nd_wg), dim3(MAT_THREADS), lds, s, a,
                     adj ? 1 : 0);
  TDX_TRY_HIP(hipGetLastError());
  if (p->n_bc_wg > 0) {
    hipLaunchKernelGGL((k_mat_bc<T, DIV, true>), dim3(p->n_bc_wg), dim3(MAT_THREADS), 0, s, a,
                       adj ? 1 : 0);
    TDX_TRY_HIP(hipGetLastError());
  }
  if (adj) {
    const int64_t eqw = (int64_t)p->n_fs * N, allw = (int64_t)p->n_slots * N;
    // the boundary slots' fields now take the boundary seeds
    if (allw > eqw) {
      hipLaunchKernelGGL((k_mat_nd_zero<T>), dim3(nd_blocks(allw - eqw)), dim3(256), 0, s,
                         F + eqw, allw - eqw);
      if (p->sc_rows_f > 0)
        hipLaunchKernelGGL((k_mat_nd_scatter<T>), dim3((p->sc_rows_f + 255) / 256), dim3(256),
                           0, s, (const int32_t *)p->sc_ptr.ptr, (const int64_t *)p->sc_key.ptr,
                           (const int32_t *)p->sc_idx.ptr, p->sc_rows_f,
                           (const T *)a.bc_seeds, F, allw, a.grad);
    }
    const int64_t gw = (int64_t)p->d.n_eq * N;
    hipLaunchKernelGGL((k_mat_nd_zero<T>), dim3(nd_blocks(gw)), dim3(256), 0, s, a.grad, gw);
    for (int sl = 0; sl < p->n_slots; ++sl) {
      const std::vector<int> &c = p->chains[sl];
      T *g = a.grad + (int64_t)c[0] * N;
      const T *src = F + (int64_t)sl * N;
      const int nax = (int)c.size() - 1;
      if (nax == 0)
        hipLaunchKernelGGL((k_mat_nd_copy<T>), dim3(nb), dim3(256), 0, s, src, g, (int64_t)N, 1);
      for (int j = nax - 1, k = 0; j >= 0; --j, ++k) {
        const int axi = c[1 + j];
        const NdAxis ax = nd_axis(p, axi);
        T *dst = j == 0 ? g : tmp[k & 1];
        hipLaunchKernelGGL((k_mat_nd_dt<T>), dim3(nb), dim3(256), 0, s, src, dst, N, ax.stride,
                           ax.n, dp, a.coef + MCOEF_AX + axi * 4 * MMAX_DP, j == 0 ? 1 : 0);
        src = dst;
      }
    }
    TDX_TRY_HIP(hipGetLastError());
    if (p->sc_rows > p->sc_rows_f) {
      const int rows = p->sc_rows - p->sc_rows_f;
      hipLaunchKernelGGL((k_mat_nd_scatter<T>), dim3((rows + 255) / 256), dim3(256), 0, s,
                         (const int32_t *)p->sc_ptr.ptr + p->sc_rows_f,
                         (const int64_t *)p->sc_key.ptr + p->sc_rows_f,
                         (const int32_t *)p->sc_idx.ptr, rows, (const T *)a.bc_seeds, F,
                         (int64_t)p->n_slots * N, a.grad);
      TDX_TRY_HIP(hipGetLastError());
    }
    if (a.fused) {
      // Adam over every point: the frame kernel with an empty rectangle on
      // a single row of N points
      MatArgs<T> b = a;
      b.n0 = 1;
      b.n1 = N;
      const int64_t tot = gw;
      hipLaunchKernelGGL((k_mat_frame_adam<T>), dim3(nd_blocks(tot)), dim3(256), 0, s, b, 0, 0,
                         0, 0, (const T *)nullptr, 0, (T *)nullptr);
      TDX_TRY_HIP(hipGetLastError());
    }
  }
  if (loss) {
    const int n = p->n_nd_wg + p->n_bc_wg;
    const int g = std::min(256, (n + 4095) / 4096);
    T *mid = a.lp_mid;
    if (g > 1) {
      hipLaunchKernelGGL((k_mat_loss<T>), dim3(g), dim3(256), 0, s, a.lp, n, mid);
      TDX_TRY_HIP(hipGetLastError());
      hipLaunchKernelGGL((k_mat_loss<T>), dim3(1), dim3(256), 0, s, (const T *)mid, g, loss);
    } else {
      hipLaunchKernelGGL((k_mat_loss<T>), dim3(1), dim3(256), 0, s, a.lp, n, loss);
    }
    TDX_TRY_HIP(hipGetLastError());
  }
  return TDX_OK;
}

template <typename T>
int mat_dispatch(const tdx_mat_plan *p, MatArgs<T> a, bool adj, T *loss, hipStream_t s) {
  if (p->nd) {
    if (p->div_mode == 0) return mat_run_nd<T, 0>(p, a, adj, loss, s);
    if (p->div_mode == 2) return mat_run_nd<T, 2>(p, a, adj, loss, s);
    return mat_run_nd<T, 1>(p, a, adj, loss, s);
  }
  if (p->div_mode == 0) return mat_run<T, 0>(p, a, adj, loss, s);
  if (p->div_mode == 2) return mat_run<T, 2>(p, a, adj, loss, s);
  return mat_run<T, 1>(p, a, adj, loss, s);
}

}  // namespace

namespace {

// Division mode of a plan: 0 product with the rounded reciprocal, 1 IEEE
// division, 2 (float) the corrected reciprocal product where it is exact for
// the steps.
int plan_div_mode(const tdx_mat_desc *d) {
  if (d->exact_div == 0) return 0;
  if (d->dtype == 0 && d->exact_div == 1 && quotient_step_exact((float)d->h[1]) &&
      (d->n0 == 1 || quotient_step_exact((float)d->h[0])) &&
      (d->n2 == 1 || quotient_step_exact((float)d->h[2])))
    return 2;
  return 1;
}

// The residual program of a float 2-D descriptor as a generator spec, with
// the tables' ranges checked.  Returns false with the reason in *why.
bool jit_spec_from_desc(const tdx_mat_desc *d, int div_mode, JitSpec *sp, std::string *why) {
  if (d->dtype != 0 || d->n0 < 2 || d->n1 < 2 || d->n2 != 1 || d->dp < 2 || d->dp > MMAX_DP ||
      d->n_eq < 1 ||
      d->n_eq > MMAX_EQ) {
    *why = "not a float 2-D grid";
    return false;
  }
  const int64_t N = (int64_t)d->n0 * d->n1;
  auto in_ints = [&](int64_t off, int64_t n) { return off >= 0 && n >= 0 && off + n <= d->n_ints; };
  auto in_reals = [&](int64_t off, int64_t n) { return off >= 0 && n >= 0 && off + n <= d->n_reals; };
  auto in_data = [&](int64_t off, int64_t n) { return off >= 0 && n >= 0 && off + n <= d->n_data; };
  auto real_at = [&](int64_t k) { return (double)((const float *)d->reals)[k]; };
  sp->n_eq = d->n_eq;
  sp->dp = d->dp;
  sp->div_mode = div_mode;
  sp->N = N;
  for (int i = 0; i < 8; ++i) {
    sp->ab[i] = i < d->dp ? (float)d->alpha_back[i] : 0.0f;
    sp->af[i] = i < d->dp ? (float)d->alpha_forw[i] : 0.0f;
  }
  for (int ax = 0; ax < 2; ++ax) {
    sp->h2[ax] = (float)(2.0 * d->h[ax]);
    sp->ih2[ax] = (float)(1.0 / (2.0 * d->h[ax]));
    for (int i = 0; i < 8; ++i) {
      // the transposed products ab/(2h), af/(2h) rounded once, as in the
      // general kernel's coefficient table
      sp->tb2[ax][i] = i < d->dp ? (float)((double)d->alpha_back[i] / (2.0 * d->h[ax])) : 0.0f;
      sp->tf2[ax][i] = i < d->dp ? (float)((double)d->alpha_forw[i] / (2.0 * d->h[ax])) : 0.0f;
    }
  }
  int cnt0 = 0, cnt1 = 0, n_fs = 0;
  sp->eqs.clear();
  if (!in_ints(d->eq_i, (int64_t)d->n_eq * TDX_MEQ_INTS)) {
    *why = "eq table out of range";
    return false;
  }
  for (int q = 0; q < d->n_eq; ++q) {
    const int32_t *E = d->ints + d->eq_i + q * TDX_MEQ_INTS;
    if (E[0] < 1 || !in_ints(E[1], (int64_t)E[0] * TDX_MTERM_INTS) || !in_reals(E[2], 1)) {
      *why = "term table out of range";
      return false;
    }
    JitEq eq;
    eq.lam = real_at(E[2]);
    for (int t = 0; t < E[0]; ++t) {
      const int32_t *Tm = d->ints + E[1] + t * TDX_MTERM_INTS;
      if (Tm[0] < 1 || Tm[0] > MMAX_TF || !in_ints(Tm[1], (int64_t)Tm[0] * TDX_MFACTOR_INTS) ||
          !((Tm[2] == 0 && in_reals(Tm[3], 1)) || (Tm[2] == 1 && in_data(Tm[3], N)))) {
        *why = "term record out of range";
        return false;
      }
      JitTerm tm;
      tm.coeff_kind = Tm[2];
      tm.coeff = Tm[2] == 0 ? real_at(Tm[3]) : 0.0;
      tm.data_off = Tm[2] == 0 ? 0 : Tm[3];
      for (int f = 0; f < Tm[0]; ++f) {
        const int32_t *F = d->ints + Tm[1] + f * TDX_MFACTOR_INTS;
        if (F[0] < 0 || F[0] >= d->n_eq || F[4] < 0 || F[4] > 3 ||
            (F[1] != 0 && !in_reals(F[3], 1)) || ++n_fs > MMAX_FS) {
          *why = "factor record out of range";
          return false;
        }
        JitFactor jf;
        jf.var = F[0];
        jf.nax = F[4];
        int c0 = 0, c1 = 0;
        for (int k = 0; k < 3; ++k) {
          jf.ax[k] = k < F[4] ? F[5 + k] : 0;
          if (k < F[4] && F[5 + k] != 0 && F[5 + k] != 1) {
            *why = "axis out of range";
            return false;
          }
          if (k < F[4]) (F[5 + k] == 0 ? c0 : c1)++;
        }
        cnt0 = std::max(cnt0, c0);
        cnt1 = std::max(cnt1, c1);
        jf.pow_kind = F[1] != 0 ? 1 : 0;
        jf.pow_int = F[2];
        jf.pow_real = F[1] != 0 ? real_at(F[3]) : 0.0;
        tm.f.push_back(jf);
      }
      eq.terms.push_back(tm);
    }
    sp->eqs.push_back(eq);
  }
  sp->R0 = cnt0 * (d->dp - 1);
  sp->R1 = cnt1 * (d->dp - 1);
  return true;
}

// Sets up the specialized interior kernel of a float 2-D plan (see the
// specialize field of tdx_mat_desc): the interior tiles of the plan's tiling
// -- those whose halo lies inside the grid, so every stencil they apply or
// transpose is central -- form a rectangle of output points that the
// generated kernel evaluates; the general kernel keeps the other tiles.
// Returns an error only when specialization was required (specialize = 1).
int mat_specialize(tdx_mat_plan *p, const tdx_mat_desc *d) {
  if (d->dtype != 0 || d->n0 == 1 || d->n2 != 1 || d->specialize < 0) return TDX_OK;
  const bool required = d->specialize > 0;
  const int T0 = p->T0, T1 = p->T1, R0 = p->R0, R1 = p->R1;
  // tile t along an axis is interior when t*T >= 2R and t*T + T + 2R <= n
  const int ta0 = (2 * R0 + T0 - 1) / T0, ta1 = (2 * R1 + T1 - 1) / T1;
  const int tb0 = d->n0 - T0 - 2 * R0 >= 0 ? (d->n0 - T0 - 2 * R0) / T0 + 1 : 0;
  const int tb1 = d->n1 - T1 - 2 * R1 >= 0 ? (d->n1 - T1 - 2 * R1) / T1 + 1 : 0;
  if (tb0 <= ta0 || tb1 <= ta1) return TDX_OK;
  const int64_t pts = (int64_t)(tb0 - ta0) * T0 * (tb1 - ta1) * T1;
  if (!required && pts < (1LL << 20)) return TDX_OK;
  JitSpec sp;
  std::string why;
  if (!jit_spec_from_desc(d, p->div_mode, &sp, &why)) {
    set_error("%s", why.c_str());
    return TDX_EINVAL;
  }
  const std::string src = mat_jit_source(sp, &why);
  if (src.empty()) {
    if (!required) return TDX_OK;
    set_error("no specialized kernel for this program: %s", why.c_str());
    return TDX_EINVAL;
  }
  const int rc = mat_jit_compile(src, &p->jit);
  if (rc) {
    if (required) return rc;
    clear_error();  // the general kernel serves the plan
    return TDX_OK;
  }
  // a generated kernel that needs scratch (register spills) is not taken
  // unless asked for: the general kernel serves such programs better
  auto spills = [](hipFunction_t f) {
    int bytes = 0;
    return hipFuncGetAttribute(&bytes, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, f) == hipSuccess &&
           bytes > 0;
  };
  if (!required && (spills(p->jit.fn_adj) || spills(p->jit.fn_step))) return TDX_OK;
  // geometry: a wavefront owns 64 - 4 R1 columns of a band of rows
  p->jr_lo = ta0 * T0;
  p->jr_hi = tb0 * T0;
  p->jc_lo = ta1 * T1;
  p->jc_hi = tb1 * T1;
  const int OW = 64 - 4 * R1;
  p->j_nstrips = (p->jc_hi - p->jc_lo + OW - 1) / OW;
  const int rows = p->jr_hi - p->jr_lo;
  // Rows per band, per kernel variant: the wavefronts run in rounds of as
  // many as are resident at once (the variant's occupancy), each round as
  // long as a band plus its halo and pipeline rows, so the band height that
  // minimizes rounds x (rows + 4 R0 + 8) is taken (on a 16384^2 grid the
  // model orders the band heights 64..1024 as the measured times do).
  int dev = 0, cus = 0;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus <= 0)
    cus = 256;
  const hipFunction_t fns[3] = {p->jit.fn_fwd, p->jit.fn_adj, p->jit.fn_step};
  p->j_waves_max = 0;
  for (int var = 0; var < 3; ++var) {
    int hb = 0;
    if (d->band_rows > 0) {
      hb = std::min(d->band_rows, rows);
    } else {
      int per_cu = 0;
      if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fns[var], 256, 0) !=
              hipSuccess ||
          per_cu <= 0)
        per_cu = 2;
      const int64_t resident = (int64_t)per_cu * 4 * cus;
      int64_t best = -1;
      for (int cand : {32, 48, 64, 96, 128, 192, 256, 384}) {
        const int h = std::min(cand, rows);
        const int64_t waves = (int64_t)p->j_nstrips * ((rows + h - 1) / h);
        const int64_t cost = ((waves + resident - 1) / resident) * (h + 4 * R0 + 8);
        if (best < 0 || cost < best) {  // ties: the shorter band (measured faster)
          best = cost;
          hb = h;
        }
      }
    }
    p->j_hb[var] = hb;
    p->j_nbands[var] = (rows + hb - 1) / hb;
    p->j_waves[var] = p->j_nstrips * p->j_nbands[var];
    p->j_waves_max = std::max(p->j_waves_max, p->j_waves[var]);
  }
  // the general kernel's tiles: every tile outside the interior rectangle
  std::vector<int32_t> list;
  for (int t0 = 0; t0 < p->tiles0; ++t0)
    for (int t1 = 0; t1 < p->tiles1; ++t1)
      if (t0 < ta0 || t0 >= tb0 || t1 < ta1 || t1 >= tb1) list.push_back(t0 * p->tiles1 + t1);
  if (!list.empty()) {
    const int urc = p->gen_list.upload(list.data(), list.size() * 4);
    if (urc) return urc;
  }
  p->n_gen_wg = (int)list.size();
  // the fused step finalizes the rectangle inside the generated kernel, so
  // no boundary adjoint may add to a point of it afterwards
  p->jit_fuse = 1;
  for (int j = 0; j < d->n_touch; ++j) {
    const int64_t pos = d->ints[d->bct_pos_i + j] % sp.N;
    const int64_t r = pos / d->n1, c = pos - (pos / d->n1) * d->n1;
    if (r >= p->jr_lo && r < p->jr_hi && c >= p->jc_lo && c < p->jc_hi) p->jit_fuse = 0;
  }
  p->jit_on = 1;
  return TDX_OK;
}

// Full-grid plan set-up: factor slots of the boundary operators (one set
// per operator term table, shared by the entries that use it), each bc
// entry's first slot, the scatter CSR of the boundary seeds (keys: field
// word slot * N + pos, then n_slots * N + gradient word; the seed order of
// k_mat_bc), the residual grid.
int nd_setup(tdx_mat_plan *p, const tdx_mat_desc *d) {
  const int64_t N = p->N;
  std::map<int, int> first_slot;
  std::vector<int32_t> bc_slot(std::max(d->n_bc, 1), -1);
  for (int e = 0; e < d->n_bc; ++e) {
    const int32_t *B = d->ints + d->bc_i + e * TDX_MBC_INTS;
    if (B[2] == 0) continue;
    auto it = first_slot.find(B[3]);
    if (it == first_slot.end()) {
      const int base = (int)p->chains.size();
      for (int t = 0; t < B[2]; ++t) {
        const int32_t *Tm = d->ints + B[3] + t * TDX_MTERM_INTS;
        for (int f = 0; f < Tm[0]; ++f) {
          const int32_t *F = d->ints + Tm[1] + f * TDX_MFACTOR_INTS;
          std::vector<int> ch{F[0]};
          for (int k = 0; k < F[4]; ++k) ch.push_back(F[5 + k]);
          p->chains.push_back(ch);
        }
      }
      it = first_slot.emplace(B[3], base).first;
    }
    bc_slot[e] = it->second;
  }
  p->n_slots = (int)p->chains.size();
  const int64_t allw = (int64_t)p->n_slots * N;
  std::map<int64_t, std::vector<int32_t>> rows;
  for (int e = 0; e < d->n_bc; ++e) {
    const int32_t *B = d->ints + d->bc_i + e * TDX_MBC_INTS;
    int sidx = B[8];
    const int var = B[4];
    if (B[9] > 0) rows[allw + var * N + d->ints[B[1]]].push_back(sidx++);
    for (int s = 0; s < B[0]; ++s) {
      const int pos = d->ints[B[1] + s];
      if (B[2] == 0) {
        rows[allw + var * N + pos].push_back(sidx++);
        continue;
      }
      int k = 0;
      for (int t = 0; t < B[2]; ++t) {
        const int32_t *Tm = d->ints + B[3] + t * TDX_MTERM_INTS;
        for (int f = 0; f < Tm[0]; ++f, ++k) rows[(int64_t)(bc_slot[e] + k) * N + pos].push_back(sidx++);
      }
    }
  }
  std::vector<int32_t> ptr{0}, idx;
  std::vector<int64_t> key;
  p->sc_rows_f = 0;
  for (const auto &kv : rows) {
    key.push_back(kv.first);
    idx.insert(idx.end(), kv.second.begin(), kv.second.end());
    ptr.push_back((int32_t)idx.size());
    if (kv.first < allw) ++p->sc_rows_f;
  }
  p->sc_rows = (int)key.size();
  if (idx.empty()) idx.push_back(0);
  if (key.empty()) key.push_back(0);
  int rc = p->bc_slot_tab.upload(bc_slot.data(), bc_slot.size() * 4);
  if (!rc) rc = p->sc_ptr.upload(ptr.data(), ptr.size() * 4);
  if (!rc) rc = p->sc_key.upload(key.data(), key.size() * 8);
  if (!rc) rc = p->sc_idx.upload(idx.data(), idx.size() * 4);
  if (rc) return rc;
  p->n_nd_wg = (int)std::max<int64_t>(1, std::min<int64_t>((N + MAT_THREADS - 1) / MAT_THREADS, 2048));
  p->n_gen_wg = p->n_nd_wg;
  p->n_tile_wg = 0;
  return TDX_OK;
}

}  // namespace

extern "C" int tdx_mat_plan_create(const tdx_mat_desc *d, tdx_mat_plan **out) {
  TDX_REQUIRE(d && out, "null argument");
  *out = nullptr;
  TDX_REQUIRE(d->dtype == 0 || d->dtype == 1, "dtype must be 0 or 1");
  TDX_REQUIRE(d->n_eq >= 1 && d->n_eq <= MMAX_EQ, "n_eq must be 1..%d", MMAX_EQ);
  TDX_REQUIRE(d->n0 >= 1 && d->n1 >= 2 && d->n2 >= 1, "bad grid shape");
  TDX_REQUIRE(d->n2 == 1 || d->n0 >= 2, "a 3-D grid needs axis 0 of length >= 2");
  TDX_REQUIRE(d->dp >= 2 && d->dp <= MMAX_DP, "derivative_points must be 2..%d", MMAX_DP);
  TDX_REQUIRE(d->tile_rows >= 0 && d->tile_rows <= 1024, "tile_rows must be 0..1024");
  TDX_REQUIRE(d->exact_div >= 0 && d->exact_div <= 2, "exact_div must be 0, 1 or 2");
  TDX_REQUIRE(d->band_rows >= 0 && d->band_rows <= (1 << 20), "band_rows must be 0..2^20");
  auto in_ints = [&](int64_t off, int64_t n) { return off >= 0 && n >= 0 && off + n <= d->n_ints; };
  auto in_reals = [&](int64_t off, int64_t n) { return off >= 0 && n >= 0 && off + n <= d->n_reals; };
  auto in_data = [&](int64_t off, int64_t n) { return off >= 0 && n >= 0 && off + n <= d->n_data; };
  const int64_t N = (int64_t)d->n0 * d->n1 * d->n2;
  TDX_REQUIRE(N * d->n_eq < (1LL << 31), "grid too large");
  const int n_axes = d->n2 > 1 ? 3 : 2;
  const int ext[3] = {d->n0, d->n1, d->n2};

  tdx_mat_plan *p = new tdx_mat_plan();
  p->d = *d;
  p->N = N;
  for (int i = 0; i < d->dp; ++i) {
    p->ab[i] = d->alpha_back[i];
    p->af[i] = d->alpha_forw[i];
  }
  int cnt0 = 0, cnt1 = 0, maxc = 0;
  p->n_fs = 0;
  std::vector<int32_t> fs_rec;
  // flattened residual program (see MatArgs)
  std::vector<int32_t> prog_i, prog_t, prog_f;
  std::vector<double> prog_r;
  auto real_at = [&](int64_t k) {
    return d->dtype == 0 ? (double)((const float *)d->reals)[k] : ((const double *)d->reals)[k];
  };
  auto fail = [&](const char *msg) {
    delete p;
    set_error("%s", msg);
    return TDX_EINVAL;
  };
  if (!in_ints(d->eq_i, (int64_t)d->n_eq * TDX_MEQ_INTS)) return fail("eq table out of range");
  int n_terms_before = 0;
  for (int q = 0; q < d->n_eq; ++q) {
    const int32_t *E = d->ints + d->eq_i + q * TDX_MEQ_INTS;
    if (E[0] < 1 || !in_ints(E[1], (int64_t)E[0] * TDX_MTERM_INTS) || !in_reals(E[2], 1))
      return fail("term table out of range");
    prog_i.push_back(E[0]);
    prog_i.push_back(n_terms_before);
    n_terms_before += E[0];
    prog_r.push_back(real_at(E[2]));
  }
  for (int q = 0; q < d->n_eq; ++q) {
    const int32_t *E = d->ints + d->eq_i + q * TDX_MEQ_INTS;
    for (int t = 0; t < E[0]; ++t) {
      const int32_t *Tm = d->ints + E[1] + t * TDX_MTERM_INTS;
      if (Tm[0] < 1 || !in_ints(Tm[1], (int64_t)Tm[0] * TDX_MFACTOR_INTS))
        return fail("factor table out of range");
      if (Tm[0] > MMAX_TF) return fail("at most 4 factors per term");
      if (!((Tm[2] == 0 && in_reals(Tm[3], 1)) || (Tm[2] == 1 && in_data(Tm[3], N))))
        return fail("coefficient out of range");
      prog_t.push_back(Tm[0]);
      prog_t.push_back(Tm[2]);
      if (Tm[2] == 0) {
        prog_t.push_back((int)prog_r.size());
        prog_r.push_back(real_at(Tm[3]));
      } else {
        prog_t.push_back(Tm[3]);
      }
      prog_t.push_back(p->n_fs);
      for (int f = 0; f < Tm[0]; ++f) {
        const int32_t *F = d->ints + Tm[1] + f * TDX_MFACTOR_INTS;
        if (F[0] < 0 || F[0] >= d->n_eq) return fail("factor variable out of range");
        if (F[1] != 0 && !in_reals(F[3], 1)) return fail("power out of range");
        if (F[4] < 0 || F[4] > TDX_MAT_MAX_CHAIN)
          return fail("derivative order must be 0..8 per factor");
        if (p->n_fs >= MMAX_FS) return fail("too many factors");
        int c0 = 0, c1 = 0;
        for (int k = 0; k < F[4]; ++k) {
          const int ax = F[5 + k];
          if (ax < 0 || ax >= n_axes) return fail("axis out of range");
          if (ext[ax] == 1) return fail("derivative along a length-1 axis");
          if (ext[ax] < 2 * d->dp - 1) return fail("axis too short for the stencil");
          if (ax == 0) ++c0; else ++c1;
        }
        cnt0 = std::max(cnt0, c0);
        cnt1 = std::max(cnt1, c1);
        maxc = std::max(maxc, (int)F[4]);
        prog_f.push_back(F[1]);
        prog_f.push_back(F[2]);
        if (F[1] != 0) {
          prog_f.push_back((int)prog_r.size());
          prog_r.push_back(real_at(F[3]));
        } else {
          prog_f.push_back(0);
        }
        fs_rec.push_back(F[0]);
        fs_rec.push_back(F[4]);
        for (int k = 0; k < TDX_MAT_MAX_CHAIN; ++k) fs_rec.push_back(k < F[4] ? F[5 + k] : 0);
        std::vector<int> ch{F[0]};
        for (int k = 0; k < F[4]; ++k) ch.push_back(F[5 + k]);
        p->chains.push_back(ch);
        ++p->n_fs;
      }
    }
  }
  // boundary tables
  if (d->n_bc > 0) {
    if (!in_ints(d->bc_i, (int64_t)d->n_bc * TDX_MBC_INTS)) return fail("bc table out of range");
    for (int e = 0; e < d->n_bc; ++e) {
      const int32_t *B = d->ints + d->bc_i + e * TDX_MBC_INTS;
      if (B[0] < 1 || !in_ints(B[1], B[0])) return fail("bc positions out of range");
      for (int s = 0; s < B[0]; ++s)
        if (d->ints[B[1] + s] < 0 || d->ints[B[1] + s] >= N) return fail("bc position outside the grid");
      if (B[4] < 0 || B[4] >= d->n_eq) return fail("bc variable out of range");
      if (!in_data(B[5], 1) || !in_reals(B[6], 1) || !in_reals(B[7], 1)) return fail("bc weights");
      int nseed = B[2] == 0 ? B[0] : 0;
      if (B[2] > 0) {
        if (!in_ints(B[3], (int64_t)B[2] * TDX_MTERM_INTS)) return fail("bc term table");
        int nf = 0;
        for (int t = 0; t < B[2]; ++t) {
          const int32_t *Tm = d->ints + B[3] + t * TDX_MTERM_INTS;
          if (Tm[0] < 1 || !in_ints(Tm[1], (int64_t)Tm[0] * TDX_MFACTOR_INTS)) return fail("bc factors");
          if (!((Tm[2] == 0 && in_reals(Tm[3], 1)) || (Tm[2] == 1 && in_data(Tm[3], N))))
            return fail("bc coefficient out of range");
          for (int f = 0; f < Tm[0]; ++f) {
            const int32_t *F = d->ints + Tm[1] + f * TDX_MFACTOR_INTS;
            if (F[0] < 0 || F[0] >= d->n_eq || F[4] < 0 || F[4] > TDX_MAT_MAX_CHAIN)
              return fail("bc factor");
            if (F[1] != 0 && !in_reals(F[3], 1)) return fail("bc power");
            for (int k = 0; k < F[4]; ++k) {
              const int ax = F[5 + k];
              if (ax < 0 || ax >= n_axes || ext[ax] == 1) return fail("bc axis");
              if (ext[ax] < 2 * d->dp - 1) return fail("bc axis too short for the stencil");
            }
            maxc = std::max(maxc, (int)F[4]);
          }
          nf += Tm[0];
        }
        nseed = nf * B[0];
      }
      if (B[9] != 0) {
        // Robin: one point set, an operator, {n_beta, alpha_r} in range
        if (B[9] < 1 || !in_ints(B[9] - 1, 2) || B[0] != 1 || B[2] < 1) return fail("bc robin record");
        const int32_t *Rb = d->ints + B[9] - 1;
        if (Rb[0] < 0 || !in_reals(Rb[1], 1 + (int64_t)Rb[0])) return fail("bc robin coefficients");
        nseed += 1;
      }
      if (B[8] < 0 || B[8] + nseed > d->n_bc_seeds) return fail("bc seed slots out of range");
    }
    if (!in_ints(d->bct_ptr_i, (int64_t)d->n_touch + 1) || !in_ints(d->bct_pos_i, d->n_touch))
      return fail("bc adjoint table out of range");
    const int nnz = d->n_touch > 0 ? d->ints[d->bct_ptr_i + d->n_touch] : 0;
    if (!in_ints(d->bct_seed_i, nnz) || !in_ints(d->bct_w_i, nnz)) return fail("bc adjoint nnz");
    for (int j = 0; j < d->n_touch; ++j) {
      const int a0 = d->ints[d->bct_ptr_i + j], a1 = d->ints[d->bct_ptr_i + j + 1];
      if (a0 < 0 || a1 < a0 || a1 > nnz) return fail("bc adjoint row pointers");
      const int pos = d->ints[d->bct_pos_i + j];
      if (pos < 0 || pos >= N * d->n_eq) return fail("bc adjoint position");
    }
    for (int k = 0; k < nnz; ++k) {
      const int sidx = d->ints[d->bct_seed_i + k], widx = d->ints[d->bct_w_i + k];
      if (sidx < 0 || sidx >= d->n_bc_seeds || !in_reals(widx, 1)) return fail("bc adjoint entry");
    }
  }
  // equation records, then term records, then factor records
  p->pterm0 = (int)prog_i.size();
  prog_i.insert(prog_i.end(), prog_t.begin(), prog_t.end());
  p->pfac0 = (int)prog_i.size();
  prog_i.insert(prog_i.end(), prog_f.begin(), prog_f.end());
  if (prog_r.size() % 2) prog_r.push_back(0.0);  // keep the ints 8-byte aligned
  p->n_pint = (int)prog_i.size();
  p->n_preal = (int)prog_r.size();
  p->R0 = cnt0 * (d->dp - 1);
  p->R1 = cnt1 * (d->dp - 1);
  const size_t ts = d->dtype == 0 ? 4 : 8;
  // Tile shape.  A row of the tile with its halo is E1 = 64 * 2^k values
  // wide (one coalesced wavefront access per 64), at least 32 of them
  // output; 1-D grids use one row of 1024.  The number of rows is the
  // largest that keeps the tile within about 48 KB of LDS (several
  // workgroups per CU), with at least 8 output rows.  Programs with long
  // reach and many factors, whose smallest such tile exceeds the 160 KB of
  // a CU, take rows of 32 or 16 values instead (at least 8 output columns).
  const size_t words = d->n_eq + p->n_fs + 2;  // U (then G), F, A, B per point
  const size_t fixed = (2 * MAT_THREADS + MCOEF_N + prog_r.size()) * ts + prog_i.size() * 4;
  auto lds_for = [&](int E0, int E1) { return (size_t)E0 * E1 * words * ts + fixed; };
  auto set_tpr = [&]() {
    p->tpr_log = 0;
    while ((1 << (p->tpr_log + 1)) <= std::min(p->E1, MAT_THREADS)) ++p->tpr_log;
    return MAT_THREADS >> p->tpr_log;  // tile rows per sweep of the workgroup
  };
  auto min_rows = [&](int rps) { return (4 * p->R0 + 8 + rps - 1) / rps * rps; };
  p->nd = (d->n2 > 1 || maxc > 3) ? 1 : 0;
  if (p->nd) {
    p->E0 = p->E1 = 1;
  } else if (d->n0 == 1) {
    p->E0 = 1;
    p->E1 = 1024;
    if (p->E1 - 4 * p->R1 < 256) p->E1 = 2048;
    set_tpr();
  } else {
    p->E1 = 64;
    while (p->E1 - 4 * p->R1 < 32) p->E1 *= 2;
    int rps = set_tpr();
    if (d->tile_rows > 0) {
      // requested output rows, rounded up so the tile rows fill whole sweeps
      p->E0 = (4 * p->R0 + d->tile_rows + rps - 1) / rps * rps;
    } else {
      int e0 = (int)((48 * 1024 - std::min<size_t>(fixed, 32 * 1024)) / ((size_t)p->E1 * words * ts));
      e0 = std::min(e0, d->n0 + 4 * p->R0);
      e0 = e0 / rps * rps;
      if (e0 - 4 * p->R0 < 8) e0 = min_rows(rps);
      p->E0 = e0;
      while (lds_for(p->E0, p->E1) > 160 * 1024 && p->E1 > 16 && p->E1 / 2 - 4 * p->R1 >= 8) {
        p->E1 /= 2;
        rps = set_tpr();
        p->E0 = min_rows(rps);
      }
    }
  }
  p->T0 = p->E0 - 4 * p->R0;
  p->T1 = p->E1 - 4 * p->R1;
  // a reach that no LDS tile holds: the full-grid passes serve the plan
  if (!p->nd && (p->T0 < 1 || p->T1 < 1 || lds_for(p->E0, p->E1) > 160 * 1024)) p->nd = 1;
  if (p->nd) {
    p->E0 = p->E1 = p->T0 = p->T1 = 1;
    p->R0 = p->R1 = 0;
  }
  p->div_mode = plan_div_mode(d);
  p->tile_lds = lds_for(p->E0, p->E1);
  p->tile_lds_fwd = p->tile_lds;
  p->tiles0 = (d->n0 + p->T0 - 1) / p->T0;
  p->tiles1 = (d->n1 + p->T1 - 1) / p->T1;
  p->n_tile_wg = p->tiles0 * p->tiles1;
  p->n_bc_wg = (d->n_bc + MAT_THREADS - 1) / MAT_THREADS;
  p->jit_on = 0;
  p->jit_fuse = 0;
  p->n_gen_wg = p->n_tile_wg;

  // stencil coefficients and their transposed products, in T
  std::vector<double> coef(MCOEF_N, 0.0);
  for (int i = 0; i < d->dp; ++i) {
    coef[i] = p->ab[i];
    coef[MMAX_DP + i] = p->af[i];
    for (int ax = 0; ax < 3; ++ax) {
      double *c = coef.data() + MCOEF_AX + ax * 4 * MMAX_DP;
      const double h = d->h[ax] > 0 ? d->h[ax] : 1.0;  // unused axes: any finite value
      c[i] = p->ab[i] / (2.0 * h);
      c[MMAX_DP + i] = p->af[i] / (2.0 * h);
      c[2 * MMAX_DP + i] = p->ab[i] / h;
      c[3 * MMAX_DP + i] = p->af[i] / h;
    }
  }
  std::vector<float> coef_f(coef.begin(), coef.end());
  if (fs_rec.empty()) fs_rec.assign(MFS_INTS, 0);
  int rc = p->ints.upload(d->ints, (size_t)std::max<int64_t>(d->n_ints, 1) * 4);
  if (!rc) rc = p->fs_tab.upload(fs_rec.data(), fs_rec.size() * 4);
  if (!rc) rc = p->pint_tab.upload(prog_i.data(), prog_i.size() * 4);
  if (!rc) {
    std::vector<float> prf(prog_r.begin(), prog_r.end());
    rc = d->dtype == 0 ? p->preal_tab.upload(prf.data(), prf.size() * 4)
                       : p->preal_tab.upload(prog_r.data(), prog_r.size() * 8);
  }
  if (!rc)
    rc = d->dtype == 0 ? p->coef_tab.upload(coef_f.data(), coef_f.size() * 4)
                       : p->coef_tab.upload(coef.data(), coef.size() * 8);
  if (!rc) rc = p->reals.upload(d->reals, (size_t)d->n_reals * ts);
  if (!rc) rc = p->data.upload(d->data, (size_t)std::max<int64_t>(d->n_data, 1) * ts);
  if (rc) {
    tdx_mat_plan_destroy(p);
    return rc;
  }
  rc = p->nd ? nd_setup(p, d) : mat_specialize(p, d);
  if (rc) {
    tdx_mat_plan_destroy(p);
    return rc;
  }
  p->d.ints = nullptr;
  p->d.reals = nullptr;
  p->d.data = nullptr;
  p->d.alpha_back = nullptr;
  p->d.alpha_forw = nullptr;
  size_t off = 0;
  p->off_lp = off;
  const int n_pairs = p->n_gen_wg + p->n_bc_wg + (p->jit_on ? p->j_waves_max : 0);
  off = align_up(off + (size_t)n_pairs * 2 * ts, 256);
  p->off_seeds = off;
  off = align_up(off + (size_t)std::max(d->n_bc_seeds, 1) * ts, 256);
  p->off_one = off;
  off = align_up(off + ts, 256);
  p->off_mid = off;
  off = align_up(off + 2 * 256 * ts, 256);
  p->off_grad = off;
  off = align_up(off + (size_t)N * d->n_eq * ts, 256);
  p->off_adam = off;
  off = align_up(off + 8 * ts, 256);
  if (p->nd) {
    p->off_fields = off;
    off = align_up(off + (size_t)p->n_slots * N * ts, 256);
    p->off_tmp = off;
    off = align_up(off + 2 * (size_t)N * ts, 256);
  }
  p->ws_bytes = off;
  *out = p;
  return TDX_OK;
}

extern "C" int64_t tdx_quotient_check(float h, int mode) {
  if (!(std::isfinite(h) && h > 0.0f) || (mode != 0 && mode != 1)) {
    set_error("tdx_quotient_check: h must be a positive finite float, mode 0 or 1");
    return -1;
  }
  return quotient_mismatches(h, mode, false);
}

extern "C" void tdx_mat_plan_destroy(tdx_mat_plan *p) {
  if (!p) return;
  p->bc_slot_tab.release();
  p->sc_ptr.release();
  p->sc_key.release();
  p->sc_idx.release();
  p->ints.release();
  p->gen_list.release();
  p->reals.release();
  p->data.release();
  p->fs_tab.release();
  p->coef_tab.release();
  p->pint_tab.release();
  p->preal_tab.release();
  delete p;
}

extern "C" size_t tdx_mat_workspace_bytes(const tdx_mat_plan *p) { return p ? p->ws_bytes : 0; }

extern "C" int64_t tdx_mat_specialized_source(const tdx_mat_desc *d, char *buf, int64_t cap,
                                              int compile) {
  TDX_REQUIRE(d, "null argument");
  TDX_REQUIRE(d->exact_div >= 0 && d->exact_div <= 2, "exact_div must be 0, 1 or 2");
  TDX_REQUIRE(d->alpha_back && d->alpha_forw && d->ints && d->reals, "null table");
  JitSpec sp;
  std::string why;
  if (!jit_spec_from_desc(d, plan_div_mode(d), &sp, &why)) {
    set_error("%s", why.c_str());
    return 0;
  }
  const std::string src = mat_jit_source(sp, &why);
  if (src.empty()) {
    set_error("%s", why.c_str());
    return 0;
  }
  if (compile) {
    std::vector<char> code;
    const int rc = mat_jit_build(src, &code);
    if (rc) return rc;
  }
  if (buf && cap > 0) {
    const size_t n = std::min<size_t>(src.size(), (size_t)(cap - 1));
    std::memcpy(buf, src.data(), n);
    buf[n] = '\0';
  }
  return (int64_t)src.size();
}

extern "C" int64_t tdx_mat_specialized_points(const tdx_mat_plan *p) {
  if (!p || !p->jit_on) return 0;
  return (int64_t)(p->jr_hi - p->jr_lo) * (p->jc_hi - p->jc_lo);
}

extern "C" int tdx_mat_forward(const tdx_mat_plan *p, const void *u, void *loss, void *op,
                               void *bval, void *ws, tdx_stream_t stream) {
  TDX_REQUIRE(p && u && ws, "null argument");
  hipStream_t s = (hipStream_t)stream;
  if (p->d.dtype == 0) {
    MatArgs<float> a = mat_args<float>(p, u, ws);
    a.op = (float *)op;
    a.bval = (float *)bval;
    return mat_dispatch<float>(p, a, false, (float *)loss, s);
  }
  MatArgs<double> a = mat_args<double>(p, u, ws);
  a.op = (double *)op;
  a.bval = (double *)bval;
  return mat_dispatch<double>(p, a, false, (double *)loss, s);
}

extern "C" int tdx_mat_backward(const tdx_mat_plan *p, const void *u, const void *dloss,
                                void *grad, void *ws, tdx_stream_t stream) {
  TDX_REQUIRE(p && u && dloss && grad && ws, "null argument");
  hipStream_t s = (hipStream_t)stream;
  if (p->d.dtype == 0) {
    MatArgs<float> a = mat_args<float>(p, u, ws);
    a.grad = (float *)grad;
    a.dloss = (const float *)dloss;
    return mat_dispatch<float>(p, a, true, nullptr, s);
  }
  MatArgs<double> a = mat_args<double>(p, u, ws);
  a.grad = (double *)grad;
  a.dloss = (const double *)dloss;
  return mat_dispatch<double>(p, a, true, nullptr, s);
}

extern "C" int tdx_mat_train_step_to(const tdx_mat_plan *p, const void *u_in, void *u_out,
                                     void *m, void *v, int64_t *step, double lr, double b1,
                                     double b2, double eps, void *loss, void *ws,
                                     tdx_stream_t stream) {
  TDX_REQUIRE(p && u_in && u_out && m && v && step && ws, "null argument");
  TDX_REQUIRE(u_in != u_out, "u_out must be a different buffer from u_in");
  hipStream_t s = (hipStream_t)stream;
  const AdamHyper h{lr, b1, b2, eps};
  char *w = (char *)ws;
  int rc;
  if (p->d.dtype == 0) {
    MatArgs<float> a = mat_args<float>(p, u_in, ws);
    a.grad = (float *)(w + p->off_grad);
    a.fused = 1;
    a.u_out = (float *)u_out;
    a.m = (float *)m;
    a.v = (float *)v;
    a.adam = (const float *)(w + p->off_adam);
    hipLaunchKernelGGL((k_mat_step_prep<float>), dim3(1), dim3(64), 0, s,
                       (float *)(w + p->off_one), step, h, (float *)(w + p->off_adam));
    rc = mat_dispatch<float>(p, a, true, (float *)loss, s);
  } else {
    MatArgs<double> a = mat_args<double>(p, u_in, ws);
    a.grad = (double *)(w + p->off_grad);
    a.fused = 1;
    a.u_out = (double *)u_out;
    a.m = (double *)m;
    a.v = (double *)v;
    a.adam = (const double *)(w + p->off_adam);
    hipLaunchKernelGGL((k_mat_step_prep<double>), dim3(1), dim3(64), 0, s,
                       (double *)(w + p->off_one), step, h, (double *)(w + p->off_adam));
    rc = mat_dispatch<double>(p, a, true, (double *)loss, s);
  }
  return rc;
}

extern "C" int tdx_mat_train_step(const tdx_mat_plan *p, void *u, void *m, void *v,
                                  int64_t *step, double lr, double b1, double b2, double eps,
                                  void *loss, void *ws, tdx_stream_t stream) {
  TDX_REQUIRE(p && u && m && v && step && ws, "null argument");
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = (int64_t)p->d.n_eq * p->N;
  int rc;
  if (p->d.dtype == 0) {
    MatArgs<float> a = mat_args<float>(p, u, ws);
    a.grad = (float *)((char *)ws + p->off_grad);
    hipLaunchKernelGGL((k_mat_one<float>), dim3(1), dim3(64), 0, s, (float *)((char *)ws + p->off_one));
    rc = mat_dispatch<float>(p, a, true, (float *)loss, s);
  } else {
    MatArgs<double> a = mat_args<double>(p, u, ws);
    a.grad = (double *)((char *)ws + p->off_grad);
    hipLaunchKernelGGL((k_mat_one<double>), dim3(1), dim3(64), 0, s, (double *)((char *)ws + p->off_one));
    rc = mat_dispatch<double>(p, a, true, (double *)loss, s);
  }
  if (rc) return rc;
  rc = tdx_adam(p->d.dtype, n, u, (char *)ws + p->off_grad, m, v, step, lr, b1, b2, eps, stream);
  if (rc) return rc;
  return tdx_step_increment(step, stream);
}
