// 'autograd' mode: exact derivatives of the tanh MLP by jets.
//
// The reference differentiates the network with nested
// torch.autograd.grad(create_graph=True) calls, one chain per derivative
// term (Derivative_autograd._nn_autograd, tedeous/derivative.py:73-98), and
// evaluates the operator on every grid point (tedeous/eval.py:117-123).
// Here each point carries a jet -- the value and every partial derivative
// the program reads -- through the layers in one pass:
//   linear layer:  z_a = W h_a (+ b for the value), per jet component a;
//   tanh layer:    h_a = sum over the set partitions P of a of
//                  tanh^(|P|)(z_0) * prod_{B in P} z_B    (Faa di Bruno).
// tanh^(n) is a polynomial in t = tanh(z_0): P_0 = t, P_{n+1} =
// (1 - t^2) P_n'(t).  The adjoint of that map, the transposed linear maps
// and a deterministic chunked reduction of the weight gradients make the
// backward pass; loss and seeds come from the residual program at the
// network's output jets.
//
// Layout in the workspace (T = float or double, J jets, P points):
//   IN0[P][J][d_in]       first layer's input jets (x, unit vectors, zeros)
//   Z_l, A_l[P][J][H]     pre-activation and activation jets, hidden l
//   U[P][J][n_out]        output jets
//   SEED, DU[P][J][n_out] dloss/dU (unscaled) and its dloss-scaled copy
//   DZ_l[P][J][H]         dloss/dZ_l
//   partials of the weight gradients, loss partials, gradient, one.

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>

#include "tdx_common.h"
#include "tdx_nn_dev.h"
#include "tdx_ad_dev.h"
#include "tdx_ad_jit.h"
#include "tdx_nn_jit.h"

using namespace tdx;

namespace {

constexpr int AD_THREADS = 256;
#ifndef TDX_AD_WG_CHUNK
#define TDX_AD_WG_CHUNK 256
#endif
constexpr int AD_WG_CHUNK = TDX_AD_WG_CHUNK;  // rows (point, jet) per reduction chunk

// Copies the jet and partition tables into LDS (the epilogues read them for
// every point and unit); returns a view with the same offsets into the copy.
__device__ __forceinline__ Tables tables_to_lds(const Tables &t, int32_t *lds) {
  const int nj = 2 * t.J, np = t.J + 1, nr = TDX_AD_PART_INTS * t.n_parts;
  const int no = t.J + 1 + 3 * t.n_occ;
  for (int k = threadIdx.x; k < nj + np + nr + no; k += AD_THREADS)
    lds[k] = k < nj ? t.ints[t.jet_i + k]
                    : (k < nj + np ? t.ints[t.pptr_i + k - nj]
                                   : (k < nj + np + nr ? t.ints[t.part_i + k - nj - np]
                                                       : t.occ[k - nj - np - nr]));
  Tables l = t;
  l.ints = lds;
  l.jet_i = 0;
  l.pptr_i = nj;
  l.part_i = nj + np;
  l.occ = lds + nj + np + nr;
  l.tptr_i = 0;
  l.tent_i = t.J + 1;
  return l;
}
__host__ __device__ inline size_t tables_lds_bytes(int J, int n_parts, int n_occ) {
  return (size_t)(2 * J + J + 1 + TDX_AD_PART_INTS * n_parts + J + 1 + 3 * n_occ) * 4;
}

// Copies rows [0, rows) of a row-major [*][ncols] array (row r at src +
// r * ncols, ncols <= 256) into LDS with leading dimension ld, zeros for
// rows r >= valid.  A row's columns go over a power-of-two span of threads
// (no division), several rows a sweep; eight independent loads per thread
// are issued before their stores, so a workgroup waits for HBM once per
// batch instead of once per element.
__device__ __forceinline__ int col_span_log(int ncols) {
  int cl = 0;
  while ((1 << cl) < ncols) ++cl;
  return cl;
}

template <typename T>
__device__ __forceinline__ void stage_rows(T *dst, int ld, const T *src, int rows, int valid,
                                           int ncols) {
  constexpr int NB = 8;
  const int cl = col_span_log(ncols);
  const int c = threadIdx.x & ((1 << cl) - 1), rs = AD_THREADS >> cl;
  for (int rb = threadIdx.x >> cl; rb < rows; rb += rs * NB) {
    T v[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      const int r = rb + u * rs;
      v[u] = (r < valid && c < ncols) ? src[(size_t)r * ncols + c] : T(0);
    }
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      const int r = rb + u * rs;
      if (r < rows && c < ncols) dst[r * ld + c] = v[u];
    }
  }
}

// ---------------------------------------------------------------------------
// forward: one layer
// ---------------------------------------------------------------------------
// A workgroup takes PB points, i.e. R = PB * J rows (point, jet) of the
// layer's input jets, staged in LDS.  The linear map of every row is one
// GEMM, Z[R][dout] = IN[R][din] * W^T, on 16x16x4 MFMA tiles (A operand
// from LDS, B operand from the transposed weights W^T[din][dout] in HBM/L2,
// 64-byte rows per k).  The epilogue adds the bias to the value row and
// forms the activation jets per (point, unit) by Faa di Bruno's formula.
template <typename T>
struct FwdArgs {
  int layer, din, dout, act, J;
  int64_t P;
  int pb, rpad;    // points per workgroup, rows rounded up to 16
  const T *x;      // layer 0: points [P][din]
  const T *in;     // layer > 0: input jets [P][J][din]
  const T *WT, *b; // [din][dout], [dout]
  T *in0;          // layer 0: input jets written here
  T *zout;         // pre-activation jets (act) [P][J][dout]
  T *out;          // activation jets (act) or output jets [P][J][dout]
  Tables tb;
};

template <typename T, int KM, bool SPLIT>
__global__ __launch_bounds__(AD_THREADS) void k_ad_fwd(FwdArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int J = a.J, din = a.din, dout = a.dout;
  const int ldi = din + 1, ldz = dout + 1;
  T *IN = reinterpret_cast<T *>(smem_raw);       // [rpad][ldi]
  T *Z = IN + (size_t)a.rpad * ldi;              // [rpad][ldz]
  const Tables tb = tables_to_lds(a.tb, reinterpret_cast<int32_t *>(Z + (size_t)a.rpad * ldz));
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t p0 = (int64_t)blockIdx.x * a.pb;
  const int npts = (int)std::min<int64_t>(a.pb, a.P - p0);
  const int R = npts * J;
  __syncthreads();
  if (a.layer == 0) {
    for (int k = tid; k < a.rpad * din; k += AD_THREADS) {
      const int r = k / din, j = k - (k / din) * din;
      T v = T(0);
      if (r < R) {
        const int pp = r / J, al = r - (r / J) * J;
        const int ord = tb.ints[tb.jet_i + 2 * al];
        const int ax = tb.ints[tb.jet_i + 2 * al + 1];
        v = al == 0 ? a.x[(p0 + pp) * din + j] : (ord == 1 && ax == j ? T(1) : T(0));
        a.in0[p0 * J * din + k] = v;
      }
      IN[r * ldi + j] = v;
    }
  } else {
    stage_rows(IN, ldi, a.in + p0 * J * din, a.rpad, R, din);
  }
  __syncthreads();
  // GEMM: each wavefront takes column tiles of 16 outputs, holds their B
  // fragments (W^T, all k) in registers and sweeps the row tiles
  // SPLIT (fewer column tiles than wavefronts): (column tile, row-tile
  // group) pairs over the wavefronts, so every wavefront has products
  const int ntc = (dout + 15) >> 4;
  const int nrg = SPLIT ? max(1, min(a.rpad >> 4, (AD_THREADS / 64) / ntc)) : 1;
  const int lr = lane & 15, lk = lane >> 4;
  for (int pc = wave; pc < ntc * nrg; pc += AD_THREADS / 64) {
    const int c = SPLIT ? pc / nrg : pc, rg = SPLIT ? pc - (pc / nrg) * nrg : 0;
    const int n = c * 16 + lr;
    T bf[KM];
#pragma unroll
    for (int u = 0; u < KM; ++u) {
      const int k = 4 * u + lk;
      bf[u] = (4 * u < din && k < din && n < dout) ? a.WT[(size_t)k * dout + n] : T(0);
    }
    for (int r0 = 16 * rg; r0 < a.rpad; r0 += 16 * nrg) {
      typename V4<T>::type acc = {T(0), T(0), T(0), T(0)};
      const T *row = IN + (r0 + lr) * ldi + lk;
#pragma unroll
      for (int u = 0; u < KM; ++u) {
        if (4 * u < din) {
          const T av = 4 * u + lk < din ? row[4 * u] : T(0);
          acc = mma16(av, bf[u], acc);
        }
      }
      if (n < dout) {
#pragma unroll
        for (int i = 0; i < 4; ++i) Z[(r0 + crow<T>(lk, i)) * ldz + n] = acc[i];
      }
    }
  }
  __syncthreads();
  // a uniform loop: every lane runs the same number of rounds
  for (int base = 0; base < npts * dout; base += AD_THREADS) {
    const int item = base + tid;
    if (item >= npts * dout) continue;
    const int pp = item / dout, i = item - (item / dout) * dout;
    const int64_t p = p0 + pp;
    T *zl = Z + (size_t)pp * J * ldz;
    zl[i] = zl[i] + a.b[i];
    if (!a.act) {
      for (int al = 0; al < J; ++al) a.out[((size_t)p * J + al) * dout + i] = zl[al * ldz + i];
      continue;
    }
    for (int al = 0; al < J; ++al) a.zout[((size_t)p * J + al) * dout + i] = zl[al * ldz + i];
    const TanhDerivs<T> s = tanh_derivs(zl[i]);
    a.out[((size_t)p * J) * dout + i] = s.s0;
    for (int al = 1; al < J; ++al) {
      const int k0 = tb.ints[tb.pptr_i + al], k1 = tb.ints[tb.pptr_i + al + 1];
      T h = T(0);
      for (int k = k0; k < k1; ++k) {
        const int32_t *pr = tb.ints + tb.part_i + TDX_AD_PART_INTS * k;
        const int m = pr[0];
        T prod = deriv_n(s.s1, s.s2, s.s3, s.s4, s.s5, m);
#pragma unroll
        for (int q = 0; q < TDX_AD_MAX_ORDER; ++q)
          if (q < m) prod = prod * zl[pr[1 + q] * ldz + i];
        h = (k == k0) ? prod : h + prod;
      }
      a.out[((size_t)p * J + al) * dout + i] = h;
    }
  }
}

// W^T of one layer: WT[k][n] = W[n][k]
template <typename T>
__global__ void k_ad_transpose(const T *W, int dout, int din, T *WT) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= dout * din) return;
  const int n = e / din, k = e - (e / din) * din;
  WT[(size_t)k * dout + n] = W[e];
}

// ---------------------------------------------------------------------------
// residual program, loss and seeds
// ---------------------------------------------------------------------------
template <typename T>
struct ResArgs {
  int J, n_out, n_groups;
  int64_t n_entries;
  const int64_t *gstart;  // entry prefix per group [n_groups + 1]
  const int32_t *ints;
  const T *reals, *data;
  int groups_i;
  const T *U;
  T *seed, *resid, *lp;
  // seed multipliers of group 0's equations, [n_eq][N] (null: none)
  const T *mult;
  int mult_n;
};

template <typename T>
__global__ __launch_bounds__(AD_THREADS) void k_ad_resid(ResArgs<T> a) {
  __shared__ T red[2 * AD_THREADS];
  const int tid = threadIdx.x;
  const int64_t ge = (int64_t)blockIdx.x * AD_THREADS + tid;
  T lsum = T(0), lnsum = T(0);
  if (ge < a.n_entries) {
    int g = 0;
    while (g + 1 < a.n_groups && a.gstart[g + 1] <= ge) ++g;
    const int e = (int)(ge - a.gstart[g]);
    const int32_t *G = a.ints + a.groups_i + TDX_AD_GROUP_INTS * g;
    const int n_sets = G[1], set_i = G[2], n_eq = G[3], eq_i = G[4];
    const int row = a.J * a.n_out;
    for (int s = 0; s < n_sets; ++s) {
      T *sd = a.seed + (size_t)(a.ints[set_i + s] + e) * row;
      for (int k = 0; k < row; ++k) sd[k] = T(0);
    }
    auto factor_value = [&](const int32_t *F) {
      const int n_st = F[0], st_i = F[1], var = F[2];
      T fv = T(0);
      for (int k = 0; k < n_st; ++k) {
        const int32_t *S = a.ints + st_i + TDX_AD_STENCIL_INTS * k;
        const int64_t pt = a.ints[set_i + S[0]] + e;
        const T v = a.U[((size_t)pt * a.J + S[1]) * a.n_out + var];
        const T w = a.reals[S[2]];
        fv = (k == 0) ? w * v : fv + w * v;
      }
      return fv;
    };
    for (int q = 0; q < n_eq; ++q) {
      const int32_t *E = a.ints + eq_i + TDX_AD_EQ_INTS * q;
      const int n_terms = E[0], term_i = E[1];
      const T lw = a.reals[E[2]], nw = a.reals[E[3]], sw = a.reals[E[6]];
      T r = T(0);
      for (int t = 0; t < n_terms; ++t) {
        const int32_t *Tr = a.ints + term_i + TDX_AD_TERM_INTS * t;
        const T c = Tr[2] == 0 ? a.reals[Tr[3]] : a.data[Tr[3] + e];
        T prod = T(1);
        if (term_has_chain(a.ints, Tr[1], Tr[0])) {
          T fv[AD_MAXF], dfv[AD_MAXF];
          for (int f = 0; f < AD_MAXF; ++f)
            fv[f] = f < Tr[0] ? factor_value(a.ints + Tr[1] + TDX_AD_FACTOR_INTS * f) : T(0);
          prod = chain_term(a.ints, Tr[1], Tr[0], a.reals, fv, dfv);
        } else {
          for (int f = 0; f < Tr[0]; ++f) {
            const int32_t *F = a.ints + Tr[1] + TDX_AD_FACTOR_INTS * f;
            prod = prod * fpow(F[3], F[4], a.reals, factor_value(F));
          }
        }
        const T tv = c * prod;
        r = (t == 0) ? tv : r + tv;
      }
      if (a.resid) a.resid[E[5] + e] = r;
      const T diff = E[4] >= 0 ? r - a.data[E[4] + e] : r;
      const T d2 = diff * diff;
      lsum += lw * d2;
      lnsum += nw * d2;
      const T gq = T(2) * sw * diff * ((a.mult && g == 0) ? a.mult[q * a.mult_n + e] : T(1));
      for (int t = 0; t < n_terms; ++t) {
        const int32_t *Tr = a.ints + term_i + TDX_AD_TERM_INTS * t;
        const int nf = Tr[0];
        const T c = Tr[2] == 0 ? a.reals[Tr[3]] : a.data[Tr[3] + e];
        const bool chain = term_has_chain(a.ints, Tr[1], nf);
        T fv[AD_MAXF], pw[AD_MAXF], dfv[AD_MAXF];
        for (int f = 0; f < AD_MAXF; ++f) {
          fv[f] = T(0);
          pw[f] = T(1);
          if (f < nf) {
            const int32_t *F = a.ints + Tr[1] + TDX_AD_FACTOR_INTS * f;
            fv[f] = factor_value(F);
            if (!chain) pw[f] = fpow(F[3], F[4], a.reals, fv[f]);
          }
        }
        if (chain) (void)chain_term(a.ints, Tr[1], nf, a.reals, fv, dfv);
        for (int f = 0; f < nf; ++f) {
          const int32_t *F = a.ints + Tr[1] + TDX_AD_FACTOR_INTS * f;
          T dv = chain ? gq * c * dfv[f] : gq * c * fdpow(F[3], F[4], a.reals, fv[f]);
          for (int f2 = 0; f2 < nf && !chain; ++f2)
            if (f2 != f) dv = dv * pw[f2];
          const int n_st = F[0], st_i = F[1], var = F[2];
          for (int k = 0; k < n_st; ++k) {
            const int32_t *S = a.ints + st_i + TDX_AD_STENCIL_INTS * k;
            const int64_t pt = a.ints[set_i + S[0]] + e;
            a.seed[((size_t)pt * a.J + S[1]) * a.n_out + var] += dv * a.reals[S[2]];
          }
        }
      }
    }
  }
  red[tid] = lsum;
  red[AD_THREADS + tid] = lnsum;
  __syncthreads();
  for (int st = AD_THREADS / 2; st > 0; st >>= 1) {
    if (tid < st) {
      red[tid] += red[tid + st];
      red[AD_THREADS + tid] += red[AD_THREADS + tid + st];
    }
    __syncthreads();
  }
  if (tid == 0) {
    a.lp[2 * blockIdx.x] = red[0];
    a.lp[2 * blockIdx.x + 1] = red[AD_THREADS];
  }
}

// fixed-order sum of the loss partials of n blocks
template <typename T>
__global__ __launch_bounds__(AD_THREADS) void k_ad_loss(const T *lp, int n, T *loss) {
  __shared__ T red[2 * AD_THREADS];
  const int tid = threadIdx.x;
  T s0 = T(0), s1 = T(0);
  for (int k = tid; k < n; k += AD_THREADS) {
    s0 += lp[2 * k];
    s1 += lp[2 * k + 1];
  }
  red[tid] = s0;
  red[AD_THREADS + tid] = s1;
  __syncthreads();
  for (int st = AD_THREADS / 2; st > 0; st >>= 1) {
    if (tid < st) {
      red[tid] += red[tid + st];
      red[AD_THREADS + tid] += red[AD_THREADS + tid + st];
    }
    __syncthreads();
  }
  if (tid == 0) {
    loss[0] = red[0];
    loss[1] = red[AD_THREADS];
  }
}

template <typename T>
__global__ void k_ad_scale(const T *seed, const T *dloss, T *du, int64_t n) {
  const T dl = *dloss;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    du[i] = seed[i] * dl;
}

template <typename T>
__global__ void k_ad_one(T *x) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *x = T(1);
}

// Causal loss (tedeous/losses.py:137-182).  Thread k walks column k of the
// N = nt x K points of group 0 (point p = i K + k): res = sum_eq r_eq[p]^2,
// w = exp(-tol * (sum of res over the earlier slices of the column)).
template <typename T>
__global__ void k_ad_causal_w(const T *resid, int n_eq, int N, int nt, int K, T tol, T *w,
                              T *wr) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  T cum = T(0);
  for (int i = 0; i < nt; ++i) {
    const int p = i * K + k;
    T res = T(0);
    for (int q = 0; q < n_eq; ++q) {
      const T r = resid[(int64_t)q * N + p];
      res = (q == 0) ? r * r : res + r * r;
    }
    const T wv = exp(-tol * cum);
    w[p] = wv;
    wr[p] = wv * res;
    cum = cum + res;
  }
}

// The seeds of group 0's points (points 0 .. N-1, row of J * n_out) times w.
template <typename T>
__global__ void k_ad_causal_seeds(T *seed, const T *w, int N, int row) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= (int64_t)N * row) return;
  seed[i] *= w[i / row];
}

// Weak form (tedeous/eval.py:13-52, 195-221).  R[q][g] = sum over run g
// of c_i r_q[i]^2, one thread per (run, equation), points in order.
template <typename T>
__global__ void k_ad_weak_runs(const T *resid, int n_eq, int N, int G, const int32_t *gstart,
                               const T *c, T *R) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= G * n_eq) return;
  const int q = t / G, g = t - (t / G) * G;
  T s = T(0);
  for (int i = gstart[g]; i < gstart[g + 1]; ++i) {
    const T r = resid[(int64_t)q * N + i];
    s += c[i] * (r * r);
  }
  R[t] = s;
}

// Seed multipliers M[q][i] = c_i (one level) or 2 x_g R[q][g] c_i (two),
// one workgroup per (run, equation).
template <typename T>
__global__ void k_ad_weak_mult(int levels, int N, int G, const int32_t *gstart, const T *c,
                               const T *x, const T *R, T *M) {
  const int g = blockIdx.x % G, q = blockIdx.x / G;
  const T f = levels == 2 ? T(2) * x[g] * R[q * G + g] : T(1);
  for (int i = gstart[g] + threadIdx.x; i < gstart[g + 1]; i += blockDim.x)
    M[(int64_t)q * N + i] = f * c[i];
}

// sum_q lambda_q W_q to the loss, sum_q W_q to the normalized loss, with
// W_q = sum_g R[q][g] (one level: one run) or sum_g x_g R[q][g]^2.
template <typename T>
__global__ __launch_bounds__(AD_THREADS) void k_ad_weak_loss(int levels, int n_eq, int G,
                                                             const T *x, const T *lam,
                                                             const T *R, T *loss) {
  __shared__ T red[AD_THREADS];
  T tot = T(0), totn = T(0);
  for (int q = 0; q < n_eq; ++q) {
    T s = T(0);
    for (int g = threadIdx.x; g < G; g += AD_THREADS) {
      const T v = R[q * G + g];
      s += levels == 2 ? x[g] * (v * v) : v;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = AD_THREADS / 2; st > 0; st >>= 1) {
      if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
      __syncthreads();
    }
    const T W = red[0];
    __syncthreads();
    tot = tot + lam[q] * W;
    totn = totn + W;
  }
  if (threadIdx.x == 0) {
    loss[0] = loss[0] + tot;
    loss[1] = loss[1] + totn;
  }
}

// mean(w * res) in a fixed order, added to both loss values.
template <typename T>
__global__ __launch_bounds__(AD_THREADS) void k_ad_causal_loss(const T *wr, int N, T *loss) {
  __shared__ T red[AD_THREADS];
  T s = T(0);
  for (int p = threadIdx.x; p < N; p += AD_THREADS) s += wr[p];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = AD_THREADS / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const T m = red[0] / T(N);
    loss[0] = loss[0] + m;
    loss[1] = loss[1] + m;
  }
}

// ---------------------------------------------------------------------------
// backward: one hidden layer
// ---------------------------------------------------------------------------
// dloss/d(activation jets) of PB points: dA[R][H] = DN[R][dn] * Wn[dn][H]
// on MFMA tiles (A operand: the next layer's pre-activation adjoints staged
// in LDS; B operand: Wn rows from HBM/L2), then the adjoint of the tanh
// jet map per (point, unit) into DZ.
template <typename T>
struct BwdArgs {
  int H, dn, J;
  int64_t P;
  int pb, rpad;
  const T *Wn;   // next layer's weights [dn][H]
  const T *DN;   // dloss/d(next pre-activation jets) [P][J][dn]
  const T *Z;    // this layer's pre-activation jets [P][J][H]
  T *DZ;         // out: dloss/dZ [P][J][H]
  Tables tb;
};

template <typename T, int KM, bool SPLIT>
__global__ __launch_bounds__(AD_THREADS) void k_ad_bwd(BwdArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int J = a.J, H = a.H, dn = a.dn;
  const int ldd = dn + 1, ldh = H + 1;
  T *DN = reinterpret_cast<T *>(smem_raw);        // [rpad][ldd]
  T *DA = DN + (size_t)a.rpad * ldd;              // [rpad][ldh]
  T *ZS = DA + (size_t)a.rpad * ldh;              // [rpad][ldh]
  const Tables tb = tables_to_lds(a.tb, reinterpret_cast<int32_t *>(ZS + (size_t)a.rpad * ldh));
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t p0 = (int64_t)blockIdx.x * a.pb;
  const int npts = (int)std::min<int64_t>(a.pb, a.P - p0);
  const int R = npts * J;
  stage_rows(DN, ldd, a.DN + p0 * J * dn, a.rpad, R, dn);
  stage_rows(ZS, ldh, a.Z + p0 * J * H, R, R, H);
  __syncthreads();
  const int ntc = (H + 15) >> 4;
  const int nrg = SPLIT ? max(1, min(a.rpad >> 4, (AD_THREADS / 64) / ntc)) : 1;
  const int lr = lane & 15, lk = lane >> 4;
  for (int pc = wave; pc < ntc * nrg; pc += AD_THREADS / 64) {
    const int c = SPLIT ? pc / nrg : pc, rg = SPLIT ? pc - (pc / nrg) * nrg : 0;
    const int n = c * 16 + lr;
    T bf[KM];
#pragma unroll
    for (int u = 0; u < KM; ++u) {
      const int k = 4 * u + lk;
      bf[u] = (4 * u < dn && k < dn && n < H) ? a.Wn[(size_t)k * H + n] : T(0);
    }
    for (int r0 = 16 * rg; r0 < a.rpad; r0 += 16 * nrg) {
      typename V4<T>::type acc = {T(0), T(0), T(0), T(0)};
      const T *row = DN + (r0 + lr) * ldd + lk;
#pragma unroll
      for (int u = 0; u < KM; ++u) {
        if (4 * u < dn) {
          const T av = 4 * u + lk < dn ? row[4 * u] : T(0);
          acc = mma16(av, bf[u], acc);
        }
      }
      if (n < H) {
#pragma unroll
        for (int i = 0; i < 4; ++i) DA[(r0 + crow<T>(lk, i)) * ldh + n] = acc[i];
      }
    }
  }
  __syncthreads();
  for (int base = 0; base < npts * H; base += AD_THREADS) {
    const int item = base + tid;
    if (item >= npts * H) continue;
    const int pp = item / H, j = item - (item / H) * H;
    const int64_t p = p0 + pp;
    const T *zg = ZS + (size_t)pp * J * ldh + j;  // z of jet al at zg[al * ldh]
    const T *da = DA + (size_t)pp * J * ldh + j;
    T *dzo = a.DZ + (size_t)p * J * H + j;
    const TanhDerivs<T> s = tanh_derivs(zg[0]);
    // the value's adjoint: every jet's map depends on z_0 through tanh^(m)
    T dz0 = da[0] * s.s1;
    for (int al = 1; al < J; ++al) {
      const T d = da[al * ldh];
      const int k0 = tb.ints[tb.pptr_i + al], k1 = tb.ints[tb.pptr_i + al + 1];
      for (int k = k0; k < k1; ++k) {
        const int32_t *pr = tb.ints + tb.part_i + TDX_AD_PART_INTS * k;
        const int m = pr[0];
        T all = deriv_n(s.s1, s.s2, s.s3, s.s4, s.s5, m + 1);
#pragma unroll
        for (int q = 0; q < TDX_AD_MAX_ORDER; ++q)
          if (q < m) all = all * zg[pr[1 + q] * ldh];
        dz0 = dz0 + d * all;
      }
    }
    dzo[0] = dz0;
    // jet b's adjoint: over the blocks where b occurs, tanh^(m) times the
    // other blocks of that partition
    for (int b = 1; b < J; ++b) {
      T acc = T(0);
      const int t0 = tb.occ[tb.tptr_i + b], t1 = tb.occ[tb.tptr_i + b + 1];
      for (int t = t0; t < t1; ++t) {
        const int32_t *oc = tb.occ + tb.tent_i + 3 * t;
        const int al = oc[0], q = oc[2];
        const int32_t *pr = tb.ints + tb.part_i + TDX_AD_PART_INTS * oc[1];
        const int m = pr[0];
        T v = da[al * ldh] * deriv_n(s.s1, s.s2, s.s3, s.s4, s.s5, m);
#pragma unroll
        for (int q2 = 0; q2 < TDX_AD_MAX_ORDER; ++q2)
          if (q2 < m && q2 != q) v = v * zg[pr[1 + q2] * ldh];
        acc = acc + v;
      }
      dzo[(size_t)b * H] = acc;
    }
  }
}

// ---------------------------------------------------------------------------
// weight gradients: dW[i][j] = sum_rows D[r][i] S[r][j], db[i] = sum over
// value rows of D[r][i]; rows r = (point, jet).  A workgroup reduces a
// chunk of rows for up to 32 output tiles of 16 x 16 on MFMA (8 per
// wavefront); the chunks' partials are summed in a fixed order after.
// ---------------------------------------------------------------------------
template <typename T>
struct WgArgs {
  int din, dout, J;
  int64_t K;      // rows
  const T *D;     // [K][dout]
  const T *S;     // [K][din]
  T *part;        // [n_chunks][dout * (din + 1)]
};

constexpr int AD_WG_TPW = 8;   // output tiles per wavefront

template <typename T>
__global__ __launch_bounds__(AD_THREADS) void k_ad_wgrad(WgArgs<T> a) {
  constexpr int RT = sizeof(T) == 4 ? 16 : 8;   // rows per LDS stage (33 KB)
  __shared__ T Ds[RT * 257];
  __shared__ T Ss[RT * 258];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int din = a.din, dout = a.dout, d1 = din + 1, ne = dout * d1;
  const int ntn = (d1 + 15) >> 4, nt = ((dout + 15) >> 4) * ntn;
  const int64_t r0 = (int64_t)blockIdx.x * AD_WG_CHUNK;
  const int64_t r1 = std::min<int64_t>(a.K, r0 + AD_WG_CHUNK);
  const int t0 = blockIdx.y * (AD_THREADS / 64) * AD_WG_TPW;
  const int lr = lane & 15, lk = lane >> 4;
  typename V4<T>::type acc[AD_WG_TPW];
#pragma unroll
  for (int u = 0; u < AD_WG_TPW; ++u) acc[u] = {T(0), T(0), T(0), T(0)};
  for (int64_t rt = r0; rt < r1; rt += RT) {
    const int nr = (int)std::min<int64_t>(RT, r1 - rt);
    __syncthreads();
    stage_rows(Ds, 257, a.D + rt * dout, RT, nr, dout);
    stage_rows(Ss, 258, a.S + rt * din, RT, nr, din);
    // the bias column: 1 on value rows
    if (tid < RT) Ss[tid * 258 + din] = tid < nr && ((rt + tid) % a.J) == 0 ? T(1) : T(0);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < AD_WG_TPW; ++u) {
      const int t = t0 + wave + (AD_THREADS / 64) * u;
      if (t >= nt) continue;
      const int i = (t / ntn) * 16 + lr, j = (t - (t / ntn) * ntn) * 16 + lr;
#pragma unroll
      for (int k0 = 0; k0 < RT; k0 += 4) {
        const int r = k0 + lk;
        const T av = i < dout ? Ds[r * 257 + i] : T(0);
        const T bv = j < d1 ? Ss[r * 258 + j] : T(0);
        acc[u] = mma16(av, bv, acc[u]);
      }
    }
  }
#pragma unroll
  for (int u = 0; u < AD_WG_TPW; ++u) {
    const int t = t0 + wave + (AD_THREADS / 64) * u;
    if (t >= nt) continue;
    const int j = (t - (t / ntn) * ntn) * 16 + lr;
    if (j >= d1) continue;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int i = (t / ntn) * 16 + crow<T>(lk, q);
      if (i < dout) a.part[(size_t)blockIdx.x * ne + i * d1 + j] = acc[u][q];
    }
  }
}

// Sum of the chunks' partials: 64 elements per workgroup, each wavefront
// sums a quarter of the chunks in order, then the quarters are added in order.
template <typename T>
__global__ __launch_bounds__(AD_THREADS) void k_ad_wreduce(const T *part, int n_chunks, int din,
                                                           int dout, T *gw, T *gb) {
  __shared__ T red[AD_THREADS];
  const int ne = dout * (din + 1);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int e = blockIdx.x * 64 + lane;
  const int q = (n_chunks + 3) / 4;
  const int c0 = wave * q, c1 = min(n_chunks, c0 + q);
  // eight independent sums (loads in flight), combined in a fixed order
  T acc[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) acc[u] = T(0);
  if (e < ne) {
    int c = c0;
    for (; c + 8 <= c1; c += 8) {
#pragma unroll
      for (int u = 0; u < 8; ++u) acc[u] += part[(size_t)(c + u) * ne + e];
    }
    for (; c < c1; ++c) acc[0] += part[(size_t)c * ne + e];
  }
  const T s = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
  red[threadIdx.x] = s;
  __syncthreads();
  if (wave != 0 || e >= ne) return;
  const T tot = ((red[lane] + red[64 + lane]) + red[128 + lane]) + red[192 + lane];
  const int i = e / (din + 1), j = e - (e / (din + 1)) * (din + 1);
  if (j < din)
    gw[i * din + j] = tot;
  else
    gb[i] = tot;
}

// ---------------------------------------------------------------------------
// per-residual Jacobian (natural-gradient Gram assembly)
// ---------------------------------------------------------------------------
// The reference forms J with one torch.autograd.grad per residual
// (tedeous/optimizers/ngd.py:57-78).  One thread owns one (group, entry):
// per equation it forms the seeds dr/dU(set, jet, var) of the residual
// program with dr = 1 (the k_ad_resid seeds), and for every set with a
// non-zero seed pushes that point's jets through the network and the jet
// maps' adjoints back, adding seed * dU/dparams into its own column r of
// jt[P][n_resid].  Per-thread state lives in LDS, thread-fastest:
// pre-activation and activation jets of every hidden layer, two adjoint
// buffers, the output jets of every set and the seeds.
template <typename T>
struct JacArgs {
  int d_in, n_out, H, L, J, n_groups, groups_i, max_sets;
  const int32_t *ints;
  const T *reals, *data, *points, *params;
  const int64_t *gstart;
  Tables tb;
  const int64_t *off_w, *off_b;   // parameter offsets per layer (device)
  T *jt;
  int64_t n_resid;
};

template <typename T>
__global__ void k_ad_jac(JacArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *smem = reinterpret_cast<T *>(smem_raw);
  const int B = blockDim.x, t = threadIdx.x;
  const int J = a.J, H = a.H, L = a.L, NL = L - 1, no = a.n_out, din = a.d_in;
  const int JH = J * H;
  // LDS per thread (element k at [k * B + t]): Z[NL][J][H] | A[NL][J][H] |
  // D0[J][H] | D1[J][H] | U[max_sets][J][no] | S[max_sets][J][no] | X[J][din]
  T *Zs = smem;
  T *As = Zs + (size_t)NL * JH * B;
  T *D0 = As + (size_t)NL * JH * B;
  T *D1 = D0 + (size_t)JH * B;
  T *Us = D1 + (size_t)JH * B;
  T *Ss = Us + (size_t)a.max_sets * J * no * B;
  T *Xs = Ss + (size_t)a.max_sets * J * no * B;
#define LV(base, k) base[(size_t)(k) * B + t]
  const int64_t ge = (int64_t)blockIdx.x * B + t;
  const int64_t total = a.gstart[a.n_groups];
  if (ge >= total) return;
  int g = 0;
  while (g + 1 < a.n_groups && a.gstart[g + 1] <= ge) ++g;
  const int e = (int)(ge - a.gstart[g]);
  const int32_t *G = a.ints + a.groups_i + TDX_AD_GROUP_INTS * g;
  const int n_sets = G[1], set_i = G[2], n_eq = G[3], eq_i = G[4];
  const Tables &tb = a.tb;
  const T *W = a.params;

  // jets of point pt through the network: Z, A of every hidden layer, and
  // (with u_out) the output jets
  auto forward = [&](int64_t pt, T *u_out) {
    for (int al = 0; al < J; ++al) {
      const int ord = tb.ints[tb.jet_i + 2 * al], ax = tb.ints[tb.jet_i + 2 * al + 1];
      for (int j = 0; j < din; ++j)
        LV(Xs, al * din + j) = al == 0 ? a.points[pt * din + j] : (ord == 1 && ax == j ? T(1) : T(0));
    }
    for (int l = 0; l < NL; ++l) {
      const int dl = l == 0 ? din : H;
      const T *Wl = W + a.off_w[l], *bl = W + a.off_b[l];
      T *Z = Zs + (size_t)l * JH * B, *A = As + (size_t)l * JH * B;
      const T *In = l == 0 ? Xs : As + (size_t)(l - 1) * JH * B;
      for (int i = 0; i < H; ++i) {
        for (int al = 0; al < J; ++al) {
          T z = al == 0 ? bl[i] : T(0);
          for (int j = 0; j < dl; ++j) z = fma(Wl[i * dl + j], LV(In, al * dl + j), z);
          LV(Z, al * H + i) = z;
        }
        const TanhDerivs<T> sd = tanh_derivs(LV(Z, i));
        LV(A, i) = sd.s0;
        for (int al = 1; al < J; ++al) {
          const int k0 = tb.ints[tb.pptr_i + al], k1 = tb.ints[tb.pptr_i + al + 1];
          T h = T(0);
          for (int k = k0; k < k1; ++k) {
            const int32_t *pr = tb.ints + tb.part_i + TDX_AD_PART_INTS * k;
            const int m = pr[0];
            T prod = deriv_n(sd.s1, sd.s2, sd.s3, sd.s4, sd.s5, m);
            for (int q = 0; q < m; ++q) prod = prod * LV(Z, pr[1 + q] * H + i);
            h = (k == k0) ? prod : h + prod;
          }
          LV(A, al * H + i) = h;
        }
      }
    }
    if (u_out) {
      const T *Wl = W + a.off_w[L - 1], *bl = W + a.off_b[L - 1];
      const T *In = As + (size_t)(NL - 1) * JH * B;
      for (int o = 0; o < no; ++o)
        for (int al = 0; al < J; ++al) {
          T u = al == 0 ? bl[o] : T(0);
          for (int j = 0; j < H; ++j) u = fma(Wl[o * H + j], LV(In, al * H + j), u);
          LV(u_out, al * no + o) = u;
        }
    }
  };

  for (int s = 0; s < n_sets; ++s) forward(a.ints[set_i + s] + e, Us + (size_t)s * J * no * B);

  for (int q = 0; q < n_eq; ++q) {
    const int32_t *E = a.ints + eq_i + TDX_AD_EQ_INTS * q;
    const int n_terms = E[0], term_i = E[1];
    if (E[5] < 0) continue;
    T *col = a.jt + E[5] + e;
    for (int k = 0; k < n_sets * J * no; ++k) LV(Ss, k) = T(0);
    auto factor_value = [&](const int32_t *F) {
      T fv = T(0);
      for (int k = 0; k < F[0]; ++k) {
        const int32_t *S = a.ints + F[1] + TDX_AD_STENCIL_INTS * k;
        const T v = LV(Us, (S[0] * J + S[1]) * no + F[2]);
        fv = (k == 0) ? a.reals[S[2]] * v : fv + a.reals[S[2]] * v;
      }
      return fv;
    };
    // dr/dU through every factor of every term
    for (int tt = 0; tt < n_terms; ++tt) {
      const int32_t *Tr = a.ints + term_i + TDX_AD_TERM_INTS * tt;
      const int nf = Tr[0];
      const T c = Tr[2] == 0 ? a.reals[Tr[3]] : a.data[Tr[3] + e];
      const bool chain = term_has_chain(a.ints, Tr[1], nf);
      T fv[AD_MAXF], pw[AD_MAXF], dfv[AD_MAXF];
      for (int f = 0; f < AD_MAXF; ++f) {
        fv[f] = T(0);
        pw[f] = T(1);
        if (f < nf) {
          const int32_t *F = a.ints + Tr[1] + TDX_AD_FACTOR_INTS * f;
          fv[f] = factor_value(F);
          if (!chain) pw[f] = fpow(F[3], F[4], a.reals, fv[f]);
        }
      }
      if (chain) (void)chain_term(a.ints, Tr[1], nf, a.reals, fv, dfv);
      for (int f = 0; f < nf; ++f) {
        const int32_t *F = a.ints + Tr[1] + TDX_AD_FACTOR_INTS * f;
        T dv = chain ? c * dfv[f] : c * fdpow(F[3], F[4], a.reals, fv[f]);
        for (int f2 = 0; f2 < nf && !chain; ++f2)
          if (f2 != f) dv = dv * pw[f2];
        for (int k = 0; k < F[0]; ++k) {
          const int32_t *S = a.ints + F[1] + TDX_AD_STENCIL_INTS * k;
          LV(Ss, (S[0] * J + S[1]) * no + F[2]) += dv * a.reals[S[2]];
        }
      }
    }
    auto add = [&](int64_t pidx, T v) { col[pidx * a.n_resid] += v; };
    for (int s = 0; s < n_sets; ++s) {
      const T *Sd = Ss + (size_t)s * J * no * B;
      bool any = false;
      for (int k = 0; k < J * no; ++k) any |= LV(Sd, k) != T(0);
      if (!any) continue;
      forward(a.ints[set_i + s] + e, nullptr);
      // output layer: dW[o][j] += sum_al dU[al][o] A[al][j]; dA = W^T dU
      {
        const T *Wl = W + a.off_w[L - 1];
        const T *Al = As + (size_t)(NL - 1) * JH * B;
        for (int o = 0; o < no; ++o) {
          for (int j = 0; j < H; ++j) {
            T g = T(0);
            for (int al = 0; al < J; ++al) g = fma(LV(Sd, al * no + o), LV(Al, al * H + j), g);
            add(a.off_w[L - 1] + o * H + j, g);
          }
          add(a.off_b[L - 1] + o, LV(Sd, o));
        }
        for (int al = 0; al < J; ++al)
          for (int j = 0; j < H; ++j) {
            T g = T(0);
            for (int o = 0; o < no; ++o) g = fma(Wl[o * H + j], LV(Sd, al * no + o), g);
            LV(D0, al * H + j) = g;
          }
      }
      T *dA = D0, *dZ = D1;
      for (int l = NL - 1; l >= 0; --l) {
        const T *Z = Zs + (size_t)l * JH * B;
        // adjoint of the tanh jet map (unit i): dZ from dA
        for (int i = 0; i < H; ++i) {
          const TanhDerivs<T> sd = tanh_derivs(LV(Z, i));
          T dz0 = LV(dA, i) * sd.s1;
          for (int al = 1; al < J; ++al) {
            const T dd = LV(dA, al * H + i);
            const int k0 = tb.ints[tb.pptr_i + al], k1 = tb.ints[tb.pptr_i + al + 1];
            for (int k = k0; k < k1; ++k) {
              const int32_t *pr = tb.ints + tb.part_i + TDX_AD_PART_INTS * k;
              const int m = pr[0];
              T all = deriv_n(sd.s1, sd.s2, sd.s3, sd.s4, sd.s5, m + 1);
              for (int q2 = 0; q2 < m; ++q2) all = all * LV(Z, pr[1 + q2] * H + i);
              dz0 = dz0 + dd * all;
            }
          }
          LV(dZ, i) = dz0;
          for (int b = 1; b < J; ++b) {
            T acc = T(0);
            const int t0 = tb.occ[tb.tptr_i + b], t1 = tb.occ[tb.tptr_i + b + 1];
            for (int tt = t0; tt < t1; ++tt) {
              const int32_t *oc = tb.occ + tb.tent_i + 3 * tt;
              const int al = oc[0], qq = oc[2];
              const int32_t *pr = tb.ints + tb.part_i + TDX_AD_PART_INTS * oc[1];
              const int m = pr[0];
              T v = LV(dA, al * H + i) * deriv_n(sd.s1, sd.s2, sd.s3, sd.s4, sd.s5, m);
              for (int q2 = 0; q2 < m; ++q2)
                if (q2 != qq) v = v * LV(Z, pr[1 + q2] * H + i);
              acc = acc + v;
            }
            LV(dZ, b * H + i) = acc;
          }
        }
        // linear layer l: dW[i][j] += sum_al dZ[al][i] In[al][j]; db += dZ[0]
        const int dl = l == 0 ? din : H;
        const T *In = l == 0 ? Xs : As + (size_t)(l - 1) * JH * B;
        const T *Wl = W + a.off_w[l];
        for (int i = 0; i < H; ++i) {
          for (int j = 0; j < dl; ++j) {
            T g = T(0);
            for (int al = 0; al < J; ++al) g = fma(LV(dZ, al * H + i), LV(In, al * dl + j), g);
            add(a.off_w[l] + i * dl + j, g);
          }
          add(a.off_b[l] + i, LV(dZ, i));
        }
        if (l > 0) {
          // dA of the layer below: W^T dZ
          for (int al = 0; al < J; ++al)
            for (int j = 0; j < H; ++j) {
              T g = T(0);
              for (int i = 0; i < H; ++i) g = fma(Wl[i * H + j], LV(dZ, al * H + i), g);
              LV(dA, al * H + j) = g;
            }
        }
      }
    }
  }
#undef LV
}

template <typename T>
__global__ void k_ad_jac_zero(T *jt, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    jt[i] = T(0);
}


// k_ad_step: the fused training step with the jet maps and the residual
// program interpreted from the plan's tables.
template <int NB, int NHM>
__global__ __launch_bounds__(256) void k_ad_step(ADStepArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  ad_step_body<NB, NHM, InterpJets, InterpADResid>(a, reinterpret_cast<float *>(smem_raw));
}

// Image of the network for k_ad_step (one instance) and the Adam step
// counter advanced (the step's Adam update reads the new value).
template <int NB>
__global__ void k_ad_pack(NNArgs<float> a, int64_t *step) {
  constexpr int HP = 16 * NB;
  if (blockIdx.x == 0 && threadIdx.x == 0) *step += 1;
  const int IMG = image_words(HP, a.NH, a.d_in, a.n_out);
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < IMG; t += gridDim.x * blockDim.x)
    a.packed[t] = image_value<float, NB>(a, a.params, t);
}

// Fixed-order sums of the items' gradient partials and loss pairs, and the
// Adam update at step *step (already advanced).  Workgroup 0 forms the loss
// pair; the others one parameter per thread.
__global__ __launch_bounds__(256) void k_ad_step_sum(const float *partial, const float *lp,
                                                     int n_items, int64_t P, float *params,
                                                     float *m, float *v, const int64_t *step,
                                                     AdamHyper h, float *loss) {
  if (blockIdx.x == 0) {
    __shared__ float r[512];
    float s0 = 0.f, s1 = 0.f;
    for (int k = threadIdx.x; k < n_items; k += 256) {
      s0 += lp[2 * k];
      s1 += lp[2 * k + 1];
    }
    r[threadIdx.x] = s0;
    r[256 + threadIdx.x] = s1;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
      if ((int)threadIdx.x < st) {
        r[threadIdx.x] += r[threadIdx.x + st];
        r[256 + threadIdx.x] += r[256 + threadIdx.x + st];
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      loss[0] = r[0];
      loss[1] = r[256];
    }
    return;
  }
  const int64_t p = (int64_t)(blockIdx.x - 1) * 256 + threadIdx.x;
  if (p >= P) return;
  float acc[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) acc[u] = 0.f;
  int k = 0;
  for (; k + 8 <= n_items; k += 8) {
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] += partial[(int64_t)(k + u) * P + p];
  }
  for (; k < n_items; ++k) acc[0] += partial[(int64_t)k * P + p];
  const float g = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
  float pv = params[p], mv = m[p], vv = v[p];
  adam_apply(pv, g, mv, vv, adam_scalars<float>(*step, h));
  params[p] = pv;
  m[p] = mv;
  v[p] = vv;
}

}  // namespace

// ---------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------
struct tdx_ad_plan {
  int dtype, d_in, n_out, H, L, J;
  int64_t P, n_entries, n_resid, n_params;
  int n_groups, groups_i, jet_i, pptr_i, part_i, n_parts, n_occ;
  // causal loss: tolerance, slices and points of group 0 (N = nt K), its
  // equations; workspace for residuals, weights and w * res
  int causal = 0, c_nt = 0, c_K = 0, c_N = 0, c_neq = 0;
  double causal_tol = 0;
  size_t off_resid = 0, off_cw = 0, off_cwr = 0;
  // weak form: levels (0 none), runs, group 0's points and equations, table
  // offsets; workspace for the runs' sums and the seed multipliers
  int weak = 0, w_G = 0, w_N = 0, w_neq = 0;
  int64_t w_d = 0, w_i = 0;
  size_t off_wR = 0, off_wM = 0;
  DevTable ints, reals, data, points, gstart, occ;
  std::vector<int64_t> off_w, off_b;
  DevTable offs;     // off_w | off_b on the device (the Jacobian kernel)
  int max_sets = 0;  // most point sets of a group
  std::vector<int> din, dout;
  // launch shapes per layer
  std::vector<int> f_pb, f_rpad, b_pb, b_rpad;  // points per workgroup, rows rounded to 16
  std::vector<size_t> f_lds, b_lds;
  int n_res_blocks;
  int64_t n_chunks;
  // workspace offsets (bytes)
  size_t off_in0, off_u, off_seed, off_du, off_lp, off_part, off_grad, off_one, total;
  std::vector<size_t> off_z, off_a, off_dz, off_wt;
  // fused training step (k_ad_step): 16-column blocks, H x H layers of the
  // instantiation, LDS bytes, work items (group, first entry, end entry)
  // and its workspace: network image, pre-activation scratch, gradient
  // partials and loss pairs per item
  int fused = 0, f_nb = 0, f_nhm = 0, n_fitems = 0;
  size_t step_lds = 0;
  DevTable fitems;
  hipFunction_t step_fn = nullptr;  // the kernel compiled with generated code, or null
  size_t off_img = 0, off_zs = 0, off_fpart = 0, off_flp = 0;
};

namespace {

int rows16(int pb, int J) { return (pb * J + 15) / 16 * 16; }
size_t ad_fwd_lds(int din, int dout, int J, int pb, size_t ts) {
  return (size_t)rows16(pb, J) * ((din + 1) + (dout + 1)) * ts;
}
size_t ad_bwd_lds(int H, int dn, int J, int pb, size_t ts) {
  return (size_t)rows16(pb, J) * ((dn + 1) + 2 * (H + 1)) * ts;
}

// LDS budgets of the layer kernels' workgroups
#ifndef TDX_AD_LDS_FWD
#define TDX_AD_LDS_FWD (40 * 1024)
#endif
#ifndef TDX_AD_LDS_BWD
#define TDX_AD_LDS_BWD (40 * 1024)
#endif

// points per workgroup: at most 64, within the budget when one point fits,
// else within 160 KB; 0 if one point does not fit
template <typename F>
int pick_pb(F lds, size_t budget) {
  int pb = 64;
  while (pb > 1 && lds(pb) > budget) --pb;
  if (lds(pb) > 160 * 1024) return 0;
  return pb;
}

int fail(const char *msg) {
  set_error("autograd plan: %s", msg);
  return TDX_EINVAL;
}

}  // namespace

namespace {

// Checks every table of the descriptor (record ranges, jets, partitions,
// groups, equations, options); on success the entry prefix per group and
// the most point sets of a group.
int validate_ad(const tdx_ad_desc *d, std::vector<int64_t> *gstart_out, int *max_sets_out) {
  if (d->dtype != 0 && d->dtype != 1) return fail("dtype must be 0 or 1");
  if (d->d_in < 1 || d->n_out < 1 || d->width < 1 || d->n_layers < 2) return fail("network shape");
  if (d->width > 256 || d->n_out > 256 || d->d_in > 256) return fail("widths above 256");
  if (d->n_jets < 1 || d->n_points < 1 || d->n_groups < 1) return fail("empty problem");
  if (!d->ints || !d->reals || !d->points || d->n_ints < 1) return fail("null table");
  const int64_t NI = d->n_ints;
  const int32_t *I = d->ints;
  auto in_ints = [&](int64_t i, int64_t n) { return i >= 0 && n >= 0 && i + n <= NI; };
  auto in_reals = [&](int64_t i) { return i >= 0 && i < d->n_reals; };
  auto in_data = [&](int64_t i, int64_t n) { return i >= 0 && n >= 0 && i + n <= d->n_data; };
  const int J = d->n_jets;
  // jets and partitions
  if (!in_ints(d->jet_i, 2 * (int64_t)J) || !in_ints(d->pptr_i, J + 1)) return fail("jet table");
  if (I[d->jet_i] != 0) return fail("jet 0 must be the value");
  for (int j = 0; j < J; ++j) {
    const int ord = I[d->jet_i + 2 * j], ax = I[d->jet_i + 2 * j + 1];
    if (ord < 0 || ord > TDX_AD_MAX_ORDER) return fail("jet order above 4");
    if (ord == 1 && (ax < 0 || ax >= d->d_in)) return fail("jet axis");
    if (j > 0 && ord == 0) return fail("only jet 0 has order 0");
    const int k0 = I[d->pptr_i + j], k1 = I[d->pptr_i + j + 1];
    if (k0 < 0 || k1 < k0) return fail("partition pointers");
    if (j == 0 && k1 != k0) return fail("the value has no partitions");
    if (j > 0 && k1 == k0) return fail("a derivative jet needs partitions");
    if (!in_ints(d->part_i + (int64_t)TDX_AD_PART_INTS * k0, (int64_t)TDX_AD_PART_INTS * (k1 - k0)))
      return fail("partition records");
    for (int k = k0; k < k1; ++k) {
      const int32_t *pr = I + d->part_i + TDX_AD_PART_INTS * k;
      if (pr[0] < 1 || pr[0] > ord) return fail("partition block count");
      for (int q = 0; q < pr[0]; ++q) {
        const int b = pr[1 + q];
        if (b < 1 || b >= J || I[d->jet_i + 2 * b] >= ord + (pr[0] == 1 ? 1 : 0))
          return fail("partition block jet");
        if (pr[0] == 1 && b != j) return fail("one-block partition must be the jet itself");
      }
    }
  }
  // groups
  if (!in_ints(d->groups_i, (int64_t)TDX_AD_GROUP_INTS * d->n_groups)) return fail("group table");
  std::vector<int64_t> gstart(d->n_groups + 1, 0);
  std::vector<char> owned(d->n_points, 0);
  int max_sets = 0;
  int64_t n_resid_used = 0;
  for (int g = 0; g < d->n_groups; ++g) {
    const int32_t *G = I + d->groups_i + TDX_AD_GROUP_INTS * g;
    const int ne = G[0], ns = G[1], set_i = G[2], neq = G[3], eq_i = G[4];
    if (ne < 0 || ns < 1 || neq < 0) return fail("group record");
    if (!in_ints(set_i, ns) || !in_ints(eq_i, (int64_t)TDX_AD_EQ_INTS * neq)) return fail("group ranges");
    for (int s = 0; s < ns; ++s) {
      const int64_t off = I[set_i + s];
      if (off < 0 || off + ne > d->n_points) return fail("set points out of range");
      for (int e = 0; e < ne; ++e) {
        if (owned[off + e]) return fail("a point belongs to two entries");
        owned[off + e] = 1;
      }
    }
    for (int q = 0; q < neq; ++q) {
      const int32_t *E = I + eq_i + TDX_AD_EQ_INTS * q;
      const int nt = E[0], ti = E[1];
      if (nt < 1 || !in_ints(ti, (int64_t)TDX_AD_TERM_INTS * nt)) return fail("equation terms");
      if (!in_reals(E[2]) || !in_reals(E[3]) || !in_reals(E[6])) return fail("equation weights");
      if (E[4] >= 0 && !in_data(E[4], ne)) return fail("equation target");
      if (E[5] < 0 || E[5] + (int64_t)ne > d->n_resid) return fail("residual offset");
      n_resid_used = std::max<int64_t>(n_resid_used, E[5] + (int64_t)ne);
      for (int t = 0; t < nt; ++t) {
        const int32_t *Tr = I + ti + TDX_AD_TERM_INTS * t;
        const int nf = Tr[0], fi = Tr[1];
        if (nf < 1 || nf > AD_MAXF || !in_ints(fi, (int64_t)TDX_AD_FACTOR_INTS * nf))
          return fail("term factors (1 to 4 per term)");
        if (Tr[2] == 0 ? !in_reals(Tr[3]) : (Tr[2] != 1 || !in_data(Tr[3], ne)))
          return fail("term coefficient");
        for (int f = 0; f < nf; ++f) {
          const int32_t *F = I + fi + TDX_AD_FACTOR_INTS * f;
          const int nst = F[0], si = F[1], var = F[2];
          if (nst < 1 || !in_ints(si, (int64_t)TDX_AD_STENCIL_INTS * nst)) return fail("factor stencil");
          if (var < 0 || var >= d->n_out) return fail("factor variable");
          if (F[3] == 2) {
            // callable power: a chain of one-argument operations
            if (!in_ints(F[4], 1)) return fail("factor power chain");
            const int n = I[F[4]];
            if (n < 0 || n > TDX_AD_MAX_CHAIN || !in_ints(F[4] + 1, 2 * (int64_t)n))
              return fail("factor power chain length");
            for (int k = 0; k < n; ++k) {
              const int op = I[F[4] + 1 + 2 * k], arg = I[F[4] + 2 + 2 * k];
              if (op < TDX_AD_OP_ADD || op > TDX_AD_OP_SQRT) return fail("factor power chain op");
              if ((op <= TDX_AD_OP_RDIV || op == TDX_AD_OP_POWR) && !in_reals(arg))
                return fail("factor power chain constant");
            }
          } else if (F[3] == 1 ? !in_reals(F[4]) : F[3] != 0) {
            return fail("factor power");
          }
          for (int k = 0; k < nst; ++k) {
            const int32_t *S = I + si + TDX_AD_STENCIL_INTS * k;
            if (S[0] < 0 || S[0] >= ns || S[1] < 0 || S[1] >= J || !in_reals(S[2]))
              return fail("stencil entry");
          }
        }
      }
    }
    gstart[g + 1] = gstart[g] + ne;
    max_sets = std::max(max_sets, ns);
  }
  if (gstart[d->n_groups] < 1) return fail("no entries");
  if (d->weak_levels != 0) {
    const int32_t *G0 = I + d->groups_i;
    const int N = G0[0], G = d->weak_runs, neq = G0[3];
    if (d->weak_levels != 1 && d->weak_levels != 2) return fail("weak form: 1 or 2 levels");
    if (d->causal_tol != 0.0) return fail("a weak form with a causal loss");
    if (G < 1 || (d->weak_levels == 1 && G != 1)) return fail("weak form: runs");
    if (G0[1] != 1 || I[G0[2]] != 0) return fail("weak form: group 0 must hold points 0 .. N-1");
    if (!in_data(d->weak_d, (int64_t)N + G + neq)) return fail("weak form: data range");
    if (!in_ints(d->weak_i, G + 1)) return fail("weak form: run table range");
    const int32_t *gs = I + d->weak_i;
    if (gs[0] != 0 || gs[G] != N) return fail("weak form: runs must cover group 0");
    for (int g = 0; g < G; ++g)
      if (gs[g + 1] <= gs[g]) return fail("weak form: empty run");
    for (int q = 0; q < neq; ++q)
      if (I[G0[4] + TDX_AD_EQ_INTS * q + 5] != q * N)
        return fail("weak form: group 0's residuals must lead the residual output");
  }
  if (d->causal_tol != 0.0) {
    const int32_t *G0 = I + d->groups_i;
    if (!std::isfinite(d->causal_tol)) return fail("causal tol must be finite");
    if (G0[1] != 1 || I[G0[2]] != 0) return fail("causal loss: group 0 must hold points 0 .. N-1");
    if (d->causal_nt < 1 || G0[0] % d->causal_nt != 0) return fail("causal slices must divide group 0");
    for (int q = 0; q < G0[3]; ++q)
      if (I[G0[4] + TDX_AD_EQ_INTS * q + 5] != q * G0[0])
        return fail("causal loss: group 0's residuals must lead the residual output");
  }
  if (gstart[d->n_groups] > (int64_t)1 << 31) return fail("more than 2^31 entries");

  *gstart_out = gstart;
  *max_sets_out = max_sets;
  return TDX_OK;
}

// Where each jet occurs as a block of a partition: tptr[J + 1], then
// (alpha, partition, block) entries.
std::vector<int32_t> ad_occurrences(const tdx_ad_desc *d) {
  const int32_t *I = d->ints;
  const int J = d->n_jets;
  std::vector<std::vector<int32_t>> occ_of(J);
  for (int al = 1; al < J; ++al)
    for (int k = I[d->pptr_i + al]; k < I[d->pptr_i + al + 1]; ++k) {
      const int32_t *pr = I + d->part_i + TDX_AD_PART_INTS * k;
      for (int q = 0; q < pr[0]; ++q) occ_of[pr[1 + q]].insert(occ_of[pr[1 + q]].end(), {al, k, q});
    }
  std::vector<int32_t> occ(J + 1, 0);
  for (int b = 0; b < J; ++b) occ[b + 1] = occ[b] + (int)occ_of[b].size() / 3;
  for (int b = 0; b < J; ++b) occ.insert(occ.end(), occ_of[b].begin(), occ_of[b].end());
  return occ;
}

// Whether the fused step (k_ad_step) serves a validated descriptor, and its
// instantiation: 16-column blocks and H x H layers.
bool ad_fused_ok(const tdx_ad_desc *d, int *nb_out, int *nhm_out) {
  const int J = d->n_jets, NH = d->n_layers - 2;
  const int nb = (d->width + 15) / 16;
  const int nhm = nb <= 7 ? 4 : 3;
  bool ok = d->dtype == 0 && d->causal_tol == 0.0 && d->weak_levels == 0 && J <= ADF_MAXJ &&
            d->n_out <= MAXO && d->d_in <= MAXD && nb <= 8 && NH <= nhm &&
            ad_step_lds_bytes(16 * nb) <= 160 * 1024;
  const char *env = std::getenv("TDX_AD_FUSED");
  if (env && env[0] == '0') ok = false;
  for (int g = 0; ok && g < d->n_groups; ++g)
    if (d->ints[d->groups_i + TDX_AD_GROUP_INTS * g + 1] * J > 16) ok = false;
  *nb_out = nb;
  *nhm_out = nhm;
  return ok;
}

AdJitSpec ad_jit_spec(const tdx_ad_desc *d, const std::vector<int32_t> &occ, int nb, int nhm) {
  AdJitSpec sp;
  sp.ints = d->ints;
  sp.n_ints = d->n_ints;
  sp.J = d->n_jets;
  sp.pptr_i = d->pptr_i;
  sp.part_i = d->part_i;
  sp.groups_i = d->groups_i;
  sp.n_groups = d->n_groups;
  sp.n_out = d->n_out;
  sp.occ = occ;
  sp.NB = nb;
  sp.NHM = nhm;
  return sp;
}

}  // namespace

extern "C" int tdx_ad_plan_create(const tdx_ad_desc *d, tdx_ad_plan **out) {
  TDX_REQUIRE(d && out, "null argument");
  *out = nullptr;
  std::vector<int64_t> gstart;
  int max_sets = 0;
  {
    const int rc = validate_ad(d, &gstart, &max_sets);
    if (rc != TDX_OK) return rc;
  }
  const int32_t *I = d->ints;
  const int J = d->n_jets;
  const size_t ts = d->dtype == 0 ? 4 : 8;
  auto *p = new tdx_ad_plan();
  p->dtype = d->dtype;
  p->d_in = d->d_in;
  p->n_out = d->n_out;
  p->H = d->width;
  p->L = d->n_layers;
  p->J = J;
  p->P = d->n_points;
  p->n_entries = gstart[d->n_groups];
  p->n_resid = d->n_resid;
  p->max_sets = max_sets;
  p->n_groups = d->n_groups;
  p->groups_i = d->groups_i;
  p->jet_i = d->jet_i;
  p->pptr_i = d->pptr_i;
  p->part_i = d->part_i;
  p->n_parts = I[d->pptr_i + J];
  // where each jet occurs as a block of a partition: (alpha, part, block)
  const std::vector<int32_t> occ = ad_occurrences(d);
  p->n_occ = occ[J];
  const size_t tab = tables_lds_bytes(J, p->n_parts, p->n_occ);
  int64_t off = 0;
  for (int l = 0; l < p->L; ++l) {
    const int di = l == 0 ? p->d_in : p->H, dd = l == p->L - 1 ? p->n_out : p->H;
    p->din.push_back(di);
    p->dout.push_back(dd);
    p->off_w.push_back(off);
    off += (int64_t)di * dd;
    p->off_b.push_back(off);
    off += dd;
  }
  p->n_params = off;
  for (int l = 0; l < p->L; ++l) {
    const bool act = l < p->L - 1;
    const int fpb = pick_pb([&](int n) { return ad_fwd_lds(p->din[l], p->dout[l], J, n, ts) + tab; },
                            TDX_AD_LDS_FWD);
    if (fpb < 1) {
      delete p;
      return fail("layer too wide for LDS");
    }
    p->f_pb.push_back(fpb);
    p->f_rpad.push_back(rows16(fpb, J));
    p->f_lds.push_back(ad_fwd_lds(p->din[l], p->dout[l], J, fpb, ts) + tab);
    if (act) {
      const int dn = p->dout[l + 1];
      const int bpb = pick_pb([&](int n) { return ad_bwd_lds(p->H, dn, J, n, ts) + tab; },
                              TDX_AD_LDS_BWD);
      if (bpb < 1) {
        delete p;
        return fail("layer too wide for LDS");
      }
      p->b_pb.push_back(bpb);
      p->b_rpad.push_back(rows16(bpb, J));
      p->b_lds.push_back(ad_bwd_lds(p->H, dn, J, bpb, ts) + tab);
    }
  }
  p->n_res_blocks = (int)((p->n_entries + AD_THREADS - 1) / AD_THREADS);
  const int64_t K = p->P * J;
  p->n_chunks = (K + AD_WG_CHUNK - 1) / AD_WG_CHUNK;

  // workspace
  size_t w = 0;
  auto take = [&](size_t bytes) {
    const size_t o = w;
    w = align_up(w + bytes, 256);
    return o;
  };
  const size_t jet_row = (size_t)p->P * J * ts;
  p->off_in0 = take(jet_row * p->d_in);
  for (int l = 0; l < p->L - 1; ++l) {
    p->off_z.push_back(take(jet_row * p->H));
    p->off_a.push_back(take(jet_row * p->H));
    p->off_dz.push_back(take(jet_row * p->H));
  }
  p->off_u = take(jet_row * p->n_out);
  p->off_seed = take(jet_row * p->n_out);
  p->off_du = take(jet_row * p->n_out);
  p->off_lp = take((size_t)2 * p->n_res_blocks * ts);
  int max_ne = 0;
  for (int l = 0; l < p->L; ++l) max_ne = std::max(max_ne, p->dout[l] * (p->din[l] + 1));
  p->off_part = take((size_t)p->n_chunks * max_ne * ts);
  p->off_grad = take((size_t)p->n_params * ts);
  for (int l = 0; l < p->L; ++l) p->off_wt.push_back(take((size_t)p->din[l] * p->dout[l] * ts));
  p->off_one = take(ts);
  if (d->causal_tol != 0.0) {
    const int32_t *G0 = I + d->groups_i;
    p->causal = 1;
    p->causal_tol = d->causal_tol;
    p->c_N = G0[0];
    p->c_neq = G0[3];
    p->c_nt = d->causal_nt;
    p->c_K = p->c_N / p->c_nt;
    p->off_resid = take((size_t)std::max<int64_t>(d->n_resid, 1) * ts);
    p->off_cw = take((size_t)p->c_N * ts);
    p->off_cwr = take((size_t)p->c_N * ts);
  }
  if (d->weak_levels != 0) {
    const int32_t *G0 = I + d->groups_i;
    p->weak = d->weak_levels;
    p->w_N = G0[0];
    p->w_neq = G0[3];
    p->w_G = d->weak_runs;
    p->w_d = d->weak_d;
    p->w_i = d->weak_i;
    if (!p->off_resid) p->off_resid = take((size_t)std::max<int64_t>(d->n_resid, 1) * ts);
    p->off_wR = take((size_t)p->w_G * p->w_neq * ts);
    p->off_wM = take((size_t)p->w_N * p->w_neq * ts);
  }
  // fused training step: float, no causal or weak terms, at most 16 jets,
  // 128 hidden units and 4 H x H layers (3 at 128 units), whole entries of
  // every group within one wavefront's 16 rows
  std::vector<int4> fitems;
  {
    const int NH = p->L - 2;
    int nb = 1, nhm = 4;
    const bool ok = ad_fused_ok(d, &nb, &nhm);
    int64_t total_tiles = 0;
    for (int g = 0; ok && g < d->n_groups; ++g) {
      const int32_t *G = I + d->groups_i + TDX_AD_GROUP_INTS * g;
      if (G[0] > 0) total_tiles += (G[0] + 4 * (16 / (G[1] * J)) - 1) / (4 * (16 / (G[1] * J)));
    }
    const size_t lds = ad_step_lds_bytes(16 * nb);
    if (ok) {
      // about one work item per compute unit (the kernel runs one
      // workgroup per compute unit); TDX_AD_TILES sets tiles per item
      int dev = 0, ncu = 256;
      if (hipGetDevice(&dev) != hipSuccess ||
          hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
          ncu < 1)
        ncu = 256;
      int64_t tpi = (total_tiles + ncu - 1) / ncu;
      const char *te = std::getenv("TDX_AD_TILES");
      if (te && std::atoi(te) > 0) tpi = std::atoi(te);
      tpi = std::max<int64_t>(tpi, 1);
      for (int g = 0; g < d->n_groups; ++g) {
        const int32_t *G = I + d->groups_i + TDX_AD_GROUP_INTS * g;
        const int ne = G[0];
        if (ne < 1) continue;
        const int E = 4 * (16 / (G[1] * J));
        const int64_t tiles = (ne + E - 1) / E;
        const int64_t n_it = (tiles + tpi - 1) / tpi;
        for (int64_t k = 0; k < n_it; ++k) {
          const int64_t ta = k * tiles / n_it, tb = (k + 1) * tiles / n_it;
          fitems.push_back(make_int4(g, (int)(ta * E), (int)std::min<int64_t>(ne, tb * E), 0));
        }
      }
      p->fused = 1;
      p->f_nb = nb;
      p->f_nhm = nhm;
      p->step_lds = lds;
      p->n_fitems = (int)fitems.size();
      const int HP = 16 * nb;
      p->off_img = take((size_t)image_words(HP, NH, p->d_in, p->n_out) * 4);
      p->off_zs = take((size_t)p->n_fitems * nhm * nb * 1024 * 4);
      p->off_fpart = take((size_t)p->n_fitems * p->n_params * 4);
      p->off_flp = take((size_t)2 * p->n_fitems * 4);
      // generated jet maps and residual programs for plans of at least 2^17
      // (point, jet) rows (TDX_AD_JIT: 0 never, 1 always); the interpreting
      // kernel if compiling fails
      const char *je = std::getenv("TDX_AD_JIT");
      const bool jit = je ? std::atoi(je) != 0 : p->P * J >= (1 << 17);
      if (jit) {
        const std::string src = ad_step_jit_source(ad_jit_spec(d, occ, nb, nhm));
        if (nn_fused_jit_load(src, &p->step_fn, "tdx_ad_step_jit") != TDX_OK) {
          p->step_fn = nullptr;
          clear_error();
        } else {
          (void)hipFuncSetAttribute(reinterpret_cast<const void *>(p->step_fn),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
          (void)hipGetLastError();
        }
      }
    }
  }
  p->total = w;

  // tables: reals and data in the working precision
  int rc = p->ints.upload(d->ints, (size_t)d->n_ints * 4);
  if (rc == TDX_OK) {
    if (d->dtype == 0) {
      std::vector<float> r(d->reals, d->reals + d->n_reals);
      rc = p->reals.upload(r.data(), r.size() * 4);
    } else {
      rc = p->reals.upload(d->reals, (size_t)d->n_reals * 8);
    }
  }
  if (rc == TDX_OK && d->n_data > 0) rc = p->data.upload(d->data, (size_t)d->n_data * ts);
  if (rc == TDX_OK) rc = p->points.upload(d->points, (size_t)d->n_points * d->d_in * ts);
  if (rc == TDX_OK) rc = p->gstart.upload(gstart.data(), gstart.size() * 8);
  if (rc == TDX_OK) rc = p->occ.upload(occ.data(), occ.size() * 4);
  if (rc == TDX_OK && p->fused) rc = p->fitems.upload(fitems.data(), fitems.size() * sizeof(int4));
  if (rc == TDX_OK) {
    std::vector<int64_t> offs(p->off_w);
    offs.insert(offs.end(), p->off_b.begin(), p->off_b.end());
    rc = p->offs.upload(offs.data(), offs.size() * 8);
  }
  if (rc != TDX_OK) {
    tdx_ad_plan_destroy(p);
    return rc;
  }
  *out = p;
  return TDX_OK;
}

extern "C" void tdx_ad_plan_destroy(tdx_ad_plan *p) {
  if (!p) return;
  p->ints.release();
  p->reals.release();
  p->data.release();
  p->points.release();
  p->gstart.release();
  p->occ.release();
  p->offs.release();
  p->fitems.release();
  delete p;
}

extern "C" int64_t tdx_ad_param_count(const tdx_ad_plan *p) { return p ? p->n_params : -1; }
extern "C" size_t tdx_ad_workspace_bytes(const tdx_ad_plan *p) { return p ? p->total : 0; }

namespace {

template <typename T>
T *wsp(void *ws, size_t off) {
  return reinterpret_cast<T *>(static_cast<unsigned char *>(ws) + off);
}

template <typename T>
Tables tables_of(const tdx_ad_plan *p) {
  Tables t;
  t.ints = (const int32_t *)p->ints.ptr;
  t.jet_i = p->jet_i;
  t.pptr_i = p->pptr_i;
  t.part_i = p->part_i;
  t.J = p->J;
  t.n_parts = p->n_parts;
  t.occ = (const int32_t *)p->occ.ptr;
  t.n_occ = p->n_occ;
  t.tptr_i = 0;
  t.tent_i = p->J + 1;
  return t;
}

template <typename T>
int ad_forward(const tdx_ad_plan *p, const T *params, T *loss, T *resid, void *ws, hipStream_t s) {
  raise_lds_limit({(const void *)k_ad_fwd<T, 4, false>,  (const void *)k_ad_fwd<T, 32, false>,
                   (const void *)k_ad_fwd<T, 64, false>, (const void *)k_ad_bwd<T, 4, false>,
                   (const void *)k_ad_bwd<T, 32, false>, (const void *)k_ad_bwd<T, 64, false>,
                   (const void *)k_ad_fwd<T, 4, true>,   (const void *)k_ad_fwd<T, 32, true>,
                   (const void *)k_ad_fwd<T, 64, true>,  (const void *)k_ad_bwd<T, 4, true>,
                   (const void *)k_ad_bwd<T, 32, true>,  (const void *)k_ad_bwd<T, 64, true>});
  const int J = p->J;
  for (int l = 0; l < p->L; ++l) {
    const int ne = p->din[l] * p->dout[l];
    hipLaunchKernelGGL((k_ad_transpose<T>), dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s,
                       params + p->off_w[l], p->dout[l], p->din[l], wsp<T>(ws, p->off_wt[l]));
    TDX_TRY_HIP(hipGetLastError());
  }
  for (int l = 0; l < p->L; ++l) {
    FwdArgs<T> a;
    a.layer = l;
    a.din = p->din[l];
    a.dout = p->dout[l];
    a.act = l < p->L - 1;
    a.J = J;
    a.P = p->P;
    a.pb = p->f_pb[l];
    a.rpad = p->f_rpad[l];
    a.x = (const T *)p->points.ptr;
    a.in = l == 0 ? nullptr : wsp<T>(ws, p->off_a[l - 1]);
    a.WT = wsp<T>(ws, p->off_wt[l]);
    a.b = params + p->off_b[l];
    a.in0 = wsp<T>(ws, p->off_in0);
    a.zout = a.act ? wsp<T>(ws, p->off_z[l]) : nullptr;
    a.out = a.act ? wsp<T>(ws, p->off_a[l]) : wsp<T>(ws, p->off_u);
    a.tb = tables_of<T>(p);
    const unsigned blocks = (unsigned)((p->P + a.pb - 1) / a.pb);
    // B fragments held in registers: k-steps of 4 over the input width;
    // layers of fewer than four column tiles share the row tiles out
    const bool split = a.dout <= 48;
    const size_t lds = p->f_lds[l];
    if (a.din <= 16) {
      if (split)
        hipLaunchKernelGGL((k_ad_fwd<T, 4, true>), dim3(blocks), dim3(AD_THREADS), lds, s, a);
      else
        hipLaunchKernelGGL((k_ad_fwd<T, 4, false>), dim3(blocks), dim3(AD_THREADS), lds, s, a);
    } else if (a.din <= 128) {
      if (split)
        hipLaunchKernelGGL((k_ad_fwd<T, 32, true>), dim3(blocks), dim3(AD_THREADS), lds, s, a);
      else
        hipLaunchKernelGGL((k_ad_fwd<T, 32, false>), dim3(blocks), dim3(AD_THREADS), lds, s, a);
    } else {
      if (split)
        hipLaunchKernelGGL((k_ad_fwd<T, 64, true>), dim3(blocks), dim3(AD_THREADS), lds, s, a);
      else
        hipLaunchKernelGGL((k_ad_fwd<T, 64, false>), dim3(blocks), dim3(AD_THREADS), lds, s, a);
    }
    TDX_TRY_HIP(hipGetLastError());
  }
  ResArgs<T> r;
  r.J = J;
  r.n_out = p->n_out;
  r.n_groups = p->n_groups;
  r.n_entries = p->n_entries;
  r.gstart = (const int64_t *)p->gstart.ptr;
  r.ints = (const int32_t *)p->ints.ptr;
  r.reals = (const T *)p->reals.ptr;
  r.data = (const T *)p->data.ptr;
  r.groups_i = p->groups_i;
  r.U = wsp<T>(ws, p->off_u);
  r.seed = wsp<T>(ws, p->off_seed);
  // the causal loss and the weak form read group 0's residuals
  r.resid = ((p->causal || p->weak) && !resid) ? wsp<T>(ws, p->off_resid) : resid;
  r.lp = wsp<T>(ws, p->off_lp);
  r.mult = nullptr;
  r.mult_n = 0;
  hipLaunchKernelGGL((k_ad_resid<T>), dim3(p->n_res_blocks), dim3(AD_THREADS), 0, s, r);
  TDX_TRY_HIP(hipGetLastError());
  if (p->weak) {
    // runs' sums from the residuals, each equation's seed multipliers, then
    // the residual program again with them (group 0 carries no other loss)
    const T *c = (const T *)p->data.ptr + p->w_d;
    const T *x = c + p->w_N;
    const int32_t *gs = (const int32_t *)p->ints.ptr + p->w_i;
    T *R = wsp<T>(ws, p->off_wR), *M = wsp<T>(ws, p->off_wM);
    hipLaunchKernelGGL((k_ad_weak_runs<T>), dim3((p->w_G * p->w_neq + 255) / 256), dim3(256), 0,
                       s, (const T *)r.resid, p->w_neq, p->w_N, p->w_G, gs, c, R);
    TDX_TRY_HIP(hipGetLastError());
    hipLaunchKernelGGL((k_ad_weak_mult<T>), dim3(p->w_G * p->w_neq), dim3(256), 0, s, p->weak,
                       p->w_N, p->w_G, gs, c, x, (const T *)R, M);
    TDX_TRY_HIP(hipGetLastError());
    ResArgs<T> r2 = r;
    r2.mult = M;
    r2.mult_n = p->w_N;
    hipLaunchKernelGGL((k_ad_resid<T>), dim3(p->n_res_blocks), dim3(AD_THREADS), 0, s, r2);
    TDX_TRY_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL((k_ad_loss<T>), dim3(1), dim3(AD_THREADS), 0, s, (const T *)r.lp,
                     p->n_res_blocks, loss);
  TDX_TRY_HIP(hipGetLastError());
  if (p->weak) {
    const T *x = (const T *)p->data.ptr + p->w_d + p->w_N, *lam = x + p->w_G;
    hipLaunchKernelGGL((k_ad_weak_loss<T>), dim3(1), dim3(AD_THREADS), 0, s, p->weak, p->w_neq,
                       p->w_G, x, lam, (const T *)wsp<T>(ws, p->off_wR), loss);
    TDX_TRY_HIP(hipGetLastError());
  }
  if (p->causal) {
    T *w = wsp<T>(ws, p->off_cw), *wr = wsp<T>(ws, p->off_cwr);
    hipLaunchKernelGGL((k_ad_causal_w<T>), dim3((p->c_K + 255) / 256), dim3(256), 0, s,
                       (const T *)r.resid, p->c_neq, p->c_N, p->c_nt, p->c_K, T(p->causal_tol), w,
                       wr);
    TDX_TRY_HIP(hipGetLastError());
    const int row = J * p->n_out;
    hipLaunchKernelGGL((k_ad_causal_seeds<T>), dim3((unsigned)(((int64_t)p->c_N * row + 255) / 256)),
                       dim3(256), 0, s, r.seed, (const T *)w, p->c_N, row);
    TDX_TRY_HIP(hipGetLastError());
    hipLaunchKernelGGL((k_ad_causal_loss<T>), dim3(1), dim3(AD_THREADS), 0, s, (const T *)wr,
                       p->c_N, loss);
    TDX_TRY_HIP(hipGetLastError());
  }
  return TDX_OK;
}

template <typename T>
int ad_backward(const tdx_ad_plan *p, const T *params, const T *dloss, T *grad, void *ws,
                hipStream_t s) {
  const int J = p->J;
  const int64_t nu = p->P * J * p->n_out;
  T *du = wsp<T>(ws, p->off_du);
  hipLaunchKernelGGL((k_ad_scale<T>), dim3((unsigned)std::min<int64_t>(4096, (nu + 255) / 256)),
                     dim3(256), 0, s, (const T *)wsp<T>(ws, p->off_seed), dloss, du, nu);
  TDX_TRY_HIP(hipGetLastError());
  for (int l = p->L - 2; l >= 0; --l) {
    BwdArgs<T> b;
    b.H = p->H;
    b.dn = p->dout[l + 1];
    b.J = J;
    b.P = p->P;
    b.pb = p->b_pb[l];
    b.rpad = p->b_rpad[l];
    b.Wn = params + p->off_w[l + 1];
    b.DN = l == p->L - 2 ? du : wsp<T>(ws, p->off_dz[l + 1]);
    b.Z = wsp<T>(ws, p->off_z[l]);
    b.DZ = wsp<T>(ws, p->off_dz[l]);
    b.tb = tables_of<T>(p);
    const unsigned blocks = (unsigned)((p->P + b.pb - 1) / b.pb);
    const bool split = b.H <= 48;
    const size_t lds = p->b_lds[l];
    if (b.dn <= 16) {
      if (split)
        hipLaunchKernelGGL((k_ad_bwd<T, 4, true>), dim3(blocks), dim3(AD_THREADS), lds, s, b);
      else
        hipLaunchKernelGGL((k_ad_bwd<T, 4, false>), dim3(blocks), dim3(AD_THREADS), lds, s, b);
    } else if (b.dn <= 128) {
      if (split)
        hipLaunchKernelGGL((k_ad_bwd<T, 32, true>), dim3(blocks), dim3(AD_THREADS), lds, s, b);
      else
        hipLaunchKernelGGL((k_ad_bwd<T, 32, false>), dim3(blocks), dim3(AD_THREADS), lds, s, b);
    } else {
      if (split)
        hipLaunchKernelGGL((k_ad_bwd<T, 64, true>), dim3(blocks), dim3(AD_THREADS), lds, s, b);
      else
        hipLaunchKernelGGL((k_ad_bwd<T, 64, false>), dim3(blocks), dim3(AD_THREADS), lds, s, b);
    }
    TDX_TRY_HIP(hipGetLastError());
  }
  T *part = wsp<T>(ws, p->off_part);
  for (int l = 0; l < p->L; ++l) {
    WgArgs<T> w;
    w.din = p->din[l];
    w.dout = p->dout[l];
    w.J = J;
    w.K = p->P * J;
    w.D = l == p->L - 1 ? du : wsp<T>(ws, p->off_dz[l]);
    w.S = l == 0 ? wsp<T>(ws, p->off_in0) : wsp<T>(ws, p->off_a[l - 1]);
    w.part = part;
    const int ne = w.dout * (w.din + 1);
    const int nt = ((w.dout + 15) / 16) * ((w.din + 1 + 15) / 16);
    const int per_block = (AD_THREADS / 64) * AD_WG_TPW;
    const unsigned ytiles = (unsigned)((nt + per_block - 1) / per_block);
    hipLaunchKernelGGL((k_ad_wgrad<T>), dim3((unsigned)p->n_chunks, ytiles), dim3(AD_THREADS), 0,
                       s, w);
    TDX_TRY_HIP(hipGetLastError());
    hipLaunchKernelGGL((k_ad_wreduce<T>), dim3((unsigned)((ne + 63) / 64)), dim3(AD_THREADS), 0, s,
                       (const T *)part, (int)p->n_chunks, w.din, w.dout, grad + p->off_w[l],
                       grad + p->off_b[l]);
    TDX_TRY_HIP(hipGetLastError());
  }
  return TDX_OK;
}

}  // namespace

extern "C" int tdx_ad_forward(const tdx_ad_plan *p, const void *params, void *loss, void *resid,
                              void *ws, tdx_stream_t stream) {
  TDX_REQUIRE(p && params && loss && ws, "null argument");
  hipStream_t s = (hipStream_t)stream;
  if (p->dtype == 0)
    return ad_forward<float>(p, (const float *)params, (float *)loss, (float *)resid, ws, s);
  return ad_forward<double>(p, (const double *)params, (double *)loss, (double *)resid, ws, s);
}

extern "C" int tdx_ad_backward(const tdx_ad_plan *p, const void *params, const void *dloss,
                               void *dparams, void *ws, tdx_stream_t stream) {
  TDX_REQUIRE(p && params && dloss && dparams && ws, "null argument");
  hipStream_t s = (hipStream_t)stream;
  if (p->dtype == 0)
    return ad_backward<float>(p, (const float *)params, (const float *)dloss, (float *)dparams, ws, s);
  return ad_backward<double>(p, (const double *)params, (const double *)dloss, (double *)dparams,
                             ws, s);
}

namespace {

template <typename T>
int ad_jacobian(const tdx_ad_plan *p, const T *params, T *jt, hipStream_t s) {
  const int NL = p->L - 1, J = p->J, H = p->H, no = p->n_out;
  const size_t words = (size_t)(2 * NL + 2) * J * H + (size_t)2 * p->max_sets * J * no +
                       (size_t)J * p->d_in;
  int B = 64;
  while (B > 1 && words * sizeof(T) * B > (160 * 1024)) B >>= 1;
  TDX_REQUIRE(words * sizeof(T) * B <= (160 * 1024), "network too large for the Jacobian kernel");
  raise_lds_limit({(const void *)k_ad_jac<T>});
  const int64_t n = p->n_params * p->n_resid;
  hipLaunchKernelGGL((k_ad_jac_zero<T>), dim3((unsigned)std::min<int64_t>((n + 255) / 256, 65536)),
                     dim3(256), 0, s, jt, n);
  TDX_TRY_HIP(hipGetLastError());
  JacArgs<T> a;
  a.d_in = p->d_in;
  a.n_out = no;
  a.H = H;
  a.L = p->L;
  a.J = J;
  a.n_groups = p->n_groups;
  a.groups_i = p->groups_i;
  a.max_sets = p->max_sets;
  a.ints = (const int32_t *)p->ints.ptr;
  a.reals = (const T *)p->reals.ptr;
  a.data = (const T *)p->data.ptr;
  a.points = (const T *)p->points.ptr;
  a.params = params;
  a.gstart = (const int64_t *)p->gstart.ptr;
  a.tb = tables_of<T>(p);
  a.off_w = (const int64_t *)p->offs.ptr;
  a.off_b = a.off_w + p->L;
  a.jt = jt;
  a.n_resid = p->n_resid;
  const unsigned blocks = (unsigned)((p->n_entries + B - 1) / B);
  hipLaunchKernelGGL((k_ad_jac<T>), dim3(blocks), dim3(B), words * sizeof(T) * B, s, a);
  TDX_TRY_HIP(hipGetLastError());
  return TDX_OK;
}

}  // namespace

extern "C" int tdx_ad_jacobian(const tdx_ad_plan *p, const void *params, void *jt,
                               tdx_stream_t stream) {
  TDX_REQUIRE(p && params && jt, "null argument");
  TDX_REQUIRE(!p->weak, "the Jacobian of a weak-form plan is not served");
  hipStream_t s = (hipStream_t)stream;
  if (p->dtype == 0) return ad_jacobian<float>(p, (const float *)params, (float *)jt, s);
  return ad_jacobian<double>(p, (const double *)params, (double *)jt, s);
}

namespace {

template <int NB, int NHM>
int ad_step_launch(const tdx_ad_plan *p, const ADStepArgs &a, hipStream_t s) {
  if (p->step_fn) {
    ADStepArgs arg = a;
    void *args[] = {&arg};
    TDX_TRY_HIP(hipModuleLaunchKernel(p->step_fn, (unsigned)p->n_fitems, 1, 1, 256, 1, 1,
                                      (unsigned)p->step_lds, s, args, nullptr));
    return TDX_OK;
  }
  raise_lds_limit({(const void *)k_ad_step<NB, NHM>});
  hipLaunchKernelGGL((k_ad_step<NB, NHM>), dim3((unsigned)p->n_fitems), dim3(256), p->step_lds, s, a);
  TDX_TRY_HIP(hipGetLastError());
  return TDX_OK;
}

template <int NB>
int ad_pack_launch(const NNArgs<float> &nn, int64_t *step, hipStream_t s) {
  const int IMG = image_words(16 * NB, nn.NH, nn.d_in, nn.n_out);
  hipLaunchKernelGGL((k_ad_pack<NB>), dim3((unsigned)((IMG + 255) / 256)), dim3(256), 0, s, nn,
                     step);
  TDX_TRY_HIP(hipGetLastError());
  return TDX_OK;
}

// The fused training step: three launches (image and step counter, the
// step kernel, gradient sums with the Adam update and the loss pair).
int ad_fused_step(const tdx_ad_plan *p, float *params, float *m, float *v, int64_t *step,
                  const AdamHyper &h, float *loss, void *ws, hipStream_t s) {
  NNArgs<float> nn{};
  nn.d_in = p->d_in;
  nn.n_out = p->n_out;
  nn.H = p->H;
  nn.NH = p->L - 2;
  nn.P = (int)p->n_params;
  nn.n_inst = 1;
  nn.params = params;
  nn.packed = wsp<float>(ws, p->off_img);
  int rc;
  switch (p->f_nb) {
    case 1: rc = ad_pack_launch<1>(nn, step, s); break;
    case 2: rc = ad_pack_launch<2>(nn, step, s); break;
    case 3: rc = ad_pack_launch<3>(nn, step, s); break;
    case 4: rc = ad_pack_launch<4>(nn, step, s); break;
    case 5: rc = ad_pack_launch<5>(nn, step, s); break;
    case 6: rc = ad_pack_launch<6>(nn, step, s); break;
    case 7: rc = ad_pack_launch<7>(nn, step, s); break;
    default: rc = ad_pack_launch<8>(nn, step, s); break;
  }
  if (rc != TDX_OK) return rc;
  ADStepArgs a;
  a.nn = nn;
  a.ints = (const int32_t *)p->ints.ptr;
  a.reals = (const float *)p->reals.ptr;
  a.data = (const float *)p->data.ptr;
  a.points = (const float *)p->points.ptr;
  a.groups_i = p->groups_i;
  a.J = p->J;
  a.tb = tables_of<float>(p);
  a.items = (const int4 *)p->fitems.ptr;
  a.zs = wsp<float>(ws, p->off_zs);
  a.partial = wsp<float>(ws, p->off_fpart);
  a.lp = wsp<float>(ws, p->off_flp);
  switch (p->f_nb) {
    case 1: rc = ad_step_launch<1, 4>(p, a, s); break;
    case 2: rc = ad_step_launch<2, 4>(p, a, s); break;
    case 3: rc = ad_step_launch<3, 4>(p, a, s); break;
    case 4: rc = ad_step_launch<4, 4>(p, a, s); break;
    case 5: rc = ad_step_launch<5, 4>(p, a, s); break;
    case 6: rc = ad_step_launch<6, 4>(p, a, s); break;
    case 7: rc = ad_step_launch<7, 4>(p, a, s); break;
    default: rc = ad_step_launch<8, 3>(p, a, s); break;
  }
  if (rc != TDX_OK) return rc;
  hipLaunchKernelGGL(k_ad_step_sum, dim3((unsigned)(1 + (p->n_params + 255) / 256)), dim3(256), 0,
                     s, (const float *)a.partial, (const float *)a.lp, p->n_fitems,
                     (int64_t)p->n_params, params, m, v, (const int64_t *)step, h, loss);
  TDX_TRY_HIP(hipGetLastError());
  return TDX_OK;
}

}  // namespace

extern "C" int tdx_ad_step_kernel(const tdx_ad_plan *p) {
  return p ? (p->fused ? (p->step_fn ? 2 : 1) : 0) : -1;
}

extern "C" int64_t tdx_ad_step_source(const tdx_ad_desc *d, char *buf, int64_t cap, int compile) {
  TDX_REQUIRE(d, "null argument");
  if (d->dtype != 0 && d->dtype != 1) return fail("dtype must be 0 or 1");
  if (d->d_in < 1 || d->n_out < 1 || d->width < 1 || d->n_layers < 2) return fail("network shape");
  if (d->width > 256 || d->n_out > 256 || d->d_in > 256) return fail("widths above 256");
  if (d->n_jets < 1 || d->n_points < 1 || d->n_groups < 1) return fail("empty problem");
  if (!d->ints || !d->reals || !d->points || d->n_ints < 1) return fail("null table");
  std::vector<int64_t> gstart;
  int max_sets = 0;
  const int rc = validate_ad(d, &gstart, &max_sets);
  if (rc != TDX_OK) return rc;
  int nb = 1, nhm = 4;
  if (!ad_fused_ok(d, &nb, &nhm)) {
    set_error("the network or options leave the fused step out");
    return 0;
  }
  const std::string src = ad_step_jit_source(ad_jit_spec(d, ad_occurrences(d), nb, nhm));
  if (compile) {
    std::vector<char> code;
    const int crc = nn_fused_jit_build(src, &code, "gfx950");
    if (crc != TDX_OK) return crc;
  }
  if (buf && cap > 0) {
    const int64_t n = std::min<int64_t>(cap - 1, (int64_t)src.size());
    std::memcpy(buf, src.data(), (size_t)n);
    buf[n] = 0;
  }
  return (int64_t)src.size();
}

extern "C" int tdx_ad_train_step(const tdx_ad_plan *p, void *params, void *m, void *v,
                                 int64_t *step, double lr, double b1, double b2, double eps,
                                 void *loss, void *ws, tdx_stream_t stream) {
  TDX_REQUIRE(p && params && m && v && step && loss && ws, "null argument");
  hipStream_t s = (hipStream_t)stream;
  if (p->fused)
    return ad_fused_step(p, (float *)params, (float *)m, (float *)v, step,
                         AdamHyper{lr, b1, b2, eps}, (float *)loss, ws, s);
  int rc;
  void *grad = static_cast<unsigned char *>(ws) + p->off_grad;
  void *one = static_cast<unsigned char *>(ws) + p->off_one;
  if (p->dtype == 0) {
    hipLaunchKernelGGL((k_ad_one<float>), dim3(1), dim3(64), 0, s, (float *)one);
    rc = ad_forward<float>(p, (const float *)params, (float *)loss, nullptr, ws, s);
    if (rc == TDX_OK)
      rc = ad_backward<float>(p, (const float *)params, (const float *)one, (float *)grad, ws, s);
  } else {
    hipLaunchKernelGGL((k_ad_one<double>), dim3(1), dim3(64), 0, s, (double *)one);
    rc = ad_forward<double>(p, (const double *)params, (double *)loss, nullptr, ws, s);
    if (rc == TDX_OK)
      rc = ad_backward<double>(p, (const double *)params, (const double *)one, (double *)grad, ws, s);
  }
  if (rc != TDX_OK) return rc;
  rc = tdx_adam(p->dtype, p->n_params, params, grad, m, v, step, lr, b1, b2, eps, stream);
  if (rc != TDX_OK) return rc;
  return tdx_step_increment(step, stream);
}
