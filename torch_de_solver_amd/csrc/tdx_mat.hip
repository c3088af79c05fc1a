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

#include <algorithm>
#include <map>

namespace tdx {

constexpr int MMAX_DP = 8;
constexpr int MMAX_EQ = 4;
constexpr int MMAX_FS = 16;   // factor slots over all terms of all equations
constexpr int MAT_THREADS = 256;

template <typename T>
struct MatArgs {
  int n_eq, n0, n1, dp;
  int64_t N;  // n0 * n1
  T h[2], h2[2], ih[2], ih2[2];
  T ab[MMAX_DP], af[MMAX_DP];
  const int32_t *ints;
  const T *reals;
  const T *data;
  int eq_i;
  // factor slots (flattened over equations and terms), built on the host
  int n_fs;
  int fs_var[MMAX_FS], fs_nax[MMAX_FS], fs_ax[MMAX_FS][3];
  int R0, R1;  // reach per axis
  int T0, T1;  // tile
  int tiles1;  // tiles along axis 1
  // bc
  int n_bc, bc_i;
  int n_touch, bct_ptr_i, bct_pos_i, bct_seed_i, bct_w_i;
  // buffers
  const T *u;
  T *grad;     // may be null (forward only)
  T *op;       // may be null
  T *bval;     // may be null
  T *lp;       // 2 per tile workgroup then 2 per bc workgroup
  T *bc_seeds;
  const T *dloss;
  int n_tile_wg;
};

template <typename T, bool EXACT>
__device__ __forceinline__ T qdiv(T x, T h, T inv) {
  if (EXACT) return x / h;
  return x * inv;
}

// One application of D along `axis` at local point p of an LDS field whose
// row stride is E1; g is the global index along the axis, n its length.
template <typename T, bool EXACT>
__device__ __forceinline__ T dstencil(const MatArgs<T> &a, const T *src, int p, int stride, int g,
                                      int n, int axis) {
  const int dp = a.dp;
  if (g < dp - 1) {
    T f = T(0);
    for (int i = 0; i < dp; ++i) f = f + src[p + i * stride] * a.af[i];
    return qdiv<T, EXACT>(f, a.h[axis], a.ih[axis]);
  }
  T b = T(0);
  for (int i = 0; i < dp; ++i) b = b + src[p + (i - (dp - 1)) * stride] * a.ab[i];
  if (g >= n - dp + 1) return qdiv<T, EXACT>(b, a.h[axis], a.ih[axis]);
  T f = T(0);
  for (int i = 0; i < dp; ++i) f = f + src[p + i * stride] * a.af[i];
  return qdiv<T, EXACT>(b + f, a.h2[axis], a.ih2[axis]);
}

// Transpose of D at global index g: sum over the rows k whose stencil reads g.
template <typename T, bool EXACT>
__device__ __forceinline__ T dtstencil(const MatArgs<T> &a, const T *src, int p, int stride, int g,
                                       int n, int axis) {
  const int dp = a.dp;
  T acc = T(0);
  // backward labels: row k = g + (dp-1) - i reads g with ab[i]
  for (int i = 0; i < dp; ++i) {
    const int k = g + (dp - 1) - i;
    if (k < 0 || k >= n || k < dp - 1) continue;  // front rows use forward labels only
    const T v = src[p + ((dp - 1) - i) * stride] * a.ab[i];
    acc += (k >= n - dp + 1) ? qdiv<T, EXACT>(v, a.h[axis], a.ih[axis])
                             : qdiv<T, EXACT>(v, a.h2[axis], a.ih2[axis]);
  }
  // forward labels: row k = g - i reads g with af[i]
  for (int i = 0; i < dp; ++i) {
    const int k = g - i;
    if (k < 0 || k >= n) continue;
    if (k >= dp - 1 && k >= n - dp + 1) continue;  // back rows use backward labels only
    const T v = src[p - i * stride] * a.af[i];
    acc += (k < dp - 1) ? qdiv<T, EXACT>(v, a.h[axis], a.ih[axis])
                        : qdiv<T, EXACT>(v, a.h2[axis], a.ih2[axis]);
  }
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

struct Box {
  int lo0, hi0, lo1, hi1;
};

// ---------------------------------------------------------------------------
// tiled residual + loss + adjoint
// ---------------------------------------------------------------------------
template <typename T, bool EXACT, bool ADJ>
__global__ __launch_bounds__(MAT_THREADS) void k_mat_tile(MatArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *smem = reinterpret_cast<T *>(smem_raw);
  const int t0 = blockIdx.x / a.tiles1, t1 = blockIdx.x - (blockIdx.x / a.tiles1) * a.tiles1;
  const int E0 = a.T0 + 4 * a.R0, E1 = a.T1 + 4 * a.R1;
  const int S = E0 * E1;
  const int base0 = t0 * a.T0 - 2 * a.R0, base1 = t1 * a.T1 - 2 * a.R1;
  T *U = smem;                    // n_eq * S
  T *F = U + a.n_eq * S;          // n_fs * S (values, then seeds)
  T *A = F + a.n_fs * S;          // scratch
  T *B = A + S;                   // scratch
  T *G = B + S;                   // n_eq * S gradient
  T *red = G + (ADJ ? a.n_eq * S : 0);
  const int tid = threadIdx.x;

  // load u with halo (zero outside the grid)
  for (int v = 0; v < a.n_eq; ++v)
    for (int p = tid; p < S; p += MAT_THREADS) {
      const int i0 = p / E1, i1 = p - (p / E1) * E1;
      const int g0 = base0 + i0, g1 = base1 + i1;
      T val = T(0);
      if (g0 >= 0 && g0 < a.n0 && g1 >= 0 && g1 < a.n1) val = a.u[v * a.N + (int64_t)g0 * a.n1 + g1];
      U[v * S + p] = val;
      if (ADJ) G[v * S + p] = T(0);
    }
  __syncthreads();

  // forward chains
  for (int fs = 0; fs < a.n_fs; ++fs) {
    const int nax = a.fs_nax[fs];
    const T *src = U + a.fs_var[fs] * S;
    Box bx{0, E0, 0, E1};
    if (nax == 0) {
      for (int p = tid; p < S; p += MAT_THREADS) F[fs * S + p] = src[p];
    }
    for (int j = 0; j < nax; ++j) {
      const int ax = a.fs_ax[fs][j];
      T *dst = (j == nax - 1) ? F + fs * S : ((j & 1) ? B : A);
      if (ax == 0) { bx.lo0 += a.dp - 1; bx.hi0 -= a.dp - 1; }
      else { bx.lo1 += a.dp - 1; bx.hi1 -= a.dp - 1; }
      const int w1 = bx.hi1 - bx.lo1, cnt = (bx.hi0 - bx.lo0) * w1;
      for (int k = tid; k < cnt; k += MAT_THREADS) {
        const int i0 = bx.lo0 + k / w1, i1 = bx.lo1 + (k - (k / w1) * w1);
        const int g0 = base0 + i0, g1 = base1 + i1;
        const int p = i0 * E1 + i1;
        T val = T(0);
        if (g0 >= 0 && g0 < a.n0 && g1 >= 0 && g1 < a.n1)
          val = (ax == 0) ? dstencil<T, EXACT>(a, src, p, E1, g0, a.n0, 0)
                          : dstencil<T, EXACT>(a, src, p, 1, g1, a.n1, 1);
        dst[p] = val;
      }
      __syncthreads();
      src = dst;
    }
  }
  __syncthreads();

  // residual, loss, seeds on the tile plus R
  T lsum = T(0), lnsum = T(0);
  {
    const Box rb{2 * a.R0 - a.R0, 2 * a.R0 + a.T0 + a.R0, 2 * a.R1 - a.R1, 2 * a.R1 + a.T1 + a.R1};
    const int w1 = rb.hi1 - rb.lo1, cnt = (rb.hi0 - rb.lo0) * w1;
    const T dl = ADJ ? *a.dloss : T(0);
    for (int k = tid; k < cnt; k += MAT_THREADS) {
      const int i0 = rb.lo0 + k / w1, i1 = rb.lo1 + (k - (k / w1) * w1);
      const int g0 = base0 + i0, g1 = base1 + i1;
      const int p = i0 * E1 + i1;
      const bool in_grid = g0 >= 0 && g0 < a.n0 && g1 >= 0 && g1 < a.n1;
      const bool in_tile = i0 >= 2 * a.R0 && i0 < 2 * a.R0 + a.T0 && i1 >= 2 * a.R1 &&
                           i1 < 2 * a.R1 + a.T1;
      const int64_t gp = (int64_t)g0 * a.n1 + g1;
      T seed[MMAX_FS];
#pragma unroll
      for (int q = 0; q < MMAX_FS; ++q) seed[q] = T(0);
      if (in_grid) {
        int fs0 = 0;
        for (int eq = 0; eq < a.n_eq; ++eq) {
          const int32_t *E = a.ints + a.eq_i + eq * TDX_MEQ_INTS;
          const int n_terms = E[0];
          T r = T(0);
          int fs = fs0;
          for (int t = 0; t < n_terms; ++t) {
            const int32_t *Tm = a.ints + E[1] + t * TDX_MTERM_INTS;
            T der = T(1);
            for (int f = 0; f < Tm[0]; ++f, ++fs) {
              const int32_t *Fc = a.ints + Tm[1] + f * TDX_MFACTOR_INTS;
              der = der * mfpow(a, Fc, F[fs * S + p]);
            }
            const T c = Tm[2] == 0 ? a.reals[Tm[3]] : a.data[Tm[3] + gp];
            const T tv = c * der;
            r = (t == 0) ? tv : r + tv;
          }
          const T lam_n = a.reals[E[2]];
          if (in_tile) {
            lsum += lam_n * (r * r);
            lnsum += (r * r) / T(a.N);
            if (a.op) a.op[eq * a.N + gp] = r;
          }
          if (ADJ) {
            const T g = T(2) * lam_n * r * dl;
            fs = fs0;
            for (int t = 0; t < n_terms; ++t) {
              const int32_t *Tm = a.ints + E[1] + t * TDX_MTERM_INTS;
              const int nf = Tm[0];
              const T c = Tm[2] == 0 ? a.reals[Tm[3]] : a.data[Tm[3] + gp];
              for (int f = 0; f < nf; ++f) {
                const int32_t *Fc = a.ints + Tm[1] + f * TDX_MFACTOR_INTS;
                T d = c * g * mdfpow(a, Fc, F[(fs + f) * S + p]);
                for (int f2 = 0; f2 < nf; ++f2)
                  if (f2 != f)
                    d *= mfpow(a, a.ints + Tm[1] + f2 * TDX_MFACTOR_INTS, F[(fs + f2) * S + p]);
#pragma unroll
                for (int q = 0; q < MMAX_FS; ++q)
                  if (q == fs + f) seed[q] = d;
              }
              fs += nf;
            }
          }
          fs0 = fs;
        }
      }
      // only this thread reads F at p in this pass, so the seeds can
      // overwrite the factor values in place
      if (ADJ) {
        for (int q = 0; q < a.n_fs; ++q) F[q * S + p] = seed[q];
      }
    }
  }
  __syncthreads();

  if (ADJ) {
    // transposed chains back to the tile
    for (int fs = 0; fs < a.n_fs; ++fs) {
      const int nax = a.fs_nax[fs];
      const T *src = F + fs * S;
      // seeds are valid on the tile plus R
      Box bx{a.R0, 3 * a.R0 + a.T0, a.R1, 3 * a.R1 + a.T1};
      for (int j = nax - 1; j >= 0; --j) {
        const int ax = a.fs_ax[fs][j];
        T *dst = ((nax - 1 - j) & 1) ? B : A;
        if (ax == 0) { bx.lo0 += a.dp - 1; bx.hi0 -= a.dp - 1; }
        else { bx.lo1 += a.dp - 1; bx.hi1 -= a.dp - 1; }
        const int w1 = bx.hi1 - bx.lo1, cnt = (bx.hi0 - bx.lo0) * w1;
        for (int k = tid; k < cnt; k += MAT_THREADS) {
          const int i0 = bx.lo0 + k / w1, i1 = bx.lo1 + (k - (k / w1) * w1);
          const int g0 = base0 + i0, g1 = base1 + i1;
          const int p = i0 * E1 + i1;
          T val = T(0);
          if (g0 >= 0 && g0 < a.n0 && g1 >= 0 && g1 < a.n1)
            val = (ax == 0) ? dtstencil<T, EXACT>(a, src, p, E1, g0, a.n0, 0)
                            : dtstencil<T, EXACT>(a, src, p, 1, g1, a.n1, 1);
          dst[p] = val;
        }
        __syncthreads();
        src = dst;
      }
      // accumulate on the tile
      T *Gv = G + a.fs_var[fs] * S;
      for (int k = tid; k < a.T0 * a.T1; k += MAT_THREADS) {
        const int i0 = 2 * a.R0 + k / a.T1, i1 = 2 * a.R1 + (k - (k / a.T1) * a.T1);
        const int p = i0 * E1 + i1;
        Gv[p] += src[p];
      }
      __syncthreads();
    }
    for (int v = 0; v < a.n_eq; ++v)
      for (int k = tid; k < a.T0 * a.T1; k += MAT_THREADS) {
        const int i0 = 2 * a.R0 + k / a.T1, i1 = 2 * a.R1 + (k - (k / a.T1) * a.T1);
        const int g0 = base0 + i0, g1 = base1 + i1;
        if (g0 < a.n0 && g1 < a.n1)
          a.grad[v * a.N + (int64_t)g0 * a.n1 + g1] = G[v * S + i0 * E1 + i1];
      }
  }

  // deterministic workgroup sum of the loss partials
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
    a.lp[2 * blockIdx.x] = red[0];
    a.lp[2 * blockIdx.x + 1] = red[MAT_THREADS];
  }
}

// ---------------------------------------------------------------------------
// boundary entries: pointwise derivative chains straight from HBM
// ---------------------------------------------------------------------------
template <typename T, bool EXACT>
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
    return point_chain<T, EXACT>(a, plane, h0, h1, axes, depth - 1);
  };
  if (g < dp - 1) {
    T f = T(0);
    for (int i = 0; i < dp; ++i) f = f + at(i) * a.af[i];
    return qdiv<T, EXACT>(f, a.h[ax], a.ih[ax]);
  }
  T b = T(0);
  for (int i = 0; i < dp; ++i) b = b + at(i - (dp - 1)) * a.ab[i];
  if (g >= n - dp + 1) return qdiv<T, EXACT>(b, a.h[ax], a.ih[ax]);
  T f = T(0);
  for (int i = 0; i < dp; ++i) f = f + at(i) * a.af[i];
  return qdiv<T, EXACT>(b + f, a.h2[ax], a.ih2[ax]);
}

template <typename T, bool EXACT>
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
    T val = T(0);
    for (int s = 0; s < n_sets; ++s) {
      const int pos = a.ints[pos_i + s];
      const int g0 = pos / a.n1, g1 = pos - (pos / a.n1) * a.n1;
      T v;
      if (n_terms == 0) {
        v = a.u[var * a.N + pos];
      } else {
        T r = T(0);
        for (int t = 0; t < n_terms; ++t) {
          const int32_t *Tm = a.ints + term_i + t * TDX_MTERM_INTS;
          T der = T(1);
          for (int f = 0; f < Tm[0]; ++f) {
            const int32_t *Fc = a.ints + Tm[1] + f * TDX_MFACTOR_INTS;
            const int ax[3] = {Fc[5], Fc[6], Fc[7]};
            const T fv = point_chain<T, EXACT>(a, a.u + Fc[0] * a.N, g0, g1, ax, Fc[4]);
            der = der * mfpow(a, Fc, fv);
          }
          const T c = Tm[2] == 0 ? a.reals[Tm[3]] : a.data[Tm[3] + pos];
          const T tv = c * der;
          r = (t == 0) ? tv : r + tv;
        }
        v = r;
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
      for (int s = 0; s < n_sets; ++s) {
        const T sg = s == 0 ? dr : -dr;
        if (n_terms == 0) {
          a.bc_seeds[sidx++] = sg;
          continue;
        }
        const int pos = a.ints[pos_i + s];
        const int g0 = pos / a.n1, g1 = pos - (pos / a.n1) * a.n1;
        for (int t = 0; t < n_terms; ++t) {
          const int32_t *Tm = a.ints + term_i + t * TDX_MTERM_INTS;
          const int nf = Tm[0];
          const T c = Tm[2] == 0 ? a.reals[Tm[3]] : a.data[Tm[3] + pos];
          for (int f = 0; f < nf; ++f) {
            const int32_t *Fc = a.ints + Tm[1] + f * TDX_MFACTOR_INTS;
            const int ax[3] = {Fc[5], Fc[6], Fc[7]};
            T d = c * sg *
                  mdfpow(a, Fc, point_chain<T, EXACT>(a, a.u + Fc[0] * a.N, g0, g1, ax, Fc[4]));
            for (int f2 = 0; f2 < nf; ++f2) {
              if (f2 == f) continue;
              const int32_t *F2 = a.ints + Tm[1] + f2 * TDX_MFACTOR_INTS;
              const int ax2[3] = {F2[5], F2[6], F2[7]};
              d *= mfpow(a, F2, point_chain<T, EXACT>(a, a.u + F2[0] * a.N, g0, g1, ax2, F2[4]));
            }
            a.bc_seeds[sidx++] = d;
          }
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

template <typename T>
__global__ void k_mat_loss(const T *lp, int n, T *loss) {
  __shared__ T red[2 * 256];
  T s0 = T(0), s1 = T(0);
  for (int k = threadIdx.x; k < n; k += 256) {
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
    loss[0] = red[0];
    loss[1] = red[256];
  }
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
  int fs_var[MMAX_FS], fs_nax[MMAX_FS], fs_ax[MMAX_FS][3];
  int R0, R1, T0, T1, tiles0, tiles1, n_tile_wg, n_bc_wg;
  size_t tile_lds, tile_lds_fwd;
  DevTable ints, reals, data;
  size_t off_lp, off_seeds, off_one, off_grad, ws_bytes;
  double ab[MMAX_DP], af[MMAX_DP];
};


namespace {

template <typename T>
MatArgs<T> mat_args(const tdx_mat_plan *p, const void *u, void *ws) {
  MatArgs<T> a;
  const tdx_mat_desc &d = p->d;
  a.n_eq = d.n_eq;
  a.n0 = d.n0;
  a.n1 = d.n1;
  a.dp = d.dp;
  a.N = (int64_t)d.n0 * d.n1;
  for (int k = 0; k < 2; ++k) {
    a.h[k] = T(d.h[k]);
    a.h2[k] = T(2.0 * d.h[k]);
    a.ih[k] = T(1.0 / d.h[k]);
    a.ih2[k] = T(1.0 / (2.0 * d.h[k]));
  }
  for (int i = 0; i < MMAX_DP; ++i) {
    a.ab[i] = i < d.dp ? T(p->ab[i]) : T(0);
    a.af[i] = i < d.dp ? T(p->af[i]) : T(0);
  }
  a.ints = (const int32_t *)p->ints.ptr;
  a.reals = (const T *)p->reals.ptr;
  a.data = (const T *)p->data.ptr;
  a.eq_i = d.eq_i;
  a.n_fs = p->n_fs;
  for (int q = 0; q < MMAX_FS; ++q) {
    a.fs_var[q] = p->fs_var[q];
    a.fs_nax[q] = p->fs_nax[q];
    for (int k = 0; k < 3; ++k) a.fs_ax[q][k] = p->fs_ax[q][k];
  }
  a.R0 = p->R0;
  a.R1 = p->R1;
  a.T0 = p->T0;
  a.T1 = p->T1;
  a.tiles1 = p->tiles1;
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
  a.bc_seeds = (T *)(w + p->off_seeds);
  a.dloss = (const T *)(w + p->off_one);
  a.n_tile_wg = p->n_tile_wg;
  return a;
}

template <typename T, bool EXACT>
int mat_run(const tdx_mat_plan *p, MatArgs<T> a, bool adj, T *loss, hipStream_t s) {
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute((const void *)k_mat_tile<T, EXACT, true>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    (void)hipFuncSetAttribute((const void *)k_mat_tile<T, EXACT, false>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    attr = true;
  }
  if (adj)
    hipLaunchKernelGGL((k_mat_tile<T, EXACT, true>), dim3(p->n_tile_wg), dim3(MAT_THREADS),
                       p->tile_lds, s, a);
  else
    hipLaunchKernelGGL((k_mat_tile<T, EXACT, false>), dim3(p->n_tile_wg), dim3(MAT_THREADS),
                       p->tile_lds_fwd, s, a);
  TDX_TRY_HIP(hipGetLastError());
  if (p->n_bc_wg > 0) {
    hipLaunchKernelGGL((k_mat_bc<T, EXACT>), dim3(p->n_bc_wg), dim3(MAT_THREADS), 0, s, a,
                       adj ? 1 : 0);
    TDX_TRY_HIP(hipGetLastError());
    if (adj && p->d.n_touch > 0) {
      hipLaunchKernelGGL((k_mat_bc_adj<T>), dim3((p->d.n_touch + 255) / 256), dim3(256), 0, s,
                         a);
      TDX_TRY_HIP(hipGetLastError());
    }
  }
  if (loss) {
    hipLaunchKernelGGL((k_mat_loss<T>), dim3(1), dim3(256), 0, s, a.lp,
                       p->n_tile_wg + p->n_bc_wg, loss);
    TDX_TRY_HIP(hipGetLastError());
  }
  return TDX_OK;
}

template <typename T>
int mat_dispatch(const tdx_mat_plan *p, MatArgs<T> a, bool adj, T *loss, hipStream_t s) {
  if (p->d.exact_div) return mat_run<T, true>(p, a, adj, loss, s);
  return mat_run<T, false>(p, a, adj, loss, s);
}

}  // namespace

extern "C" int tdx_mat_plan_create(const tdx_mat_desc *d, tdx_mat_plan **out) {
  TDX_REQUIRE(d && out, "null argument");
  *out = nullptr;
  TDX_REQUIRE(d->dtype == 0 || d->dtype == 1, "dtype must be 0 or 1");
  TDX_REQUIRE(d->n_eq >= 1 && d->n_eq <= MMAX_EQ, "n_eq must be 1..%d", MMAX_EQ);
  TDX_REQUIRE(d->n0 >= 1 && d->n1 >= 2, "bad grid shape");
  TDX_REQUIRE(d->dp >= 2 && d->dp <= MMAX_DP, "derivative_points must be 2..%d", MMAX_DP);
  auto in_ints = [&](int64_t off, int64_t n) { return off >= 0 && n >= 0 && off + n <= d->n_ints; };
  auto in_reals = [&](int64_t off, int64_t n) { return off >= 0 && n >= 0 && off + n <= d->n_reals; };
  auto in_data = [&](int64_t off, int64_t n) { return off >= 0 && n >= 0 && off + n <= d->n_data; };
  const int64_t N = (int64_t)d->n0 * d->n1;
  TDX_REQUIRE(N * d->n_eq < (1LL << 31), "grid too large");

  tdx_mat_plan *p = new tdx_mat_plan();
  p->d = *d;
  for (int i = 0; i < d->dp; ++i) {
    p->ab[i] = d->alpha_back[i];
    p->af[i] = d->alpha_forw[i];
  }
  int cnt0 = 0, cnt1 = 0;
  p->n_fs = 0;
  auto fail = [&](const char *msg) {
    delete p;
    set_error("%s", msg);
    return TDX_EINVAL;
  };
  if (!in_ints(d->eq_i, (int64_t)d->n_eq * TDX_MEQ_INTS)) return fail("eq table out of range");
  for (int q = 0; q < d->n_eq; ++q) {
    const int32_t *E = d->ints + d->eq_i + q * TDX_MEQ_INTS;
    if (E[0] < 1 || !in_ints(E[1], (int64_t)E[0] * TDX_MTERM_INTS) || !in_reals(E[2], 1))
      return fail("term table out of range");
    for (int t = 0; t < E[0]; ++t) {
      const int32_t *Tm = d->ints + E[1] + t * TDX_MTERM_INTS;
      if (Tm[0] < 1 || !in_ints(Tm[1], (int64_t)Tm[0] * TDX_MFACTOR_INTS))
        return fail("factor table out of range");
      if (!((Tm[2] == 0 && in_reals(Tm[3], 1)) || (Tm[2] == 1 && in_data(Tm[3], N))))
        return fail("coefficient out of range");
      for (int f = 0; f < Tm[0]; ++f) {
        const int32_t *F = d->ints + Tm[1] + f * TDX_MFACTOR_INTS;
        if (F[0] < 0 || F[0] >= d->n_eq) return fail("factor variable out of range");
        if (F[1] != 0 && !in_reals(F[3], 1)) return fail("power out of range");
        if (F[4] < 0 || F[4] > 3) return fail("derivative order must be 0..3");
        if (p->n_fs >= MMAX_FS) return fail("too many factors");
        int c0 = 0, c1 = 0;
        for (int k = 0; k < F[4]; ++k) {
          if (F[5 + k] != 0 && F[5 + k] != 1) return fail("axis out of range");
          if (F[5 + k] == 0) ++c0; else ++c1;
        }
        if (d->n0 == 1 && c0 > 0) return fail("derivative along a length-1 axis");
        if (c0 > 0 && d->n0 < 2 * d->dp - 1) return fail("axis 0 too short for the stencil");
        if (c1 > 0 && d->n1 < 2 * d->dp - 1) return fail("axis 1 too short for the stencil");
        cnt0 = std::max(cnt0, c0);
        cnt1 = std::max(cnt1, c1);
        p->fs_var[p->n_fs] = F[0];
        p->fs_nax[p->n_fs] = F[4];
        for (int k = 0; k < 3; ++k) p->fs_ax[p->n_fs][k] = F[5 + k];
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
            if (F[0] < 0 || F[0] >= d->n_eq || F[4] < 0 || F[4] > 3) return fail("bc factor");
            if (F[1] != 0 && !in_reals(F[3], 1)) return fail("bc power");
            for (int k = 0; k < F[4]; ++k) {
              if (F[5 + k] != 0 && F[5 + k] != 1) return fail("bc axis");
              if (F[5 + k] == 0 && d->n0 < 2 * d->dp - 1) return fail("bc axis 0 too short");
              if (F[5 + k] == 1 && d->n1 < 2 * d->dp - 1) return fail("bc axis 1 too short");
            }
          }
          nf += Tm[0];
        }
        nseed = nf * B[0];
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
  p->R0 = cnt0 * (d->dp - 1);
  p->R1 = cnt1 * (d->dp - 1);
  const size_t ts = d->dtype == 0 ? 4 : 8;
  // tile shape: 1-D grids use one long row
  if (d->n0 == 1) {
    p->T0 = 1;
    p->T1 = 1024;
  } else {
    p->T0 = 16;
    p->T1 = 64;
  }
  auto lds_for = [&](int T0, int T1, bool adj) {
    const size_t S = (size_t)(T0 + 4 * p->R0) * (T1 + 4 * p->R1);
    return (S * (d->n_eq + p->n_fs + 2 + (adj ? d->n_eq : 0)) + 2 * MAT_THREADS) * ts;
  };
  while (lds_for(p->T0, p->T1, true) > 160 * 1024 && p->T1 > 16) p->T1 /= 2;
  while (lds_for(p->T0, p->T1, true) > 160 * 1024 && p->T0 > 1) p->T0 /= 2;
  if (lds_for(p->T0, p->T1, true) > 160 * 1024) return fail("stencil reach too large for LDS");
  p->tile_lds = lds_for(p->T0, p->T1, true);
  p->tile_lds_fwd = lds_for(p->T0, p->T1, false);
  p->tiles0 = (d->n0 + p->T0 - 1) / p->T0;
  p->tiles1 = (d->n1 + p->T1 - 1) / p->T1;
  p->n_tile_wg = p->tiles0 * p->tiles1;
  p->n_bc_wg = (d->n_bc + MAT_THREADS - 1) / MAT_THREADS;

  int rc = p->ints.upload(d->ints, (size_t)std::max<int64_t>(d->n_ints, 1) * 4);
  if (!rc) rc = p->reals.upload(d->reals, (size_t)d->n_reals * ts);
  if (!rc) rc = p->data.upload(d->data, (size_t)std::max<int64_t>(d->n_data, 1) * ts);
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
  off = align_up(off + (size_t)(p->n_tile_wg + p->n_bc_wg) * 2 * ts, 256);
  p->off_seeds = off;
  off = align_up(off + (size_t)std::max(d->n_bc_seeds, 1) * ts, 256);
  p->off_one = off;
  off = align_up(off + ts, 256);
  p->off_grad = off;
  off = align_up(off + (size_t)N * d->n_eq * ts, 256);
  p->ws_bytes = off;
  *out = p;
  return TDX_OK;
}

extern "C" void tdx_mat_plan_destroy(tdx_mat_plan *p) {
  if (!p) return;
  p->ints.release();
  p->reals.release();
  p->data.release();
  delete p;
}

extern "C" size_t tdx_mat_workspace_bytes(const tdx_mat_plan *p) { return p ? p->ws_bytes : 0; }

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

extern "C" int tdx_mat_train_step(const tdx_mat_plan *p, void *u, void *m, void *v,
                                  int64_t *step, double lr, double b1, double b2, double eps,
                                  void *loss, void *ws, tdx_stream_t stream) {
  TDX_REQUIRE(p && u && m && v && step && ws, "null argument");
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = (int64_t)p->d.n_eq * p->d.n0 * p->d.n1;
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
