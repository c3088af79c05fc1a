// Device code of the 'NN' kernels that the runtime-compiled fused kernel
// shares with the library: argument block, network image and layer
// helpers, the residual program interpreter, and the fused training-step
// body.  Include after tdx_dev.h (and, in the library, tdx.h for the table
// record sizes TDX_*_INTS); self-contained otherwise, so that hipRTC
// compiles this text as it is.
#ifndef TDX_NN_DEV_H
#define TDX_NN_DEV_H

namespace tdx {

constexpr int MAXD = 4;   // network inputs
constexpr int MAXO = 4;   // network outputs
constexpr int MAXF = 4;   // factors per term
constexpr int K1_WAVES = 4;

template <typename T>
struct NNArgs {
  // architecture
  int d_in, n_out, H, NH, P;  // NH = hidden H x H layers
  // tables
  const int32_t *groups;
  const int32_t *ints;
  const T *reals;
  const T *data;
  // work lists
  const int2 *k1_items;       // (group, entry start) per forward workgroup
  const int4 *k2_items;       // (group, row start, row end, inst) per backward workgroup
  const int32_t *inst_k1_ptr, *inst_k1_idx;
  const int32_t *inst_k2_ptr, *inst_k2_idx;
  int n_inst;
  // buffers
  const T *params;            // n_inst * P
  T *packed;                  // n_inst packed network images
  T *seeds;                   // n_rows * n_out
  T *lp;                      // 2 per forward workgroup
  T *partial;                 // P per backward workgroup
  T *resid;
  const T *dloss;
  int acc_lds;                // backward accumulates in LDS
  int tile_rows;
  // seed multipliers of group 0's equations, [n_eq][N] (null: none)
  const T *seed_mult;
  int seed_mult_n;
  // network values per seed row [n_rows][n_out]: u_out receives them
  // (null: not kept); with u_in the forward kernel reads them instead of
  // evaluating the network (null: evaluate)
  T *u_out;
  const T *u_in;
  // persistent launches: work item counts, the largest slot count times
  // n_out (LDS stride of the forward kernel), and the fused step's unit
  // seed scale and Adam step counter (null: not touched)
  int n_k1, n_k2, so_max;
  const int32_t *k1_bounds;   // persistent forward: item range of each workgroup
  T *one_out;
  int64_t *step_inc;
  // fused step: (group, entry start, entry end, forward item of the start)
  // per workgroup, and each group's residual program as a self-contained
  // table pair (ints offset, ints, reals offset, reals)
  const int4 *k3_items;
  const int4 *prog_idx;
  const int32_t *prog_ints;
  const T *prog_reals;
  const int32_t *prog_kind;   // generated program of each group (run-time compiled kernel)
};

template <typename T>
struct Layout {
  int d_in, n_out, H, NH;
  __device__ __host__ int oW1() const { return 0; }
  __device__ __host__ int ob1() const { return H * d_in; }
  __device__ __host__ int oWh(int j) const { return H * d_in + H + j * (H * H + H); }
  __device__ __host__ int obh(int j) const { return oWh(j) + H * H; }
  __device__ __host__ int oWL() const { return H * d_in + H + NH * (H * H + H); }
  __device__ __host__ int obL() const { return oWL() + n_out * H; }
  __device__ __host__ int count() const { return obL() + n_out; }
};

// ---------------------------------------------------------------------------
// packed network image
// ---------------------------------------------------------------------------
// One image per instance, in the workspace, rebuilt from the parameters by
// k_pack at the start of every forward pass (HP = 16 NB):
//   [Wf: NH x HP*HP][small: SW][Wb: NH x HP*HP]
// Wf and Wb hold each hidden layer's W_j as MFMA operand images, laid out so
// that lane l reads the four values of k-steps i = 0..3 of block (xb, mb)
// with one 128-bit load:
//   Wf[j][ob][mb][lane][i] = W_j[16*ob + lane%16][16*mb + crow(lane/16, i)]
//   Wb[j][ib][mb][lane][i] = W_j[16*mb + crow(lane/16, i)][16*ib + lane%16]
// The small arrays are indexed by accumulator position p = 16 mb + 4 q + i,
// which holds hidden feature f = 16 mb + crow(q, i), zero for f >= H:
//   b1[HP] | W1[d_in][HP] | bh[NH][HP] | WL[n_out][HP] | bL[n_out, padded to 4]
// so a lane's four features of one block are one 128-bit load, and padding
// features evaluate to tanh(0) = 0 without a guard.  The forward kernel
// stages the prefix without Wb in LDS, the backward kernel all of it.
__host__ __device__ inline int small_words(int HP, int NH, int d_in, int n_out) {
  return HP * (1 + d_in + NH + n_out) + ((n_out + 3) & ~3);
}
__host__ __device__ inline int image_words(int HP, int NH, int d_in, int n_out) {
  return 2 * NH * HP * HP + small_words(HP, NH, d_in, n_out);
}
// Words of the LDS copy (in T units).
__host__ __device__ inline int stage_words(int HP, int NH, int d_in, int n_out, bool with_wb) {
  return NH * HP * HP * (with_wb ? 2 : 1) + small_words(HP, NH, d_in, n_out);
}

// Hidden feature held at accumulator position p of a block.
template <typename T>
__device__ __forceinline__ int feature_at(int p) {
  return (p & ~15) + crow<T>((p & 15) >> 2, p & 3);
}

// Word `off` of an instance's packed image, from its parameters W (torch
// layout).
template <typename T, int NB>
__device__ __forceinline__ T image_value(const NNArgs<T> &a, const T *W, int off) {
  constexpr int HP = 16 * NB;
  const int NH = a.NH, H = a.H, d = a.d_in, no = a.n_out;
  const int img = HP * HP;
  const int SW = small_words(HP, NH, d, no);
  const Layout<T> L{d, no, H, NH};
  {
    T v = T(0);
    if (off < NH * img || off >= NH * img + SW) {
      const int which = off < NH * img ? 0 : 1;
      if (which) off -= NH * img + SW;
      const int j = off / img;
      int r = off - j * img;
      const int i = r & 3;
      const int lane = (r >> 2) & 63;
      r >>= 8;
      const int mb = r % NB, xb = r / NB;
      int out, in;
      if (which == 0) {
        out = 16 * xb + (lane & 15);
        in = 16 * mb + crow<T>(lane >> 4, i);
      } else {
        out = 16 * mb + crow<T>(lane >> 4, i);
        in = 16 * xb + (lane & 15);
      }
      if (out < H && in < H) v = W[L.oWh(j) + out * H + in];
    } else {
      const int sp = off - NH * img;
      const int blk = sp / HP, f = feature_at<T>(sp - blk * HP);
      if (blk == 0) {
        if (f < H) v = W[L.ob1() + f];
      } else if (blk < 1 + d) {
        if (f < H) v = W[L.oW1() + f * d + (blk - 1)];
      } else if (blk < 1 + d + NH) {
        if (f < H) v = W[L.obh(blk - 1 - d) + f];
      } else if (blk < 1 + d + NH + no) {
        if (f < H) v = W[L.oWL() + (blk - 1 - d - NH) * H + f];
      } else {
        const int o = sp - HP * (1 + d + NH + no);
        if (o < no) v = W[L.obL() + o];
      }
    }
    return v;
  }
}

// ---------------------------------------------------------------------------
// network view: the packed image in global memory or its copy in LDS
// ---------------------------------------------------------------------------
template <typename T, int NB>
struct Net {
  const T *Wf, *Wb;          // layer j image at Wf + j*HP*HP
  const T *b1, *W1, *bh, *WL, *bL;  // small arrays, row stride HP
  Layout<T> L;
};

template <typename T, int NB>
__device__ __forceinline__ Net<T, NB> net_at(const NNArgs<T> &a, const T *base) {
  constexpr int HP = 16 * NB;
  Net<T, NB> n;
  n.L = Layout<T>{a.d_in, a.n_out, a.H, a.NH};
  n.Wf = base;
  const T *sm = base + a.NH * HP * HP;
  n.b1 = sm;
  n.W1 = sm + HP;
  n.bh = n.W1 + a.d_in * HP;
  n.WL = n.bh + a.NH * HP;
  n.bL = n.WL + a.n_out * HP;
  n.Wb = sm + small_words(HP, a.NH, a.d_in, a.n_out);
  return n;
}

template <typename T, int NB>
__device__ __forceinline__ Net<T, NB> make_net(const NNArgs<T> &a, int inst) {
  constexpr int HP = 16 * NB;
  return net_at<T, NB>(a, a.packed + (int64_t)inst * image_words(HP, a.NH, a.d_in, a.n_out));
}

// Cooperative copy of one instance's image (without Wb unless with_wb) into
// LDS at `dst`; returns the network view on the copy.  Ends with a
// workgroup barrier.
template <typename T, int NB>
__device__ Net<T, NB> stage_net(const NNArgs<T> &a, int inst, T *dst, bool with_wb) {
  constexpr int HP = 16 * NB;
  typedef typename V4<T>::type v4;
  const T *src = a.packed + (int64_t)inst * image_words(HP, a.NH, a.d_in, a.n_out);
  const int n4 = stage_words(HP, a.NH, a.d_in, a.n_out, with_wb) / 4;
  const v4 *s4 = reinterpret_cast<const v4 *>(src);
  v4 *d4 = reinterpret_cast<v4 *>(dst);
  for (int k = threadIdx.x; k < n4; k += blockDim.x) d4[k] = s4[k];
  __syncthreads();
  return net_at<T, NB>(a, dst);
}

// Four consecutive values of a small array at accumulator position p.
template <typename T>
__device__ __forceinline__ typename V4<T>::type ld4(const T *p) {
  return *reinterpret_cast<const typename V4<T>::type *>(p);
}

// Layer 1 (VALU): act[mb][i] = tanh(b1[f] + sum_a x[a] W1[f][a]).
template <typename T, int NB>
__device__ __forceinline__ void layer_in(const Net<T, NB> &n, const T *x, int q,
                                         T (&act)[NB][4]) {
  constexpr int HP = 16 * NB;
  typedef typename V4<T>::type v4;
  const int d = n.L.d_in;
#pragma unroll
  for (int mb = 0; mb < NB; ++mb) {
    v4 z = ld4(n.b1 + 16 * mb + 4 * q);
#pragma unroll
    for (int k = 0; k < MAXD; ++k)
      if (k < d) {
        const v4 w = ld4(n.W1 + k * HP + 16 * mb + 4 * q);
#pragma unroll
        for (int i = 0; i < 4; ++i) z[i] = fma(x[k], w[i], z[i]);
      }
#pragma unroll
    for (int i = 0; i < 4; ++i) act[mb][i] = dtanh(z[i]);
  }
}

// acc[ob] += W_j (rows of block ob) x act, W given as an operand image.
template <typename T, int NB>
__device__ __forceinline__ void mma_layer(const T *img, int lane, const T (&b)[NB][4],
                                          typename V4<T>::type (&acc)[NB]) {
  typedef typename V4<T>::type v4;
  const v4 *wp = reinterpret_cast<const v4 *>(img) + lane;
#pragma unroll
  for (int mb = 0; mb < NB; ++mb)
#pragma unroll
    for (int ob = 0; ob < NB; ++ob) {
      const v4 w = wp[(ob * NB + mb) * 64];
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[ob] = mma16(w[i], b[mb][i], acc[ob]);
    }
}

// Hidden layer j (MFMA, the accumulator starts at the bias):
// act <- tanh(b_j + W_j act).
template <typename T, int NB>
__device__ __forceinline__ void layer_hidden(const Net<T, NB> &n, int j, int lane,
                                             T (&act)[NB][4]) {
  constexpr int HP = 16 * NB;
  typedef typename V4<T>::type v4;
  const int q = lane >> 4;
  v4 acc[NB];
#pragma unroll
  for (int ob = 0; ob < NB; ++ob) acc[ob] = ld4(n.bh + j * HP + 16 * ob + 4 * q);
  mma_layer<T, NB>(n.Wf + j * HP * HP, lane, act, acc);
#pragma unroll
  for (int ob = 0; ob < NB; ++ob)
#pragma unroll
    for (int i = 0; i < 4; ++i) act[ob][i] = dtanh(acc[ob][i]);
}

// Output layer (VALU + cross-quarter shuffle): u[o] for the lane's row.
template <typename T, int NB>
__device__ __forceinline__ T layer_out(const Net<T, NB> &n, int o, int q,
                                       const T (&act)[NB][4]) {
  constexpr int HP = 16 * NB;
  typedef typename V4<T>::type v4;
  T s = T(0);
#pragma unroll
  for (int mb = 0; mb < NB; ++mb) {
    const v4 w = ld4(n.WL + o * HP + 16 * mb + 4 * q);
#pragma unroll
    for (int i = 0; i < 4; ++i) s = fma(act[mb][i], w[i], s);
  }
  s += __shfl_xor(s, 16);
  s += __shfl_xor(s, 32);
  return s + n.bL[o];
}

// Access to a residual program's tables by word index: records (i), real
// values (r), per-entry data (d), the entry's network values (u) and its
// seed accumulators (sd_add).  PtrProg reads them through pointers.
template <typename T>
struct PtrProg {
  const int32_t *ints;
  const T *reals;
  const T *data;
  const T *uv;
  T *sd;
  __device__ __forceinline__ int i(int k) const { return ints[k]; }
  __device__ __forceinline__ T r(int k) const { return reals[k]; }
  __device__ __forceinline__ T d(int64_t k) const { return data[k]; }
  __device__ __forceinline__ T u(int k) const { return uv[k]; }
  __device__ __forceinline__ void sd_add(int k, T v) const { sd[k] += v; }
};

// The residual program of one entry e of a group: for each equation
// (records at word eq_i) the residual r = sum over terms of coefficient x
// product of factor powers, each factor a stencil sum of network values
// u(slot * n_out + var); r goes to resid[E[3] + e] when resid is given, the
// loss terms lambda_d (r - target)^2 and lambda_n (r - target)^2 are added
// to lsum and lnsum, and the seeds d/du(slot, var) of 2 w (r - target)
// (times mult[eq * mult_n + e] when mult is given) are added through
// sd_add (accumulators zeroed by the caller) through every factor of every
// term: d(prod_j f_j^p_j)/d f_k = p_k f_k^(p_k-1) * prefix_k * suffix_k.
// A term with a callable power (factor kind 2, its operation chain at word
// F + 2 of the factor record): the reference's running product P = g(P * x)
// at a callable factor, P = P * x ** p at the others, P from 1
// (tedeous/derivative.py:54-55); returns P and writes dP/dx of each factor
// into dx (zero beyond nf).
template <typename T, typename A>
__device__ __forceinline__ T nn_chain_term(const A &g, int fi, int nf, const T (&x)[MAXF],
                                           T (&dx)[MAXF]) {
  T P = T(1), G[MAXF], M[MAXF];
#pragma unroll
  for (int f = 0; f < MAXF; ++f) {
    G[f] = T(0);
    M[f] = T(1);
    if (f < nf) {
      const int F = fi + f * TDX_FACTOR_INTS;
      const int kind = g.i(F + 1);
      if (kind == 2) {
        T v, d;
        chain_apply(g, g.i(F + 2), P * x[f], v, d);
        G[f] = d * P;
        M[f] = d * x[f];
        P = v;
      } else {
        T pw, dp;
        if (kind == 0) {
          const int p = g.i(F + 2);
          pw = ipow(x[f], p);
          dp = p == 0 ? T(0) : T(p) * ipow(x[f], p - 1);
        } else {
          const T p = g.r(g.i(F + 5));
          pw = dpow(x[f], p);
          dp = p * dpow(x[f], p - T(1));
        }
        G[f] = P * dp;
        M[f] = pw;
        P = P * pw;
      }
    }
  }
  T gg = T(1);
#pragma unroll
  for (int f = MAXF - 1; f >= 0; --f) {
    dx[f] = gg * G[f];
    gg = gg * M[f];
  }
  return P;
}
template <typename A>
__device__ __forceinline__ bool nn_term_has_chain(const A &g, int fi, int nf) {
  bool c = false;
  for (int f = 0; f < nf; ++f) c |= g.i(fi + f * TDX_FACTOR_INTS + 1) == 2;
  return c;
}

// CHAIN: whether terms may hold callable powers (the fused kernels' tables
// never do: such plans run the separate kernels).
template <typename T, typename A, bool CHAIN = true>
__device__ __forceinline__ void entry_program(A &&g, int eq_i, int n_eq, int e, int n_out,
                                              T &lsum, T &lnsum, T *resid, const T *mult,
                                              int mult_n) {
  // factor power x**p and its derivative (F: the factor record's word)
  auto fpw = [&](int F, T x) {
    return g.i(F + 1) == 0 ? ipow(x, g.i(F + 2)) : dpow(x, g.r(g.i(F + 5)));
  };
  auto dfpw = [&](int F, T x) {
    if (g.i(F + 1) == 0) {
      const int p = g.i(F + 2);
      return p == 0 ? T(0) : T(p) * ipow(x, p - 1);
    }
    const T p = g.r(g.i(F + 5));
    return p * dpow(x, p - T(1));
  };
  auto coeff = [&](int Tm) {
    return g.i(Tm + 2) == 0 ? g.r(g.i(Tm + 3)) : g.d((int64_t)g.i(Tm + 3) + e);
  };
  auto stencil_sum = [&](int F) {
    const int var = g.i(F), ns = g.i(F + 3), si = g.i(F + 4);
    T gs = T(0);
    for (int k = 0; k < ns; ++k) {
      const int S = si + k * TDX_STENCIL_INTS;
      gs = gs + g.u(g.i(S) * n_out + var) * g.r(g.i(S + 1));
    }
    return gs;
  };
  for (int eq = 0; eq < n_eq; ++eq) {
    const int E = eq_i + eq * TDX_EQ_INTS;
    const int n_terms = g.i(E), term_i = g.i(E + 1);
    T r = T(0);
    for (int t = 0; t < n_terms; ++t) {
      const int Tm = term_i + t * TDX_TERM_INTS;
      const int nf = g.i(Tm), fi = g.i(Tm + 1);
      T der = T(1);
      if (CHAIN && nn_term_has_chain(g, fi, nf)) {
        T x[MAXF], dx[MAXF];
#pragma unroll
        for (int f = 0; f < MAXF; ++f) x[f] = f < nf ? stencil_sum(fi + f * TDX_FACTOR_INTS) : T(0);
        der = nn_chain_term(g, fi, nf, x, dx);
      } else {
        for (int f = 0; f < nf; ++f) {
          const int F = fi + f * TDX_FACTOR_INTS;
          der = der * fpw(F, stencil_sum(F));
        }
      }
      const T tv = coeff(Tm) * der;
      r = (t == 0) ? tv : r + tv;
    }
    const int tg = g.i(E + 2), ro = g.i(E + 3);
    const T target = tg >= 0 ? g.d((int64_t)tg + e) : T(0);
    if (ro >= 0 && resid) resid[ro + e] = r;
    const T diff = r - target;
    const T sq = diff * diff;
    const T lam_d = g.r(g.i(E + 4));
    lsum += lam_d * sq;
    lnsum += g.r(g.i(E + 5)) * sq;
    T dr = T(2) * g.r(g.i(E + 6)) * diff;
    if (mult) dr *= mult[(int64_t)eq * mult_n + e];
    for (int t = 0; t < n_terms; ++t) {
      const int Tm = term_i + t * TDX_TERM_INTS;
      const int nf = g.i(Tm), fi = g.i(Tm + 1);
      T gsv[MAXF], pw[MAXF], pre[MAXF + 1], suf[MAXF + 1], dch[MAXF];
      const bool chain = CHAIN && nn_term_has_chain(g, fi, nf);
#pragma unroll
      for (int f = 0; f < MAXF; ++f) {
        gsv[f] = T(0);
        pw[f] = T(1);
        if (f < nf) {
          const int F = fi + f * TDX_FACTOR_INTS;
          gsv[f] = stencil_sum(F);
          if (!chain) pw[f] = fpw(F, gsv[f]);
        }
      }
      if (chain) (void)nn_chain_term(g, fi, nf, gsv, dch);
      pre[0] = T(1);
#pragma unroll
      for (int f = 0; f < MAXF; ++f) pre[f + 1] = pre[f] * pw[f];
      suf[MAXF] = T(1);
#pragma unroll
      for (int f = MAXF - 1; f >= 0; --f) suf[f] = suf[f + 1] * pw[f];
      const T cdr = coeff(Tm) * dr;
#pragma unroll
      for (int f = 0; f < MAXF; ++f) {
        if (f < nf) {
          const int F = fi + f * TDX_FACTOR_INTS;
          const T gf = chain ? cdr * dch[f] : cdr * dfpw(F, gsv[f]) * pre[f] * suf[f + 1];
          const int var = g.i(F), ns = g.i(F + 3), si = g.i(F + 4);
          for (int k = 0; k < ns; ++k) {
            const int S = si + k * TDX_STENCIL_INTS;
            g.sd_add(g.i(S) * n_out + var, gf * g.r(g.i(S + 1)));
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// wavefront-private transposes
// ---------------------------------------------------------------------------
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wavefront-private transpose buffers are feature-major, [feature][row],
// with a row stride of 20 words: the per-lane writes (16 rows x 4 quarters)
// and the 128-bit operand reads (4 consecutive row positions per lane) both
// touch 64 distinct banks.
constexpr int WRS = 20;
template <int HP>
constexpr int wave_ls() { return WRS * HP / 16; }

// ---------------------------------------------------------------------------
// k_fused: forward, residual program, adjoint and weight gradients in one
// pass (the fused training step of narrow float networks)
// ---------------------------------------------------------------------------
// A workgroup of two wavefronts takes a range of 64-entry blocks of one
// group and walks it in rounds of 16 entries.  In round r wavefront w
// evaluates the network at its own slots s = w', w' + 2, ... with w' = w xor
// (r mod 2) (16 rows per slot, one MFMA tile) and keeps every activation of those tiles in registers; the network
// values go to LDS, where both wavefronts read all slots of the 16 entries.
// Each wavefront then runs the residual program of the 16 entries itself
// (its four lane quarters alike, so every lane of a row holds the row's
// seeds) and back-propagates its own slots' seeds at once.  So the network
// is evaluated once per step, not again inside the adjoint, and no seed
// goes through HBM.  The residual program is a policy: the library's kernel
// interprets the group's tables from an LDS copy (InterpResid); the kernel
// compiled at plan creation runs straight-line code generated from them.
// The per-entry arithmetic is the forward kernel's and the wavefront-private
// backward's, operation for operation: values, seeds and loss pairs (the
// forward kernel's pairwise tree over each 64-entry block) are the same
// bits; the weight gradients add the same per-row terms in another order.
// Requirements (checked at plan creation): one network output, at most two
// inputs, at most 2 x 16 hidden features, at most two H x H layers, at most
// FUSE_WAVES x FUSE_SPW slots per group.
constexpr int FUSE_WAVES = 2;
constexpr int FUSE_SPW = 4;                         // slots per wavefront
constexpr int FUSE_MAX_SLOTS = FUSE_WAVES * FUSE_SPW;
static_assert(FUSE_WAVES == 2, "k_fused picks a wavefront's seeds between two slot sets");
constexpr int FUSE_PI = 512, FUSE_PR = 256;         // program words of one group

// Words before the network values: the wavefronts' transpose buffers,
// reused for the final gradient sum (P words).
__host__ __device__ inline int fused_buf_words(int HP, int P) {
  return (max(FUSE_WAVES * 2 * WRS * HP, P) + 3) & ~3;
}
// Words of the residual policy's LDS area (the interpreter's table copy and
// seed rows; the generated program uses none of it).
__host__ __device__ constexpr int fused_resid_words() {
  return FUSE_PI + FUSE_PR + FUSE_WAVES * 16 * FUSE_MAX_SLOTS;
}
// All LDS words of the fused kernel: buffers, network values (two rounds),
// the loss values of one block (loss, normalized), the residual policy's
// area and the staged network image.
__host__ __device__ inline int fused_lds_words(int HP, int NH, int d_in, int n_out, int P) {
  return fused_buf_words(HP, P) + 2 * 16 * FUSE_MAX_SLOTS + 128 + fused_resid_words() +
         stage_words(HP, NH, d_in, n_out, true);
}

// Residual policy of the library's fused kernel: the group's program tables
// (a self-contained pair, record offsets relative to it) copied to LDS once
// per workgroup and interpreted by entry_program; seeds gathered in an LDS
// row per entry, then returned in registers.
struct InterpResid {
  const int32_t *pints;
  const float *preals;
  const float *data;
  float *sd;
  int n_eq;
  // Copies the tables of group g; the caller's next workgroup barrier makes
  // them visible.  area: fused_resid_words() words of LDS.
  __device__ InterpResid(const NNArgs<float> &a, int g, int neq, float *area, int w) {
    int32_t *pi_l = reinterpret_cast<int32_t *>(area);
    float *pr_l = area + FUSE_PI;
    const int4 pi = a.prog_idx[g];
    for (int k = threadIdx.x; k < pi.y; k += blockDim.x) pi_l[k] = a.prog_ints[pi.x + k];
    for (int k = threadIdx.x; k < pi.w; k += blockDim.x) pr_l[k] = a.prog_reals[pi.z + k];
    pints = pi_l;
    preals = pr_l;
    data = a.data;
    sd = area + FUSE_PI + FUSE_PR + w * 16 * FUSE_MAX_SLOTS;
    n_eq = neq;
  }
  // Entry e (valid) of the row: loss values into ls/lns, the seeds of all
  // slots into sdv (zero where none); u: the row's network values.
  __device__ __forceinline__ void run(int e, bool valid, int row, const float *u,
                                      float (&sdv)[FUSE_MAX_SLOTS], float &ls, float &lns) {
    float *sdr = sd + row * FUSE_MAX_SLOTS;
#pragma unroll
    for (int s = 0; s < FUSE_MAX_SLOTS; ++s) sdr[s] = 0.0f;
    if (valid)
      entry_program<float, PtrProg<float>, false>(PtrProg<float>{pints, preals, data, u, sdr}, 0,
                                                  n_eq, e, 1, ls, lns, (float *)nullptr,
                                                  (const float *)nullptr, 0);
#pragma unroll
    for (int s = 0; s < FUSE_MAX_SLOTS; ++s) sdv[s] = sdr[s];
  }
};

// The fused kernel's body (launch: 64 FUSE_WAVES threads per workgroup,
// fused_lds_words(...) words of dynamic LDS, one workgroup per work item)
// with the residual policy R.
template <int NB, int NH, typename R>
__device__ __forceinline__ void fused_body(const NNArgs<float> &a, float *smem) {
  typedef float T;
  constexpr int HP = 16 * NB;
  constexpr int FB = WRS * HP;
  constexpr int NJ = NH > 0 ? NH : 1;
  typedef typename V4<T>::type v4;
  const int4 item = a.k3_items[blockIdx.x];
  const int32_t *G = a.groups + item.x * TDX_GROUP_INTS;
  const int n_slots = G[2], slot_base_i = G[3], slot_off_r = G[4], n_eq = G[5];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = lane & 15, q = lane >> 4;
  const int d_in = a.d_in, H = a.H, P = a.P;
  const Layout<T> L{d_in, 1, H, NH};

  T *dbuf = smem + w * 2 * FB, *abuf = dbuf + FB;
  T *u_all = smem + fused_buf_words(HP, P);               // [2][16][FUSE_MAX_SLOTS]
  T *lbuf = u_all + 2 * 16 * FUSE_MAX_SLOTS;              // [2][64]
  T *rarea = lbuf + 128;
  T *stage = rarea + fused_resid_words();
  R resid(a, item.x, n_eq, rarea, w);
  const Net<T, NB> net = stage_net<T, NB>(a, G[0], stage, true);  // ends with a barrier
  const T dl_scale = *a.dloss;

  // gradient accumulators, as in the wavefront-private backward (one output
  // and at most two inputs: the first and last layers on the VALU)
  v4 dwacc[NJ][NB * NB];
  T bhacc[NJ][NB];
  T wlacc[NB][4], w1acc[NB][4][2], b1acc[NB][4];
  T blacc = T(0);
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
#pragma unroll
    for (int b = 0; b < NB * NB; ++b) dwacc[j][b] = v4{0, 0, 0, 0};
#pragma unroll
    for (int b = 0; b < NB; ++b) bhacc[j][b] = T(0);
  }
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      wlacc[b][i] = T(0);
      w1acc[b][i][0] = T(0);
      w1acc[b][i][1] = T(0);
      b1acc[b][i] = T(0);
    }

  const int e_end = item.z;
  int blk = item.w;
  int rnd = 0;
  for (int e0 = item.y; e0 < e_end; e0 += 16, ++rnd) {
    const int e = e0 + row;
    const bool valid = e < e_end;
    const int ec = valid ? e : e_end - 1;
    T *ub = u_all + (rnd & 1) * 16 * FUSE_MAX_SLOTS;
    // slot set of this wavefront in this round (w ^ round parity): with an
    // odd slot count the larger set alternates between the two wavefronts,
    // which evens out their work between the round barriers
    const int ws = w ^ (rnd & 1);

    // forward at the wavefront's slots, every activation kept
    T act[FUSE_SPW][NH + 1][NB][4];
    T x[FUSE_SPW][2];
#pragma unroll
    for (int k = 0; k < FUSE_SPW; ++k) {
      const int s = ws + FUSE_WAVES * k;
      if (s < n_slots) {
        const int64_t base = a.ints[slot_base_i + s];
        T xs[MAXD];
#pragma unroll
        for (int c = 0; c < MAXD; ++c)
          xs[c] = (c < d_in) ? a.data[base + (int64_t)ec * d_in + c] +
                                   a.reals[slot_off_r + s * d_in + c]
                             : T(0);
        x[k][0] = xs[0];
        x[k][1] = xs[1];
        layer_in<T, NB>(net, xs, q, act[k][0]);
#pragma unroll
        for (int j = 1; j <= NH; ++j) {
#pragma unroll
          for (int mb = 0; mb < NB; ++mb)
#pragma unroll
            for (int i = 0; i < 4; ++i) act[k][j][mb][i] = act[k][j - 1][mb][i];
          layer_hidden<T, NB>(net, j - 1, lane, act[k][j]);
        }
        const T u = layer_out<T, NB>(net, 0, q, act[k][NH]);
        if (q == 0) ub[row * FUSE_MAX_SLOTS + s] = u;
      }
    }
    __syncthreads();

    // residual program of the 16 entries; loss values of the block
    T sdv[FUSE_MAX_SLOTS];
    T ls = T(0), lns = T(0);
    resid.run(e, valid, row, ub + row * FUSE_MAX_SLOTS, sdv, ls, lns);
    if (w == 0) {
      if ((e0 & 63) == 0) {
        lbuf[lane] = T(0);
        lbuf[64 + lane] = T(0);
      }
      if (q == 0 && valid) {
        lbuf[(e0 & 63) + row] = ls;
        lbuf[64 + (e0 & 63) + row] = lns;
      }
      if (((e0 + 16) & 63) == 0 || e0 + 16 >= e_end) {
        // the forward kernel's pairwise tree over the block's 64 entries
        T v = lbuf[lane], vn = lbuf[64 + lane];
#pragma unroll
        for (int st = 32; st > 0; st >>= 1) {
          v += __shfl_down(v, st);
          vn += __shfl_down(vn, st);
        }
        if (lane == 0) {
          a.lp[2 * blk] = v;
          a.lp[2 * blk + 1] = vn;
        }
        ++blk;
      }
    }

    // adjoint of the wavefront's slots
#pragma unroll
    for (int k = 0; k < FUSE_SPW; ++k) {
      const int s = ws + FUSE_WAVES * k;
      if (s >= n_slots) continue;
      const T sk = ws == 0 ? sdv[FUSE_WAVES * k] : sdv[FUSE_WAVES * k + 1];
      const T so = valid ? sk * dl_scale : T(0);
      if (q == 0) blacc += so;
      T dl[NB][4];
#pragma unroll
      for (int mb = 0; mb < NB; ++mb) {
        const v4 wl = ld4(net.WL + 16 * mb + 4 * q);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          wlacc[mb][i] = fma(so, act[k][NH][mb][i], wlacc[mb][i]);
          const T av = act[k][NH][mb][i];
          dl[mb][i] = fma(wl[i], so, T(0)) * (T(1) - av * av);
        }
      }
#pragma unroll
      for (int j = NH; j >= 1; --j) {
        const int jj = j - 1;
#pragma unroll
        for (int mb = 0; mb < NB; ++mb)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int f = 16 * mb + crow<T>(q, i);
            dbuf[f * WRS + row] = dl[mb][i];
            abuf[f * WRS + row] = act[k][j - 1][mb][i];
          }
        wave_sync();
        {
          v4 av[NB], dv[NB];
#pragma unroll
          for (int xb = 0; xb < NB; ++xb) {
            dv[xb] = *reinterpret_cast<const v4 *>(dbuf + (16 * xb + row) * WRS + 4 * q);
            av[xb] = *reinterpret_cast<const v4 *>(abuf + (16 * xb + row) * WRS + 4 * q);
          }
#pragma unroll
          for (int ob = 0; ob < NB; ++ob) {
            bhacc[jj][ob] += (dv[ob][0] + dv[ob][1]) + (dv[ob][2] + dv[ob][3]);
#pragma unroll
            for (int ib = 0; ib < NB; ++ib) {
              v4 c = dwacc[jj][ob * NB + ib];
#pragma unroll
              for (int kb = 0; kb < 4; ++kb) c = mma16(dv[ob][kb], av[ib][kb], c);
              dwacc[jj][ob * NB + ib] = c;
            }
          }
        }
        wave_sync();
        // delta_{j-1} = (W_jj^T delta_j) * (1 - act_{j-1}^2)
        v4 c[NB];
#pragma unroll
        for (int ib = 0; ib < NB; ++ib) c[ib] = v4{0, 0, 0, 0};
        mma_layer<T, NB>(net.Wb + jj * HP * HP, lane, dl, c);
#pragma unroll
        for (int ib = 0; ib < NB; ++ib)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const T av = act[k][j - 1][ib][i];
            dl[ib][i] = c[ib][i] * (T(1) - av * av);
          }
      }
      // first layer: per-lane sums over the rows, as the backward kernel
#pragma unroll
      for (int mb = 0; mb < NB; ++mb)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const T dv = dl[mb][i];
          b1acc[mb][i] += dv;
          w1acc[mb][i][0] = fma(dv, x[k][0], w1acc[mb][i][0]);
          w1acc[mb][i][1] = fma(dv, x[k][1], w1acc[mb][i][1]);
        }
    }
  }

  // fixed-order sum of the two wavefronts' gradients (the transpose buffers
  // hold it), then the workgroup's partial
  __syncthreads();
  T *cp = smem;
  for (int ww = 0; ww < FUSE_WAVES; ++ww) {
    if (w == ww) {
      const bool first = ww == 0;
      auto put = [&](int pos, T v) { cp[pos] = first ? v : cp[pos] + v; };
#pragma unroll
      for (int ob = 0; ob < NB; ++ob)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int out = 16 * ob + crow<T>(q, i);
          if (out >= H) continue;
#pragma unroll
          for (int jj = 0; jj < NH; ++jj)
#pragma unroll
            for (int ib = 0; ib < NB; ++ib) {
              const int in = 16 * ib + row;
              if (in < H) put(L.oWh(jj) + out * H + in, dwacc[jj][ob * NB + ib][i]);
            }
        }
#pragma unroll
      for (int ob = 0; ob < NB; ++ob)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int out = 16 * ob + crow<T>(q, i);
#pragma unroll
          for (int c = 0; c < 2; ++c) {
            const T v = rowsum16(w1acc[ob][i][c]);
            if (row == 0 && out < H && c < d_in) put(L.oW1() + out * d_in + c, v);
          }
          const T vb = rowsum16(b1acc[ob][i]);
          if (row == 0 && out < H) put(L.ob1() + out, vb);
          const T vl = rowsum16(wlacc[ob][i]);
          if (row == 0 && out < H) put(L.oWL() + out, vl);
        }
#pragma unroll
      for (int jj = 0; jj < NH; ++jj)
#pragma unroll
        for (int xb = 0; xb < NB; ++xb) {
          T v = bhacc[jj][xb];
          v += __shfl_xor(v, 16);
          v += __shfl_xor(v, 32);
          const int f = 16 * xb + row;
          if (q == 0 && f < H) put(L.obh(jj) + f, v);
        }
      const T bl = rowsum16(blacc);
      if (row == 0 && q == 0) put(L.obL(), bl);
    }
    __syncthreads();
  }
  T *out = a.partial + (int64_t)blockIdx.x * P;
  for (int k = threadIdx.x; k < P; k += blockDim.x) out[k] = cp[k];
}

// ---------------------------------------------------------------------------
// block-owning backward (wide float networks)
// ---------------------------------------------------------------------------
// Block-owning backward for wide float networks (HP = 16 NB up to 128).
// A workgroup of 4 wavefronts walks tiles of 64 rows, 16 per wavefront.
// Each wavefront recomputes the forward pass of its 16 rows and stores every
// layer's activations of the tile in feature-major LDS buffers
// ([layer][feature][row], row stride 68: 128-bit operand loads and the
// row-wise stores both free of bank conflicts); the deltas of the current
// layer go to one more such buffer.  The weight gradient dW = delta^T act
// over the 64 rows is computed block by block: wavefront w owns the output
// blocks ob = w, w + 4, ... of every layer and keeps their sums in
// accumulator registers over the whole work item, so no gradient is read
// back or combined between wavefronts until the end.
constexpr int CTRS = 68;  // row stride of the feature-major tile buffers

// acc[ob] += W (rows of block ob) x b where b is this lane's row trow of a
// feature-major LDS buffer (stride CTRS); one input block at a time, so only
// that block's four operands and the NB accumulators are live.
template <typename T, int NB>
__device__ __forceinline__ void mma_layer_fm(const T *img, int lane, const T *src, int trow,
                                             typename V4<T>::type (&acc)[NB]) {
  typedef typename V4<T>::type v4;
  const int q = lane >> 4;
  const v4 *wp = reinterpret_cast<const v4 *>(img) + lane;
  // two input blocks per round, the second block's weights loaded with the
  // first's: their latency is paid once per pair
  auto block = [&](int mb, const v4 (&w)[NB]) {
    T b[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) b[i] = src[(16 * mb + crow<T>(q, i)) * CTRS + trow];
#pragma unroll
    for (int ob = 0; ob < NB; ++ob)
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[ob] = mma16(w[ob][i], b[i], acc[ob]);
  };
#pragma unroll 1
  for (int mb = 0; mb < NB; mb += 2) {
    v4 wa[NB], wb[NB];
#pragma unroll
    for (int ob = 0; ob < NB; ++ob) wa[ob] = wp[(ob * NB + mb) * 64];
    if (mb + 1 < NB) {
#pragma unroll
      for (int ob = 0; ob < NB; ++ob) wb[ob] = wp[(ob * NB + mb + 1) * 64];
    }
    block(mb, wa);
    if (mb + 1 < NB) block(mb + 1, wb);
  }
}

__host__ __device__ inline size_t block_lds_words(int HP, int NH) {
  return (size_t)(NH + 1) * HP * CTRS + (size_t)(MAXD + MAXO) * CTRS + 16;
}
// With the seeds formed in the kernel (FS): the tile's network values and
// seeds, [64 rows][n_out] each, and the group's residual program (its
// self-contained table pair).
__host__ __device__ inline size_t block_fs_lds_words(int HP, int NH) {
  return block_lds_words(HP, NH) + 2 * 64 * MAXO + FUSE_PI + FUSE_PR;
}

// Residual policy of the library's block-owning kernel: the group's
// program tables (a self-contained pair) copied to LDS and interpreted by
// entry_program.  g < 0: no program (the adjoint of a forward pass).
template <typename T>
struct BlockInterp {
  const int32_t *pints;
  const T *preals;
  const T *data;
  __device__ BlockInterp(const NNArgs<T> &a, int g, int32_t *pi_l, T *pr_l) {
    pints = pi_l;
    preals = pr_l;
    data = a.data;
    if (g < 0) return;
    const int4 pi = a.prog_idx[g];
    for (int k = threadIdx.x; k < pi.y; k += blockDim.x) pi_l[k] = a.prog_ints[pi.x + k];
    for (int k = threadIdx.x; k < pi.w; k += blockDim.x) pr_l[k] = a.prog_reals[pi.z + k];
  }
  // Entry e of the group: loss values into ls/lns, seeds added into sd
  // (zeroed by the caller); u: the entry's network values.
  __device__ __forceinline__ void run(int n_eq, int e, int n_out, const T *u, T *sd, T &ls,
                                      T &lns) {
    entry_program<T, PtrProg<T>, false>(PtrProg<T>{pints, preals, data, u, sd}, 0, n_eq, e, n_out, ls, lns,
                     (T *)nullptr, (const T *)nullptr, 0);
  }
};

// FS = false: the adjoint of a forward pass whose seeds are in the
// workspace; work item (group, row start, row end, instance) over the
// group's slot-major rows.
// FS = true (the training step without a separate forward kernel): work
// item (group, entry start, entry end, instance); a 64-row tile holds
// E = 64 / S whole entries of the group's S slots, entry-major (row
// el * S + s: entry el, slot s; rows from E * S on are padding).  After the
// forward recompute every row's network values go to LDS, wavefront 0 runs
// the residual program of the tile's entries (one per lane; loss values
// kept per lane over the item) and writes their seeds, which the adjoint
// then reads from LDS; the item's loss pair (fixed-order lane sum) goes to
// lp[item].  R: the residual policy (BlockInterp, or the generated programs
// of a run-time compiled kernel).  Launch: 256 threads per workgroup,
// block_lds_words / block_fs_lds_words words of dynamic LDS, one workgroup
// per work item.
template <typename T, int NB, int NH, bool FS, typename R>
__device__ __forceinline__ void block_body(const NNArgs<T> &a, T *smem) {
  constexpr int HP = 16 * NB;
  constexpr int OBW = (NB + 3) / 4;  // output blocks owned per wavefront
  constexpr int TR = 64;
  typedef typename V4<T>::type v4;
  // [NH+1][HP][CTRS] activations; the delta of layer j replaces layer j's
  // activations once their last readers are done (dW_{j+1} and delta_j)
  T *actb = smem;
  T *xbuf = actb + (NH + 1) * HP * CTRS;   // [MAXD][CTRS] coordinates
  T *sbuf = xbuf + MAXD * CTRS;            // [MAXO][CTRS] seeds
  T *red = sbuf + MAXO * CTRS;             // 4 wavefront partials of the output bias
  T *ubuf = red + 16;                      // FS: [64][n_out] network values
  T *sdb = ubuf + 64 * MAXO;               // FS: [64][n_out] seeds
  int32_t *pints = reinterpret_cast<int32_t *>(sdb + 64 * MAXO);  // FS: program
  T *preals = reinterpret_cast<T *>(pints + FUSE_PI);
  const int4 item = a.k2_items[blockIdx.x];
  const int32_t *G = a.groups + item.x * TDX_GROUP_INTS;
  const int n_entries = G[1], n_slots = G[2];
  const int slot_base_i = G[3], slot_off_r = G[4];
  const int n_eq = G[5], eq_i = G[6];
  const int64_t row_base = G[7];
  const int E = FS ? 64 / n_slots : 0;    // entries per tile
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = lane & 15, q = lane >> 4;
  const int trow = 16 * w + row;  // this lane's row within the tile
  const int H = a.H, d_in = a.d_in, n_out = a.n_out, P = a.P;
  const Net<T, NB> net = make_net<T, NB>(a, item.w);
  const Layout<T> &L = net.L;
  const T dl_scale = *a.dloss;
  const int col_b1 = d_in, col_wl = d_in + 1;

  v4 dwacc[NH][OBW][NB];
  v4 sacc[OBW];
  T bhacc[NH][OBW];
#pragma unroll
  for (int j = 0; j < NH; ++j)
#pragma unroll
    for (int u = 0; u < OBW; ++u) {
      bhacc[j][u] = T(0);
#pragma unroll
      for (int ib = 0; ib < NB; ++ib) dwacc[j][u][ib] = v4{0, 0, 0, 0};
    }
#pragma unroll
  for (int u = 0; u < OBW; ++u) sacc[u] = v4{0, 0, 0, 0};
  T blacc = T(0);
  T lsum = T(0), lnsum = T(0);  // FS: loss values of wavefront 0's entry lanes
  // FS: the residual program of the group (the interpreter copies its tables
  // to LDS, visible after the tile loop's first barrier)
  R resid(a, FS ? item.x : -1, pints, preals);

  auto layer_buf = [&](int j) { return actb + j * HP * CTRS; };
  // own-row store of a layer's values held as accumulator rows
  auto store_rows = [&](T *buf, const T (&v)[NB][4]) {
#pragma unroll
    for (int mb = 0; mb < NB; ++mb)
#pragma unroll
      for (int i = 0; i < 4; ++i) buf[(16 * mb + crow<T>(q, i)) * CTRS + trow] = v[mb][i];
  };

  const int64_t r_begin = item.y, r_end = item.z;
  for (int64_t t0 = r_begin; t0 < r_end; t0 += (FS ? E : TR)) {
    bool valid;
    int s, e;
    int64_t rc = 0;
    if constexpr (FS) {
      // t0: the tile's first entry; this lane's row is entry el, slot s
      const int el = trow / n_slots;
      s = trow - el * n_slots;
      valid = el < E && t0 + el < r_end;
      e = (int)(valid ? t0 + el : r_end - 1);
    } else {
      const int64_t r = t0 + trow;
      valid = r < r_end;
      rc = valid ? r : r_begin;
      const int lr = (int)(rc - row_base);
      s = lr / n_entries;
      e = lr - s * n_entries;
    }
    T x[MAXD];
    const int64_t base = a.ints[slot_base_i + s];
#pragma unroll
    for (int k = 0; k < MAXD; ++k)
      x[k] = (k < d_in) ? a.data[base + (int64_t)e * d_in + k] + a.reals[slot_off_r + s * d_in + k]
                        : T(0);
    if (q == 0) {
#pragma unroll
      for (int k = 0; k < MAXD; ++k) xbuf[k * CTRS + trow] = x[k];
    }

    // forward recompute of this wavefront's 16 rows, every layer to LDS
    T act[NB][4];
    layer_in<T, NB>(net, x, q, act);
    store_rows(layer_buf(0), act);
#pragma unroll 1
    for (int j = 1; j <= NH; ++j) {
      v4 acc[NB];
#pragma unroll
      for (int ob = 0; ob < NB; ++ob) acc[ob] = ld4(net.bh + (j - 1) * HP + 16 * ob + 4 * q);
      wave_sync();  // the wavefront's own rows of layer j-1 are in LDS
      mma_layer_fm<T, NB>(net.Wf + (j - 1) * HP * HP, lane, layer_buf(j - 1), trow, acc);
#pragma unroll
      for (int ob = 0; ob < NB; ++ob)
#pragma unroll
        for (int i = 0; i < 4; ++i) act[ob][i] = dtanh(acc[ob][i]);
      store_rows(layer_buf(j), act);
    }

    if constexpr (FS) {
      // network values of the tile's rows, then the residual program of its
      // entries (wavefront 0, one entry per lane) writes their seeds
#pragma unroll
      for (int o = 0; o < MAXO; ++o)
        if (o < n_out) {
          const T u = layer_out<T, NB>(net, o, q, act);
          if (q == 0) ubuf[trow * n_out + o] = u;
        }
      __syncthreads();
      if (w == 0 && lane < E) {
        const int so_n = n_slots * n_out;
        T *sdr = sdb + lane * so_n;
        for (int k = 0; k < so_n; ++k) sdr[k] = T(0);
        if (t0 + lane < r_end)
          resid.run(n_eq, (int)(t0 + lane), n_out, ubuf + lane * so_n, sdr, lsum, lnsum);
      }
      __syncthreads();
    }

    // seeds and the delta of the last hidden activation (act holds it)
    T dl[NB][4];
#pragma unroll
    for (int mb = 0; mb < NB; ++mb)
#pragma unroll
      for (int i = 0; i < 4; ++i) dl[mb][i] = T(0);
#pragma unroll
    for (int o = 0; o < MAXO; ++o) {
      if (o < n_out) {
        T sd = T(0);
        if constexpr (FS) {
          sd = sdb[trow * n_out + o];
        } else {
          sd = a.seeds[rc * n_out + o];
        }
        const T so = valid ? sd * dl_scale : T(0);
        if (o == q) blacc += so;
        if (q == 0) sbuf[o * CTRS + trow] = so;
#pragma unroll
        for (int mb = 0; mb < NB; ++mb) {
          const v4 wl = ld4(net.WL + o * HP + 16 * mb + 4 * q);
#pragma unroll
          for (int i = 0; i < 4; ++i) dl[mb][i] = fma(wl[i], so, dl[mb][i]);
        }
      }
    }
#pragma unroll
    for (int mb = 0; mb < NB; ++mb)
#pragma unroll
      for (int i = 0; i < 4; ++i) dl[mb][i] *= (T(1) - act[mb][i] * act[mb][i]);
    __syncthreads();
    // output layer W: columns col_wl + o of the owned small-term tiles,
    // B = the seed column o (zero elsewhere)
    {
      const T *aN = layer_buf(NH);
      const int o = row - col_wl;
#pragma unroll
      for (int u = 0; u < OBW; ++u) {
        const int ob = w + 4 * u;
        if (ob < NB) {
#pragma unroll
          for (int kk = 0; kk < TR / 16; ++kk) {
            const v4 av = *reinterpret_cast<const v4 *>(aN + (16 * ob + row) * CTRS + 16 * kk + 4 * q);
            const v4 sv = (o >= 0 && o < n_out)
                              ? *reinterpret_cast<const v4 *>(sbuf + o * CTRS + 16 * kk + 4 * q)
                              : v4{0, 0, 0, 0};
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) sacc[u] = mma16(av[kb], sv[kb], sacc[u]);
          }
        }
      }
    }
    __syncthreads();  // the last activations' readers are done

#pragma unroll
    for (int j = NH; j >= 1; --j) {
      const int jj = j - 1;
      const T *ain = layer_buf(j - 1);
      T *dbuf = layer_buf(j);  // delta_j over act_j (read for the last time above)
      store_rows(dbuf, dl);
      __syncthreads();
      // owned blocks of dW_jj over the 64 rows; hidden bias from the delta
      // operand (each lane holds 4 rows of one feature per row group)
#pragma unroll
      for (int u = 0; u < OBW; ++u) {
        const int ob = w + 4 * u;
        if (ob < NB) {
          // row groups outermost (not unrolled): only one group's operands
          // are live; each block still sums its rows in the same order
#pragma unroll 1
          for (int kk = 0; kk < TR / 16; ++kk) {
            const v4 dv = *reinterpret_cast<const v4 *>(dbuf + (16 * ob + row) * CTRS + 16 * kk + 4 * q);
            bhacc[jj][u] += (dv[0] + dv[1]) + (dv[2] + dv[3]);
#pragma unroll
            for (int ib = 0; ib < NB; ++ib) {
              const v4 av = *reinterpret_cast<const v4 *>(ain + (16 * ib + row) * CTRS + 16 * kk + 4 * q);
              v4 c = dwacc[jj][u][ib];
#pragma unroll
              for (int kb = 0; kb < 4; ++kb) c = mma16(dv[kb], av[kb], c);
              dwacc[jj][u][ib] = c;
            }
          }
        }
      }
      // delta_{j-1} = (W_jj^T delta_j) * (1 - act_{j-1}^2) for this
      // wavefront's rows; act_{j-1} of the own rows comes back from LDS
      v4 c[NB];
#pragma unroll
      for (int ib = 0; ib < NB; ++ib) c[ib] = v4{0, 0, 0, 0};
      mma_layer_fm<T, NB>(net.Wb + jj * HP * HP, lane, dbuf, trow, c);
#pragma unroll
      for (int ib = 0; ib < NB; ++ib)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const T av = ain[(16 * ib + crow<T>(q, i)) * CTRS + trow];
          dl[ib][i] = c[ib][i] * (T(1) - av * av);
        }
      __syncthreads();
    }
    // first layer: columns c < d_in of the small-term tiles (B = coordinate
    // c) and column d_in (B = 1); delta_0 over act_0
    T *dbuf = layer_buf(0);
    store_rows(dbuf, dl);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < OBW; ++u) {
      const int ob = w + 4 * u;
      if (ob < NB) {
#pragma unroll
        for (int kk = 0; kk < TR / 16; ++kk) {
          const v4 dv = *reinterpret_cast<const v4 *>(dbuf + (16 * ob + row) * CTRS + 16 * kk + 4 * q);
          const v4 xv = row < d_in
                            ? *reinterpret_cast<const v4 *>(xbuf + row * CTRS + 16 * kk + 4 * q)
                            : (row == col_b1 ? v4{1, 1, 1, 1} : v4{0, 0, 0, 0});
#pragma unroll
          for (int kb = 0; kb < 4; ++kb) sacc[u] = mma16(dv[kb], xv[kb], sacc[u]);
        }
      }
    }
    __syncthreads();
  }

  // each wavefront writes the blocks it owns; the output bias is the only
  // sum over wavefronts (fixed order)
  T *out = a.partial + (int64_t)blockIdx.x * P;
#pragma unroll
  for (int u = 0; u < OBW; ++u) {
    const int ob = w + 4 * u;
    if (ob < NB) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int of = 16 * ob + crow<T>(q, i);
        if (of >= H) continue;
#pragma unroll
        for (int jj = 0; jj < NH; ++jj)
#pragma unroll
          for (int ib = 0; ib < NB; ++ib) {
            const int in = 16 * ib + row;
            if (in < H) out[L.oWh(jj) + of * H + in] = dwacc[jj][u][ib][i];
          }
        const T sv = sacc[u][i];
        if (row < d_in) out[L.oW1() + of * d_in + row] = sv;
        else if (row == col_b1) out[L.ob1() + of] = sv;
        else if (row < col_wl + n_out) out[L.oWL() + (row - col_wl) * H + of] = sv;
      }
#pragma unroll
      for (int jj = 0; jj < NH; ++jj) {
        T v = bhacc[jj][u];
        v += __shfl_xor(v, 16);
        v += __shfl_xor(v, 32);
        const int f = 16 * ob + row;
        if (q == 0 && f < H) out[L.obh(jj) + f] = v;
      }
    }
  }
  const T bl = rowsum16(blacc);
  if (row == 0) red[4 * w + q] = bl;
  __syncthreads();
  if ((int)threadIdx.x < n_out) {
    const int o = threadIdx.x;
    out[L.obL() + o] = ((red[o] + red[4 + o]) + red[8 + o]) + red[12 + o];
  }
  if constexpr (FS) {
    if (w == 0) {
#pragma unroll
      for (int st = 32; st > 0; st >>= 1) {
        lsum += __shfl_down(lsum, st);
        lnsum += __shfl_down(lnsum, st);
      }
      if (lane == 0) {
        a.lp[2 * blockIdx.x] = lsum;
        a.lp[2 * blockIdx.x + 1] = lnsum;
      }
    }
  }
}

}  // namespace tdx

#endif  // TDX_NN_DEV_H
