// 'NN' mode: residual/loss forward and adjoint of a tanh MLP evaluated at
// finite-difference-shifted copies of the collocation points.
//
// What the reference does per step (tedeous/derivative.py:30-58,
// tedeous/eval.py:143-193, 283-461, tedeous/losses.py:84-135): for every
// factor of every term it evaluates the network on a materialised shifted
// copy of the grid, combines the copies with the stencil weights, multiplies
// factors and coefficient, sums terms, squares and averages; autograd then
// differentiates the whole graph.
//
// Here:
//   k_pack      packs hidden-layer weights into MFMA operand order.
//   k_forward   one wavefront per 16 entries of a group: for each slot, the
//               shifted inputs are formed in registers and pushed through
//               the network (hidden layers on 16x16x4 MFMA, tanh on VALU);
//               the residual program is then evaluated per entry, squared
//               into the loss, and the seeds dL/du(slot) are stored.
//   k_backward  tiles of 16*NW rows (row = one (slot, entry) pair): the
//               forward is recomputed with activations kept in LDS, the
//               seeds are propagated back through the network, and the
//               weight gradients are accumulated with MFMA (K = rows) into
//               a per-work-item partial; no atomics.
//   k_reduce    sums the partials of each instance in a fixed order and
//               either stores dparams or applies the Adam update.
//   k_loss      sums per-workgroup loss partials in a fixed order.
//
// Register layout of activations (hidden width padded to HP = 16*NB): lane
// l of a wavefront holds row l%16; quarter q = l/16 holds the features
// f = 16*mb + crow(q, i) in register [mb][i].  A hidden layer is computed
// transposed, Z^T = W * A^T, so that the accumulator of one layer is
// directly the B operand of the next (k-step (mb, i) contracts over the
// features held in register [mb][i]); the weights are packed to match.

#include "tdx_common.h"
#include "tdx_nn_dev.h"
#include "tdx_nn_jit.h"

#include <algorithm>
#include <cstdlib>
#include <map>
#include <cmath>
#include <cstring>

namespace tdx {

// Packs every instance's image; in a fused step it also sets the adjoint's
// unit seed scale and advances the Adam step counter (a.one_out,
// a.step_inc), which saves two launches.
template <typename T, int NB>
__global__ void k_pack(NNArgs<T> a) {
  constexpr int HP = 16 * NB;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (a.one_out) *a.one_out = T(1);
    if (a.step_inc) *a.step_inc += 1;
  }
  const int IMG = image_words(HP, a.NH, a.d_in, a.n_out);
  const int64_t total = (int64_t)a.n_inst * IMG;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int inst = (int)(t / IMG);
    const int off = (int)(t - (int64_t)inst * IMG);
    a.packed[t] = image_value<T, NB>(a, a.params + (int64_t)inst * a.P, off);
  }
}

// The same LDS copy as stage_net, gathered from the instance's parameters
// (no packed image in global memory needed).  Ends with a workgroup
// barrier.
template <typename T, int NB>
__device__ void gather_image(const NNArgs<T> &a, int inst, T *dst, bool with_wb) {
  constexpr int HP = 16 * NB;
  const T *W = a.params + (int64_t)inst * a.P;
  const int n = stage_words(HP, a.NH, a.d_in, a.n_out, with_wb);
  for (int k = threadIdx.x; k < n; k += blockDim.x) dst[k] = image_value<T, NB>(a, W, k);
  __syncthreads();
}

template <typename T>
__device__ __forceinline__ T coeff_of(const NNArgs<T> &a, const int32_t *Tm, int e) {
  return Tm[2] == 0 ? a.reals[Tm[3]] : a.data[Tm[3] + e];
}

template <typename T>
__device__ __forceinline__ T fpow(const NNArgs<T> &a, const int32_t *F, T x) {
  return F[1] == 0 ? ipow(x, F[2]) : dpow(x, a.reals[F[5]]);
}

// d(x**p)/dx
template <typename T>
__device__ __forceinline__ T dfpow(const NNArgs<T> &a, const int32_t *F, T x) {
  if (F[1] == 0) {
    const int p = F[2];
    return p == 0 ? T(0) : T(p) * ipow(x, p - 1);
  }
  const T p = a.reals[F[5]];
  return p * dpow(x, p - T(1));
}

// ---------------------------------------------------------------------------
// k_forward
// ---------------------------------------------------------------------------
// LDS regions of the forward kernel: network values and seeds of the 64
// entries (stride so_max words per entry), then 512 words of reduction
// space, then the staged network image.
template <typename T>
__device__ __forceinline__ T *forward_stage_area(const NNArgs<T> &a, T *smem) {
  return smem + 2 * K1_WAVES * 16 * a.so_max + 512;
}

// One forward work item (64 entries of one group): network at every slot,
// residual program, loss partials (lp[2 it], lp[2 it + 1]) and seeds.
template <typename T, int NB, bool STAGE>
__device__ __forceinline__ void forward_item(const NNArgs<T> &a, int it, const Net<T, NB> &net,
                                             T *smem) {
  const int2 item = a.k1_items[it];
  const int32_t *G = a.groups + item.x * TDX_GROUP_INTS;
  const int n_entries = G[1], n_slots = G[2];
  const int slot_base_i = G[3], slot_off_r = G[4], n_eq = G[5], eq_i = G[6];
  const int64_t row_base = G[7];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = lane & 15, q = lane >> 4;
  const int e = item.y + w * 16 + row;
  const bool valid = e < n_entries;
  const int ec = valid ? e : n_entries - 1;
  const int SO = n_slots * a.n_out;
  T *u_lds = smem + (w * 16 + row) * SO;
  T *sd_lds = smem + K1_WAVES * 16 * a.so_max + (w * 16 + row) * SO;
  T *red = smem + 2 * K1_WAVES * 16 * a.so_max;

  if (a.u_in) {
    // network values given: the residual program runs on them
    if (q == 0)
      for (int s = 0; s < n_slots; ++s)
        for (int o = 0; o < a.n_out; ++o)
          u_lds[s * a.n_out + o] = a.u_in[(row_base + (int64_t)s * n_entries + ec) * a.n_out + o];
  } else {
  T act[NB][4];
  for (int s = 0; s < n_slots; ++s) {
    T x[MAXD];
    const int64_t base = a.ints[slot_base_i + s];
#pragma unroll
    for (int k = 0; k < MAXD; ++k)
      x[k] = (k < a.d_in) ? a.data[base + (int64_t)ec * a.d_in + k] +
                                a.reals[slot_off_r + s * a.d_in + k]
                          : T(0);
    layer_in<T, NB>(net, x, q, act);
    for (int j = 0; j < a.NH; ++j) layer_hidden<T, NB>(net, j, lane, act);
    for (int o = 0; o < a.n_out; ++o) {
      const T u = layer_out<T, NB>(net, o, q, act);
      if (q == 0) u_lds[s * a.n_out + o] = u;
      if (a.u_out && q == 0 && valid)
        a.u_out[(row_base + (int64_t)s * n_entries + e) * a.n_out + o] = u;
    }
  }
  }
  __syncthreads();

  // residual, loss and seeds of the workgroup's 64 entries: wavefront 0,
  // one entry per lane (the other wavefronts' SIMDs are free meanwhile)
  T lsum = T(0), lnsum = T(0);
  if (w == 0) {
    const int e = item.y + lane;
    const bool valid = e < n_entries;
    T *u_lds = smem + lane * SO;
    T *sd_lds = smem + K1_WAVES * 16 * a.so_max + lane * SO;
    for (int k = 0; k < SO; ++k) sd_lds[k] = T(0);
    if (valid)
      entry_program<T>(PtrProg<T>{a.ints, a.reals, a.data, u_lds, sd_lds}, eq_i, n_eq, e,
                       a.n_out, lsum, lnsum, a.resid,
                       (a.seed_mult && item.x == 0) ? a.seed_mult : nullptr, a.seed_mult_n);
  }
  __syncthreads();
  // store seeds: row (slot s, entry e) = row_base + s*n_entries + e
  if (valid && q == 0) {
    for (int s = 0; s < n_slots; ++s)
      for (int o = 0; o < a.n_out; ++o)
        a.seeds[(row_base + (int64_t)s * n_entries + e) * a.n_out + o] =
            sd_lds[s * a.n_out + o];
  }
  // deterministic workgroup sum of the loss partials
  red[threadIdx.x] = lsum;
  red[256 + threadIdx.x] = lnsum;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      red[threadIdx.x] += red[threadIdx.x + st];
      red[256 + threadIdx.x] += red[256 + threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    a.lp[2 * it] = red[0];
    a.lp[2 * it + 1] = red[256];
  }
}

// k_forward: one work item per workgroup, or (PERSIST) a persistent grid
// whose workgroup g takes the contiguous items [k1_bounds[g],
// k1_bounds[g + 1]) (ranges of equal cost, slots x entries, set at plan
// creation),
// gathering its instance's network into LDS from the parameters when the
// instance changes (no per-item staging).  The workgroup that holds an
// instance's first item also writes that instance's packed image for the
// backward kernel, so no separate packing launch is needed.  The persistent
// launch also sets the adjoint's unit seed scale and advances the Adam step
// counter of a fused step (a.one_out, a.step_inc).
template <typename T, int NB, bool STAGE, bool PERSIST>
__global__ __launch_bounds__(256) void k_forward(NNArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *smem = reinterpret_cast<T *>(smem_raw);
  if constexpr (PERSIST) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      if (a.one_out) *a.one_out = T(1);
      if (a.step_inc) *a.step_inc += 1;
    }
    const int lo = a.k1_bounds[blockIdx.x], hi = a.k1_bounds[blockIdx.x + 1];
    int cur = -1;
    T *stage = forward_stage_area(a, smem);
    const Net<T, NB> net = net_at<T, NB>(a, stage);
    for (int it = lo; it < hi; ++it) {
      const int inst = a.groups[a.k1_items[it].x * TDX_GROUP_INTS];
      __syncthreads();
      if (!a.u_in && inst != cur) {
        gather_image<T, NB>(a, inst, stage, false);
        cur = inst;
        if (it == a.inst_k1_idx[a.inst_k1_ptr[inst]]) {
          constexpr int HP = 16 * NB;
          const int IMG = image_words(HP, a.NH, a.d_in, a.n_out);
          const T *W = a.params + (int64_t)inst * a.P;
          T *dst = a.packed + (int64_t)inst * IMG;
          for (int k = threadIdx.x; k < IMG; k += blockDim.x) dst[k] = image_value<T, NB>(a, W, k);
        }
      }
      forward_item<T, NB, STAGE>(a, it, net, smem);
    }
  } else {
    Net<T, NB> net;
    if (!a.u_in) {
      const int inst = a.groups[a.k1_items[blockIdx.x].x * TDX_GROUP_INTS];
      if constexpr (STAGE)
        net = stage_net<T, NB>(a, inst, forward_stage_area(a, smem), false);
      else
        net = make_net<T, NB>(a, inst);
    }
    forward_item<T, NB, STAGE>(a, blockIdx.x, net, smem);
  }
}

// ---------------------------------------------------------------------------
// k_backward
// ---------------------------------------------------------------------------
template <typename T, int NB>
__global__ __launch_bounds__(256) void k_backward(NNArgs<T> a) {
  constexpr int HP = 16 * NB;
  constexpr int LS = HP + 4;  // LDS row stride
  typedef typename V4<T>::type v4;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *smem = reinterpret_cast<T *>(smem_raw);
  const int4 item = a.k2_items[blockIdx.x];
  const int32_t *G = a.groups + item.x * TDX_GROUP_INTS;
  const int n_entries = G[1];
  const int slot_base_i = G[3], slot_off_r = G[4];
  const int64_t row_base = G[7];
  const int inst = item.w;
  const int NW = blockDim.x >> 6;
  const int TR = 16 * NW;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = lane & 15, q = lane >> 4;
  const int trow = w * 16 + row;
  const int NACT = a.NH + 1;
  const int H = a.H, d_in = a.d_in, n_out = a.n_out;

  T *act_l = smem;                         // [NACT][TR][LS]
  T *dst = act_l + (int64_t)NACT * TR * LS;  // [TR][LS]
  T *xs = dst + TR * LS;                     // [TR][MAXD]
  T *ss = xs + TR * MAXD;                    // [TR][MAXO]
  T *acc = a.acc_lds ? ss + TR * MAXO : a.partial + (int64_t)blockIdx.x * a.P;

  for (int k = threadIdx.x; k < a.P; k += blockDim.x) acc[k] = T(0);
  __syncthreads();

  const Net<T, NB> net = make_net<T, NB>(a, inst);
  const Layout<T> &L = net.L;
  const T dl = *a.dloss;
  const int64_t r_begin = item.y, r_end = item.z;

  for (int64_t r0 = r_begin; r0 < r_end; r0 += TR) {
    const int64_t r = r0 + trow;
    const bool valid = r < r_end;
    const int64_t rc = valid ? r : r_begin;
    const int lr = (int)(rc - row_base);
    const int s = lr / n_entries;
    const int e = lr - s * n_entries;
    T x[MAXD], so[MAXO];
    const int64_t base = a.ints[slot_base_i + s];
#pragma unroll
    for (int k = 0; k < MAXD; ++k)
      x[k] = (k < d_in) ? a.data[base + (int64_t)e * d_in + k] +
                              a.reals[slot_off_r + s * d_in + k]
                        : T(0);
#pragma unroll
    for (int o = 0; o < MAXO; ++o)
      so[o] = (valid && o < n_out) ? a.seeds[rc * n_out + o] * dl : T(0);
    if (q == 0) {
#pragma unroll
      for (int k = 0; k < MAXD; ++k) xs[trow * MAXD + k] = x[k];
#pragma unroll
      for (int o = 0; o < MAXO; ++o) ss[trow * MAXO + o] = so[o];
    }

    // forward recompute, activations to LDS
    T act[NB][4];
    layer_in<T, NB>(net, x, q, act);
    for (int j = 0; j <= a.NH; ++j) {
      if (j > 0) layer_hidden<T, NB>(net, j - 1, lane, act);
      T *dstl = act_l + ((int64_t)j * TR + trow) * LS;
#pragma unroll
      for (int mb = 0; mb < NB; ++mb)
#pragma unroll
        for (int i = 0; i < 4; ++i) dstl[16 * mb + crow<T>(q, i)] = act[mb][i];
    }
    __syncthreads();

    // output layer: dW_L, db_L (VALU over rows); delta of the last hidden act
    {
      const T *aN = act_l + (int64_t)a.NH * TR * LS;
      for (int t = threadIdx.x; t < n_out * H; t += blockDim.x) {
        const int o = t / H, k = t - o * H;
        T sum = T(0);
        for (int rr = 0; rr < TR; ++rr) sum = fma(ss[rr * MAXO + o], aN[rr * LS + k], sum);
        acc[L.oWL() + t] += sum;
      }
      for (int o = threadIdx.x; o < n_out; o += blockDim.x) {
        T sum = T(0);
        for (int rr = 0; rr < TR; ++rr) sum += ss[rr * MAXO + o];
        acc[L.obL() + o] += sum;
      }
    }
    T dl_r[NB][4];
    {
      const T *aN = act_l + ((int64_t)a.NH * TR + trow) * LS;
#pragma unroll
      for (int mb = 0; mb < NB; ++mb)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int f = 16 * mb + crow<T>(q, i);
          T g = T(0);
          if (f < H) {
            for (int o = 0; o < n_out; ++o) g = fma(net.WL[o * HP + 16 * mb + 4 * q + i], so[o], g);
            const T av = aN[f];
            g *= (T(1) - av * av);
          }
          dl_r[mb][i] = g;
          dst[trow * LS + f] = g;
        }
    }
    __syncthreads();

    for (int j = a.NH; j >= 1; --j) {
      const int jj = j - 1;  // hidden weight layer mapping act[j-1] -> act[j]
      const T *ain = act_l + (int64_t)(j - 1) * TR * LS;
      // dW_jj = delta_j^T act_{j-1}, blocks owned round-robin by wavefronts
      for (int blk = w; blk < NB * NB; blk += NW) {
        const int ob = blk / NB, ib = blk - ob * NB;
        v4 c = v4{0, 0, 0, 0};
        for (int kb = 0; kb < TR / 4; ++kb) {
          const int rr = 4 * kb + q;
          c = mma16(dst[rr * LS + 16 * ob + row], ain[rr * LS + 16 * ib + row], c);
        }
        const int in = 16 * ib + row;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int out = 16 * ob + crow<T>(q, i);
          if (out < H && in < H) acc[L.oWh(jj) + out * H + in] += c[i];
        }
      }
      for (int k = threadIdx.x; k < H; k += blockDim.x) {
        T sum = T(0);
        for (int rr = 0; rr < TR; ++rr) sum += dst[rr * LS + k];
        acc[L.obh(jj) + k] += sum;
      }
      // delta_{j-1} = (W_jj^T delta_j) * (1 - act_{j-1}^2)
      T dn[NB][4];
      {
        v4 c[NB];
#pragma unroll
        for (int ib = 0; ib < NB; ++ib) c[ib] = v4{0, 0, 0, 0};
        mma_layer<T, NB>(net.Wb + jj * HP * HP, lane, dl_r, c);
        const T *arow = ain + trow * LS;
#pragma unroll
        for (int ib = 0; ib < NB; ++ib)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int f = 16 * ib + crow<T>(q, i);
            const T av = arow[f];
            dn[ib][i] = (f < H) ? c[ib][i] * (T(1) - av * av) : T(0);
          }
      }
      __syncthreads();
#pragma unroll
      for (int mb = 0; mb < NB; ++mb)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          dl_r[mb][i] = dn[mb][i];
          dst[trow * LS + 16 * mb + crow<T>(q, i)] = dn[mb][i];
        }
      __syncthreads();
    }
    // first layer: dW_1 = delta_0^T x, db_1
    for (int t = threadIdx.x; t < H * d_in; t += blockDim.x) {
      const int k = t / d_in, c = t - k * d_in;
      T sum = T(0);
      for (int rr = 0; rr < TR; ++rr) sum = fma(dst[rr * LS + k], xs[rr * MAXD + c], sum);
      acc[L.oW1() + t] += sum;
    }
    for (int k = threadIdx.x; k < H; k += blockDim.x) {
      T sum = T(0);
      for (int rr = 0; rr < TR; ++rr) sum += dst[rr * LS + k];
      acc[L.ob1() + k] += sum;
    }
    __syncthreads();
  }
  if (a.acc_lds) {
    T *out = a.partial + (int64_t)blockIdx.x * a.P;
    for (int k = threadIdx.x; k < a.P; k += blockDim.x) out[k] = acc[k];
  }
}

// ---------------------------------------------------------------------------
// k_backward_w: wavefront-private backward for narrow networks
// ---------------------------------------------------------------------------
// Each wavefront walks its own 16-row groups of the work item with no
// workgroup barrier: activations and deltas stay in registers, the sums over
// rows of the bias/first/last-layer gradients are butterfly shuffles over the
// 16 lanes that hold different rows, and the hidden weight gradients are
// MFMAs with K = the wavefront's 16 rows, after a transpose through a
// wavefront-private LDS buffer.  All gradient accumulators are registers;
// at the end the wavefronts are summed in a fixed order (deterministic).

// acc[t][ob] += W (rows of block ob) x b[t] for NT independent tiles; each
// weight operand is loaded once and used by every tile.
template <typename T, int NB, int NT>
__device__ __forceinline__ void mma_layer_nt(const T *img, int lane, const T (&b)[NT][NB][4],
                                             typename V4<T>::type (&acc)[NT][NB]) {
  typedef typename V4<T>::type v4;
  const v4 *wp = reinterpret_cast<const v4 *>(img) + lane;
#pragma unroll
  for (int mb = 0; mb < NB; ++mb)
#pragma unroll
    for (int ob = 0; ob < NB; ++ob) {
      const v4 w = wp[(ob * NB + mb) * 64];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t][ob] = mma16(w[i], b[t][mb][i], acc[t][ob]);
    }
}

// Wavefront-private backward: each wavefront walks 16-row tiles of its work
// item, NT tiles at a time (independent instruction streams, so one tile's
// tanh and transposes overlap the other's matrix products).  Per tile: the
// forward recompute keeps every activation in registers; the adjoint runs
// layer by layer; weight gradients accumulate in MFMA tiles (registers) over
// the whole item, and the workgroup's wavefronts add theirs in a fixed order
// at the end.
// Wavefronts per SIMD the register allocation must allow: 3 for float
// networks of two 32-wide hidden layers (fits without spilling), else the
// compiler's choice.
constexpr int backward_w_min_waves(int bytes, int NB, int NH, int NT) {
  return (bytes == 4 && NB == 2 && NH <= 2 && NT == 1) ? 3 : 1;
}

// Words of the wavefront buffers (then the final sum) of k_backward_w; the
// staged network image follows them.
template <int NB, int NT>
__device__ __forceinline__ int backward_w_buffer_words(int NW, int P) {
  constexpr int FB = WRS * 16 * NB;
  return max(NW * NT * (2 * FB + 2 * 16 * WRS), P);
}

// One backward work item (a range of rows of one group) of the
// wavefront-private backward: its gradient partial goes to partial[it].
template <typename T, int NB, int NH, int NT>
__device__ __forceinline__ void backward_w_item(const NNArgs<T> &a, int it,
                                                const Net<T, NB> &net, T *smem) {
  constexpr int HP = 16 * NB;
  constexpr int FB = WRS * HP;  // words of one feature-major buffer
  typedef typename V4<T>::type v4;
  const int4 item = a.k2_items[it];
  const int32_t *G = a.groups + item.x * TDX_GROUP_INTS;
  const int n_entries = G[1];
  const int slot_base_i = G[3], slot_off_r = G[4];
  const int64_t row_base = G[7];
  const int NW = blockDim.x >> 6;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = lane & 15, q = lane >> 4;
  const int H = a.H, d_in = a.d_in, n_out = a.n_out, P = a.P;
  // wavefront-private feature-major buffers per tile: deltas, activations,
  // and the seeds/coordinates of the 16 rows ([column][row], stride WRS)
  const int WB = 2 * FB + 2 * 16 * WRS;
  T *dbuf[NT], *abuf[NT], *sbuf[NT], *xbuf[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    dbuf[t] = smem + (w * NT + t) * WB;
    abuf[t] = dbuf[t] + FB;
    sbuf[t] = abuf[t] + FB;          // seeds, one column per output
    xbuf[t] = sbuf[t] + 16 * WRS;    // coordinates, one column per input
  }
  const Layout<T> &L = net.L;
  const T dl_scale = *a.dloss;

  // gradient accumulators (MFMA tiles; row m = crow(q, i) is the hidden
  // feature 16 ob + m, column n = row):
  //   dwacc[jj][ob*NB+ib]: hidden W ( out, in )
  //   sacc[ob]: the first and last layers' terms in one tile, by column --
  //     n < d_in first W, d_in first bias, d_in + 1 + o output W row o.
  //     Each product accumulated into the tile has a B operand that is zero
  //     outside its own columns, so the columns add exactly what separate
  //     tiles would.
  //   bhacc[jj][xb]: hidden bias of feature 16 xb + row over the rows
  //     4q .. 4q+3 of each tile, summed over the quarters at the end
  //   blacc: output bias o = q, summed over the row lanes at the end
  v4 dwacc[NH > 0 ? NH : 1][NB * NB];
  v4 sacc[NB];
  T bhacc[NH > 0 ? NH : 1][NB];
#pragma unroll
  for (int j = 0; j < (NH > 0 ? NH : 1); ++j) {
#pragma unroll
    for (int b = 0; b < NB * NB; ++b) dwacc[j][b] = v4{0, 0, 0, 0};
#pragma unroll
    for (int b = 0; b < NB; ++b) bhacc[j][b] = T(0);
  }
#pragma unroll
  for (int b = 0; b < NB; ++b) sacc[b] = v4{0, 0, 0, 0};
  T blacc = T(0);
  const int col_b1 = d_in, col_wl = d_in + 1;
  // one output: its weight gradient per lane (feature 16 mb + crow(q, i) over
  // the lane's rows) on the VALU, summed over the row lanes at the end,
  // instead of a tile column of matrix products
  const bool wl_valu = n_out == 1;
  // up to two inputs: the first layer's weight and bias gradients likewise
  const bool w1_valu = d_in <= 2;
  T wlacc[NB][4], w1acc[NB][4][2], b1acc[NB][4];
#pragma unroll
  for (int b = 0; b < NB; ++b)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      wlacc[b][i] = T(0);
      w1acc[b][i][0] = T(0);
      w1acc[b][i][1] = T(0);
      b1acc[b][i] = T(0);
    }

  const int64_t r_begin = item.y, r_end = item.z;
  for (int64_t g0 = r_begin + 16 * NT * w; g0 < r_end; g0 += 16 * NT * NW) {
    // rows g0 + 16 t + row; (slot, entry) of a row: one division per
    // wavefront iteration, then a bounded walk (rows are slot-major)
    bool valid[NT];
    int64_t rc[NT];
    T x[NT][MAXD];
    {
      const int lr0 = (int)(g0 - row_base);
      int s = lr0 / n_entries;
      int e = lr0 - s * n_entries;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int64_t r = g0 + 16 * t + row;
        valid[t] = r < r_end;
        rc[t] = valid[t] ? r : r_begin;
        int et = e + (valid[t] ? 16 * t + row : 0), st = s;
        while (et >= n_entries) {
          et -= n_entries;
          ++st;
        }
        const int64_t base = a.ints[slot_base_i + st];
#pragma unroll
        for (int k = 0; k < MAXD; ++k)
          x[t][k] = (k < d_in) ? a.data[base + (int64_t)et * d_in + k] +
                                     a.reals[slot_off_r + st * d_in + k]
                               : T(0);
        if (q == 0) {
#pragma unroll
          for (int k = 0; k < MAXD; ++k)
            if (k < d_in) xbuf[t][k * WRS + row] = x[t][k];
        }
      }
    }

    // forward recompute, every activation kept
    T act[NH + 1][NT][NB][4];
#pragma unroll
    for (int t = 0; t < NT; ++t) layer_in<T, NB>(net, x[t], q, act[0][t]);
#pragma unroll
    for (int j = 1; j <= NH; ++j) {
      v4 acc[NT][NB];
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int ob = 0; ob < NB; ++ob) acc[t][ob] = ld4(net.bh + (j - 1) * HP + 16 * ob + 4 * q);
      mma_layer_nt<T, NB, NT>(net.Wf + (j - 1) * HP * HP, lane, act[j - 1], acc);
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int ob = 0; ob < NB; ++ob)
#pragma unroll
          for (int i = 0; i < 4; ++i) act[j][t][ob][i] = dtanh(acc[t][ob][i]);
    }

    // seeds; delta of the last hidden activation
    T dl[NT][NB][4];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
#pragma unroll
      for (int mb = 0; mb < NB; ++mb)
#pragma unroll
        for (int i = 0; i < 4; ++i) dl[t][mb][i] = T(0);
#pragma unroll
      for (int o = 0; o < MAXO; ++o) {
        if (o < n_out) {
          const T so = valid[t] ? a.seeds[rc[t] * n_out + o] * dl_scale : T(0);
          if (o == q) blacc += so;
          if (wl_valu) {
#pragma unroll
            for (int mb = 0; mb < NB; ++mb)
#pragma unroll
              for (int i = 0; i < 4; ++i) wlacc[mb][i] = fma(so, act[NH][t][mb][i], wlacc[mb][i]);
          } else if (q == 0) {
            sbuf[t][o * WRS + row] = so;
          }
#pragma unroll
          for (int mb = 0; mb < NB; ++mb) {
            const v4 wl = ld4(net.WL + o * HP + 16 * mb + 4 * q);
#pragma unroll
            for (int i = 0; i < 4; ++i) dl[t][mb][i] = fma(wl[i], so, dl[t][mb][i]);
          }
        }
      }
#pragma unroll
      for (int mb = 0; mb < NB; ++mb)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const T av = act[NH][t][mb][i];
          dl[t][mb][i] *= (T(1) - av * av);
          if (!wl_valu) abuf[t][(16 * mb + crow<T>(q, i)) * WRS + row] = av;
        }
    }
    wave_sync();
    // output layer: dW_L[o][f] = sum_rows s_o act_NH[f]; k-step kb of quarter
    // q takes row position 4q + kb in both operands
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      if (wl_valu) break;
      const int o = row - col_wl;
      const v4 sv = (o >= 0 && o < n_out)
                        ? *reinterpret_cast<const v4 *>(sbuf[t] + o * WRS + 4 * q)
                        : v4{0, 0, 0, 0};
#pragma unroll
      for (int ob = 0; ob < NB; ++ob) {
        const v4 av = *reinterpret_cast<const v4 *>(abuf[t] + (16 * ob + row) * WRS + 4 * q);
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) sacc[ob] = mma16(av[kb], sv[kb], sacc[ob]);
      }
    }
    wave_sync();

#pragma unroll
    for (int j = NH; j >= 1; --j) {
      const int jj = j - 1;
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int mb = 0; mb < NB; ++mb)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int f = 16 * mb + crow<T>(q, i);
            dbuf[t][f * WRS + row] = dl[t][mb][i];
            abuf[t][f * WRS + row] = act[j - 1][t][mb][i];
          }
      wave_sync();
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        v4 av[NB], dv[NB];
#pragma unroll
        for (int xb = 0; xb < NB; ++xb) {
          dv[xb] = *reinterpret_cast<const v4 *>(dbuf[t] + (16 * xb + row) * WRS + 4 * q);
          av[xb] = *reinterpret_cast<const v4 *>(abuf[t] + (16 * xb + row) * WRS + 4 * q);
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
      v4 c[NT][NB];
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int ib = 0; ib < NB; ++ib) c[t][ib] = v4{0, 0, 0, 0};
      mma_layer_nt<T, NB, NT>(net.Wb + jj * HP * HP, lane, dl, c);
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int ib = 0; ib < NB; ++ib)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const T av = act[j - 1][t][ib][i];
            dl[t][ib][i] = c[t][ib][i] * (T(1) - av * av);
          }
    }
    // first layer: dW_1[f][c] = sum_rows delta_0[f] x_c, bias in column d_in
    if (w1_valu) {
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int mb = 0; mb < NB; ++mb)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const T dv = dl[t][mb][i];
            b1acc[mb][i] += dv;
            w1acc[mb][i][0] = fma(dv, x[t][0], w1acc[mb][i][0]);
            if (MAXD > 1) w1acc[mb][i][1] = fma(dv, x[t][MAXD > 1 ? 1 : 0], w1acc[mb][i][1]);
          }
      continue;
    }
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int mb = 0; mb < NB; ++mb)
#pragma unroll
        for (int i = 0; i < 4; ++i) dbuf[t][(16 * mb + crow<T>(q, i)) * WRS + row] = dl[t][mb][i];
    wave_sync();
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const v4 xv = row < d_in ? *reinterpret_cast<const v4 *>(xbuf[t] + row * WRS + 4 * q)
                               : (row == col_b1 ? v4{1, 1, 1, 1} : v4{0, 0, 0, 0});
#pragma unroll
      for (int ob = 0; ob < NB; ++ob) {
        const v4 dv = *reinterpret_cast<const v4 *>(dbuf[t] + (16 * ob + row) * WRS + 4 * q);
#pragma unroll
        for (int kb = 0; kb < 4; ++kb) sacc[ob] = mma16(dv[kb], xv[kb], sacc[ob]);
      }
    }
    wave_sync();
  }

  // fixed-order sum over the wavefronts into one gradient vector laid out
  // like the parameters; it reuses the wavefront buffers
  __syncthreads();
  T *cp = smem;
  for (int ww = 0; ww < NW; ++ww) {
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
          const T sv = sacc[ob][i];
          if (!w1_valu && row < d_in) put(L.oW1() + out * d_in + row, sv);
          else if (!w1_valu && row == col_b1) put(L.ob1() + out, sv);
          else if (!wl_valu && row >= col_wl && row < col_wl + n_out)
            put(L.oWL() + (row - col_wl) * H + out, sv);
        }
      if (w1_valu) {
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
          }
      }
      if (wl_valu) {
#pragma unroll
        for (int ob = 0; ob < NB; ++ob)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const T v = rowsum16(wlacc[ob][i]);
            const int out = 16 * ob + crow<T>(q, i);
            if (row == 0 && out < H) put(L.oWL() + out, v);
          }
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
      if (row == 0 && q < n_out) put(L.obL() + q, bl);
    }
    __syncthreads();
  }
  T *out = a.partial + (int64_t)it * P;
  for (int k = threadIdx.x; k < P; k += blockDim.x) out[k] = cp[k];
}

// k_backward_w: one work item per workgroup.
template <typename T, int NB, int NH, int NT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(
    backward_w_min_waves(sizeof(T), NB, NH, NT)))) void k_backward_w(NNArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *smem = reinterpret_cast<T *>(smem_raw);
  T *stage = smem + backward_w_buffer_words<NB, NT>(blockDim.x >> 6, a.P);
  const Net<T, NB> net = stage_net<T, NB>(a, a.k2_items[blockIdx.x].w, stage, true);
  backward_w_item<T, NB, NH, NT>(a, blockIdx.x, net, smem);
}

// k_fused: the fused training step with the interpreted residual program.
template <int NB, int NH>
__global__ __launch_bounds__(64 * FUSE_WAVES) __attribute__((amdgpu_waves_per_eu(2))) void
k_fused(NNArgs<float> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  fused_body<NB, NH, InterpResid>(a, reinterpret_cast<float *>(smem_raw));
}

// k_backward_c: the block-owning backward (body: block_body in
// tdx_nn_dev.h) with the interpreted residual program where it forms its
// own seeds.
template <typename T, int NB, int NH, bool FS>
__global__ __launch_bounds__(256) void k_backward_c(NNArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  block_body<T, NB, NH, FS, BlockInterp<T>>(a, reinterpret_cast<T *>(smem_raw));
}

// ---------------------------------------------------------------------------
// reductions and the optimizer
// ---------------------------------------------------------------------------
// Causal loss (tedeous/losses.py:137-182).  Thread k walks column k of the
// N = nt x K points of group 0 (point p = i K + k): res = sum_eq r_eq[p]^2,
// w = exp(-tol * (sum of res over the earlier slices of the column)),
// stores w and w * res.
template <typename T>
__global__ void k_causal_w(const T *resid, int n_eq, int N, int nt, int K, T tol, T *w, T *wr) {
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

// Seeds of group 0's rows (slot-major: row s N + p) times w[p].
template <typename T>
__global__ void k_causal_seeds(T *seeds, const T *w, int N, int n_slots, int n_out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t tot = (int64_t)n_slots * N * n_out;
  if (i >= tot) return;
  const int p = (int)((i / n_out) % N);
  seeds[i] *= w[p];
}

// Weak form (tedeous/eval.py:13-52, 195-221).  R[q][g] = sum over run g of
// c_i r_q[i]^2, one thread per (run, equation), points in order.
template <typename T>
__global__ void k_weak_runs(const T *resid, int n_eq, int N, int G, const int32_t *gstart,
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

// Seed multipliers of group 0's points: dW/dr_i = 2 m_i r_i with m_i = c_i
// on one level and 2 x_g R_g c_i on two (one equation); one workgroup per run.
template <typename T>
__global__ void k_weak_mult(int levels, const int32_t *gstart, const T *c, const T *x,
                            const T *R, T *m) {
  const int g = blockIdx.x;
  const T f = levels == 2 ? T(2) * x[g] * R[g] : T(1);
  for (int i = gstart[g] + threadIdx.x; i < gstart[g + 1]; i += blockDim.x) m[i] = f * c[i];
}

// Per-equation multipliers M[q][i] = 2 x_g R[q][g] c_i (two levels, several
// equations); one workgroup per (run, equation).
template <typename T>
__global__ void k_weak_mult_eq(int N, int G, const int32_t *gstart, const T *c, const T *x,
                               const T *R, T *M) {
  const int g = blockIdx.x % G, q = blockIdx.x / G;
  const T f = T(2) * x[g] * R[q * G + g];
  for (int i = gstart[g] + threadIdx.x; i < gstart[g + 1]; i += blockDim.x)
    M[(int64_t)q * N + i] = f * c[i];
}

// sum_q lambda_q W_q added to the loss, sum_q W_q to the normalized loss;
// W_q = R[q][0] on one level, sum_g x_g R[q][g]^2 on two (fixed order).
template <typename T>
__global__ void k_weak_loss(int levels, int n_eq, int G, const T *x, const T *lam, const T *R,
                            T *loss) {
  __shared__ T red[256];
  T tot = T(0), totn = T(0);
  for (int q = 0; q < n_eq; ++q) {
    T s = T(0);
    for (int g = threadIdx.x; g < G; g += 256) {
      const T v = R[q * G + g];
      s += levels == 2 ? x[g] * (v * v) : v;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
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
__global__ void k_causal_loss(const T *wr, int N, T *loss) {
  __shared__ T red[256];
  T s = T(0);
  for (int p = threadIdx.x; p < N; p += 256) s += wr[p];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const T m = red[0] / T(N);
    loss[0] = loss[0] + m;
    loss[1] = loss[1] + m;
  }
}

// Loss pair of instance i: fixed-order sum of its forward items' partials
// (a workgroup of 256 threads).
template <typename T>
__device__ __forceinline__ void loss_of_instance(const NNArgs<T> &a, int i, T *loss) {
  __shared__ T red[512];
  T s0 = T(0), s1 = T(0);
  for (int k = a.inst_k1_ptr[i] + threadIdx.x; k < a.inst_k1_ptr[i + 1]; k += blockDim.x) {
    const int wg = a.inst_k1_idx[k];
    s0 += a.lp[2 * wg];
    s1 += a.lp[2 * wg + 1];
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
    loss[2 * i] = red[0];
    loss[2 * i + 1] = red[256];
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_loss(NNArgs<T> a, T *loss) {
  loss_of_instance(a, blockIdx.x, loss);
}

template <typename T>
__device__ __forceinline__ void adam_update(T &p, T g, T &m, T &v, int64_t t,
                                            const AdamHyper &h) {
  adam_apply(p, g, m, v, adam_scalars<T>(t, h));
}

// dparams (mode 0) or the Adam update (mode 1) from the per-item partials
// Gradient sums and the Adam update.  The first loss_blocks workgroups
// instead form the loss pairs of the instances (a fused step's k_loss);
// with step_done the step counter already holds this step's number.
template <typename T>
__global__ __launch_bounds__(256) void k_reduce(NNArgs<T> a, int mode, T *dparams, T *params,
                                                T *m, T *v, const int64_t *step, AdamHyper h,
                                                T *loss, int loss_blocks, int step_done) {
  if ((int)blockIdx.x < loss_blocks) {
    loss_of_instance(a, blockIdx.x, loss);
    return;
  }
  const int64_t tot = (int64_t)a.n_inst * a.P;
  const int64_t idx = (blockIdx.x - loss_blocks) * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= tot) return;
  const int inst = (int)(idx / a.P);
  const int p = (int)(idx - (int64_t)inst * a.P);
  // eight independent sums (loads in flight), combined in a fixed order
  T acc[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) acc[u] = T(0);
  int k = a.inst_k2_ptr[inst];
  const int k1 = a.inst_k2_ptr[inst + 1];
  for (; k + 8 <= k1; k += 8) {
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] += a.partial[(int64_t)a.inst_k2_idx[k + u] * a.P + p];
  }
  for (; k < k1; ++k) acc[0] += a.partial[(int64_t)a.inst_k2_idx[k] * a.P + p];
  const T g = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
  if (mode == 0) {
    dparams[idx] = g;
  } else {
    T pv = params[idx], mv = m[idx], vv = v[idx];
    adam_update(pv, g, mv, vv, step_done ? *step : *step + 1, h);
    params[idx] = pv;
    m[idx] = mv;
    v[idx] = vv;
  }
}

// First stage of the gradient sum of plans with few instances (too few
// parameters to keep enough loads in flight with one thread per parameter):
// slice sl of instance i's work items -- a contiguous share of its list, in
// list order -- summed per parameter into out[(sl * n_inst + i) * P + p];
// k_reduce then sums the slices in order.
template <typename T>
__global__ __launch_bounds__(256) void k_reduce_slices(NNArgs<T> a, int n_slices, T *out) {
  const int64_t tot = (int64_t)a.n_inst * a.P;
  const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (idx >= tot) return;
  const int sl = blockIdx.y;
  const int inst = (int)(idx / a.P);
  const int p = (int)(idx - (int64_t)inst * a.P);
  const int k0 = a.inst_k2_ptr[inst], n = a.inst_k2_ptr[inst + 1] - k0;
  int k = k0 + (int)((int64_t)n * sl / n_slices);
  const int k1 = k0 + (int)((int64_t)n * (sl + 1) / n_slices);
  T acc[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) acc[u] = T(0);
  for (; k + 8 <= k1; k += 8) {
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] += a.partial[(int64_t)a.inst_k2_idx[k + u] * a.P + p];
  }
  for (; k < k1; ++k) acc[0] += a.partial[(int64_t)a.inst_k2_idx[k] * a.P + p];
  out[((int64_t)sl * a.n_inst + inst) * a.P + p] =
      ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
}

template <typename T>
__global__ __launch_bounds__(256) void k_adam(int64_t n, T *p, const T *g, T *m, T *v,
                                              const int64_t *step, AdamHyper h, int vec) {
  typedef typename V4<T>::type v4;
  const AdamScalars<T> sc = adam_scalars<T>(*step + 1, h);
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t done = 0;
  if (vec) {
    // 4 consecutive values per access; every pointer is 4-value aligned
    const int64_t n4 = n / 4;
    for (int64_t i = t0; i < n4; i += stride) {
      v4 pv = reinterpret_cast<const v4 *>(p)[i];
      const v4 gv = reinterpret_cast<const v4 *>(g)[i];
      v4 mv = reinterpret_cast<const v4 *>(m)[i];
      v4 vv = reinterpret_cast<const v4 *>(v)[i];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        T pk = pv[k], mk = mv[k], vk = vv[k];
        adam_apply(pk, gv[k], mk, vk, sc);
        pv[k] = pk;
        mv[k] = mk;
        vv[k] = vk;
      }
      reinterpret_cast<v4 *>(p)[i] = pv;
      reinterpret_cast<v4 *>(m)[i] = mv;
      reinterpret_cast<v4 *>(v)[i] = vv;
    }
    done = n4 * 4;
  }
  for (int64_t i = done + t0; i < n; i += stride) {
    T pv = p[i], mv = m[i], vv = v[i];
    adam_apply(pv, g[i], mv, vv, sc);
    p[i] = pv;
    m[i] = mv;
    v[i] = vv;
  }
}

__global__ void k_step_inc(int64_t *step) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *step += 1;
}

template <typename T>
__global__ void k_set_one(T *x) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *x = T(1);
}

// ---------------------------------------------------------------------------
// k_jac: per-residual Jacobian (natural-gradient Gram assembly)
// ---------------------------------------------------------------------------
// The reference forms J row by row with one torch.autograd.grad per residual
// (tedeous/optimizers/ngd.py:57-78).  Here one thread owns one (group,
// entry): it evaluates the network at every slot of its entry, forms, per
// equation, the seeds dr/du(slot, var) of the residual program (the
// k_forward seeds with dr = 1), and for every slot with a non-zero seed runs
// the network forward again and back with those seeds, adding
// seed * du/dparams into its own column r of jt[P][n_resid] (coalesced
// across the threads of a workgroup; no two threads share a column).
// Activations, deltas, values and seeds live in LDS, thread-fastest.
template <typename T>
__global__ void k_jac_zero(T *jt, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    jt[i] = T(0);
}

template <typename T>
__global__ void k_jac(NNArgs<T> a, T *jt, int64_t n_resid) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *smem = reinterpret_cast<T *>(smem_raw);
  const int B = blockDim.x, t = threadIdx.x;
  const int2 item = a.k1_items[blockIdx.x];
  const int32_t *G = a.groups + item.x * TDX_GROUP_INTS;
  const int n_entries = G[1], n_slots = G[2];
  const int slot_base_i = G[3], slot_off_r = G[4], n_eq = G[5], eq_i = G[6];
  const int H = a.H, NH = a.NH, d = a.d_in, no = a.n_out;
  const int SO = n_slots * no;
  const Layout<T> L{d, no, H, NH};
  const T *W = a.params;
  // per-thread LDS arrays, element k of thread t at [k * B + t]
  T *act = smem;                         // (NH + 1) x H
  T *dA = act + (int64_t)(NH + 1) * H * B;   // H
  T *dB = dA + (int64_t)H * B;               // H
  T *uv = dB + (int64_t)H * B;               // SO network values
  T *sd = uv + (int64_t)SO * B;              // SO seeds
#define ACT(l, i) act[((l) * H + (i)) * B + t]
  for (int e = item.y + t; e < item.y + 16 * K1_WAVES && e < n_entries; e += B) {
    // inputs and network values of every slot
    auto input = [&](int s, T (&x)[MAXD]) {
      const int64_t base = a.ints[slot_base_i + s];
#pragma unroll
      for (int k = 0; k < MAXD; ++k)
        x[k] = k < d ? a.data[base + (int64_t)e * d + k] + a.reals[slot_off_r + s * d + k] : T(0);
    };
    auto forward = [&](const T (&x)[MAXD]) {
      for (int i = 0; i < H; ++i) {
        T z = W[L.ob1() + i];
#pragma unroll
        for (int k = 0; k < MAXD; ++k)
          if (k < d) z = fma(W[L.oW1() + i * d + k], x[k], z);
        ACT(0, i) = dtanh(z);
      }
      for (int j = 0; j < NH; ++j)
        for (int i = 0; i < H; ++i) {
          T z = W[L.obh(j) + i];
          for (int k = 0; k < H; ++k) z = fma(W[L.oWh(j) + i * H + k], ACT(j, k), z);
          ACT(j + 1, i) = dtanh(z);
        }
    };
    for (int s = 0; s < n_slots; ++s) {
      T x[MAXD];
      input(s, x);
      forward(x);
      for (int o = 0; o < no; ++o) {
        T u = W[L.obL() + o];
        for (int k = 0; k < H; ++k) u = fma(W[L.oWL() + o * H + k], ACT(NH, k), u);
        uv[(s * no + o) * B + t] = u;
      }
    }
    for (int eq = 0; eq < n_eq; ++eq) {
      const int32_t *E = a.ints + eq_i + eq * TDX_EQ_INTS;
      const int64_t r = (int64_t)E[3] + e;
      if (E[3] < 0) continue;
      for (int k = 0; k < SO; ++k) sd[k * B + t] = T(0);
      // dr/du(slot, var) through every factor of every term
      for (int tt = 0; tt < E[0]; ++tt) {
        const int32_t *Tm = a.ints + E[1] + tt * TDX_TERM_INTS;
        const int nf = Tm[0];
        const PtrTab<T> tab{a.ints, a.reals};
        const bool chain = nn_term_has_chain(tab, Tm[1], nf);
        T gsv[MAXF], pw[MAXF], pre[MAXF + 1], suf[MAXF + 1], dch[MAXF];
#pragma unroll
        for (int f = 0; f < MAXF; ++f) {
          gsv[f] = T(0);
          pw[f] = T(1);
          if (f < nf) {
            const int32_t *F = a.ints + Tm[1] + f * TDX_FACTOR_INTS;
            T gs = T(0);
            for (int k = 0; k < F[3]; ++k) {
              const int32_t *S = a.ints + F[4] + k * TDX_STENCIL_INTS;
              gs = gs + uv[(S[0] * no + F[0]) * B + t] * a.reals[S[1]];
            }
            gsv[f] = gs;
            if (!chain) pw[f] = fpow(a, F, gs);
          }
        }
        if (chain) (void)nn_chain_term(tab, Tm[1], nf, gsv, dch);
        pre[0] = T(1);
#pragma unroll
        for (int f = 0; f < MAXF; ++f) pre[f + 1] = pre[f] * pw[f];
        suf[MAXF] = T(1);
#pragma unroll
        for (int f = MAXF - 1; f >= 0; --f) suf[f] = suf[f + 1] * pw[f];
        const T c = coeff_of(a, Tm, e);
#pragma unroll
        for (int f = 0; f < MAXF; ++f) {
          if (f < nf) {
            const int32_t *F = a.ints + Tm[1] + f * TDX_FACTOR_INTS;
            const T g = chain ? c * dch[f] : c * dfpow(a, F, gsv[f]) * pre[f] * suf[f + 1];
            for (int k = 0; k < F[3]; ++k) {
              const int32_t *S = a.ints + F[4] + k * TDX_STENCIL_INTS;
              sd[(S[0] * no + F[0]) * B + t] += g * a.reals[S[1]];
            }
          }
        }
      }
      // seed * du/dparams of every slot with a non-zero seed
      T *col = jt + r;
      auto add = [&](int p, T v) { col[(int64_t)p * n_resid] += v; };
      for (int s = 0; s < n_slots; ++s) {
        bool any = false;
        for (int o = 0; o < no; ++o) any |= sd[(s * no + o) * B + t] != T(0);
        if (!any) continue;
        T x[MAXD];
        input(s, x);
        forward(x);
        // output layer
        for (int o = 0; o < no; ++o) {
          const T so = sd[(s * no + o) * B + t];
          for (int k = 0; k < H; ++k) add(L.oWL() + o * H + k, so * ACT(NH, k));
          add(L.obL() + o, so);
        }
        T *dc = dA, *dn = dB;
        for (int k = 0; k < H; ++k) {
          T g = T(0);
          for (int o = 0; o < no; ++o) g = fma(W[L.oWL() + o * H + k], sd[(s * no + o) * B + t], g);
          const T av = ACT(NH, k);
          dc[k * B + t] = g * (T(1) - av * av);
        }
        // hidden layers, last to first: layer j maps ACT(j) to ACT(j + 1)
        for (int j = NH - 1; j >= 0; --j) {
          for (int i = 0; i < H; ++i) {
            const T di = dc[i * B + t];
            for (int k = 0; k < H; ++k) add(L.oWh(j) + i * H + k, di * ACT(j, k));
            add(L.obh(j) + i, di);
          }
          for (int k = 0; k < H; ++k) {
            T g = T(0);
            for (int i = 0; i < H; ++i) g = fma(W[L.oWh(j) + i * H + k], dc[i * B + t], g);
            const T av = ACT(j, k);
            dn[k * B + t] = g * (T(1) - av * av);
          }
          T *tmp = dc;
          dc = dn;
          dn = tmp;
        }
        // first layer
        for (int i = 0; i < H; ++i) {
          const T di = dc[i * B + t];
#pragma unroll
          for (int k = 0; k < MAXD; ++k)
            if (k < d) add(L.oW1() + i * d + k, di * x[k]);
          add(L.ob1() + i, di);
        }
      }
    }
  }
#undef ACT
}

}  // namespace tdx

// ===========================================================================
// host side
// ===========================================================================
using namespace tdx;

struct tdx_nn_plan;
static size_t wave_private_lds(const tdx_nn_plan *p, const tdx_nn_desc *d);

struct tdx_nn_plan {
  int dtype, d_in, n_out, H, n_layers, NH, NB, HP, P, n_inst;
  int n_groups;
  int max_slots;      // largest slot count of a group
  int64_t n_resid;    // length of the residual output
  // persistent forward: a grid of CUs x resident workgroups (sized at the
  // first launch) that gathers the networks into LDS from the parameters
  // and writes the packed images; a fused step is then three launches
  int persist = 0;
  int g1 = 0;           // workgroups of the persistent forward
  DevTable k1_bounds;   // their item ranges (g1 + 1 ints)
  int64_t n_rows;
  int n_k1, n_k2;
  int k2_waves;
  int k1_stage;
  int k2_wave_private;
  int k2_nt;          // tiles per wavefront iteration of the wave-private backward
  int k2_block;       // block-owning backward (wide float networks)
  // fused training step (k_fused): its work items, their per-instance
  // lists, the groups' residual programs and the kernel's LDS bytes
  int fused = 0;
  int n_k3 = 0;
  size_t k3_lds = 0;
  DevTable k3_items, ik3p, ik3i, prog_idx, prog_ints, prog_reals, prog_kind;
  // the fused kernel compiled with the plan's generated residual programs
  // (null: the interpreted kernel)
  hipFunction_t fused_fn = nullptr;
  // two-stage gradient sum (plans with few instances): slices per instance,
  // their list (instance i: items s * n_inst + i) and workspace offset
  int red_slices = 1;
  DevTable ik5p, ik5i;
  size_t off_red = 0;
  // training step of wide float networks: the block-owning backward forms
  // its own seeds (FS work items: tiles of whole entries)
  int k2_fs = 0;
  int n_k4 = 0;
  size_t k4_lds = 0;
  DevTable k4_items, ik4p, ik4i;
  // that kernel compiled with the plan's generated residual programs (null:
  // interpreted)
  hipFunction_t block_fn = nullptr;
  int acc_lds;
  size_t k1_lds, k2_lds;
  DevTable groups, ints, reals, data, k1_items, k2_items, ik1p, ik1i, ik2p, ik2i;
  // workspace layout (byte offsets)
  size_t off_packed, off_seeds, off_lp, off_partial, off_one, ws_bytes;
  // causal loss: tolerance, slices, points of group 0 (N = nt * K), its
  // equations and slots; workspace for residuals, weights and w * res
  int causal = 0;
  double causal_tol = 0;
  int c_nt = 0, c_K = 0, c_N = 0, c_neq = 0, c_slots = 0;
  size_t off_resid = 0, off_cw = 0, off_cwr = 0;
  // weak form: levels (0 none), runs, group 0's points and equations, table
  // offsets (data: c_i, x_g, lambda_eq; ints: run starts); workspace for
  // the runs' sums and the points' seed multipliers
  int weak = 0, w_G = 0, w_N = 0, w_neq = 0;
  int64_t w_d = 0, w_i = 0;
  size_t off_wR = 0, off_wm = 0;
  // optional launch timing: event pairs per kernel kind
  int prof_cap = 0;
  int prof_n[TDX_NN_KERNEL_KINDS] = {0};
  std::vector<hipEvent_t> prof_ev[TDX_NN_KERNEL_KINDS];
};

// LDS words of the wavefront-private backward: transpose buffers of every
// tile of the 4 wavefronts (reused for the final sum, so at least P words)
// and the staged network image.
static size_t wave_private_lds(const tdx_nn_plan *p, const tdx_nn_desc *d) {
  const size_t wb = (size_t)2 * 20 * p->HP + 2 * 16 * 20;
  const size_t tr = std::max<size_t>(4 * p->k2_nt * wb, (size_t)p->P);
  return tr + stage_words(p->HP, p->NH, d->d_in, d->n_out, true);
}

namespace {
// Records an event pair around one launch when timing is on.
struct ProfScope {
  tdx_nn_plan *p;
  int kind;
  hipStream_t s;
  hipEvent_t stop = nullptr;
  ProfScope(const tdx_nn_plan *pc, int k, hipStream_t st) : p(const_cast<tdx_nn_plan *>(pc)), kind(k), s(st) {
    if (p->prof_cap > 0 && p->prof_n[kind] < p->prof_cap) {
      const int i = p->prof_n[kind]++;
      (void)hipEventRecord(p->prof_ev[kind][2 * i], s);
      stop = p->prof_ev[kind][2 * i + 1];
    }
  }
  ~ProfScope() {
    if (stop) (void)hipEventRecord(stop, s);
  }
};
}  // namespace

namespace {

template <typename T>
NNArgs<T> make_args(const tdx_nn_plan *p, const void *params, void *ws) {
  NNArgs<T> a;
  a.d_in = p->d_in;
  a.n_out = p->n_out;
  a.H = p->H;
  a.NH = p->NH;
  a.P = p->P;
  a.groups = (const int32_t *)p->groups.ptr;
  a.ints = (const int32_t *)p->ints.ptr;
  a.reals = (const T *)p->reals.ptr;
  a.data = (const T *)p->data.ptr;
  a.k1_items = (const int2 *)p->k1_items.ptr;
  a.k2_items = (const int4 *)p->k2_items.ptr;
  a.inst_k1_ptr = (const int32_t *)p->ik1p.ptr;
  a.inst_k1_idx = (const int32_t *)p->ik1i.ptr;
  a.inst_k2_ptr = (const int32_t *)p->ik2p.ptr;
  a.inst_k2_idx = (const int32_t *)p->ik2i.ptr;
  a.n_inst = p->n_inst;
  char *w = (char *)ws;
  a.params = (const T *)params;
  a.packed = w ? (T *)(w + p->off_packed) : nullptr;
  a.seeds = w ? (T *)(w + p->off_seeds) : nullptr;
  a.lp = w ? (T *)(w + p->off_lp) : nullptr;
  a.partial = w ? (T *)(w + p->off_partial) : nullptr;
  a.resid = nullptr;
  a.dloss = w ? (const T *)(w + p->off_one) : nullptr;
  a.acc_lds = p->acc_lds;
  a.tile_rows = 16 * p->k2_waves;
  a.seed_mult = nullptr;
  a.seed_mult_n = 0;
  a.u_out = nullptr;
  a.u_in = nullptr;
  a.n_k1 = p->n_k1;
  a.n_k2 = p->n_k2;
  a.so_max = p->max_slots * p->n_out;
  a.k1_bounds = (const int32_t *)p->k1_bounds.ptr;
  a.one_out = nullptr;
  a.step_inc = nullptr;
  a.k3_items = (const int4 *)p->k3_items.ptr;
  a.prog_idx = (const int4 *)p->prog_idx.ptr;
  a.prog_ints = (const int32_t *)p->prog_ints.ptr;
  a.prog_reals = (const T *)p->prog_reals.ptr;
  a.prog_kind = (const int32_t *)p->prog_kind.ptr;
  return a;
}

// Resident workgroups of the persistent forward kernel on the current
// device (CUs x workgroups per CU at the plan's LDS size), 0 if unknown.
template <typename T, int NB>
int forward_grid_size(size_t lds) {
  int dev = 0, cus = 0, per_cu = 0;
  const void *k = (const void *)k_forward<T, NB, true, true>;
  raise_lds_limit({k});
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 256, lds) != hipSuccess)
    return 0;
  return cus > 0 && per_cu > 0 ? cus * per_cu : 0;
}

template <typename T>
int forward_grid_size_nb(int NB, size_t lds) {
  switch (NB) {
    case 1: return forward_grid_size<T, 1>(lds);
    case 2: return forward_grid_size<T, 2>(lds);
    case 3: return forward_grid_size<T, 3>(lds);
    case 4: return forward_grid_size<T, 4>(lds);
    case 5: return forward_grid_size<T, 5>(lds);
    case 6: return forward_grid_size<T, 6>(lds);
    case 7: return forward_grid_size<T, 7>(lds);
    case 8: return forward_grid_size<T, 8>(lds);
    default: return 0;
  }
}

template <typename T, int NB>
int launch_pack(const tdx_nn_plan *p, const NNArgs<T> &a, hipStream_t s) {
  const int64_t total = (int64_t)p->n_inst * image_words(p->HP, p->NH, p->d_in, p->n_out);
  const int blocks = (int)std::min<int64_t>((total + 255) / 256, 65535);
  hipLaunchKernelGGL((k_pack<T, NB>), dim3(blocks), dim3(256), 0, s, a);
  TDX_TRY_HIP(hipGetLastError());
  return TDX_OK;
}

template <typename T, int NB>
int launch_forward_kernel(const tdx_nn_plan *p, const NNArgs<T> &a, hipStream_t s) {
  ProfScope ps(p, 0, s);
  if (p->persist) {
    raise_lds_limit({(const void *)k_forward<T, NB, true, true>});
    hipLaunchKernelGGL((k_forward<T, NB, true, true>), dim3(p->g1), dim3(256), p->k1_lds, s, a);
  } else if (p->k1_stage) {
    hipLaunchKernelGGL((k_forward<T, NB, true, false>), dim3(p->n_k1), dim3(256), p->k1_lds, s,
                       a);
  } else {
    hipLaunchKernelGGL((k_forward<T, NB, false, false>), dim3(p->n_k1), dim3(256), p->k1_lds, s,
                       a);
  }
  TDX_TRY_HIP(hipGetLastError());
  return TDX_OK;
}

// Forward pass: network image (the persistent forward writes it itself;
// given network values need none), forward kernel, loss pairs (unless
// fused_loss: the reduction kernel forms them), causal and weak-form
// passes.
template <typename T, int NB>
int launch_forward(const tdx_nn_plan *p, NNArgs<T> a, T *loss, T *resid, hipStream_t s,
                   void *ws, bool fused_loss = false) {
  if (!a.u_in && !p->persist) {
    int rc = launch_pack<T, NB>(p, a, s);
    if (rc) return rc;
  }
  // the causal loss and the weak form read group 0's residuals
  if ((p->causal || p->weak) && !resid) resid = (T *)((char *)ws + p->off_resid);
  a.resid = resid;
  {
    int rc = launch_forward_kernel<T, NB>(p, a, s);
    if (rc) return rc;
  }
  if (loss && !fused_loss) {
    hipLaunchKernelGGL((k_loss<T>), dim3(p->n_inst), dim3(256), 0, s, a, loss);
    TDX_TRY_HIP(hipGetLastError());
  }
  if (p->causal) {
    T *w = (T *)((char *)ws + p->off_cw), *wr = (T *)((char *)ws + p->off_cwr);
    hipLaunchKernelGGL((k_causal_w<T>), dim3((p->c_K + 255) / 256), dim3(256), 0, s,
                       (const T *)resid, p->c_neq, p->c_N, p->c_nt, p->c_K, T(p->causal_tol), w, wr);
    TDX_TRY_HIP(hipGetLastError());
    const int64_t tot = (int64_t)p->c_slots * p->c_N * p->n_out;
    hipLaunchKernelGGL((k_causal_seeds<T>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                       a.seeds, (const T *)w, p->c_N, p->c_slots, p->n_out);
    TDX_TRY_HIP(hipGetLastError());
    if (loss) {
      hipLaunchKernelGGL((k_causal_loss<T>), dim3(1), dim3(256), 0, s, (const T *)wr, p->c_N, loss);
      TDX_TRY_HIP(hipGetLastError());
    }
  }
  if (p->weak) {
    const T *c = (const T *)p->data.ptr + p->w_d;
    const T *x = c + p->w_N, *lam = x + p->w_G;
    const int32_t *gs = (const int32_t *)p->ints.ptr + p->w_i;
    T *R = (T *)((char *)ws + p->off_wR), *m = (T *)((char *)ws + p->off_wm);
    hipLaunchKernelGGL((k_weak_runs<T>), dim3((p->w_G * p->w_neq + 255) / 256), dim3(256), 0, s,
                       (const T *)resid, p->w_neq, p->w_N, p->w_G, gs, c, R);
    TDX_TRY_HIP(hipGetLastError());
    if (p->weak == 2 && p->w_neq > 1) {
      // one multiplier per equation: the forward pass again, each
      // equation's seeds of group 0 scaled by its own 2 x_g R_g c_i
      hipLaunchKernelGGL((k_weak_mult_eq<T>), dim3(p->w_G * p->w_neq), dim3(256), 0, s, p->w_N,
                         p->w_G, gs, c, x, (const T *)R, m);
      TDX_TRY_HIP(hipGetLastError());
      NNArgs<T> a2 = a;
      a2.seed_mult = m;
      a2.seed_mult_n = p->w_N;
      a2.one_out = nullptr;
      a2.step_inc = nullptr;
      int rc = launch_forward_kernel<T, NB>(p, a2, s);
      if (rc) return rc;
    } else {
      hipLaunchKernelGGL((k_weak_mult<T>), dim3(p->w_G), dim3(256), 0, s, p->weak, gs, c, x,
                         (const T *)R, m);
      TDX_TRY_HIP(hipGetLastError());
      const int64_t tot = (int64_t)p->c_slots * p->w_N * p->n_out;
      hipLaunchKernelGGL((k_causal_seeds<T>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                         s, a.seeds, (const T *)m, p->w_N, p->c_slots, p->n_out);
      TDX_TRY_HIP(hipGetLastError());
    }
    if (loss) {
      hipLaunchKernelGGL((k_weak_loss<T>), dim3(1), dim3(256), 0, s, p->weak, p->w_neq, p->w_G, x,
                         lam, (const T *)R, loss);
      TDX_TRY_HIP(hipGetLastError());
    }
  }
  return TDX_OK;
}

template <typename T, int NB, int NH, int NT>
int launch_backward_wt(const tdx_nn_plan *p, const NNArgs<T> &a, hipStream_t s) {
  raise_lds_limit({(const void *)k_backward_w<T, NB, NH, NT>});
  ProfScope ps(p, 1, s);
  hipLaunchKernelGGL((k_backward_w<T, NB, NH, NT>), dim3(p->n_k2), dim3(64 * p->k2_waves),
                     p->k2_lds, s, a);
  TDX_TRY_HIP(hipGetLastError());
  return TDX_OK;
}

// Accumulator registers of the block-owning backward (per lane).
__host__ __device__ constexpr int block_acc_regs(int NB, int NH) {
  return NH * ((NB + 3) / 4) * NB * 4;
}

template <typename T, int NB, int NH, bool FS = false>
int launch_backward_c(const tdx_nn_plan *p, const NNArgs<T> &a, hipStream_t s) {
  if constexpr (block_acc_regs(NB, NH) <= 192) {
    if (FS && p->block_fn) {
      ProfScope ps(p, 1, s);
      NNArgs<T> arg = a;
      void *args[] = {&arg};
      TDX_TRY_HIP(hipModuleLaunchKernel(p->block_fn, p->n_k4, 1, 1, 256, 1, 1,
                                        (unsigned)p->k4_lds, s, args, nullptr));
      return TDX_OK;
    }
    raise_lds_limit({(const void *)k_backward_c<T, NB, NH, FS>});
    ProfScope ps(p, 1, s);
    hipLaunchKernelGGL((k_backward_c<T, NB, NH, FS>), dim3(FS ? p->n_k4 : p->n_k2), dim3(256),
                       FS ? p->k4_lds : p->k2_lds, s, a);
    TDX_TRY_HIP(hipGetLastError());
    return TDX_OK;
  } else {
    set_error("block-owning backward: too many accumulators");
    return TDX_EUNSUPPORTED;
  }
}

template <typename T, int NB, int NH>
int launch_backward_w(const tdx_nn_plan *p, const NNArgs<T> &a, hipStream_t s) {
  if constexpr (sizeof(T) == 4) {
    if (p->k2_nt == 2) return launch_backward_wt<T, NB, NH, 2>(p, a, s);
  }
  return launch_backward_wt<T, NB, NH, 1>(p, a, s);
}

template <typename T, int NB>
int launch_backward(const tdx_nn_plan *p, NNArgs<T> a, hipStream_t s) {
  if (p->k2_wave_private) {
    if constexpr (NB <= 2) {
      switch (p->NH) {
        case 0: return launch_backward_w<T, NB, 0>(p, a, s);
        case 1: return launch_backward_w<T, NB, 1>(p, a, s);
        case 2: return launch_backward_w<T, NB, 2>(p, a, s);
        case 3: return launch_backward_w<T, NB, 3>(p, a, s);
        case 4: return launch_backward_w<T, NB, 4>(p, a, s);
        default: break;
      }
    }
    set_error("no wave-private backward for this shape");
    return TDX_EUNSUPPORTED;
  }
  if (p->k2_block) {
    if constexpr (sizeof(T) == 4 && NB >= 3) {
      switch (p->NH) {
        case 1: return launch_backward_c<T, NB, 1>(p, a, s);
        case 2: return launch_backward_c<T, NB, 2>(p, a, s);
        case 3: return launch_backward_c<T, NB, 3>(p, a, s);
        case 4: return launch_backward_c<T, NB, 4>(p, a, s);
        default: break;
      }
    }
    set_error("no block-owning backward for this shape");
    return TDX_EUNSUPPORTED;
  }
  raise_lds_limit({(const void *)k_backward<T, NB>});
  {
    ProfScope ps(p, 1, s);
    hipLaunchKernelGGL((k_backward<T, NB>), dim3(p->n_k2), dim3(64 * p->k2_waves), p->k2_lds,
                       s, a);
  }
  TDX_TRY_HIP(hipGetLastError());
  return TDX_OK;
}

template <int NB, int NH>
int launch_fused_shape(const tdx_nn_plan *p, const NNArgs<float> &a, hipStream_t s) {
  ProfScope ps(p, 1, s);
  if (p->fused_fn) {
    NNArgs<float> arg = a;
    void *args[] = {&arg};
    TDX_TRY_HIP(hipModuleLaunchKernel(p->fused_fn, p->n_k3, 1, 1, 64 * FUSE_WAVES, 1, 1,
                                      (unsigned)p->k3_lds, s, args, nullptr));
    return TDX_OK;
  }
  raise_lds_limit({(const void *)k_fused<NB, NH>});
  hipLaunchKernelGGL((k_fused<NB, NH>), dim3(p->n_k3), dim3(64 * FUSE_WAVES), p->k3_lds, s, a);
  TDX_TRY_HIP(hipGetLastError());
  return TDX_OK;
}

// The fused step's kernel (the profile counts it as the backward kind).
template <typename T, int NB>
int launch_fused(const tdx_nn_plan *p, const NNArgs<T> &a, hipStream_t s) {
  if constexpr (sizeof(T) == 4 && NB <= 2) {
    switch (p->NH) {
      case 0: return launch_fused_shape<NB, 0>(p, a, s);
      case 1: return launch_fused_shape<NB, 1>(p, a, s);
      case 2: return launch_fused_shape<NB, 2>(p, a, s);
      default: break;
    }
  }
  set_error("no fused step for this shape");
  return TDX_EUNSUPPORTED;
}

// The block-owning backward forming its own seeds (the training step of
// wide float networks without a separate forward kernel); the items and
// their lists are the FS work items.
template <typename T, int NB>
int launch_backward_fs(const tdx_nn_plan *p, NNArgs<T> a, hipStream_t s) {
  a.k2_items = (const int4 *)p->k4_items.ptr;
  if constexpr (sizeof(T) == 4 && NB >= 3) {
    switch (p->NH) {
      case 1: return launch_backward_c<T, NB, 1, true>(p, a, s);
      case 2: return launch_backward_c<T, NB, 2, true>(p, a, s);
      case 3: return launch_backward_c<T, NB, 3, true>(p, a, s);
      case 4: return launch_backward_c<T, NB, 4, true>(p, a, s);
      default: break;
    }
  }
  set_error("no block-owning step for this shape");
  return TDX_EUNSUPPORTED;
}

template <typename T>
int launch_reduce(const tdx_nn_plan *p, const NNArgs<T> &a, int mode, T *dparams, T *params,
                  T *m, T *v, const int64_t *step, AdamHyper h, hipStream_t s,
                  T *loss = nullptr, int step_done = 0) {
  const int64_t tot = (int64_t)p->n_inst * p->P;
  const int loss_blocks = loss ? p->n_inst : 0;
  ProfScope ps(p, 2, s);
  NNArgs<T> ar = a;
  if (p->red_slices > 1) {
    // the slices' sums in the workspace (a.partial is workspace + off_partial)
    T *sl = (T *)((char *)a.partial - p->off_partial + p->off_red);
    hipLaunchKernelGGL((k_reduce_slices<T>), dim3((unsigned)((tot + 255) / 256), p->red_slices),
                       dim3(256), 0, s, a, p->red_slices, sl);
    TDX_TRY_HIP(hipGetLastError());
    ar.partial = sl;
    ar.inst_k2_ptr = (const int32_t *)p->ik5p.ptr;
    ar.inst_k2_idx = (const int32_t *)p->ik5i.ptr;
  }
  hipLaunchKernelGGL((k_reduce<T>), dim3((unsigned)((tot + 255) / 256) + loss_blocks), dim3(256),
                     0, s, ar, mode, dparams, params, m, v, step, h, loss, loss_blocks, step_done);
  TDX_TRY_HIP(hipGetLastError());
  return TDX_OK;
}

// op 0: forward; 1: backward; 2: fused Adam step; 3: forward keeping the
// network values (values = u_out); 4: forward from given network values
// (values = u_in, seeds to `seeds` when given)
template <typename T>
int nn_dispatch(const tdx_nn_plan *p, int op, const void *params, void *params_mut, void *loss,
                void *resid, const void *dloss, void *dparams, void *m, void *v,
                int64_t *step, AdamHyper h, void *ws, hipStream_t s, void *values = nullptr,
                void *seeds = nullptr) {
  NNArgs<T> a = make_args<T>(p, params ? params : params_mut, ws);
  if (op == 3) a.u_out = (T *)values;
  if (op == 4) {
    a.u_in = (const T *)values;
    if (seeds) a.seeds = (T *)seeds;
  }
#define TDX_NB_CASE(NBV)                                                          \
  case NBV: {                                                                     \
    int rc = TDX_OK;                                                              \
    if (op == 0 || op == 3 || op == 4) {                                          \
      rc = launch_forward<T, NBV>(p, a, (T *)loss, (T *)resid, s, ws);            \
    } else if (op == 1) {                                                         \
      a.dloss = (const T *)dloss;                                                 \
      rc = launch_pack<T, NBV>(p, a, s);                                          \
      if (!rc) rc = launch_backward<T, NBV>(p, a, s);                             \
      if (!rc) rc = launch_reduce<T>(p, a, 0, (T *)dparams, nullptr, nullptr,     \
                                     nullptr, nullptr, h, s);                     \
    } else if (p->persist) {                                                      \
      /* three launches: the forward also sets the unit seed scale and */        \
      /* advances the step; the reduction also forms the loss pairs */           \
      a.one_out = (T *)((char *)ws + p->off_one);                                 \
      a.step_inc = step;                                                          \
      rc = launch_forward<T, NBV>(p, a, (T *)loss, nullptr, s, ws, true);         \
      a.one_out = nullptr;                                                        \
      a.step_inc = nullptr;                                                       \
      if (!rc) rc = launch_backward<T, NBV>(p, a, s);                             \
      if (!rc) rc = launch_reduce<T>(p, a, 1, nullptr, (T *)params_mut, (T *)m,  \
                                     (T *)v, step, h, s, (T *)loss, 1);           \
    } else if (p->k2_fs) {                                                        \
      /* three launches: packing (unit seed scale, step), the block-owning */    \
      /* backward forming its own seeds and loss pairs, the reduction */         \
      a.one_out = (T *)((char *)ws + p->off_one);                                 \
      a.step_inc = step;                                                          \
      rc = launch_pack<T, NBV>(p, a, s);                                          \
      a.one_out = nullptr;                                                        \
      a.step_inc = nullptr;                                                       \
      if (!rc) rc = launch_backward_fs<T, NBV>(p, a, s);                          \
      NNArgs<T> ar = a;                                                           \
      ar.inst_k2_ptr = (const int32_t *)p->ik4p.ptr;                              \
      ar.inst_k2_idx = (const int32_t *)p->ik4i.ptr;                              \
      ar.inst_k1_ptr = (const int32_t *)p->ik4p.ptr;                              \
      ar.inst_k1_idx = (const int32_t *)p->ik4i.ptr;                              \
      if (!rc) rc = launch_reduce<T>(p, ar, 1, nullptr, (T *)params_mut, (T *)m, \
                                     (T *)v, step, h, s, (T *)loss, 1);           \
    } else if (p->fused) {                                                        \
      /* three launches: packing (unit seed scale, step), the fused */           \
      /* forward and adjoint, the reduction over the fused items' */             \
      /* partials with the Adam update and the loss pairs */                     \
      a.one_out = (T *)((char *)ws + p->off_one);                                 \
      a.step_inc = step;                                                          \
      rc = launch_pack<T, NBV>(p, a, s);                                          \
      a.one_out = nullptr;                                                        \
      a.step_inc = nullptr;                                                       \
      if (!rc) rc = launch_fused<T, NBV>(p, a, s);                                \
      NNArgs<T> ar = a;                                                           \
      ar.inst_k2_ptr = (const int32_t *)p->ik3p.ptr;                              \
      ar.inst_k2_idx = (const int32_t *)p->ik3i.ptr;                              \
      if (!rc) rc = launch_reduce<T>(p, ar, 1, nullptr, (T *)params_mut, (T *)m, \
                                     (T *)v, step, h, s, (T *)loss, 1);           \
    } else if (!p->causal && !p->weak) {                                          \
      /* four launches: the packing kernel also sets the unit seed scale */      \
      /* and advances the step; the reduction also forms the loss pairs */       \
      a.one_out = (T *)((char *)ws + p->off_one);                                 \
      a.step_inc = step;                                                          \
      rc = launch_pack<T, NBV>(p, a, s);                                          \
      a.one_out = nullptr;                                                        \
      a.step_inc = nullptr;                                                       \
      if (!rc) {                                                                  \
        NNArgs<T> af = a;                                                         \
        af.u_in = nullptr;                                                        \
        rc = launch_forward_kernel<T, NBV>(p, af, s);                             \
      }                                                                           \
      if (!rc) rc = launch_backward<T, NBV>(p, a, s);                             \
      if (!rc) rc = launch_reduce<T>(p, a, 1, nullptr, (T *)params_mut, (T *)m,  \
                                     (T *)v, step, h, s, (T *)loss, 1);           \
    } else {                                                                      \
      rc = launch_forward<T, NBV>(p, a, (T *)loss, nullptr, s, ws);               \
      if (!rc) {                                                                  \
        hipLaunchKernelGGL((k_set_one<T>), dim3(1), dim3(64), 0, s,               \
                           (T *)((char *)ws + p->off_one));                       \
        rc = launch_backward<T, NBV>(p, a, s);                                    \
      }                                                                           \
      if (!rc) rc = launch_reduce<T>(p, a, 1, nullptr, (T *)params_mut, (T *)m,  \
                                     (T *)v, step, h, s);                         \
      if (!rc) {                                                                  \
        hipLaunchKernelGGL(k_step_inc, dim3(1), dim3(64), 0, s, step);            \
        TDX_TRY_HIP(hipGetLastError());                                           \
      }                                                                           \
    }                                                                             \
    return rc;                                                                    \
  }
  switch (p->NB) {
    TDX_NB_CASE(1)
    TDX_NB_CASE(2)
    TDX_NB_CASE(3)
    TDX_NB_CASE(4)
    TDX_NB_CASE(5)
    TDX_NB_CASE(6)
    TDX_NB_CASE(7)
    TDX_NB_CASE(8)
    default:
      set_error("unsupported width");
      return TDX_EUNSUPPORTED;
  }
#undef TDX_NB_CASE
}

// Host-side validation of every table index, so that a malformed table is an
// error code and never an out-of-bounds access on the device.
int validate_nn(const tdx_nn_desc *d, int64_t *n_rows_out) {
  auto in_ints = [&](int64_t off, int64_t n) { return off >= 0 && n >= 0 && off + n <= d->n_ints; };
  auto in_reals = [&](int64_t off, int64_t n) { return off >= 0 && n >= 0 && off + n <= d->n_reals; };
  auto in_data = [&](int64_t off, int64_t n) { return off >= 0 && n >= 0 && off + n <= d->n_data; };
  int64_t n_rows = 0;
  for (int g = 0; g < d->n_groups; ++g) {
    const int32_t *G = d->groups + (int64_t)g * TDX_GROUP_INTS;
    const int inst = G[0], ne = G[1], ns = G[2];
    TDX_REQUIRE(inst >= 0 && inst < d->n_inst, "group %d: bad instance", g);
    TDX_REQUIRE(ne >= 1, "group %d: no entries", g);
    TDX_REQUIRE(ns >= 1 && ns * d->n_out <= 256, "group %d: bad slot count %d", g, ns);
    TDX_REQUIRE(in_ints(G[3], ns), "group %d: slot table out of range", g);
    TDX_REQUIRE(in_reals(G[4], (int64_t)ns * d->d_in), "group %d: offsets out of range", g);
    for (int s = 0; s < ns; ++s)
      TDX_REQUIRE(in_data(d->ints[G[3] + s], (int64_t)ne * d->d_in),
                  "group %d slot %d: coordinates out of range", g, s);
    TDX_REQUIRE(G[5] >= 1 && in_ints(G[6], (int64_t)G[5] * TDX_EQ_INTS), "group %d: eq table", g);
    TDX_REQUIRE(G[7] == n_rows, "group %d: row base %d, expected %lld", g, G[7], (long long)n_rows);
    n_rows += (int64_t)ns * ne;
    for (int q = 0; q < G[5]; ++q) {
      const int32_t *E = d->ints + G[6] + q * TDX_EQ_INTS;
      TDX_REQUIRE(E[0] >= 1 && in_ints(E[1], (int64_t)E[0] * TDX_TERM_INTS), "eq table");
      TDX_REQUIRE(E[2] < 0 || in_data(E[2], ne), "target out of range");
      TDX_REQUIRE(E[3] < 0 || (E[3] + (int64_t)ne <= d->n_resid), "residual slot out of range");
      TDX_REQUIRE(in_reals(E[4], 1) && in_reals(E[5], 1) && in_reals(E[6], 1),
                  "eq weights out of range");
      for (int t = 0; t < E[0]; ++t) {
        const int32_t *Tm = d->ints + E[1] + t * TDX_TERM_INTS;
        TDX_REQUIRE(Tm[0] >= 1 && Tm[0] <= MAXF, "a term needs 1..%d factors", MAXF);
        TDX_REQUIRE(in_ints(Tm[1], (int64_t)Tm[0] * TDX_FACTOR_INTS), "factor table");
        TDX_REQUIRE((Tm[2] == 0 && in_reals(Tm[3], 1)) || (Tm[2] == 1 && in_data(Tm[3], ne)),
                    "coefficient out of range");
        for (int f = 0; f < Tm[0]; ++f) {
          const int32_t *F = d->ints + Tm[1] + f * TDX_FACTOR_INTS;
          TDX_REQUIRE(F[0] >= 0 && F[0] < d->n_out, "factor variable out of range");
          if (F[1] == 2) {
            // callable power: a chain of one-argument operations at F[2]
            TDX_REQUIRE(in_ints(F[2], 1) && d->ints[F[2]] >= 0 &&
                            d->ints[F[2]] <= TDX_AD_MAX_CHAIN &&
                            in_ints(F[2] + 1, 2 * (int64_t)d->ints[F[2]]),
                        "power chain out of range");
            for (int k = 0; k < d->ints[F[2]]; ++k) {
              const int op = d->ints[F[2] + 1 + 2 * k], arg = d->ints[F[2] + 2 + 2 * k];
              TDX_REQUIRE(op >= TDX_AD_OP_ADD && op <= TDX_AD_OP_SQRT, "power chain op");
              TDX_REQUIRE(!(op <= TDX_AD_OP_RDIV || op == TDX_AD_OP_POWR) || in_reals(arg, 1),
                          "power chain constant out of range");
            }
          } else {
            TDX_REQUIRE((F[1] == 0 || F[1] == 1) && (F[1] == 0 || in_reals(F[5], 1)),
                        "power out of range");
          }
          TDX_REQUIRE(F[3] >= 1 && in_ints(F[4], (int64_t)F[3] * TDX_STENCIL_INTS), "stencil");
          for (int k = 0; k < F[3]; ++k) {
            const int32_t *S = d->ints + F[4] + k * TDX_STENCIL_INTS;
            TDX_REQUIRE(S[0] >= 0 && S[0] < ns && in_reals(S[1], 1), "stencil entry");
          }
        }
      }
    }
  }
  *n_rows_out = n_rows;
  return TDX_OK;
}

// Whether a desc's network and options admit the fused training step
// (k_fused): float, at most 32 hidden features and two H x H layers, one
// output, at most two inputs, at most FUSE_MAX_SLOTS slots per group,
// default tiles, no persistent forward, no causal or weak-form loss.
bool fused_shape_ok(const tdx_nn_desc *d, int max_slots) {
  return d->dtype == 0 && d->width <= 32 && d->n_layers - 2 <= 2 && d->d_in <= 2 &&
         d->n_out == 1 && max_slots <= FUSE_MAX_SLOTS && d->tile_rows == 0 &&
         d->persistent == 0 && d->causal_tol == 0.0 && d->weak_levels == 0;
}

// Every group's residual program rewritten as its own table pair (record
// offsets relative to the pair, real values copied, unread words zero);
// false when one exceeds FUSE_PI record or FUSE_PR real words.
bool fused_programs(const tdx_nn_desc *d, std::vector<FusedProg> *out) {
  const float *R = (const float *)d->reals;
  out->clear();
  for (int g = 0; g < d->n_groups; ++g) {
    const int32_t *G = d->groups + (int64_t)g * TDX_GROUP_INTS;
    FusedProg fp;
    fp.n_eq = G[5];
    std::vector<int32_t> &mi = fp.ints;
    std::vector<float> &mr = fp.reals;
    mi.assign((size_t)G[5] * TDX_EQ_INTS, 0);
    auto real = [&](int32_t idx) {
      mr.push_back(R[idx]);
      return (int32_t)mr.size() - 1;
    };
    for (int q = 0; q < G[5]; ++q) {
      const int32_t *E = d->ints + G[6] + q * TDX_EQ_INTS;
      const size_t eo = (size_t)q * TDX_EQ_INTS;
      for (int k = 0; k < TDX_EQ_INTS; ++k) mi[eo + k] = E[k];
      for (int k = 4; k <= 6; ++k) mi[eo + k] = real(E[k]);
      const size_t to = mi.size();
      mi[eo + 1] = (int32_t)to;
      mi.resize(to + (size_t)E[0] * TDX_TERM_INTS);
      for (int t = 0; t < E[0]; ++t) {
        const int32_t *Tm = d->ints + E[1] + t * TDX_TERM_INTS;
        const size_t tt = to + (size_t)t * TDX_TERM_INTS;
        for (int k = 0; k < TDX_TERM_INTS; ++k) mi[tt + k] = Tm[k];
        if (Tm[2] == 0) mi[tt + 3] = real(Tm[3]);
        const size_t fo = mi.size();
        mi[tt + 1] = (int32_t)fo;
        mi.resize(fo + (size_t)Tm[0] * TDX_FACTOR_INTS);
        for (int f = 0; f < Tm[0]; ++f) {
          const int32_t *F = d->ints + Tm[1] + f * TDX_FACTOR_INTS;
          if (F[1] == 2) return false;  // callable powers: the separate kernels
          const size_t ff = fo + (size_t)f * TDX_FACTOR_INTS;
          for (int k = 0; k < TDX_FACTOR_INTS; ++k) mi[ff + k] = F[k];
          mi[ff + 5] = F[1] != 0 ? real(F[5]) : 0;
          const size_t so = mi.size();
          mi[ff + 4] = (int32_t)so;
          mi.resize(so + (size_t)F[3] * TDX_STENCIL_INTS);
          for (int k = 0; k < F[3]; ++k) {
            const int32_t *S = d->ints + F[4] + k * TDX_STENCIL_INTS;
            mi[so + 2 * k] = S[0];
            mi[so + 2 * k + 1] = real(S[1]);
          }
        }
      }
    }
    if (mi.size() > (size_t)FUSE_PI || mr.size() > (size_t)FUSE_PR) return false;
    out->push_back(std::move(fp));
  }
  return true;
}

}  // namespace

extern "C" int tdx_nn_plan_create(const tdx_nn_desc *d, tdx_nn_plan **out) {
  TDX_REQUIRE(d && out, "null argument");
  *out = nullptr;
  TDX_REQUIRE(d->dtype == 0 || d->dtype == 1, "dtype must be 0 (f32) or 1 (f64)");
  TDX_REQUIRE(d->d_in >= 1 && d->d_in <= MAXD, "d_in must be 1..%d", MAXD);
  TDX_REQUIRE(d->n_out >= 1 && d->n_out <= MAXO, "n_out must be 1..%d", MAXO);
  TDX_REQUIRE(d->width >= 1 && d->width <= 128, "width must be 1..128");
  TDX_REQUIRE(d->n_layers >= 2, "at least two Linear layers");
  TDX_REQUIRE(d->n_inst >= 1 && d->n_groups >= 1, "empty problem");
  int64_t n_rows = 0;
  int rc = validate_nn(d, &n_rows);
  if (rc) return rc;

  TDX_REQUIRE(n_rows < (1LL << 31), "too many rows");
  if (d->causal_tol != 0.0) {
    const int32_t *G0 = d->groups;
    TDX_REQUIRE(std::isfinite(d->causal_tol), "causal tol must be finite");
    TDX_REQUIRE(d->n_inst == 1, "the causal loss serves one instance");
    TDX_REQUIRE(d->causal_nt >= 1 && G0[1] % d->causal_nt == 0,
                "causal slices must divide the %d points of group 0", G0[1]);
    TDX_REQUIRE(G0[7] == 0, "group 0 must start at row 0");
    for (int q = 0; q < G0[5]; ++q)
      TDX_REQUIRE(d->ints[G0[6] + q * TDX_EQ_INTS + 3] == (int64_t)q * G0[1],
                  "causal loss: group 0's residuals must lead the residual output");
  }
  if (d->weak_levels != 0) {
    const int32_t *G0 = d->groups;
    const int N = G0[1], G = d->weak_runs, neq = G0[5];
    TDX_REQUIRE(d->weak_levels == 1 || d->weak_levels == 2, "weak form: 1 or 2 levels");
    TDX_REQUIRE(d->n_inst == 1, "the weak form serves one instance");
    TDX_REQUIRE(d->causal_tol == 0.0, "a weak form with a causal loss");
    TDX_REQUIRE(G >= 1 && (d->weak_levels == 2 || G == 1), "weak form: runs");
    TDX_REQUIRE(d->weak_d >= 0 && d->weak_d + N + G + neq <= d->n_data, "weak form: data range");
    TDX_REQUIRE(d->weak_i >= 0 && d->weak_i + G + 1 <= d->n_ints, "weak form: run table range");
    const int32_t *gs = d->ints + d->weak_i;
    TDX_REQUIRE(gs[0] == 0 && gs[G] == N, "weak form: runs must cover group 0");
    for (int g = 0; g < G; ++g) TDX_REQUIRE(gs[g + 1] > gs[g], "weak form: empty run");
    TDX_REQUIRE(G0[7] == 0, "group 0 must start at row 0");
    for (int q = 0; q < neq; ++q)
      TDX_REQUIRE(d->ints[G0[6] + q * TDX_EQ_INTS + 3] == (int64_t)q * N,
                  "weak form: group 0's residuals must lead the residual output");
  }
  tdx_nn_plan *p = new tdx_nn_plan();
  p->dtype = d->dtype;
  p->d_in = d->d_in;
  p->n_out = d->n_out;
  p->H = d->width;
  p->n_layers = d->n_layers;
  p->NH = d->n_layers - 2;
  p->NB = (d->width + 15) / 16;
  p->HP = 16 * p->NB;
  p->n_inst = d->n_inst;
  p->n_groups = d->n_groups;
  p->n_rows = n_rows;
  p->n_resid = d->n_resid;
  const int H = p->H;
  p->P = H * d->d_in + H + p->NH * (H * H + H) + d->n_out * H + d->n_out;
  const size_t ts = d->dtype == 0 ? 4 : 8;

  // forward work: 64 entries per workgroup
  std::vector<int2> k1;
  std::vector<std::vector<int>> inst_k1(d->n_inst), inst_k2(d->n_inst), inst_k3(d->n_inst);
  std::vector<int> k1_first(d->n_groups);
  for (int g = 0; g < d->n_groups; ++g) {
    const int32_t *G = d->groups + (int64_t)g * TDX_GROUP_INTS;
    k1_first[g] = (int)k1.size();
    for (int e = 0; e < G[1]; e += 16 * K1_WAVES) {
      inst_k1[G[0]].push_back((int)k1.size());
      k1.push_back(make_int2(g, e));
    }
  }
  int max_slots = 1;
  for (int g = 0; g < d->n_groups; ++g)
    max_slots = std::max(max_slots, (int)d->groups[(int64_t)g * TDX_GROUP_INTS + 2]);
  p->max_slots = max_slots;
  p->k1_lds = (2 * K1_WAVES * 16 * max_slots * d->n_out + 512) * ts;
  {
    // stage the forward operand images in LDS when they are small
    const size_t st = (size_t)stage_words(p->HP, p->NH, d->d_in, d->n_out, false) * ts;
    p->k1_stage = (p->k1_lds + st <= 64 * 1024) ? 1 : 0;
    if (p->k1_stage) p->k1_lds += st;
  }

  // backward work: tiles of 16*waves rows, chunk_tiles tiles per workgroup
  const int LS = p->HP + 4;
  const int NACT = p->NH + 1;
  int waves = d->tile_rows > 0 ? d->tile_rows / 16 : 4;
  waves = std::max(1, std::min(4, waves));
  auto lds_of = [&](int wv, bool with_acc) {
    const int TR = 16 * wv;
    size_t n = (size_t)NACT * TR * LS + (size_t)TR * LS + (size_t)TR * MAXD + (size_t)TR * MAXO;
    if (with_acc) n += p->P;
    return n * ts;
  };
  const size_t LDS_MAX = 160 * 1024;
  while (waves > 1 && lds_of(waves, false) > LDS_MAX) --waves;
  if (lds_of(waves, false) > LDS_MAX) {
    delete p;
    set_error("network too large for the backward tile (%zu bytes of LDS)", lds_of(waves, false));
    return TDX_EUNSUPPORTED;
  }
  p->k2_waves = waves;
  p->acc_lds = lds_of(waves, true) <= LDS_MAX ? 1 : 0;
  p->k2_lds = lds_of(waves, p->acc_lds != 0);
  // narrow networks: wavefront-private backward with register accumulators
  p->k2_nt = (d->dtype == 0 && d->tile_rows == 32) ? 2 : 1;
  p->k2_wave_private = (p->NB <= 2 && p->NH <= 4 && d->tile_rows >= 0 &&
                        d->d_in + 1 + d->n_out <= 16 &&
                        wave_private_lds(p, d) * ts <= LDS_MAX) ? 1 : 0;
  if (p->k2_wave_private) {
    p->k2_waves = 4;
    p->acc_lds = 1;
    p->k2_lds = wave_private_lds(p, d) * ts;
  }
  // wide float networks: block-owning backward with register accumulators
  p->k2_block = (!p->k2_wave_private && d->dtype == 0 && p->NB >= 3 && p->NB <= 8 &&
                 p->NH >= 1 && p->NH <= 4 && block_acc_regs(p->NB, p->NH) <= 192 &&
                 d->tile_rows >= 0 && d->d_in + 1 + d->n_out <= 16 &&
                 block_lds_words(p->HP, p->NH) * ts <= 160 * 1024) ? 1 : 0;
  if (p->k2_block) {
    p->k2_waves = 4;
    p->acc_lds = 1;
    p->k2_lds = block_lds_words(p->HP, p->NH) * ts;
  }
  p->persist = (d->persistent > 0 && p->k1_stage && d->causal_tol == 0.0 &&
                d->weak_levels == 0) ? 1 : 0;
  const int TR = 16 * p->k2_waves;
  const int chunk_tiles = d->chunk_tiles > 0 ? d->chunk_tiles
                        : (p->k2_wave_private ? 256 : (p->k2_block ? 8 : (p->acc_lds ? 16 : 4)));
  std::vector<int4> k2;
  for (int g = 0; g < d->n_groups; ++g) {
    const int32_t *G = d->groups + (int64_t)g * TDX_GROUP_INTS;
    const int64_t rb = G[7], nr = (int64_t)G[1] * G[2];
    for (int64_t r = 0; r < nr; r += (int64_t)TR * chunk_tiles) {
      const int64_t re = std::min(nr, r + (int64_t)TR * chunk_tiles);
      inst_k2[G[0]].push_back((int)k2.size());
      k2.push_back(make_int4(g, (int)(rb + r), (int)(rb + re), G[0]));
    }
  }
  p->n_k1 = (int)k1.size();
  p->n_k2 = (int)k2.size();
  std::vector<int32_t> ik1p(1, 0), ik1i, ik2p(1, 0), ik2i;
  for (int i = 0; i < d->n_inst; ++i) {
    for (int x : inst_k1[i]) ik1i.push_back(x);
    ik1p.push_back((int32_t)ik1i.size());
    for (int x : inst_k2[i]) ik2i.push_back(x);
    ik2p.push_back((int32_t)ik2i.size());
  }
  if (ik1i.empty()) ik1i.push_back(0);
  if (ik2i.empty()) ik2i.push_back(0);

  // fused training step for narrow float networks with one output and at
  // most two inputs: work items of chunk_tiles 64-entry blocks (default 4)
  // of one group, and each group's residual program as its own table pair
  std::vector<int4> k3, prog_idx;
  std::vector<int32_t> prog_ints, ik3p(1, 0), ik3i, prog_kind;
  std::vector<float> prog_reals;
  std::vector<FusedProg> progs;
  p->k3_lds = (size_t)fused_lds_words(p->HP, p->NH, d->d_in, d->n_out, p->P) * ts;
  bool fuse = fused_shape_ok(d, max_slots) && p->k2_wave_private && p->k3_lds <= LDS_MAX &&
              fused_programs(d, &progs);
  if (fuse) {
    for (const FusedProg &fp : progs) {
      prog_idx.push_back(make_int4((int)prog_ints.size(), (int)fp.ints.size(),
                                   (int)prog_reals.size(), (int)fp.reals.size()));
      prog_ints.insert(prog_ints.end(), fp.ints.begin(), fp.ints.end());
      prog_reals.insert(prog_reals.end(), fp.reals.begin(), fp.reals.end());
    }
  }
  if (fuse) {
    const int C = d->chunk_tiles > 0 ? d->chunk_tiles : 4;
    for (int g = 0; g < d->n_groups; ++g) {
      const int32_t *G = d->groups + (int64_t)g * TDX_GROUP_INTS;
      const int nblk = (G[1] + 63) / 64;
      for (int j = 0; j < nblk; j += C) {
        inst_k3[G[0]].push_back((int)k3.size());
        k3.push_back(make_int4(g, 64 * j, std::min(G[1], 64 * (j + C)), k1_first[g] + j));
      }
    }
    for (int i = 0; i < d->n_inst; ++i) {
      for (int x : inst_k3[i]) ik3i.push_back(x);
      ik3p.push_back((int32_t)ik3i.size());
    }
    p->fused = 1;
    p->n_k3 = (int)k3.size();
    // generated residual programs for plans of at least 2^17 network rows
    // (TDX_NN_JIT: 0 never, 1 always); the interpreted kernel if compiling
    // fails
    const char *env = std::getenv("TDX_NN_JIT");
    const bool jit = env ? std::atoi(env) != 0 : n_rows >= (1 << 17);
    if (jit) {
      std::vector<std::vector<int32_t>> keys;
      std::vector<int> key_neq;
      std::map<std::vector<int32_t>, int> kind_of;
      for (const FusedProg &fp : progs) {
        const std::vector<int32_t> key = nn_prog_key(fp);
        auto it = kind_of.find(key);
        if (it == kind_of.end()) {
          it = kind_of.emplace(key, (int)keys.size()).first;
          keys.push_back(key);
          key_neq.push_back(fp.n_eq);
        }
        prog_kind.push_back(it->second);
      }
      const std::string src = nn_fused_jit_source(keys, key_neq, p->NB, p->NH);
      if (nn_fused_jit_load(src, &p->fused_fn, "tdx_fused_jit") != TDX_OK) {
        p->fused_fn = nullptr;
        prog_kind.clear();
        clear_error();
      }
    }
  }

  // wide float networks: the training step's block-owning backward forms
  // its own seeds over tiles of whole entries (E = 64 / slots per tile),
  // chunk_tiles tiles per work item (default 8)
  std::vector<int4> k4;
  std::vector<int32_t> ik4p(1, 0), ik4i;
  std::vector<int> inst_k4_size(d->n_inst, 0);
  if (p->k2_block && !p->fused && d->dtype == 0 && max_slots <= 64 && d->tile_rows == 0 &&
      d->persistent == 0 && d->causal_tol == 0.0 && d->weak_levels == 0 &&
      block_fs_lds_words(p->HP, p->NH) * ts <= LDS_MAX && fused_programs(d, &progs)) {
    for (const FusedProg &fp : progs) {
      prog_idx.push_back(make_int4((int)prog_ints.size(), (int)fp.ints.size(),
                                   (int)prog_reals.size(), (int)fp.reals.size()));
      prog_ints.insert(prog_ints.end(), fp.ints.begin(), fp.ints.end());
      prog_reals.insert(prog_reals.end(), fp.reals.begin(), fp.reals.end());
    }
    std::vector<std::vector<int>> inst_k4(d->n_inst);
    const int C = d->chunk_tiles > 0 ? d->chunk_tiles : 8;
    for (int g = 0; g < d->n_groups; ++g) {
      const int32_t *G = d->groups + (int64_t)g * TDX_GROUP_INTS;
      const int E = 64 / G[2];
      const int ntile = (G[1] + E - 1) / E;
      for (int j = 0; j < ntile; j += C) {
        inst_k4[G[0]].push_back((int)k4.size());
        k4.push_back(make_int4(g, j * E, std::min(G[1], (j + C) * E), G[0]));
      }
    }
    for (int i = 0; i < d->n_inst; ++i) {
      for (int x : inst_k4[i]) ik4i.push_back(x);
      ik4p.push_back((int32_t)ik4i.size());
      inst_k4_size[i] = (int)inst_k4[i].size();
    }
    p->k2_fs = 1;
    p->n_k4 = (int)k4.size();
    p->k4_lds = block_fs_lds_words(p->HP, p->NH) * ts;
    // generated residual programs for plans of at least 2^17 network rows
    // (TDX_NN_JIT: 0 never, 1 always); the interpreting kernel if compiling
    // fails
    const char *env = std::getenv("TDX_NN_JIT");
    const bool jit = env ? std::atoi(env) != 0 : n_rows >= (1 << 17);
    if (jit) {
      std::vector<std::vector<int32_t>> keys;
      std::vector<int> key_neq;
      std::map<std::vector<int32_t>, int> kind_of;
      for (const FusedProg &fp : progs) {
        const std::vector<int32_t> key = nn_prog_key(fp);
        auto it = kind_of.find(key);
        if (it == kind_of.end()) {
          it = kind_of.emplace(key, (int)keys.size()).first;
          keys.push_back(key);
          key_neq.push_back(fp.n_eq);
        }
        prog_kind.push_back(it->second);
      }
      const std::string src = nn_block_jit_source(keys, key_neq, p->NB, p->NH, d->n_out);
      if (nn_fused_jit_load(src, &p->block_fn, "tdx_block_jit") != TDX_OK) {
        p->block_fn = nullptr;
        prog_kind.clear();
        clear_error();
      } else {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(p->block_fn),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        (void)hipGetLastError();
      }
    }
  }

  // two-stage gradient sum when one thread per parameter of every instance
  // is too few threads: enough slices per instance for about 2^20 threads,
  // at least 8 work items per slice (an instance of few items is summed in
  // one pass: a second launch would cost more than it saves)
  std::vector<int32_t> ik5p, ik5i;
  {
    const int64_t np = (int64_t)d->n_inst * p->P;
    int max_items = 0;
    for (int i = 0; i < d->n_inst; ++i)
      max_items = std::max({max_items, (int)inst_k2[i].size(), (int)inst_k3[i].size(),
                            (int)inst_k4_size[i]});
    const int sl = (int)std::min<int64_t>({64, ((1 << 20) + np - 1) / np, max_items / 8});
    if (np < (1 << 19) && sl >= 2) {
      p->red_slices = sl;
      for (int i = 0; i <= d->n_inst; ++i) ik5p.push_back(i * p->red_slices);
      for (int i = 0; i < d->n_inst; ++i)
        for (int sl = 0; sl < p->red_slices; ++sl) ik5i.push_back(sl * d->n_inst + i);
    }
  }

  int r2 = TDX_OK;
  if (!r2 && p->red_slices > 1) r2 = p->ik5p.upload(ik5p.data(), ik5p.size() * 4);
  if (!r2 && p->red_slices > 1) r2 = p->ik5i.upload(ik5i.data(), ik5i.size() * 4);
  if (!r2) r2 = p->groups.upload(d->groups, (size_t)d->n_groups * TDX_GROUP_INTS * 4);
  if (!r2) r2 = p->ints.upload(d->ints, (size_t)std::max<int64_t>(d->n_ints, 1) * 4);
  if (!r2) r2 = p->reals.upload(d->reals, (size_t)d->n_reals * ts);
  if (!r2) r2 = p->data.upload(d->data, (size_t)d->n_data * ts);
  if (!r2) r2 = p->k1_items.upload(k1.data(), k1.size() * sizeof(int2));
  if (!r2) r2 = p->k2_items.upload(k2.data(), k2.size() * sizeof(int4));
  if (!r2) r2 = p->ik1p.upload(ik1p.data(), ik1p.size() * 4);
  if (!r2) r2 = p->ik1i.upload(ik1i.data(), ik1i.size() * 4);
  if (!r2) r2 = p->ik2p.upload(ik2p.data(), ik2p.size() * 4);
  if (!r2) r2 = p->ik2i.upload(ik2i.data(), ik2i.size() * 4);
  if (!r2 && p->k2_fs) {
    r2 = p->k4_items.upload(k4.data(), k4.size() * sizeof(int4));
    if (!r2) r2 = p->ik4p.upload(ik4p.data(), ik4p.size() * 4);
    if (!r2) r2 = p->ik4i.upload(ik4i.data(), ik4i.size() * 4);
    if (!r2) r2 = p->prog_idx.upload(prog_idx.data(), prog_idx.size() * sizeof(int4));
    if (!r2) r2 = p->prog_ints.upload(prog_ints.data(), prog_ints.size() * 4);
    if (!r2) r2 = p->prog_reals.upload(prog_reals.data(), prog_reals.size() * 4);
    if (!r2 && !prog_kind.empty())
      r2 = p->prog_kind.upload(prog_kind.data(), prog_kind.size() * 4);
  }
  if (!r2 && p->fused) {
    r2 = p->k3_items.upload(k3.data(), k3.size() * sizeof(int4));
    if (!r2) r2 = p->ik3p.upload(ik3p.data(), ik3p.size() * 4);
    if (!r2) r2 = p->ik3i.upload(ik3i.data(), ik3i.size() * 4);
    if (!r2) r2 = p->prog_idx.upload(prog_idx.data(), prog_idx.size() * sizeof(int4));
    if (!r2) r2 = p->prog_ints.upload(prog_ints.data(), prog_ints.size() * 4);
    if (!r2) r2 = p->prog_reals.upload(prog_reals.data(), prog_reals.size() * 4);
    if (!r2 && !prog_kind.empty())
      r2 = p->prog_kind.upload(prog_kind.data(), prog_kind.size() * 4);
  }
  if (!r2 && p->persist) {
    // persistent forward: as many workgroups as the device keeps resident,
    // each taking a contiguous range of work items of about equal cost
    // (slots + 1 network evaluations per entry: the +1 stands for the
    // residual program and the reductions)
    const int g = p->dtype == 0 ? forward_grid_size_nb<float>(p->NB, p->k1_lds)
                                : forward_grid_size_nb<double>(p->NB, p->k1_lds);
    if (g <= 0) {
      p->persist = 0;
    } else {
      p->g1 = std::min(g, p->n_k1);
      std::vector<double> prefix(k1.size() + 1, 0.0);
      for (size_t i = 0; i < k1.size(); ++i) {
        const int32_t *G = d->groups + (int64_t)k1[i].x * TDX_GROUP_INTS;
        const int ent = std::min(16 * K1_WAVES, G[1] - k1[i].y);
        prefix[i + 1] = prefix[i] + (double)(G[2] + 1) * ent;
      }
      std::vector<int32_t> bounds(p->g1 + 1, 0);
      size_t i = 0;
      for (int w = 1; w < p->g1; ++w) {
        const double target = prefix.back() * w / p->g1;
        while (i < k1.size() && prefix[i] < target) ++i;
        bounds[w] = (int32_t)i;
      }
      bounds[p->g1] = (int32_t)k1.size();
      r2 = p->k1_bounds.upload(bounds.data(), bounds.size() * 4);
    }
  }
  if (r2) {
    tdx_nn_plan_destroy(p);
    return r2;
  }
  size_t off = 0;
  p->off_packed = off;
  off = align_up(off + (size_t)p->n_inst * image_words(p->HP, p->NH, d->d_in, d->n_out) * ts, 256);
  p->off_seeds = off;
  off = align_up(off + (size_t)n_rows * p->n_out * ts, 256);
  p->off_lp = off;
  off = align_up(off + (size_t)std::max(p->n_k1, p->n_k4) * 2 * ts, 256);
  p->off_partial = off;
  off = align_up(off + (size_t)std::max({p->n_k2, p->n_k3, p->n_k4}) * p->P * ts, 256);
  p->off_one = off;
  off = align_up(off + ts, 256);
  p->off_red = off;
  if (p->red_slices > 1) off = align_up(off + (size_t)p->red_slices * p->n_inst * p->P * ts, 256);
  if (d->causal_tol != 0.0) {
    const int32_t *G0 = d->groups;
    p->causal = 1;
    p->causal_tol = d->causal_tol;
    p->c_N = G0[1];
    p->c_slots = G0[2];
    p->c_neq = G0[5];
    p->c_nt = d->causal_nt;
    p->c_K = p->c_N / std::max(1, p->c_nt);
    p->off_resid = off;
    off = align_up(off + (size_t)std::max<int64_t>(d->n_resid, 1) * ts, 256);
    p->off_cw = off;
    off = align_up(off + (size_t)p->c_N * ts, 256);
    p->off_cwr = off;
    off = align_up(off + (size_t)p->c_N * ts, 256);
  }
  if (d->weak_levels != 0) {
    const int32_t *G0 = d->groups;
    p->weak = d->weak_levels;
    p->w_N = G0[1];
    p->w_G = d->weak_runs;
    p->w_neq = G0[5];
    p->c_slots = G0[2];
    p->w_d = d->weak_d;
    p->w_i = d->weak_i;
    p->off_resid = off;
    off = align_up(off + (size_t)std::max<int64_t>(d->n_resid, 1) * ts, 256);
    p->off_wR = off;
    off = align_up(off + (size_t)p->w_G * p->w_neq * ts, 256);
    p->off_wm = off;
    off = align_up(off + (size_t)p->w_N * p->w_neq * ts, 256);
  }
  p->ws_bytes = off;
  *out = p;
  return TDX_OK;
}

extern "C" int tdx_nn_profile(tdx_nn_plan *p, int capacity) {
  TDX_REQUIRE(p && capacity >= 0, "bad argument");
  for (int k = 0; k < TDX_NN_KERNEL_KINDS; ++k) {
    for (hipEvent_t e : p->prof_ev[k]) (void)hipEventDestroy(e);
    p->prof_ev[k].clear();
    p->prof_n[k] = 0;
  }
  p->prof_cap = 0;
  if (capacity == 0) return TDX_OK;
  for (int k = 0; k < TDX_NN_KERNEL_KINDS; ++k) {
    p->prof_ev[k].resize(2 * (size_t)capacity);
    for (auto &e : p->prof_ev[k]) TDX_TRY_HIP(hipEventCreate(&e));
  }
  p->prof_cap = capacity;
  return TDX_OK;
}

extern "C" int tdx_nn_profile_read(tdx_nn_plan *p, double *ms, int *count) {
  TDX_REQUIRE(p && ms && count, "null argument");
  for (int k = 0; k < TDX_NN_KERNEL_KINDS; ++k) {
    double sum = 0.0;
    for (int i = 0; i < p->prof_n[k]; ++i) {
      TDX_TRY_HIP(hipEventSynchronize(p->prof_ev[k][2 * i + 1]));
      float e = 0.f;
      TDX_TRY_HIP(hipEventElapsedTime(&e, p->prof_ev[k][2 * i], p->prof_ev[k][2 * i + 1]));
      sum += e;
    }
    ms[k] = sum;
    count[k] = p->prof_n[k];
    p->prof_n[k] = 0;
  }
  return TDX_OK;
}

extern "C" void tdx_nn_plan_destroy(tdx_nn_plan *p) {
  if (!p) return;
  (void)tdx_nn_profile(p, 0);
  DevTable *t[] = {&p->groups,   &p->ints,     &p->reals,     &p->data,     &p->k1_items,
                   &p->k2_items, &p->ik1p,     &p->ik1i,      &p->ik2p,     &p->ik2i,
                   &p->k1_bounds, &p->k3_items, &p->ik3p,     &p->ik3i,     &p->prog_idx,
                   &p->prog_ints, &p->prog_reals, &p->prog_kind, &p->k4_items, &p->ik4p,
                   &p->ik4i, &p->ik5p, &p->ik5i};
  for (DevTable *x : t) x->release();
  delete p;
}

extern "C" int64_t tdx_nn_fused_source(const tdx_nn_desc *d, char *buf, int64_t cap,
                                       int compile) {
  TDX_REQUIRE(d && d->groups && d->ints && d->reals, "null argument");
  TDX_REQUIRE(d->dtype == 0 || d->dtype == 1, "dtype must be 0 (f32) or 1 (f64)");
  TDX_REQUIRE(d->d_in >= 1 && d->d_in <= MAXD && d->n_out >= 1 && d->n_out <= MAXO &&
                  d->width >= 1 && d->width <= 128 && d->n_layers >= 2 && d->n_inst >= 1 &&
                  d->n_groups >= 1,
              "bad network shape");
  int64_t n_rows = 0;
  const int rc = validate_nn(d, &n_rows);
  if (rc) return rc;
  int max_slots = 1;
  for (int g = 0; g < d->n_groups; ++g)
    max_slots = std::max(max_slots, (int)d->groups[(int64_t)g * TDX_GROUP_INTS + 2]);
  std::vector<FusedProg> progs;
  const int NB = (d->width + 15) / 16, NH = d->n_layers - 2;
  // wide float networks: the block-owning kernel forming its own seeds
  const bool wide = d->dtype == 0 && NB >= 3 && NB <= 8 && NH >= 1 && NH <= 4 &&
                    block_acc_regs(NB, NH) <= 192 && d->d_in + 1 + d->n_out <= 16 &&
                    max_slots <= 64 && d->tile_rows == 0 && d->persistent == 0 &&
                    d->causal_tol == 0.0 && d->weak_levels == 0 &&
                    block_fs_lds_words(16 * NB, NH) * 4 <= 160 * 1024;
  if (!fused_shape_ok(d, max_slots) && !wide) {
    set_error("the network or options leave the fused step out");
    return 0;
  }
  if (!fused_programs(d, &progs)) {
    set_error("a residual program exceeds the fused kernel's table size");
    return 0;
  }
  std::vector<std::vector<int32_t>> keys;
  std::vector<int> key_neq;
  std::map<std::vector<int32_t>, int> kind_of;
  for (const FusedProg &fp : progs) {
    const std::vector<int32_t> key = nn_prog_key(fp);
    if (kind_of.emplace(key, (int)keys.size()).second) {
      keys.push_back(key);
      key_neq.push_back(fp.n_eq);
    }
  }
  const std::string src = wide ? nn_block_jit_source(keys, key_neq, NB, NH, d->n_out)
                               : nn_fused_jit_source(keys, key_neq, NB, NH);
  if (compile) {
    std::vector<char> code;
    const int crc = nn_fused_jit_build(src, &code);
    if (crc) return crc;
  }
  if (buf && cap > 0) {
    const size_t n = std::min<size_t>(src.size(), (size_t)(cap - 1));
    std::memcpy(buf, src.data(), n);
    buf[n] = '\0';
  }
  return (int64_t)src.size();
}

extern "C" int tdx_nn_step_kernel(const tdx_nn_plan *p) {
  if (!p) return -1;
  if (p->k2_fs) return p->block_fn ? 4 : 3;
  return p->fused ? (p->fused_fn ? 2 : 1) : 0;
}

extern "C" int64_t tdx_nn_param_count(const tdx_nn_plan *p) { return p ? p->P : -1; }
extern "C" size_t tdx_nn_workspace_bytes(const tdx_nn_plan *p) { return p ? p->ws_bytes : 0; }

extern "C" int tdx_nn_forward(const tdx_nn_plan *p, const void *params, void *loss, void *resid,
                              void *ws, tdx_stream_t stream) {
  TDX_REQUIRE(p && params && ws, "null argument");
  AdamHyper h{0, 0, 0, 0};
  hipStream_t s = (hipStream_t)stream;
  if (p->dtype == 0)
    return nn_dispatch<float>(p, 0, params, nullptr, loss, resid, nullptr, nullptr, nullptr,
                              nullptr, nullptr, h, ws, s);
  return nn_dispatch<double>(p, 0, params, nullptr, loss, resid, nullptr, nullptr, nullptr,
                             nullptr, nullptr, h, ws, s);
}

extern "C" int64_t tdx_nn_row_count(const tdx_nn_plan *p) { return p ? p->n_rows : -1; }

extern "C" int tdx_nn_network_values(const tdx_nn_plan *p, const void *params, void *values,
                                     void *loss, void *ws, tdx_stream_t stream) {
  TDX_REQUIRE(p && params && values && ws, "null argument");
  AdamHyper h{0, 0, 0, 0};
  hipStream_t s = (hipStream_t)stream;
  if (p->dtype == 0)
    return nn_dispatch<float>(p, 3, params, nullptr, loss, nullptr, nullptr, nullptr, nullptr,
                              nullptr, nullptr, h, ws, s, values);
  return nn_dispatch<double>(p, 3, params, nullptr, loss, nullptr, nullptr, nullptr, nullptr,
                             nullptr, nullptr, h, ws, s, values);
}

extern "C" int tdx_nn_forward_values(const tdx_nn_plan *p, const void *values, void *loss,
                                     void *resid, void *seeds, void *ws, tdx_stream_t stream) {
  TDX_REQUIRE(p && values && ws, "null argument");
  AdamHyper h{0, 0, 0, 0};
  hipStream_t s = (hipStream_t)stream;
  if (p->dtype == 0)
    return nn_dispatch<float>(p, 4, nullptr, nullptr, loss, resid, nullptr, nullptr, nullptr,
                              nullptr, nullptr, h, ws, s, const_cast<void *>(values), seeds);
  return nn_dispatch<double>(p, 4, nullptr, nullptr, loss, resid, nullptr, nullptr, nullptr,
                             nullptr, nullptr, h, ws, s, const_cast<void *>(values), seeds);
}

// LDS words per thread of k_jac.
static size_t jac_words(const tdx_nn_plan *p, int max_slots) {
  return (size_t)(p->NH + 3) * p->H + 2 * (size_t)max_slots * p->n_out;
}

extern "C" int tdx_nn_jacobian(const tdx_nn_plan *p, const void *params, void *jt,
                               tdx_stream_t stream) {
  TDX_REQUIRE(p && params && jt, "null argument");
  TDX_REQUIRE(p->n_inst == 1, "the Jacobian serves one instance per plan");
  const size_t ts = p->dtype == 0 ? 4 : 8;
  const size_t per_thread = jac_words(p, p->max_slots) * ts;
  int B = 64;
  while (B > 1 && per_thread * B > 160 * 1024) B >>= 1;
  TDX_REQUIRE(per_thread * B <= 160 * 1024, "network too wide for the Jacobian kernel");
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = (int64_t)p->P * p->n_resid;
  const unsigned zb = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 65536));
  auto run = [&](auto tag) -> int {
    using T = decltype(tag);
    NNArgs<T> a = make_args<T>(p, params, nullptr);
    hipLaunchKernelGGL((k_jac_zero<T>), dim3(zb), dim3(256), 0, s, (T *)jt, n);
    TDX_TRY_HIP(hipGetLastError());
    raise_lds_limit({(const void *)k_jac<T>});
    hipLaunchKernelGGL((k_jac<T>), dim3(p->n_k1), dim3(B), per_thread * B, s, a, (T *)jt,
                       (int64_t)p->n_resid);
    TDX_TRY_HIP(hipGetLastError());
    return TDX_OK;
  };
  return p->dtype == 0 ? run(0.0f) : run(0.0);
}

extern "C" int tdx_nn_backward(const tdx_nn_plan *p, const void *params, const void *dloss,
                               void *dparams, void *ws, tdx_stream_t stream) {
  TDX_REQUIRE(p && params && dloss && dparams && ws, "null argument");
  AdamHyper h{0, 0, 0, 0};
  hipStream_t s = (hipStream_t)stream;
  if (p->dtype == 0)
    return nn_dispatch<float>(p, 1, params, nullptr, nullptr, nullptr, dloss, dparams, nullptr,
                              nullptr, nullptr, h, ws, s);
  return nn_dispatch<double>(p, 1, params, nullptr, nullptr, nullptr, dloss, dparams, nullptr,
                             nullptr, nullptr, h, ws, s);
}

extern "C" int tdx_nn_train_step(const tdx_nn_plan *p, void *params, void *m, void *v,
                                 int64_t *step, double lr, double b1, double b2, double eps,
                                 void *loss, void *ws, tdx_stream_t stream) {
  TDX_REQUIRE(p && params && m && v && step && ws, "null argument");
  AdamHyper h{lr, b1, b2, eps};
  hipStream_t s = (hipStream_t)stream;
  if (p->dtype == 0)
    return nn_dispatch<float>(p, 2, nullptr, params, loss, nullptr, nullptr, nullptr, m, v,
                              step, h, ws, s);
  return nn_dispatch<double>(p, 2, nullptr, params, loss, nullptr, nullptr, nullptr, m, v,
                             step, h, ws, s);
}

extern "C" int tdx_adam(int dtype, int64_t n, void *param, const void *grad, void *m, void *v,
                        const int64_t *step, double lr, double b1, double b2, double eps,
                        tdx_stream_t stream) {
  TDX_REQUIRE(param && grad && m && v && step && n >= 0, "null argument");
  if (n == 0) return TDX_OK;
  TDX_REQUIRE(dtype == 0 || dtype == 1, "dtype must be 0 or 1");
  AdamHyper h{lr, b1, b2, eps};
  hipStream_t s = (hipStream_t)stream;
  const size_t al = (dtype == 0 ? 4 : 8) * 4;
  const int vec = ((uintptr_t)param % al == 0 && (uintptr_t)grad % al == 0 &&
                   (uintptr_t)m % al == 0 && (uintptr_t)v % al == 0) ? 1 : 0;
  // about 16 vectors per thread at most: the per-step scalars are computed
  // once per thread
  const int64_t per_block = 256 * (vec ? 4 : 1);
  const unsigned blocks =
      (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + per_block - 1) / per_block,
                                                       std::max<int64_t>(4096, n / (per_block * 16))));
  if (dtype == 0)
    hipLaunchKernelGGL((k_adam<float>), dim3(blocks), dim3(256), 0, s, n, (float *)param,
                       (const float *)grad, (float *)m, (float *)v, step, h, vec);
  else
    hipLaunchKernelGGL((k_adam<double>), dim3(blocks), dim3(256), 0, s, n, (double *)param,
                       (const double *)grad, (double *)m, (double *)v, step, h, vec);
  TDX_TRY_HIP(hipGetLastError());
  return TDX_OK;
}

extern "C" int tdx_step_increment(int64_t *step, tdx_stream_t stream) {
  TDX_REQUIRE(step, "null argument");
  hipLaunchKernelGGL(k_step_inc, dim3(1), dim3(64), 0, (hipStream_t)stream, step);
  TDX_TRY_HIP(hipGetLastError());
  return TDX_OK;
}
