// Device code of the fused 'autograd' training step that the library and
// the kernel compiled at plan creation share: tanh jet derivatives, the
// argument block, the interpreted jet maps and residual program, and the
// step body.  Include after tdx_dev.h and tdx_nn_dev.h (and, in the library,
// tdx.h for the table record sizes TDX_AD_*); self-contained otherwise, so
// that hipRTC compiles this text as it is.
#ifndef TDX_AD_DEV_H
#define TDX_AD_DEV_H

namespace tdx {

constexpr int AD_MAXF = 4;  // factors per term

template <typename T>
struct TanhDerivs {
  T s0, s1, s2, s3, s4, s5;  // tanh^(n)(z), n = 0 .. 5
};

// P_1 = 1 - t^2, P_2 = -2 t P_1, P_3 = P_1 (6 t^2 - 2),
// P_4 = P_1 t (16 - 24 t^2), P_5 = P_1 (16 - 120 t^2 + 120 t^4)
template <typename T>
__device__ __forceinline__ TanhDerivs<T> tanh_derivs(T z) {
  TanhDerivs<T> d;
  const T t = dtanh(z);
  const T t2 = t * t;
  const T u = T(1) - t2;
  d.s0 = t;
  d.s1 = u;
  d.s2 = T(-2) * t * u;
  d.s3 = u * (T(6) * t2 - T(2));
  d.s4 = u * t * (T(16) - T(24) * t2);
  d.s5 = u * (T(16) + t2 * (T(-120) + T(120) * t2));
  return d;
}

// tanh^(n) for a runtime n in 1 .. 5
template <typename T>
__device__ __forceinline__ T deriv_n(const T s1, const T s2, const T s3, const T s4, const T s5,
                                    int n) {
  // value selects (a select of field addresses becomes a private-memory load)
  const T a = n >= 4 ? s4 : s2, b = n >= 4 ? s5 : s3;
  const T c = (n & 1) ? (n == 1 ? s1 : b) : a;
  return c;
}

struct Tables {
  const int32_t *ints;
  int jet_i, pptr_i, part_i;
  int J, n_parts;  // LDS copy: jet (2J) | pptr (J + 1) | parts (5 n_parts) ints
  // where each jet occurs as a block: tptr (J + 1) | entries (alpha, part, block)
  const int32_t *occ;
  int n_occ, tptr_i, tent_i;
};

template <typename T>
__device__ __forceinline__ T fpow(int kind, int ref, const T *reals, T x) {
  return kind == 0 ? ipow(x, ref) : dpow(x, reals[ref]);
}
template <typename T>
__device__ __forceinline__ T fdpow(int kind, int ref, const T *reals, T x) {
  if (kind == 0) return ref == 0 ? T(0) : T(ref) * ipow(x, ref - 1);
  const T p = reals[ref];
  return p * dpow(x, p - T(1));
}

// A callable power at ints[ref] of the plan's tables (chain_apply in
// tdx_dev.h): v its value at x, d its derivative.
template <typename T>
__device__ __forceinline__ void chain_eval(const int32_t *I, int ref, const T *R, T x, T &v,
                                           T &d) {
  chain_apply(PtrTab<T>{I, R}, ref, x, v, d);
}

// Whether a term's factors include a callable power.
__device__ __forceinline__ bool term_has_chain(const int32_t *I, int fi, int nf) {
  bool c = false;
  for (int f = 0; f < nf; ++f) c |= I[fi + TDX_AD_FACTOR_INTS * f + 3] == 2;
  return c;
}

// A term with a callable power, as the reference forms it
// (tedeous/derivative.py:124-129): the running product P = g(P * fv) at a
// callable factor, P = P * fv ** p at the others, P starting at 1.  Returns
// the final P and writes dP/dfv of every factor into dfv (zero beyond nf).
template <typename T>
__device__ __forceinline__ T chain_term(const int32_t *I, int fi, int nf, const T *R,
                                        const T (&fv)[AD_MAXF], T (&dfv)[AD_MAXF]) {
  T P = T(1);
  T G[AD_MAXF], M[AD_MAXF];
#pragma unroll
  for (int f = 0; f < AD_MAXF; ++f) {
    G[f] = T(0);
    M[f] = T(1);
    if (f < nf) {
      const int F = fi + TDX_AD_FACTOR_INTS * f;
      const int kind = I[F + 3], ref = I[F + 4];
      if (kind == 2) {
        T v, d;
        chain_eval(I, ref, R, P * fv[f], v, d);
        G[f] = d * P;
        M[f] = d * fv[f];
        P = v;
      } else {
        const T pw = fpow(kind, ref, R, fv[f]);
        G[f] = P * fdpow(kind, ref, R, fv[f]);
        M[f] = pw;
        P = P * pw;
      }
    }
  }
  T g = T(1);
#pragma unroll
  for (int f = AD_MAXF - 1; f >= 0; --f) {
    dfv[f] = g * G[f];
    g = g * M[f];
  }
  return P;
}

// ---------------------------------------------------------------------------
// fused training step (float): forward jets, residual program, adjoint and
// weight gradients of a tile of whole entries in one workgroup
// ---------------------------------------------------------------------------
// A workgroup of four wavefronts walks a range of entries of one group in
// tiles.  Each wavefront takes EW = 16 / (S J) whole entries (S point sets,
// J jets): its 16 rows are (entry, set, jet), jet fastest, rows from
// EW S J on are padding.  Per tile:
//   forward: the first layer's jets on the VALU (value rows x W1 + b1, a
//     first-order row of axis a the column W1[:, a], higher orders zero),
//     each hidden layer Z = W A (+ b on value rows) on MFMA with the
//     wavefront's rows as the B operand (feature-major LDS, the block-owning
//     kernel's layout), the tanh jet map per (point, unit) by Faa di Bruno's
//     formula from the partition tables, the output jets on the VALU; the
//     pre-activation jets of every layer but the last go to a per-workgroup
//     scratch area (each lane its own 16-byte words), the last stays in LDS;
//   residual program: one lane per entry, reading the entry's output jets
//     and writing its seeds dL/dU (the library's k_ad_resid arithmetic);
//   adjoint: dA of the last layer = seeds x W_L, the jet map's adjoint per
//     (point, unit) (k_ad_bwd's arithmetic) gives dZ; per layer from the top
//     A of the layer below is formed again from its Z (scratch), wavefront w
//     accumulates the output blocks ob = w (mod 4) of dW = dZ^T A over the
//     tile's 64 rows in accumulator registers (bias terms over value rows),
//     and dA below = W^T dZ on MFMA for the wavefront's own rows.
// The first layer's and output layer's weights share the small-term MFMA
// tiles of the block-owning kernel (B operand: the input jets, the
// value-row mask, the seeds).  After the item: each wavefront writes the
// gradient blocks it owns to partial[item], the loss pair to lp[item].
constexpr int ADF_MAXJ = 16;  // jets per point (rows of an entry: at most 16)

struct ADStepArgs {
  NNArgs<float> nn;   // d_in, n_out, H, NH, P and the packed image (instance 0)
  const int32_t *ints;
  const float *reals, *data, *points;
  int groups_i, J;
  Tables tb;
  const int4 *items;  // (group, first entry, end entry, 0)
  float *zs;          // [items][NHM][NB][256][4] pre-activation jets
  float *partial;     // [items][P]
  float *lp;          // [items][2]
};

__host__ __device__ inline size_t ad_step_lds_bytes(int HP) {
  return (size_t)(3 * HP * CTRS + (MAXD + 1 + MAXO) * CTRS + 2 * 64 * MAXO + 32) * 4;
}

// Activation jets h of the wavefront's NP points from their pre-activation
// jets z (feature-major, rows r0 + p J + alpha).  Each lane takes ADF_R
// (point, unit) items at once, so the table reads (uniform) are issued once
// for all of them and their LDS reads overlap.
constexpr int ADF_R = 5;

// Reads of the plan's read-only tables at wavefront-uniform indices, through
// the constant address space: scalar loads, which do not wait behind LDS
// traffic.
__device__ __forceinline__ int ctab(const int32_t *p, int i) {
  return ((const __attribute__((address_space(4))) int32_t *)p)[i];
}
__device__ __forceinline__ float ctabf(const float *p, int i) {
  return ((const __attribute__((address_space(4))) float *)p)[i];
}

__device__ __forceinline__ void ad_items(int kb, int n, int NP, float rNP, int r0, int J,
                                         int (&off)[ADF_R], bool (&ok)[ADF_R]) {
#pragma unroll
  for (int r = 0; r < ADF_R; ++r) {
    const int k = kb + 64 * r;
    ok[r] = k < n;
    const int kk = ok[r] ? k : kb;
    const int f = (int)((kk + 0.5f) * rNP);
    off[r] = f * CTRS + r0 + (kk - f * NP) * J;
  }
}

__device__ __forceinline__ void ad_jet_fwd(const float *zb, float *hb, int r0, int NP, float rNP,
                                           int J, int H, const Tables &tb, int lane) {
  const int n = NP * H;
  for (int kb = lane; kb < n; kb += 64 * ADF_R) {
    int off[ADF_R];
    bool ok[ADF_R];
    ad_items(kb, n, NP, rNP, r0, J, off, ok);
    TanhDerivs<float> s[ADF_R];
#pragma unroll
    for (int r = 0; r < ADF_R; ++r) s[r] = tanh_derivs(zb[off[r]]);
#pragma unroll
    for (int r = 0; r < ADF_R; ++r)
      if (ok[r]) hb[off[r]] = s[r].s0;
    for (int al = 1; al < J; ++al) {
      const int k0 = ctab(tb.ints, tb.pptr_i + al), k1 = ctab(tb.ints, tb.pptr_i + al + 1);
      float v[ADF_R];
      for (int kp = k0; kp < k1; ++kp) {
        const int pr = tb.part_i + TDX_AD_PART_INTS * kp;
        const int m = ctab(tb.ints, pr);
        float prod[ADF_R];
#pragma unroll
        for (int r = 0; r < ADF_R; ++r)
          prod[r] = deriv_n(s[r].s1, s[r].s2, s[r].s3, s[r].s4, s[r].s5, m);
#pragma unroll
        for (int qq = 0; qq < TDX_AD_MAX_ORDER; ++qq)
          if (qq < m) {
            const int bq = ctab(tb.ints, pr + 1 + qq);
#pragma unroll
            for (int r = 0; r < ADF_R; ++r) prod[r] = prod[r] * zb[off[r] + bq];
          }
#pragma unroll
        for (int r = 0; r < ADF_R; ++r) v[r] = (kp == k0) ? prod[r] : v[r] + prod[r];
      }
#pragma unroll
      for (int r = 0; r < ADF_R; ++r)
        if (ok[r]) hb[off[r] + al] = v[r];
    }
  }
}

// The jet map's adjoint for the wavefront's NP points: dL/dz into dzb from
// dL/dh in dab and the pre-activation jets zb (k_ad_bwd's arithmetic, in
// its order), ADF_R items per lane at once.
__device__ __forceinline__ void ad_jet_adj(const float *zb, const float *dab, float *dzb, int r0,
                                           int NP, float rNP, int J, int H, const Tables &tb,
                                           int lane) {
  const int n = NP * H;
  for (int kb = lane; kb < n; kb += 64 * ADF_R) {
    int off[ADF_R];
    bool ok[ADF_R];
    ad_items(kb, n, NP, rNP, r0, J, off, ok);
    TanhDerivs<float> s[ADF_R];
    float dz0[ADF_R];
#pragma unroll
    for (int r = 0; r < ADF_R; ++r) {
      s[r] = tanh_derivs(zb[off[r]]);
      dz0[r] = dab[off[r]] * s[r].s1;
    }
    for (int al = 1; al < J; ++al) {
      float da[ADF_R];
#pragma unroll
      for (int r = 0; r < ADF_R; ++r) da[r] = dab[off[r] + al];
      const int k0 = ctab(tb.ints, tb.pptr_i + al), k1 = ctab(tb.ints, tb.pptr_i + al + 1);
      for (int kp = k0; kp < k1; ++kp) {
        const int pr = tb.part_i + TDX_AD_PART_INTS * kp;
        const int m = ctab(tb.ints, pr);
        float all[ADF_R];
#pragma unroll
        for (int r = 0; r < ADF_R; ++r)
          all[r] = deriv_n(s[r].s1, s[r].s2, s[r].s3, s[r].s4, s[r].s5, m + 1);
#pragma unroll
        for (int qq = 0; qq < TDX_AD_MAX_ORDER; ++qq)
          if (qq < m) {
            const int bq = ctab(tb.ints, pr + 1 + qq);
#pragma unroll
            for (int r = 0; r < ADF_R; ++r) all[r] = all[r] * zb[off[r] + bq];
          }
#pragma unroll
        for (int r = 0; r < ADF_R; ++r) dz0[r] = dz0[r] + da[r] * all[r];
      }
    }
#pragma unroll
    for (int r = 0; r < ADF_R; ++r)
      if (ok[r]) dzb[off[r]] = dz0[r];
    for (int b = 1; b < J; ++b) {
      float acc[ADF_R];
#pragma unroll
      for (int r = 0; r < ADF_R; ++r) acc[r] = 0.f;
      const int t0 = ctab(tb.occ, tb.tptr_i + b), t1 = ctab(tb.occ, tb.tptr_i + b + 1);
      for (int t = t0; t < t1; ++t) {
        const int oc = tb.tent_i + 3 * t;
        const int al = ctab(tb.occ, oc), qo = ctab(tb.occ, oc + 2);
        const int pr = tb.part_i + TDX_AD_PART_INTS * ctab(tb.occ, oc + 1);
        const int m = ctab(tb.ints, pr);
        float v[ADF_R];
#pragma unroll
        for (int r = 0; r < ADF_R; ++r)
          v[r] = dab[off[r] + al] * deriv_n(s[r].s1, s[r].s2, s[r].s3, s[r].s4, s[r].s5, m);
#pragma unroll
        for (int q2 = 0; q2 < TDX_AD_MAX_ORDER; ++q2)
          if (q2 < m && q2 != qo) {
            const int bq = ctab(tb.ints, pr + 1 + q2);
#pragma unroll
            for (int r = 0; r < ADF_R; ++r) v[r] = v[r] * zb[off[r] + bq];
          }
#pragma unroll
        for (int r = 0; r < ADF_R; ++r) acc[r] = acc[r] + v[r];
      }
#pragma unroll
      for (int r = 0; r < ADF_R; ++r)
        if (ok[r]) dzb[off[r] + b] = acc[r];
    }
  }
}

// Residual program of entry e of group G (k_ad_resid's per-entry
// arithmetic): U and SD are the entry's output jets and seeds, row
// set * J + jet, n_out values a row.
__device__ __forceinline__ void ad_entry(const ADStepArgs &a, const int32_t *G, int e, int n_out,
                                         const float *U, float *SD, float &lsum, float &lnsum) {
  const int J = a.J, n_eq = G[3], eq_i = G[4];
  auto factor_value = [&](int F) {
    const int n_st = ctab(a.ints, F), st_i = ctab(a.ints, F + 1), var = ctab(a.ints, F + 2);
    float fv = 0.f;
    for (int k = 0; k < n_st; ++k) {
      const int S = st_i + TDX_AD_STENCIL_INTS * k;
      const float v = U[(ctab(a.ints, S) * J + ctab(a.ints, S + 1)) * n_out + var];
      const float w = ctabf(a.reals, ctab(a.ints, S + 2));
      fv = (k == 0) ? w * v : fv + w * v;
    }
    return fv;
  };
  for (int qe = 0; qe < n_eq; ++qe) {
    const int E = eq_i + TDX_AD_EQ_INTS * qe;
    const int n_terms = ctab(a.ints, E), term_i = ctab(a.ints, E + 1);
    const float lw = ctabf(a.reals, ctab(a.ints, E + 2));
    const float nw = ctabf(a.reals, ctab(a.ints, E + 3));
    const float sw = ctabf(a.reals, ctab(a.ints, E + 6));
    const int tgt = ctab(a.ints, E + 4);
    float r = 0.f;
    for (int t = 0; t < n_terms; ++t) {
      const int Tr = term_i + TDX_AD_TERM_INTS * t;
      const int nf = ctab(a.ints, Tr), fi = ctab(a.ints, Tr + 1);
      const int ck = ctab(a.ints, Tr + 2), cr = ctab(a.ints, Tr + 3);
      const float c = ck == 0 ? ctabf(a.reals, cr) : a.data[cr + e];
      float prod = 1.f;
      if (term_has_chain(a.ints, fi, nf)) {
        float fv[AD_MAXF], dfv[AD_MAXF];
#pragma unroll
        for (int f = 0; f < AD_MAXF; ++f)
          fv[f] = f < nf ? factor_value(fi + TDX_AD_FACTOR_INTS * f) : 0.f;
        prod = chain_term(a.ints, fi, nf, a.reals, fv, dfv);
      } else {
        for (int f = 0; f < nf; ++f) {
          const int F = fi + TDX_AD_FACTOR_INTS * f;
          prod = prod * fpow(ctab(a.ints, F + 3), ctab(a.ints, F + 4), a.reals, factor_value(F));
        }
      }
      const float tv = c * prod;
      r = (t == 0) ? tv : r + tv;
    }
    const float diff = tgt >= 0 ? r - a.data[tgt + e] : r;
    const float d2 = diff * diff;
    lsum += lw * d2;
    lnsum += nw * d2;
    const float gq = 2.f * sw * diff;
    for (int t = 0; t < n_terms; ++t) {
      const int Tr = term_i + TDX_AD_TERM_INTS * t;
      const int nf = ctab(a.ints, Tr), fi = ctab(a.ints, Tr + 1);
      const int ck = ctab(a.ints, Tr + 2), cr = ctab(a.ints, Tr + 3);
      const float c = ck == 0 ? ctabf(a.reals, cr) : a.data[cr + e];
      float fv[AD_MAXF], pw[AD_MAXF], dfv[AD_MAXF];
      const bool chain = term_has_chain(a.ints, fi, nf);
#pragma unroll
      for (int f = 0; f < AD_MAXF; ++f) {
        fv[f] = 0.f;
        pw[f] = 1.f;
        if (f < nf) {
          const int F = fi + TDX_AD_FACTOR_INTS * f;
          fv[f] = factor_value(F);
          if (!chain) pw[f] = fpow(ctab(a.ints, F + 3), ctab(a.ints, F + 4), a.reals, fv[f]);
        }
      }
      if (chain) (void)chain_term(a.ints, fi, nf, a.reals, fv, dfv);
#pragma unroll
      for (int f = 0; f < AD_MAXF; ++f) {
        if (f < nf) {
          const int F = fi + TDX_AD_FACTOR_INTS * f;
          float dv;
          if (chain) {
            dv = gq * c * dfv[f];
          } else {
            dv = gq * c * fdpow(ctab(a.ints, F + 3), ctab(a.ints, F + 4), a.reals, fv[f]);
#pragma unroll
            for (int f2 = 0; f2 < AD_MAXF; ++f2)
              if (f2 != f && f2 < nf) dv = dv * pw[f2];
          }
          const int n_st = ctab(a.ints, F), st_i = ctab(a.ints, F + 1), var = ctab(a.ints, F + 2);
          for (int k = 0; k < n_st; ++k) {
            const int S = st_i + TDX_AD_STENCIL_INTS * k;
            SD[(ctab(a.ints, S) * J + ctab(a.ints, S + 1)) * n_out + var] +=
                dv * ctabf(a.reals, ctab(a.ints, S + 2));
          }
        }
      }
    }
  }
}

// The step body.  JP: the jet maps (interpreted from the partition tables,
// or straight-line code generated for the plan's jets); RP: the residual
// program (interpreted, or generated per group).  Launch: 256 threads,
// ad_step_lds_bytes of dynamic LDS, one workgroup per work item.
template <int NB, int NHM, typename JP, typename RP>
__device__ __forceinline__ void ad_step_body(const ADStepArgs &a, float *smem) {
  constexpr int HP = 16 * NB;
  constexpr int OBW = (NB + 3) / 4;  // output blocks owned per wavefront
  typedef typename V4<float>::type v4;
  float *abuf = smem;                                  // [HP][CTRS] activation jets
  float *zbuf = abuf + HP * CTRS;                      // [HP][CTRS] pre-activation jets
  float *dbuf = zbuf + HP * CTRS;                      // [HP][CTRS] adjoints
  float *xbuf = dbuf + HP * CTRS;         // [MAXD + 1][CTRS] input jets, value-row mask
  float *sbuf = xbuf + (MAXD + 1) * CTRS; // [MAXO][CTRS] seeds
  float *ubuf = sbuf + MAXO * CTRS;       // [64][n_out] output jets
  float *sdb = ubuf + 64 * MAXO;          // [64][n_out] seeds
  float *red = sdb + 64 * MAXO;           // 32
  const Tables &tb = a.tb;
  for (int k = threadIdx.x; k < 3 * HP * CTRS; k += 256) abuf[k] = 0.f;

  const int4 item = a.items[blockIdx.x];
  const int32_t *G = a.ints + a.groups_i + TDX_AD_GROUP_INTS * item.x;
  const int n_sets = G[1], set_i = G[2];
  const int J = a.J, SJ = n_sets * J, EW = 16 / SJ, E = 4 * EW, NP = EW * n_sets;
  const float rNP = 1.0f / (float)NP;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row = lane & 15, q = lane >> 4;
  const int trow = 16 * w + row;
  const int H = a.nn.H, d_in = a.nn.d_in, n_out = a.nn.n_out, P = a.nn.P, NH = a.nn.NH;
  // this lane's row: entry elw of the wavefront, set s_row, jet al_row;
  // kind 0 value, 1 + a first order in axis a, -1 zero input
  const int elw = row / SJ, rem = row - elw * SJ;
  const int s_row = rem / J, al_row = rem - s_row * J;
  const bool rowok = elw < EW;
  __syncthreads();  // buffers cleared
  int kind = -1;
  if (rowok) {
    const int ord = tb.ints[tb.jet_i + 2 * al_row], ax = tb.ints[tb.jet_i + 2 * al_row + 1];
    kind = al_row == 0 ? 0 : (ord == 1 ? 1 + ax : -1);
  }
  const float vm = kind == 0 ? 1.f : 0.f;
  const int r0w = 16 * w;  // the wavefront's first row
  const Net<float, NB> net = make_net<float, NB>(a.nn, 0);
  const Layout<float> &L = net.L;
  float *zsc = a.zs + (size_t)blockIdx.x * NHM * NB * 1024;
  const int col_b1 = d_in, col_wl = d_in + 1;

  v4 dwacc[NHM][OBW][NB];
  v4 sacc[OBW];
  float bhacc[NHM][OBW];
#pragma unroll
  for (int j = 0; j < NHM; ++j)
#pragma unroll
    for (int u = 0; u < OBW; ++u) {
      bhacc[j][u] = 0.f;
#pragma unroll
      for (int ib = 0; ib < NB; ++ib) dwacc[j][u][ib] = v4{0, 0, 0, 0};
    }
#pragma unroll
  for (int u = 0; u < OBW; ++u) sacc[u] = v4{0, 0, 0, 0};
  float blacc = 0.f, lsum = 0.f, lnsum = 0.f;

  auto store_rows = [&](float *buf, const float (&v)[NB][4]) {
#pragma unroll
    for (int mb = 0; mb < NB; ++mb)
#pragma unroll
      for (int i = 0; i < 4; ++i) buf[(16 * mb + crow<float>(q, i)) * CTRS + trow] = v[mb][i];
  };
  auto zstore = [&](int j, const float (&v)[NB][4]) {
#pragma unroll
    for (int mb = 0; mb < NB; ++mb)
      *reinterpret_cast<v4 *>(zsc + ((j * NB + mb) * 256 + threadIdx.x) * 4) =
          v4{v[mb][0], v[mb][1], v[mb][2], v[mb][3]};
  };
  auto zload = [&](int j, float (&v)[NB][4]) {
#pragma unroll
    for (int mb = 0; mb < NB; ++mb) {
      const v4 t = *reinterpret_cast<const v4 *>(zsc + ((j * NB + mb) * 256 + threadIdx.x) * 4);
#pragma unroll
      for (int i = 0; i < 4; ++i) v[mb][i] = t[i];
    }
  };

  for (int64_t t0 = item.y; t0 < item.z; t0 += E) {
    const int el = w * EW + elw;
    const bool valid = rowok && t0 + el < item.z;
    const int e = (int)(valid ? t0 + el : item.z - 1);
    // input jets of this lane's row
    float in0[MAXD];
    {
      const int64_t pt = (int64_t)a.ints[set_i + (rowok ? s_row : 0)] + e;
#pragma unroll
      for (int k = 0; k < MAXD; ++k)
        in0[k] = k >= d_in ? 0.f
                           : (kind == 0 ? a.points[pt * d_in + k] : (kind == 1 + k ? 1.f : 0.f));
    }
    if (q == 0) {
#pragma unroll
      for (int k = 0; k < MAXD; ++k)
        if (k < d_in) xbuf[k * CTRS + trow] = in0[k];
      xbuf[d_in * CTRS + trow] = vm;
    }
    // first layer
    float z[NB][4];
#pragma unroll
    for (int mb = 0; mb < NB; ++mb) {
      const v4 b1 = ld4(net.b1 + 16 * mb + 4 * q);
      v4 zz = kind == 0 ? b1 : v4{0, 0, 0, 0};
#pragma unroll
      for (int k = 0; k < MAXD; ++k)
        if (k < d_in) {
          const v4 wv = ld4(net.W1 + k * HP + 16 * mb + 4 * q);
#pragma unroll
          for (int i = 0; i < 4; ++i) zz[i] = fmaf(in0[k], wv[i], zz[i]);
        }
#pragma unroll
      for (int i = 0; i < 4; ++i) z[mb][i] = zz[i];
    }
    if (NH > 0) zstore(0, z);
    store_rows(zbuf, z);
    wave_sync();
    JP::fwd(zbuf, abuf, r0w, NP, rNP, J, H, tb, lane);
    wave_sync();
#pragma unroll
    for (int j = 1; j <= NHM; ++j) {
      if (j > NH) break;
      v4 acc[NB];
#pragma unroll
      for (int ob = 0; ob < NB; ++ob)
        acc[ob] = kind == 0 ? ld4(net.bh + (j - 1) * HP + 16 * ob + 4 * q) : v4{0, 0, 0, 0};
      mma_layer_fm<float, NB>(net.Wf + (j - 1) * HP * HP, lane, abuf, trow, acc);
#pragma unroll
      for (int ob = 0; ob < NB; ++ob)
#pragma unroll
        for (int i = 0; i < 4; ++i) z[ob][i] = acc[ob][i];
      if (j < NH) zstore(j, z);
      wave_sync();
      store_rows(zbuf, z);
      wave_sync();
      JP::fwd(zbuf, abuf, r0w, NP, rNP, J, H, tb, lane);
      wave_sync();
    }
    // output jets of the wavefront's rows, then the residual program of its
    // entries (one lane each)
    {
      float act[NB][4];
#pragma unroll
      for (int mb = 0; mb < NB; ++mb)
#pragma unroll
        for (int i = 0; i < 4; ++i) act[mb][i] = abuf[(16 * mb + crow<float>(q, i)) * CTRS + trow];
#pragma unroll
      for (int o = 0; o < MAXO; ++o)
        if (o < n_out) {
          float sv = 0.f;
#pragma unroll
          for (int mb = 0; mb < NB; ++mb) {
            const v4 wl = ld4(net.WL + o * HP + 16 * mb + 4 * q);
#pragma unroll
            for (int i = 0; i < 4; ++i) sv = fmaf(act[mb][i], wl[i], sv);
          }
          sv += __shfl_xor(sv, 16);
          sv += __shfl_xor(sv, 32);
          if (kind == 0) sv += net.bL[o];
          if (q == 0) ubuf[trow * n_out + o] = sv;
        }
    }
    for (int k = lane; k < 16 * n_out; k += 64) sdb[r0w * n_out + k] = 0.f;
    wave_sync();
    if (lane < EW && t0 + w * EW + lane < item.z)
      RP::run(a, item.x, (int)(t0 + w * EW + lane), n_out, ubuf + (r0w + lane * SJ) * n_out,
              sdb + (r0w + lane * SJ) * n_out, lsum, lnsum);
    wave_sync();
    // seeds: dA of the last layer, the output layer's small terms
    {
      float dl[NB][4];
#pragma unroll
      for (int mb = 0; mb < NB; ++mb)
#pragma unroll
        for (int i = 0; i < 4; ++i) dl[mb][i] = 0.f;
#pragma unroll
      for (int o = 0; o < MAXO; ++o)
        if (o < n_out) {
          const float so = sdb[trow * n_out + o];
          if (o == q) blacc += so * vm;
          if (q == 0) sbuf[o * CTRS + trow] = so;
#pragma unroll
          for (int mb = 0; mb < NB; ++mb) {
            const v4 wl = ld4(net.WL + o * HP + 16 * mb + 4 * q);
#pragma unroll
            for (int i = 0; i < 4; ++i) dl[mb][i] = fmaf(wl[i], so, dl[mb][i]);
          }
        }
      store_rows(dbuf, dl);
    }
    __syncthreads();
    // output layer W: columns col_wl + o of the owned small-term tiles
    {
      const int o = row - col_wl;
#pragma unroll
      for (int u = 0; u < OBW; ++u) {
        const int ob = w + 4 * u;
        if (ob < NB) {
#pragma unroll
          for (int kk = 0; kk < 4; ++kk) {
            const v4 av = *reinterpret_cast<const v4 *>(abuf + (16 * ob + row) * CTRS + 16 * kk + 4 * q);
            const v4 sv = (o >= 0 && o < n_out)
                              ? *reinterpret_cast<const v4 *>(sbuf + o * CTRS + 16 * kk + 4 * q)
                              : v4{0, 0, 0, 0};
#pragma unroll
            for (int kb = 0; kb < 4; ++kb) sacc[u] = mma16(av[kb], sv[kb], sacc[u]);
          }
        }
      }
    }
    __syncthreads();
    // from here on abuf holds dZ of the current layer (all rows once a
    // barrier has passed), dbuf dA and then A of the layer below
    JP::adj(zbuf, dbuf, abuf, r0w, NP, rNP, J, H, tb, lane);
#pragma unroll
    for (int j = NHM; j >= 1; --j) {
      if (j > NH) continue;
      const int jj = j - 1;
      // A of layer j - 1 again from its pre-activation jets
      {
        float zp[NB][4];
        zload(j - 1, zp);
        store_rows(zbuf, zp);
      }
      wave_sync();
      JP::fwd(zbuf, dbuf, r0w, NP, rNP, J, H, tb, lane);
      __syncthreads();
      // owned blocks of dW_jj over the 64 rows; the bias over value rows
#pragma unroll
      for (int u = 0; u < OBW; ++u) {
        const int ob = w + 4 * u;
        if (ob < NB) {
#pragma unroll 1
          for (int kk = 0; kk < 4; ++kk) {
            const v4 dv = *reinterpret_cast<const v4 *>(abuf + (16 * ob + row) * CTRS + 16 * kk + 4 * q);
            const v4 mv = *reinterpret_cast<const v4 *>(xbuf + col_b1 * CTRS + 16 * kk + 4 * q);
            bhacc[jj][u] += (dv[0] * mv[0] + dv[1] * mv[1]) + (dv[2] * mv[2] + dv[3] * mv[3]);
#pragma unroll
            for (int ib = 0; ib < NB; ++ib) {
              const v4 av = *reinterpret_cast<const v4 *>(dbuf + (16 * ib + row) * CTRS + 16 * kk + 4 * q);
              v4 c = dwacc[jj][u][ib];
#pragma unroll
              for (int kb = 0; kb < 4; ++kb) c = mma16(dv[kb], av[kb], c);
              dwacc[jj][u][ib] = c;
            }
          }
        }
      }
      // dA of layer j - 1 = W_jj^T dZ_j for this wavefront's rows
      v4 c[NB];
#pragma unroll
      for (int ib = 0; ib < NB; ++ib) c[ib] = v4{0, 0, 0, 0};
      mma_layer_fm<float, NB>(net.Wb + jj * HP * HP, lane, abuf, trow, c);
      __syncthreads();
      {
        float dv[NB][4];
#pragma unroll
        for (int ib = 0; ib < NB; ++ib)
#pragma unroll
          for (int i = 0; i < 4; ++i) dv[ib][i] = c[ib][i];
        store_rows(dbuf, dv);
      }
      wave_sync();
      JP::adj(zbuf, dbuf, abuf, r0w, NP, rNP, J, H, tb, lane);
    }
    __syncthreads();
    // first layer: columns c < d_in of the small-term tiles (B = input jet
    // column c) and column d_in (B = the value-row mask)
#pragma unroll
    for (int u = 0; u < OBW; ++u) {
      const int ob = w + 4 * u;
      if (ob < NB) {
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
          const v4 dv = *reinterpret_cast<const v4 *>(abuf + (16 * ob + row) * CTRS + 16 * kk + 4 * q);
          const v4 xv = row <= d_in ? *reinterpret_cast<const v4 *>(xbuf + row * CTRS + 16 * kk + 4 * q)
                                    : v4{0, 0, 0, 0};
#pragma unroll
          for (int kb = 0; kb < 4; ++kb) sacc[u] = mma16(dv[kb], xv[kb], sacc[u]);
        }
      }
    }
    __syncthreads();
  }

  // each wavefront writes the blocks it owns; the output bias and the loss
  // are the only sums over wavefronts (fixed order)
  float *out = a.partial + (int64_t)blockIdx.x * P;
#pragma unroll
  for (int u = 0; u < OBW; ++u) {
    const int ob = w + 4 * u;
    if (ob < NB) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int of = 16 * ob + crow<float>(q, i);
        if (of >= H) continue;
#pragma unroll
        for (int jj = 0; jj < NHM; ++jj)
          if (jj < NH) {
#pragma unroll
            for (int ib = 0; ib < NB; ++ib) {
              const int in = 16 * ib + row;
              if (in < H) out[L.oWh(jj) + of * H + in] = dwacc[jj][u][ib][i];
            }
          }
        const float sv = sacc[u][i];
        if (row < d_in) out[L.oW1() + of * d_in + row] = sv;
        else if (row == col_b1) out[L.ob1() + of] = sv;
        else if (row < col_wl + n_out) out[L.oWL() + (row - col_wl) * H + of] = sv;
      }
#pragma unroll
      for (int jj = 0; jj < NHM; ++jj)
        if (jj < NH) {
          float v = bhacc[jj][u];
          v += __shfl_xor(v, 16);
          v += __shfl_xor(v, 32);
          const int f = 16 * ob + row;
          if (q == 0 && f < H) out[L.obh(jj) + f] = v;
        }
    }
  }
  const float bl = rowsum16(blacc);
  if (row == 0) red[4 * w + q] = bl;
#pragma unroll
  for (int st = 32; st > 0; st >>= 1) {
    lsum += __shfl_down(lsum, st);
    lnsum += __shfl_down(lnsum, st);
  }
  if (lane == 0) {
    red[16 + 2 * w] = lsum;
    red[17 + 2 * w] = lnsum;
  }
  __syncthreads();
  if ((int)threadIdx.x < n_out) {
    const int o = threadIdx.x;
    out[L.obL() + o] = ((red[o] + red[4 + o]) + red[8 + o]) + red[12 + o];
  }
  if (threadIdx.x == 0) {
    a.lp[2 * blockIdx.x] = ((red[16] + red[18]) + red[20]) + red[22];
    a.lp[2 * blockIdx.x + 1] = ((red[17] + red[19]) + red[21]) + red[23];
  }
}

// Policies of the library's kernel: the jet maps and the residual program
// interpreted from the plan's tables.
struct InterpJets {
  static __device__ __forceinline__ void fwd(const float *zb, float *hb, int r0, int NP, float rNP,
                                             int J, int H, const Tables &tb, int lane) {
    ad_jet_fwd(zb, hb, r0, NP, rNP, J, H, tb, lane);
  }
  static __device__ __forceinline__ void adj(const float *zb, const float *dab, float *dzb, int r0,
                                             int NP, float rNP, int J, int H, const Tables &tb,
                                             int lane) {
    ad_jet_adj(zb, dab, dzb, r0, NP, rNP, J, H, tb, lane);
  }
};
struct InterpADResid {
  static __device__ __forceinline__ void run(const ADStepArgs &a, int g, int e, int n_out,
                                             const float *U, float *SD, float &ls, float &lns) {
    ad_entry(a, a.ints + a.groups_i + TDX_AD_GROUP_INTS * g, e, n_out, U, SD, ls, lns);
  }
};

}  // namespace tdx

#endif  // TDX_AD_DEV_H
