/*
 * tdx.h -- C ABI of the MI355X residual/loss engine for TEDEouS-style solvers.
 *
 * The library replaces, for the 'NN', 'mat' and 'autograd' solution modes,
 * the per-step objective of the reference:
 *
 *   Solution.evaluate()             tedeous/solution.py:129-168
 *     Operator.operator_compute()   tedeous/eval.py:167-193, 223-232
 *       Derivative_NN.take_derivative   tedeous/derivative.py:30-58
 *       Derivative_mat.take_derivative  tedeous/derivative.py:293-323
 *       Derivative_autograd             tedeous/derivative.py:61-132
 *     Bounds.apply_bcs()            tedeous/eval.py:283-461 (+ dict_to_matrix 55-87)
 *     Losses._default_loss()        tedeous/losses.py:84-135
 *   loss.backward()                 tedeous/optimizers/closure.py:60
 *   torch.optim.Adam.step()         tedeous/optimizers/optimizer.py:44-45, 63-66
 *
 * Conventions (all entry points):
 *   - Plain pointers and sizes only.  Every device buffer is owned by the
 *     caller: parameters, gradients, outputs and the workspace.  A plan owns
 *     only immutable tables uploaded by its create call.
 *   - All work is enqueued on the given stream; nothing synchronises the
 *     host, nothing allocates, so a step can be captured in a HIP graph.
 *   - Return 0 on success, a negative TDX_E* code otherwise; the message of
 *     the last error on the calling thread is returned by tdx_last_error().
 *   - dtype 0 = float32 (the product path), 1 = float64 (the same kernels
 *     instantiated in double precision, used for tight parity checks).
 */
#ifndef TDX_H
#define TDX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TDX_OK 0
#define TDX_EINVAL (-1)
#define TDX_EUNSUPPORTED (-2)
#define TDX_ELAUNCH (-3)
#define TDX_ENOMEM (-4)

typedef void *tdx_stream_t; /* a hipStream_t; NULL = the null stream */

/* ABI version; bumped on any layout change of the structs below. */
int tdx_version(void);
/* Message of the last failure on the calling thread ("" if none). */
const char *tdx_last_error(void);

/* ===================================================================== */
/* 'NN' mode                                                              */
/* ===================================================================== */
/*
 * The network is the reference's tanh MLP (nn.Sequential of Linear/Tanh):
 * d_in -> width -> ... -> width -> n_out with n_layers Linear layers.
 * Parameters of one instance are stored exactly as torch's
 * net.parameters() flattened: W1[width][d_in], b1, W2[width][width], b2,
 * ..., W_L[n_out][width], b_L.  n_inst instances of the same architecture
 * are stored back to back (stride tdx_nn_param_count()).
 *
 * The residual program is a list of groups (see torch_de_solver_amd/
 * compiler.py).  Group g has n_entries entries and n_slots slots; the
 * network input of (slot s, entry e) is
 *     data[slot_base[s] + e*d_in + a] + reals[slot_off + s*d_in + a]
 * i.e. the reference's shifted grid x + fl(shift*h) (tedeous/
 * input_preprocessing.py:239-264), generated in the kernel.
 * Tables (int32 records in `ints`, working-precision values in `reals`
 * and `data`):
 *   group   [TDX_GROUP_INTS]: inst, n_entries, n_slots, slot_base_i,
 *           slot_off_r, n_eq, eq_i, row_base
 *   slot_base_i -> n_slots ints: data offset of entry 0 of the slot's base
 *   eq      [TDX_EQ_INTS]: n_terms, term_i, target_off (-1: zero target),
 *           resid_off (-1: none), lam_r (lambda/denominator: loss weight),
 *           invden_r (1/denominator: normalized-loss weight), seed_r
 *           (weight of the seeds 2 w r dr/du; lam_r except for the causal
 *           loss's operator equations)
 *   term    [TDX_TERM_INTS]: n_factors, factor_i, coeff_kind
 *           (0: reals[coeff_idx], 1: data[coeff_idx + entry]), coeff_idx
 *   factor  [TDX_FACTOR_INTS]: var, pow_kind (0: integer pow_int,
 *           1: real reals[pow_r]), pow_int, n_stencil, stencil_i, pow_r
 *   stencil [TDX_STENCIL_INTS]: slot, weight_r
 * Seed rows: row (group g, slot s, entry e) = row_base + s*n_entries + e.
 */
#define TDX_GROUP_INTS 8
#define TDX_EQ_INTS 7
#define TDX_TERM_INTS 4
#define TDX_FACTOR_INTS 6
#define TDX_STENCIL_INTS 2

typedef struct tdx_nn_desc {
  int dtype;
  int d_in, n_out, width, n_layers;
  int n_inst;
  int n_groups;
  const int32_t *groups; /* n_groups * TDX_GROUP_INTS */
  int64_t n_ints;
  const int32_t *ints;
  int64_t n_reals;
  const void *reals; /* dtype values */
  int64_t n_data;
  const void *data; /* dtype values: coordinates, coefficients, targets */
  int64_t n_resid;  /* length of the optional residual output */
  int tile_rows;    /* backward tile rows (0: default; -1: workgroup-cooperative kernel;
                       32: two 16-row tiles per wavefront iteration of the
                       wavefront-private kernel (float), otherwise one) */
  int chunk_tiles;  /* work per gradient work item (0: default): 64-row tiles of the
                       separate adjoint kernel, or 64-entry blocks of the fused
                       training-step kernel (default 4) */
  /* Causal loss (tedeous/losses.py:137-182), one instance: tol != 0 replaces
   * the operator term of group 0's equations (loss and normalized weights 0,
   * seed weight 1/N) by mean(w * res), res[p] = sum_eq r_eq[p]^2 over the
   * N points of group 0 read as causal_nt slices of N / causal_nt points,
   * w[i][k] = exp(-tol * sum_{j<i} res[j][k]) (no gradient through w). */
  double causal_tol;
  int causal_nt;
  /* Weak form (tedeous/eval.py:13-52, 195-221, tedeous/losses.py:184-228),
   * one instance: weak_levels 1 or 2 replaces the operator term of group
   * 0's equations (loss and normalized weights 0, seed weight lambda) by
   * sum_eq lambda_eq W_eq, with on one level W = sum_i c_i r_i^2 and on two
   * levels W = sum_g x_g (sum_{i in run g} c_i r_i^2)^2 (several equations
   * on two levels: the forward pass runs again with per-equation seeds),
   * c_i = data[weak_d + i] (N values: trapezoid weight times the squared
   * basis product), x_g = data[weak_d + N + g] (weak_runs values),
   * lambda_eq = data[weak_d + N + weak_runs + eq]; run g holds points
   * ints[weak_i + g] .. ints[weak_i + g + 1] - 1 of group 0. */
  int weak_levels;
  int weak_runs;
  int64_t weak_d;
  int weak_i;
  /* Training-step kernels.  0: the fused step where the plan allows it --
   * float networks of at most 32 hidden features, at most two H x H layers,
   * one output, at most two inputs, at most 8 slots per group, default
   * tile_rows, no causal or weak-form loss, every group's residual program
   * within 512 record and 256 real words: one kernel evaluates the network,
   * the residual program and the adjoint of 16-entry rounds with every
   * activation kept in registers (no forward recompute, no seeds through
   * HBM), so a step is three launches (packing, fused kernel, reduction with
   * the Adam update and the loss pairs); the loss pairs are the same bits
   * as with separate kernels, the gradient the same per-row terms summed in
   * another order.  Float networks of 48 to 128 hidden features (at most
   * 64 slots per group, same other conditions) step without a forward
   * kernel too: their block-owning adjoint, which evaluates the network
   * again anyway, takes tiles of whole entries, runs the residual program
   * on them and forms its own seeds (three launches; chunk_tiles counts
   * such tiles, default 8).  Otherwise, and with -1, separate forward and
   * adjoint kernels, one workgroup per work item.  1: separate kernels with the
   * persistent forward: when the network's forward image fits in LDS
   * and the plan has no causal or weak-form loss, the forward kernel runs
   * as a grid of CUs x resident workgroups, each walking a contiguous range
   * of equal cost of the (instance, entry tile) work items and gathering
   * its instance's network into LDS from the parameters when the instance
   * changes; the workgroup holding an instance's first work item also
   * writes the packed image the backward kernel reads, so a fused step is
   * three launches (forward, backward, reduction with the Adam update and
   * the loss pairs) instead of four (packing first), with the same result
   * bits as -1. */
  int persistent;
} tdx_nn_desc;

typedef struct tdx_nn_plan tdx_nn_plan;

/* Upload the tables and build the work lists.  Replaces the compile step
 * Solution.__init__ (tedeous/solution.py:23-88) for 'NN' mode. */
int tdx_nn_plan_create(const tdx_nn_desc *desc, tdx_nn_plan **out);
void tdx_nn_plan_destroy(tdx_nn_plan *plan);
/* Parameters per instance (torch layout). */
int64_t tdx_nn_param_count(const tdx_nn_plan *plan);
/* Bytes of caller-provided device workspace for every call below. */
size_t tdx_nn_workspace_bytes(const tdx_nn_plan *plan);

/* Residual + boundary + loss, forward.  Writes loss[2*i] (lambda-weighted)
 * and loss[2*i+1] (normalized, all lambdas 1) for every instance i, and the
 * residual values to resid (may be NULL).  Replaces Solution.evaluate()
 * (tedeous/solution.py:129-168) in 'NN' mode. */
int tdx_nn_forward(const tdx_nn_plan *plan, const void *params, void *loss,
                   void *resid, void *workspace, tdx_stream_t stream);
/* Adjoint of tdx_nn_forward: dparams = dloss[0] * d(sum_i loss_i)/dparams.
 * dloss is a device pointer to one value.  Must follow a tdx_nn_forward with
 * the same params and workspace (it reads the seeds that call stored).
 * Replaces loss.backward() (tedeous/optimizers/closure.py:60). */
int tdx_nn_backward(const tdx_nn_plan *plan, const void *params,
                    const void *dloss, void *dparams, void *workspace,
                    tdx_stream_t stream);
/* Seed rows of the plan: one per (group, slot, entry), all instances. */
int64_t tdx_nn_row_count(const tdx_nn_plan *plan);
/* tdx_nn_forward that also writes the network output of every seed row to
 * values[(row_base + slot*n_entries + entry)*n_out + o] (the value the
 * residual program reads: the reference's net(x + shift*h)[:, o] of
 * tedeous/derivative.py:48-52 and the boundary values of tedeous/eval.py:
 * 283-345).  loss may be NULL.  Splits the parity of the 'NN' path at the
 * network boundary: the network alone is compared through these values. */
int tdx_nn_network_values(const tdx_nn_plan *plan, const void *params, void *values,
                          void *loss, void *workspace, tdx_stream_t stream);
/* The residual program, boundary values, losses and seeds on given network
 * values ([n_rows][n_out], as written by tdx_nn_network_values) instead of
 * the network: the reference's Operator/Bounds/Losses (tedeous/eval.py:
 * 143-193, 283-461, tedeous/losses.py:84-135) applied to fixed outputs.
 * Writes loss (may be NULL), resid (may be NULL) and the seeds dL/du of
 * every row to `seeds` ([n_rows][n_out]) or, when NULL, to the workspace
 * for a following tdx_nn_backward.  Needs no parameters. */
int tdx_nn_forward_values(const tdx_nn_plan *plan, const void *values, void *loss,
                          void *resid, void *seeds, void *workspace, tdx_stream_t stream);
/* Jacobian of the residuals with respect to the parameters, for the natural
 * gradient's Gram matrix: jt[p * n_resid + r] = d resid[r] / d params[p]
 * (P x n_resid, the transpose of the reference's J; r indexes the residual
 * output of tdx_nn_forward, operator and boundary entries alike).  Replaces
 * the per-residual torch.autograd.grad loop of NGD.gram_factory
 * (tedeous/optimizers/ngd.py:57-78).  One instance per plan; needs no
 * workspace. */
int tdx_nn_jacobian(const tdx_nn_plan *plan, const void *params, void *jt,
                    tdx_stream_t stream);
/* One fused training step: forward, adjoint and torch.optim.Adam update
 * (no weight decay, amsgrad off) of all instances.  step is a device int64
 * holding the number of completed steps; it is incremented by the call.
 * loss receives the pre-update losses as in tdx_nn_forward. */
int tdx_nn_train_step(const tdx_nn_plan *plan, void *params, void *exp_avg,
                      void *exp_avg_sq, int64_t *step, double lr,
                      double beta1, double beta2, double eps, void *loss,
                      void *workspace, tdx_stream_t stream);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  tdx_nn_train_step with the workspace taken from a
 * file-scope device array of tdx_noalloc_scratch_bytes() bytes (64 MiB) per
 * GPU instead of a caller buffer; plans whose tdx_nn_workspace_bytes exceeds
 * it are refused with TDX_EINVAL before anything is launched.  Allocates
 * nothing (the array's device address is looked up once per device).  All
 * calls on one GPU share that GPU's array: calls that overlap on one GPU
 * (concurrent streams or threads) race; calls in stream order do not. */
size_t tdx_noalloc_scratch_bytes(void);
int tdx_nn_train_step_noalloc(const tdx_nn_plan *plan, void *params, void *exp_avg,
                              void *exp_avg_sq, int64_t *step, double lr,
                              double beta1, double beta2, double eps, void *loss,
                              tdx_stream_t stream);

/* Launch timing of the kernels of a plan, for measurement: with capacity
 * > 0 every following launch of the forward (kind 0), backward (kind 1) and
 * reduce/update (kind 2) kernels is bracketed by a pair of HIP events on its
 * stream, up to capacity launches per kind; tdx_nn_profile_read waits for
 * them and returns the summed milliseconds and launch counts per kind (and
 * resets the counts).  capacity 0 turns timing off.  Not for use under
 * graph capture.  No counterpart in the reference (measurement only). */
#define TDX_NN_KERNEL_KINDS 3
/* Source of the training-step kernel that a plan of `desc` compiles at
 * creation with its generated residual programs (float networks the fused
 * step serves, or wide ones whose adjoint forms its own seeds; see
 * tdx_nn_desc.persistent), copied into buf (cap
 * bytes, NUL-terminated; buf may be NULL); with compile != 0 also compiled
 * by hipRTC for gfx950 (no device needed).  Returns the source length, 0
 * when the plan has no fused step (reason in tdx_last_error) or a negative
 * error code.  Plans generate the programs when they hold at least 2^17
 * network rows, or always / never with
 * the environment variable TDX_NN_JIT=1 / 0, and fall back to the library's
 * interpreting kernel if compiling fails.  Diagnostic; no counterpart in the
 * reference. */
int64_t tdx_nn_fused_source(const tdx_nn_desc *desc, char *buf, int64_t cap, int compile);
/* Training-step kernels of a plan: 0 separate forward and adjoint kernels,
 * 1 the fused kernel interpreting the residual programs, 2 the fused kernel
 * compiled with the plan's generated residual programs, 3 the block-owning
 * adjoint of wide networks forming its own seeds (no forward kernel) with
 * the interpreted programs, 4 the same compiled with generated programs;
 * -1 for NULL.  Diagnostic; no counterpart in the reference. */
int tdx_nn_step_kernel(const tdx_nn_plan *plan);
int tdx_nn_profile(tdx_nn_plan *plan, int capacity);
int tdx_nn_profile_read(tdx_nn_plan *plan, double *ms, int *count);

/* Host-side check of the 'mat' plans' division: for the float step h and
 * y = RN(1/h), counts the x in [1, 2) (all 2^23 significands) for which
 * mode 0, RN(RN(x y) + (x - RN(x y) h) y) (one correction step), or mode 1,
 * RN(x y) (the reciprocal product alone), differs from the correctly rounded
 * x / h.  0 means exact; plans use mode 0 only for steps that give 0.
 * No counterpart in the reference (diagnostic, no device needed).  */
int64_t tdx_quotient_check(float h, int mode);

/* ===================================================================== */
/* 'mat' mode                                                             */
/* ===================================================================== */
/*
 * The unknown is a grid tensor u[n_eq][n_0][n_1][n_2] (1-D grids use
 * n_0 = n_2 = 1 and the single axis is axis 1; 2-D grids n_2 = 1; 3-D grids
 * have axes 0, 1, 2).  D_a is the reference's per-axis stencil
 * (tedeous/derivative.py:199-291): interior (sum_b alpha_back[b] u[i+b-dp+1]
 * + sum_f alpha_forw[f] u[i+f]) / (2 h_a), first dp-1 points forw/h_a,
 * last dp-1 points back/h_a.  A factor applies up to TDX_MAT_MAX_CHAIN
 * of them.
 * Tables:
 *   eq      [TDX_MEQ_INTS]: n_terms, term_i, lam_r (lambda/N)
 *   term    [TDX_MTERM_INTS]: n_factors, factor_i, coeff_kind
 *           (0: reals[coeff_idx], 1: data[coeff_idx + flat point]), coeff_idx
 *   factor  [TDX_MFACTOR_INTS]: var, pow_kind, pow_int, pow_r, n_axes,
 *           axis[TDX_MAT_MAX_CHAIN] (in application order)
 *   bc entry [TDX_MBC_INTS]: n_sets, pos_i (ints: n_sets flat positions),
 *           n_terms (0: the plain value u[var]), term_i, var, target_r
 *           (data index of the padded target), lam_r (lambda/max_len),
 *           invden_r (1/max_len), seed_i (first seed slot of this entry),
 *           robin (0, or 1 + ints offset of {n_beta, alpha_r}: the entry's
 *           value is alpha u[var] + sum_k beta_k (its operator), with the
 *           betas at reals alpha_r + 1 .. alpha_r + n_beta and one extra
 *           seed slot, first, for the alpha u[var] term; one point set)
 *   bc adjoint CSR over touched grid points: bct_ptr[n_touch+1],
 *           bct_pos[n_touch] (flat index into u incl. the n_eq plane),
 *           bct_seed[nnz], bct_w_r[nnz] (reals index of the weight)
 * Plans of 1-D and 2-D grids whose factors apply at most 3 stencils run
 * tiled kernels.  Other plans (3-D grids, longer chains, reaches beyond an
 * LDS tile) run every factor's chain as full-grid passes; their boundary
 * entries read the chain fields at their positions and scatter their seeds
 * back, so the bc adjoint CSR is not read (n_touch may be 0).
 */
#define TDX_MEQ_INTS 3
#define TDX_MTERM_INTS 4
#define TDX_MAT_MAX_CHAIN 8
#define TDX_MFACTOR_INTS (5 + TDX_MAT_MAX_CHAIN)
#define TDX_MBC_INTS 10

typedef struct tdx_mat_desc {
  int dtype;
  int n_eq;
  int n0, n1, n2; /* grid shape; 1-D grids: n0 = n2 = 1, 2-D grids: n2 = 1 */
  int dp;         /* derivative_points */
  double h[3];    /* step per axis (h[0] unused for 1-D, h[2] below 3-D) */
  const double *alpha_back; /* dp values */
  const double *alpha_forw; /* dp values */
  int eq_i;                 /* ints offset of the n_eq equation records */
  int n_bc;                 /* boundary entries */
  int bc_i;                 /* ints offset of the bc entry records */
  int n_bc_seeds;           /* seed slots over all bc entries */
  int n_touch;              /* grid points touched by bc adjoints */
  int bct_ptr_i, bct_pos_i, bct_seed_i, bct_w_i; /* ints offsets */
  int64_t n_ints;
  const int32_t *ints;
  int64_t n_reals;
  const void *reals;
  int64_t n_data;
  const void *data;
  int exact_div; /* division by 2h and h: 0 product with the rounded reciprocal;
                    1 correctly rounded (bit-identical fields), by a checked
                    correction step where it is exact for the step, else by
                    IEEE division; 2 correctly rounded by IEEE division */
  int tile_rows; /* output rows per tile workgroup (2-D grids); 0: chosen by the plan */
  int specialize; /* float 2-D grids: interior points by a kernel generated and
                     compiled (hipRTC) for this program at plan creation.
                     0: when the interior holds at least 2^20 points and the
                     generated kernels need no scratch;
                     1: whenever the grid has an interior (errors are
                     reported); -1: never */
  int band_rows;  /* rows per wavefront of the specialized kernel; 0: chosen
                     by the plan */
} tdx_mat_desc;

typedef struct tdx_mat_plan tdx_mat_plan;

int tdx_mat_plan_create(const tdx_mat_desc *desc, tdx_mat_plan **out);
void tdx_mat_plan_destroy(tdx_mat_plan *plan);
size_t tdx_mat_workspace_bytes(const tdx_mat_plan *plan);
/* Grid points the plan's specialized kernel covers (0: none).  Diagnostic;
 * no counterpart in the reference. */
int64_t tdx_mat_specialized_points(const tdx_mat_plan *plan);
/* Source of the kernel a plan of `desc` generates for its interior (float
 * 2-D grids), copied into buf (cap bytes, NUL-terminated; buf may be NULL);
 * with compile != 0 also compiled by hipRTC for gfx950 (no device needed).
 * Returns the source length, 0 when the program has no specialized kernel
 * (reason in tdx_last_error) or a negative error code.  Diagnostic; no
 * counterpart in the reference. */
int64_t tdx_mat_specialized_source(const tdx_mat_desc *desc, char *buf,
                                   int64_t cap, int compile);

/* Forward: loss[0] (lambda-weighted), loss[1] (normalized); op receives the
 * residual field [n_eq][n0*n1] (may be NULL); bval receives the boundary
 * values, one per bc entry (may be NULL).  Replaces Solution.evaluate()
 * (tedeous/solution.py:129-168) in 'mat' mode. */
int tdx_mat_forward(const tdx_mat_plan *plan, const void *u, void *loss,
                    void *op, void *bval, void *workspace, tdx_stream_t stream);
/* grad = dloss[0] * dloss/du (dloss: device pointer to one value).
 * Replaces loss.backward() for the grid tensor. */
int tdx_mat_backward(const tdx_mat_plan *plan, const void *u,
                     const void *dloss, void *grad, void *workspace,
                     tdx_stream_t stream);
/* Fused step: gradient and torch.optim.Adam update of u (see
 * tdx_nn_train_step for step and loss). */
int tdx_mat_train_step(const tdx_mat_plan *plan, void *u, void *exp_avg,
                       void *exp_avg_sq, int64_t *step, double lr,
                       double beta1, double beta2, double eps, void *loss,
                       void *workspace, tdx_stream_t stream);
/* The same step with the updated grid written to u_out (u_in is left
 * unchanged; exp_avg, exp_avg_sq and step are updated in place): results
 * identical to copying u_in to u_out and calling tdx_mat_train_step on it.
 * With u_in read-only, the kernel generated for the grid interior applies
 * the Adam update as it forms each gradient (24 B per point instead of 8 for
 * the gradient plus 28 for the update); callers alternate two grid buffers.
 * Replaces the same optimizer.step(closure) as tdx_mat_train_step. */
int tdx_mat_train_step_to(const tdx_mat_plan *plan, const void *u_in,
                          void *u_out, void *exp_avg, void *exp_avg_sq,
                          int64_t *step, double lr, double beta1, double beta2,
                          double eps, void *loss, void *workspace,
                          tdx_stream_t stream);

/* ===================================================================== */
/* 'autograd' mode                                                        */
/* ===================================================================== */
/*
 * The same tanh MLP and parameter layout as 'NN' mode (one instance), with
 * derivatives taken exactly instead of by finite differences.  The
 * reference differentiates the network with nested
 * torch.autograd.grad(create_graph=True) calls (Derivative_autograd,
 * tedeous/derivative.py:61-132); here every point carries a jet: the value
 * and each partial derivative (a multiset of input axes, order <= 4) the
 * program reads, pushed through the layers in one forward pass (linear
 * maps act on each jet component; tanh combines them by Faa di Bruno's
 * formula over the set partitions of each multiset).  The adjoint runs the
 * transposed maps back.
 *
 * Jets: jet 0 is the value.  ints[jet_i + 2j] = order of jet j,
 * ints[jet_i + 2j + 1] = its axis when the order is 1, else -1.  The
 * partitions of jet j are part records ints[part_i + TDX_AD_PART_INTS * k]
 * for k in [ints[pptr_i + j], ints[pptr_i + j + 1]): (number of blocks m,
 * jet of block 0, ..., block m-1; -1 pads), every block a jet of the list.
 *
 * Points: points[n_points][d_in].  Group g (TDX_AD_GROUP_INTS at
 * ints[groups_i + TDX_AD_GROUP_INTS * g]): (n_entries, n_sets, set_i, n_eq,
 * eq_i); entry e of set s is point ints[set_i + s] + e.  Every point
 * belongs to exactly one (group, set, entry).
 * Equation (TDX_AD_EQ_INTS): (n_terms, term_i, lw_r, nw_r, target_d,
 * resid_off, sw_r): value = in-order sum of terms at the entry, diff =
 * value - target with target data[target_d + e] (0 when target_d < 0);
 * reals[lw_r] * diff^2 adds to the loss, reals[nw_r] * diff^2 to the
 * normalized loss, and the seed of the value is 2 reals[sw_r] diff (lambda
 * / denominator, 1 / denominator, lambda / denominator for the default
 * loss); the value of entry e goes to resid[resid_off + e].
 * Term (TDX_AD_TERM_INTS): (n_factors, factor_i, coeff_kind, coeff_ref):
 * coefficient reals[coeff_ref] (kind 0) or data[coeff_ref + e] (kind 1)
 * times the product of its factors.
 * Factor (TDX_AD_FACTOR_INTS): (n_st, st_i, var, pow_kind, pow_ref):
 * (sum over its stencil entries of w * jet component) ** pow, pow the
 * integer pow_ref (kind 0) or reals[pow_ref] (kind 1).  Kind 2 is a callable
 * power, applied as the reference does to the running product of the term
 * (tedeous/derivative.py:124-129: P = g(P * factor), P starting at 1, the
 * other factors P = P * factor ** pow): pow_ref is the ints offset of g as a
 * chain of one-argument operations, n then n (op, arg) pairs, applied to the
 * value in order; op TDX_AD_OP_*, arg an index into reals (ADD..RDIV, POWR),
 * an integer power (POWI) or unused.
 * Stencil entry (TDX_AD_STENCIL_INTS): (set, jet, weight_r).
 */
#define TDX_AD_OP_ADD 1   /* v + c */
#define TDX_AD_OP_SUB 2   /* v - c */
#define TDX_AD_OP_RSUB 3  /* c - v */
#define TDX_AD_OP_MUL 4   /* c * v */
#define TDX_AD_OP_DIV 5   /* v / c */
#define TDX_AD_OP_RDIV 6  /* c / v */
#define TDX_AD_OP_POWI 7  /* v ** k, k the integer arg */
#define TDX_AD_OP_POWR 8  /* v ** c */
#define TDX_AD_OP_NEG 9
#define TDX_AD_OP_SIN 10
#define TDX_AD_OP_COS 11
#define TDX_AD_OP_EXP 12
#define TDX_AD_OP_LOG 13
#define TDX_AD_OP_TANH 14
#define TDX_AD_OP_SQRT 15
#define TDX_AD_MAX_CHAIN 16  /* operations of one callable power */
#define TDX_AD_MAX_ORDER 4
#define TDX_AD_PART_INTS 5
#define TDX_AD_GROUP_INTS 5
#define TDX_AD_EQ_INTS 7
#define TDX_AD_TERM_INTS 4
#define TDX_AD_FACTOR_INTS 5
#define TDX_AD_STENCIL_INTS 3

typedef struct tdx_ad_desc {
  int dtype;
  int d_in, n_out, width, n_layers;
  int n_jets, jet_i, pptr_i, part_i;
  int64_t n_points;
  const void *points;
  int n_groups, groups_i;
  int64_t n_ints;
  const int32_t *ints;
  int64_t n_reals;
  const double *reals;
  int64_t n_data;
  const void *data;
  int64_t n_resid;
  /* Causal loss (tedeous/losses.py:137-182): tol != 0 adds mean(w * res)
   * over the N points of group 0 (its equations carry zero loss weights and
   * seed weight 1/N), res[p] = sum_eq r_eq[p]^2 read as causal_nt slices of
   * N / causal_nt points, w[i][k] = exp(-tol * sum_{j<i} res[j][k]); the
   * seeds of group 0's points are scaled by w (no gradient through w). */
  double causal_tol;
  int causal_nt;
  /* Weak form (tedeous/eval.py:13-52, 195-221, tedeous/losses.py:184-228):
   * weak_levels 1 (1-D) or 2 (2-D) adds sum_eq lambda_eq W_eq over the N
   * points of group 0 (its equations carry zero loss weights and seed
   * weight lambda_eq), W_eq = sum_i c_i r_i^2 on one level and
   * sum_g x_g (sum_{i in run g} c_i r_i^2)^2 on two; c_i = data[weak_d + i],
   * x_g = data[weak_d + N + g], lambda_eq = data[weak_d + N + weak_runs + eq],
   * run g = points ints[weak_i + g] .. ints[weak_i + g + 1] - 1.  The
   * residual program runs a second time with each equation's seeds scaled
   * by dW/dr / (2 r). */
  int weak_levels;
  int weak_runs;
  int64_t weak_d;
  int weak_i;
} tdx_ad_desc;

typedef struct tdx_ad_plan tdx_ad_plan;

/* Upload the tables.  Replaces Solution.__init__ (tedeous/solution.py:23-88)
 * with Equation_autograd (tedeous/input_preprocessing.py:411-505). */
int tdx_ad_plan_create(const tdx_ad_desc *desc, tdx_ad_plan **out);
void tdx_ad_plan_destroy(tdx_ad_plan *plan);
int64_t tdx_ad_param_count(const tdx_ad_plan *plan);
size_t tdx_ad_workspace_bytes(const tdx_ad_plan *plan);
/* Forward: loss[0] (lambda-weighted), loss[1] (normalized); resid (may be
 * NULL) receives every equation's value per entry.  Replaces
 * Solution.evaluate() (tedeous/solution.py:129-168) in 'autograd' mode. */
int tdx_ad_forward(const tdx_ad_plan *plan, const void *params, void *loss,
                   void *resid, void *workspace, tdx_stream_t stream);
/* dparams = dloss[0] * dloss/dparams after a tdx_ad_forward with the same
 * params and workspace.  Replaces loss.backward()
 * (tedeous/optimizers/closure.py:60). */
int tdx_ad_backward(const tdx_ad_plan *plan, const void *params,
                    const void *dloss, void *dparams, void *workspace,
                    tdx_stream_t stream);
/* Jacobian of the residuals with respect to the parameters,
 * jt[p * n_resid + r] = d resid[r] / d params[p] (P x n_resid; r indexes the
 * residual output of tdx_ad_forward), for the natural gradient's Gram
 * matrix.  Replaces the per-residual torch.autograd.grad loop of
 * NGD.gram_factory (tedeous/optimizers/ngd.py:57-78) over the 'autograd'
 * operator (tedeous/derivative.py:61-132).  Needs no workspace; weak-form
 * plans are refused (their residuals are integrals over runs). */
int tdx_ad_jacobian(const tdx_ad_plan *plan, const void *params, void *jt,
                    tdx_stream_t stream);
/* Fused step: forward, adjoint and torch.optim.Adam update (see
 * tdx_nn_train_step for step and loss).  Float plans without causal or
 * weak terms, with at most 16 jets, 128 hidden units, 4 H x H layers (3 at
 * 113-128 units) and groups whose point sets times jets are at most 16 run
 * it as three launches around one kernel that keeps a tile's jets on chip
 * (forward, residual program, adjoint, weight gradients); other plans run
 * the layer kernels of tdx_ad_forward/tdx_ad_backward and tdx_adam.  The
 * two differ in the order of the float sums only.  Environment:
 * TDX_AD_FUSED=0 keeps the layer kernels; TDX_AD_TILES sets the tiles per
 * work item of the fused kernel (default: about one item per compute
 * unit).  Replaces Closure._closure (tedeous/optimizers/closure.py:49-64)
 * with loss.backward() through Derivative_autograd
 * (tedeous/derivative.py:61-132). */
int tdx_ad_train_step(const tdx_ad_plan *plan, void *params, void *exp_avg,
                      void *exp_avg_sq, int64_t *step, double lr, double beta1,
                      double beta2, double eps, void *loss, void *workspace,
                      tdx_stream_t stream);
/* Kernels of tdx_ad_train_step: 0 the layer kernels, 1 the fused kernel
 * interpreting the jet maps and residual programs from the tables (k_ad_step),
 * 2 the fused kernel compiled at plan creation with the plan's generated jet
 * maps and residual programs (tdx_ad_step_jit; plans of at least 2^17
 * (point, jet) rows, TDX_AD_JIT=0/1 to force); -1 for a null plan.
 * Diagnostic; no counterpart in the reference. */
int tdx_ad_step_kernel(const tdx_ad_plan *plan);
/* The generated source of a descriptor's fused step kernel (no device
 * needed): writes up to cap - 1 characters and a terminating zero into buf
 * (may be NULL) and returns the full length; with compile != 0 also
 * compiles it for gfx950 (hipRTC) and returns the error code if that fails;
 * 0 when the fused step does not serve the descriptor, a negative error code
 * for an invalid one.  Diagnostic; no counterpart in the reference. */
int64_t tdx_ad_step_source(const tdx_ad_desc *desc, char *buf, int64_t cap, int compile);

/* ===================================================================== */
/* Optimizer                                                              */
/* ===================================================================== */
/* torch.optim.Adam update of n values (weight_decay 0, amsgrad off,
 * maximize off): step is a device int64 of completed steps, read as t and
 * not modified; pass a separate tdx_step_increment after the update. */
int tdx_adam(int dtype, int64_t n, void *param, const void *grad,
             void *exp_avg, void *exp_avg_sq, const int64_t *step, double lr,
             double beta1, double beta2, double eps, tdx_stream_t stream);
int tdx_step_increment(int64_t *step, tdx_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* TDX_H */
