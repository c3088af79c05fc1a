// Plan-specialized jet maps and residual programs of the fused 'autograd'
// training step: source generation (compiled and loaded by
// nn_fused_jit_load, tdx_nn_jit.h).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace tdx {

// What the generated code is specialized on: the plan's int tables (jets,
// partitions at pptr_i/part_i, groups at groups_i and their records), the
// occurrence table of the jet maps' adjoint (tptr[J + 1] | (alpha, part,
// block) entries), the jet count, network outputs and the kernel's
// instantiation (NB column blocks, NHM H x H layers).
struct AdJitSpec {
  const int32_t *ints = nullptr;
  int64_t n_ints = 0;
  int J = 0, pptr_i = 0, part_i = 0, groups_i = 0, n_groups = 0, n_out = 1;
  std::vector<int32_t> occ;
  int NB = 1, NHM = 4;
};

// Source of the fused step kernel `tdx_ad_step_jit`: the body of
// tdx_ad_dev.h with straight-line jet maps for the plan's jets and one
// straight-line residual function per group.
std::string ad_step_jit_source(const AdJitSpec &sp);

}  // namespace tdx
