// Plan-specialized residual programs of the fused 'NN' training step:
// source generation and runtime compilation (hipRTC).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace tdx {

// A group's residual program as a self-contained table pair: record words
// (equations at word 0, offsets into these words) and real values (indexed
// by the records), as the fused kernel reads them.
struct FusedProg {
  std::vector<int32_t> ints;
  std::vector<float> reals;
  int n_eq = 1;
};

// The program's structure: its record words with the per-group data
// offsets (targets, coefficient arrays, residual slots) masked, so groups
// whose programs differ only in real values and data share generated code.
std::vector<int32_t> nn_prog_key(const FusedProg &p);

// Source of the fused kernel for networks of NB 16-column blocks and NH
// H x H layers, with one generated residual function per structure in
// `keys` (a group's kind is its key's index; n_eq per key).
std::string nn_fused_jit_source(const std::vector<std::vector<int32_t>> &keys,
                                const std::vector<int> &n_eq, int NB, int NH);

// Source of the block-owning kernel that forms its own seeds (wide float
// networks, NB blocks, NH H x H layers, n_out outputs) with one generated
// residual function per structure in `keys`.
std::string nn_block_jit_source(const std::vector<std::vector<int32_t>> &keys,
                                const std::vector<int> &n_eq, int NB, int NH, int n_out);

// Compiles `src` to a code object for `arch` (no device needed); returns 0
// or a negative code with the compiler log in the error message.
int nn_fused_jit_build(const std::string &src, std::vector<char> *code,
                       const std::string &arch = "gfx950");

// Compiles and loads `src` on the current device once per process (cached
// by device and source); returns 0 and its kernel `name`, or a negative code.
int nn_fused_jit_load(const std::string &src, hipFunction_t *fn, const char *name);

}  // namespace tdx
