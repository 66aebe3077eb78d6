// kernels.hpp -- launch-side declarations shared by the host code and kernels.hip.
#pragma once
#include <stdint.h>

namespace edenhip {

// One multi-compartment cell as seen by tree_kernel.
struct TreeCell {
    int64_t comp_begin;   // first global compartment
    int32_t n;            // compartments
    int32_t rounds_fwd;   // forward-elimination rounds
    int32_t rounds_bwd;   // back-substitution rounds
    int32_t words_off;    // offset of its task words
    int32_t n_gate;       // gate variables (one per gate of every channel distribution)
    int64_t gate_off;     // offset of its gate variables
    int32_t n_input;      // compartments with Poisson inputs
    int32_t input_off;    // offset of its input-list entries
    int32_t n_ctask;      // gated channel distributions
    int32_t ctask_off;    // offset of its channel tasks
};

// Compartment word of the tree kernel: channel count (bits 0-2), gated-channel-slot mask
// (bits 3-6), Poisson-input flag (bit 7), first channel task (bits 16-31).
constexpr uint32_t TCM_INPUT = 1u << 7;

constexpr int TREE_SLOTS = 4;        // compartments per thread



const void *tree_kernel_ptr(int tpb, bool split);
const void *point_kernel_ptr();

}  // namespace edenhip
