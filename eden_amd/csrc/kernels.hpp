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
};

const void *tree_kernel_ptr(int tpb);
const void *point_kernel_ptr();

}  // namespace edenhip
