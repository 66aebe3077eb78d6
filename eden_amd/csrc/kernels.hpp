// kernels.hpp -- launch-side declarations shared by the host code and kernels.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace edenhip {

// One multi-compartment cell as seen by tree_kernel.
struct TreeCell {
    int64_t comp_begin;   // first global compartment
    int32_t n;            // compartments
    int32_t longest;      // Hines rounds on the schedule's longest path (information)
    int32_t n_junction;   // nodes in the schedule's junction (information)
    int32_t words_off;    // offset of its schedule block (plan_tree_waves)
    int32_t words_len;    // length of that block in 32-bit words
    int32_t n_gate;       // gate variables (one per gate of every channel distribution)
    int64_t gate_off;     // offset of its gate variables
    int32_t n_input;      // compartments with Poisson inputs
    int32_t input_off;    // offset of its input-list entries
    int32_t n_ctask;      // gated channel distributions
    int32_t ctask_off;    // offset of its channel tasks
    int32_t n_cx;         // complex compartments (see TreeComplex)
    int32_t cx_off;       // offset of their records
    int32_t const_tree;   // 1: constant diagonals below the root (see tree_kernel)
    int32_t root_children;   // children of the root (BFS positions 1 .. root_children)
};

// Per-compartment flag word of tree cells: 0 for a simple compartment (one ungated channel,
// no pulse, no Poisson input, no spike event column, not recorded), else TCX_COMPLEX | index
// of its TreeComplex record within the cell (bits 0-15; the kernel keeps the compartment's
// position in bits 16-26).
constexpr uint32_t TCX_COMPLEX = 1u << 31;

// Compartment word of the tree kernel: channel count (bits 0-2), gated-channel-slot mask
// (bits 3-6), Poisson-input flag (bit 7), chemical-synapse flag (bit 8), first channel task
// (bits 16-31).
constexpr uint32_t TCM_INPUT = 1u << 7;
constexpr uint32_t TCM_SYN = 1u << 8;
constexpr uint32_t TCM_XCH = 1u << 9;    // channels run in ext_channels_kernel (chmeta CHM_EXTERNAL)

constexpr int TREE_SLOTS = 4;        // compartments per thread



// Dynamic LDS of the tree kernel, in 4-byte units: Hines node records [nmax+1] of 16 bytes
// ((diagonal, rhs), (Ui, Uj)) | sIn [nmax] float2
// (only with Poisson inputs) | sV [nbuf][nvmax+1] | sQ [ngmax] | sGT [ngmax] | sF [nbuf][nctmax]
// | sCT [nctmax] | sIL [ninmax] | schedule block [words_max] (16-byte aligned) | sCX [cxmax]
// TreeComplex (only when the records are staged in LDS) | gate table [n_gates] DGate |
// synapse types [n_stypes] DSynType; nbuf = 2 when a separate wave runs the channel phase,
// else 1. nvmax: voltage slots of the bucket's cells (1 for a constant tree, whose only
// voltage read from LDS is the root's, else its compartment count).
struct TreeLds {
    int32_t node, in, v, q, gt, f, ct, il, w, cx, gtab, stab, total;
};

__host__ __device__ inline TreeLds tree_lds(int32_t nmax, int32_t nvmax, int32_t ngmax, int32_t nctmax,
                                            int32_t ninmax, int32_t words_max, int32_t cxmax, int32_t n_gates,
                                            int32_t n_stypes, bool split) {
    // the node records first, at the start of dynamic LDS: their address is then a constant,
    // which the compiler folds into the offsets of the Hines solve's LDS instructions
    TreeLds L;
    const int32_t nbuf = split ? 2 : 1;
    L.node = 0;
    L.in = L.node + 4 * (nmax + 1);
    L.v = L.in + (ninmax > 0 ? 2 * nmax : 0);
    L.q = L.v + nbuf * (nvmax + 1);
    L.gt = L.q + ngmax;
    L.f = L.gt + ngmax;
    L.ct = L.f + 2 * nctmax;                // open fractions, double-buffered
    L.il = L.ct + nctmax;
    L.w = (L.il + ninmax + 3) & ~3;         // task words, read up to 16 bytes at a time
    L.cx = (L.w + words_max + 3) & ~3;      // 16-byte records
    L.gtab = L.cx + 20 * cxmax;             // 32-byte DGate records (multiple of 4 units)
    L.stab = L.gtab + 8 * n_gates;          // 32-byte DSynType records
    L.total = L.stab + 8 * n_stypes;
    return L;
}

// Tables of the windowed network path for networks of one-compartment cells. Lane l of
// wave b advances point cell 64 b + l; its synapses use the block-interleaved slots of
// DevTables (synapse k at syn_base + 64 k). Ring word (step, b, w, l) =
// (step & ring_mask) * ring_stride + (b * nw + w) * 64 + l, bit j: synapse 32 w + j of that
// cell receives its event at that step. State arrays come in two copies: a window reads
// copy `cur` and writes copy `cur ^ 1`, so a window can be run again.
struct NetTables {
    int32_t n_blocks;              // waves of 64 point cells
    int32_t nw;                    // ring words per cell (ceil(max synapses per cell / 32))
    int32_t ring_mask;
    int64_t ring_stride;           // words per ring step = n_blocks * nw * 64
    const uint32_t *runs;          // [NC * 4] per compartment: type run end | type << 16, 0 = none
    float *gA[2], *gB[2];          // expTwo state per slot
    uint32_t *ring;                // [R * ring_stride]
    float *V[2], *q[2];            // compartment voltage and gate state
    float *poi_tsince[2], *poi_isi[2], *poi_A[2], *poi_B[2];
    uint64_t *spike_mask;          // [NC] spikes of the window, bit = step - window start
    int32_t *spk_list;             // compartments that spiked in the window
    uint32_t *spk_count;
    const uint32_t *tgt_rw;        // per spike target (CSR spk_begin): ring word offset << 5 | bit
    int32_t *conflict;             // set when a window's result differs from step-by-step order
    unsigned long long *n_deliv;   // synaptic events received (spike-triggered), counted
};

// net: the variant for networks (synapses on tree compartments, spikes sent to recipients),
// launched one step at a time; only without a channel wave.
const void *tree_kernel_ptr(int tpb, bool split, bool net = false);
const void *net_window_kernel_ptr();
const void *net_deliver_kernel_ptr();
const void *point_kernel_ptr();
const void *ext_channels_kernel_ptr();
const void *point_quad_kernel_ptr();

}  // namespace edenhip
