// kernels.hip -- gfx950 kernels of the per-timestep loop.
//
//  point_kernel   one lane per single-compartment cell; time loop inside the launch, state in
//                 registers (a launch may cover one step, as in the network path).
//  point_quad_kernel  four lanes per uncoupled single-compartment cell, each advancing one
//                 of its gate variables beside the voltage update.
//  tree_kernel    one workgroup (TPB threads, one or more wave64s) per multi-compartment
//                 cell, popped from a work queue sorted by compartment count (largest first);
//                 compartment state in registers (4 compartments per thread), the Hines solve
//                 in LDS as rounds of independent per-node gathers. The gather at each tree
//                 node folds its children in ascending index order, which is the order EDEN's
//                 serial depth-first elimination (eden/Eden.cpp:6114-6225) applies them, so
//                 every diagonal and right-hand-side value sees the same operations in the
//                 same order as the serial chain.
#include "device_model.hpp"
#include "kernels.hpp"
#include "tree_schedule.hpp"

namespace edenhip {

__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2)))
point_kernel(DevTables T, const DGate *__restrict__ gates_g, int n_gates, const DSynType *__restrict__ stypes_g,
             int n_stypes, const int32_t *__restrict__ comps, int32_t n, int64_t s0, int64_t s1, int64_t row0,
             float dt) {
    __shared__ DGate gates[MAX_CHAN_TYPES * GATES_PER_TYPE];
    __shared__ DSynType stypes[64];
    for (int i = threadIdx.x; i < n_gates; i += blockDim.x) gates[i] = gates_g[i];
    for (int i = threadIdx.x; i < n_stypes; i += blockDim.x) stypes[i] = stypes_g[i];
    __syncthreads();
    const int32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const int64_t g = comps[idx];
    CompConst k;
    load_comp_const(T, g, k);
    float V = T.V[g];
    float q[NQ];
#pragma unroll
    for (int s = 0; s < NQ; s++) q[s] = T.q[g * NQ + s];
    const int32_t syn_base = T.syn_base[g], syn_cnt = T.syn_cnt[g];
    const bool has_syn = syn_cnt > 0;
    const bool has_poi = T.poi_begin[g + 1] > T.poi_begin[g];
    const bool has_spk = T.spk_begin[g + 1] > T.spk_begin[g];

    // time_f64 accumulates `time += dt` in double, as the reference's engine loop does
    double time = (s0 >= 1) ? T.t_row[s0 - 1] : 0.0;
    for (int64_t s = s0; s < s1; s++) {
        const bool init = s <= 0;
        const float t = (float)time;
        const int now = (int)((s + 3) & 1), nxt = now ^ 1;
        if (!init && k.rec >= 0) trace_at(T, k.rec, s - 1 - row0) = V;
        float I_ch, G_ch;
        channels_step(gates, k, V, q, init, dt, I_ch, G_ch);
        if (k.chmeta == CHM_EXTERNAL) {
            const float2 e = T.chx_ext[g];   // channels of this step (ext_channels_kernel)
            I_ch = e.x;
            G_ch = e.y;
        }
        float I_syn = 0.0f, G_syn = 0.0f;
        if (has_syn) {
            if (T.syn_ext) {
                const float2 e = T.syn_ext[g];   // generic synapses of this step (lems_syn_kernel)
                I_syn = e.x;
                G_syn = e.y;
            } else {
                synapses_step(T, stypes, syn_base, syn_cnt, V, init, t, dt, now, s, I_syn, G_syn);
            }
        }
        float I_in = pulses_current(T, k, g, t), G_in = 0.0f;
        if (has_poi) {
            const PoiState ps{T.poi_tsince, T.poi_isi, T.poi_A, T.poi_B};
            poisson_step(T, stypes, g, V, init, s, dt, I_in, G_in, ps, ps);
        }
        if (T.inp_ext && has_syn) {
            const float2 e = T.inp_ext[g];   // LEMS inputs of this step, after the native ones
            I_in += e.x;
            G_in += e.y;
        }
        const float I_int = I_ch + I_in + I_syn;
        const float G_int = G_ch + G_in + G_syn;
        float D = 1 * (G_int / k.C) + k.invrc;
        float Vn;
        if (init) {
            Vn = V;
        } else {
            Vn = V + (dt * (I_int + ((V * G_int))) / k.C);
            D = 1 + D * dt;     // one-compartment tree: no elimination, root divide only
            Vn = Vn / D;
        }
        if (V < k.vthr && k.vthr < Vn) {
            if (has_spk) send_spike(T, g, nxt);
            log_event(T, k.ev, s);
        }
        V = Vn;
        if (!init) time += (double)dt;
    }
    T.V[g] = V;
#pragma unroll
    for (int s = 0; s < NQ; s++) T.q[g * NQ + s] = q[s];
}

// Uncoupled one-compartment cells (no synapses, Poisson inputs or spike recipients), four
// lanes per cell. The lanes of a quad compute the cell's currents and voltage update
// identically (the same operations on the same values), and lane s of the quad also advances
// gate variable s. A step's currents use the present gate values (gathered from the quad's
// lanes), while the gate updates feed only the next step (eden/Eden.cpp:5330: each gate's
// present value enters the open fraction before it is updated), so the gates' rate and exp
// chains run side by side and beside the voltage update instead of one after another.
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2)))
point_quad_kernel(DevTables T, const DGate *__restrict__ gates_g, int n_gates, const int32_t *__restrict__ comps,
                  int32_t n, int64_t s0, int64_t s1, int64_t row0, float dt) {
    __shared__ DGate gates[MAX_CHAN_TYPES * GATES_PER_TYPE];
    for (int i = threadIdx.x; i < n_gates; i += blockDim.x) gates[i] = gates_g[i];
    __syncthreads();
    const int32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int32_t idx = gid >> 2, qs = gid & 3;
    if (idx >= n) return;   // the whole quad: its lanes share idx
    const int64_t g = comps[idx];
    CompConst k;
    load_comp_const(T, g, k);
    float V = T.V[g];
    float qv = T.q[g * NQ + qs];   // gate variable qs of the cell
    const int nq = (k.gmeta >> 24) & 7;
    const int nch = (k.gmeta >> 28) & 7;
    const bool mine = qs < nq;
    DGate G{};
    if (mine) {
        const int cs = (k.gmeta >> (4 * qs)) & 3;
        const int gi = (k.gmeta >> (4 * qs + 2)) & 3;
        G = gates[((k.chmeta >> (8 * cs)) & 0xFF) * GATES_PER_TYPE + gi];
    }
    // time_f64 accumulates `time += dt` in double, as the reference's engine loop does
    double time = (s0 >= 1) ? T.t_row[s0 - 1] : 0.0;
    for (int64_t s = s0; s < s1; s++) {
        const bool init = s <= 0;
        const float t = (float)time;
        if (!init && k.rec >= 0 && qs == 0) trace_at(T, k.rec, s - 1 - row0) = V;
        float q[NQ];
#pragma unroll
        for (int j = 0; j < NQ; j++) q[j] = __shfl(qv, j, 4);
        // open fractions from the present gate values, in gate order (channels_step)
        float fop[NCH];
#pragma unroll
        for (int c = 0; c < NCH; c++) fop[c] = 1.0f;
#pragma unroll
        for (int j = 0; j < NQ; j++) {
            if (j < nq) {
                const int cs = (k.gmeta >> (4 * j)) & 3;
                const int gi = (k.gmeta >> (4 * j + 2)) & 3;
                const int inst = (gates[((k.chmeta >> (8 * cs)) & 0xFF) * GATES_PER_TYPE + gi].forms >> 16) & 0xFF;
#pragma unroll
                for (int c = 0; c < NCH; c++) {
                    if (c == cs) {
                        float pw = fop[c];
                        for (int i = 0; i < inst; i++) pw = pw * q[j];
                        fop[c] = pw;
                    }
                }
            }
        }
        // this lane's gate variable for the next step
        const float nv = mine ? gate_next(G, V, qv, init, dt) : qv;
        float I_ch = 0.0f, G_ch = 0.0f;
#pragma unroll
        for (int c = 0; c < NCH; c++) {
            if (c < nch) {
                const float Gscaled = k.g[c] * fop[c];
                const float iDensity = Gscaled * (k.e[c] - V) * 1e-5f;
                const float gDensity = (Gscaled) * 1e-5f;
                I_ch += (iDensity * k.A);
                G_ch += (gDensity * k.A);
            }
        }
        const float I_in = pulses_current(T, k, g, t), G_in = 0.0f;
        const float I_syn = 0.0f, G_syn = 0.0f;
        const float I_int = I_ch + I_in + I_syn;
        const float G_int = G_ch + G_in + G_syn;
        float D = 1 * (G_int / k.C) + k.invrc;
        float Vn;
        if (init) {
            Vn = V;
        } else {
            Vn = V + (dt * (I_int + ((V * G_int))) / k.C);
            D = 1 + D * dt;
            Vn = Vn / D;
        }
        if (V < k.vthr && k.vthr < Vn && qs == 0) log_event(T, k.ev, s);
        V = Vn;
        qv = nv;
        if (!init) time += (double)dt;
    }
    if (qs == 0) T.V[g] = V;
    T.q[g * NQ + qs] = qv;
}

// LDS-only workgroup barrier: waits for this wave's LDS operations, then (for more than one
// wave) the hardware barrier. Outstanding global loads/stores (the trace write, the next
// step's time prefetch) are not drained, unlike __syncthreads().
// A single wave needs no wait at all: LDS executes one wave's instructions in issue order,
// so a later read sees an earlier write; only the compiler must not reorder across it.
template <int TPB>
__device__ __forceinline__ void lds_barrier() {
    if constexpr (TPB <= 64) {
        __builtin_amdgcn_wave_barrier();
        asm volatile("" ::: "memory");
    } else {
        asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    }
}

// Open fraction of every channel of a compartment from its gate values (in gate order,
// each multiplied in `instances` times, eden/Eden.cpp:5679-5721), then the channel current
// and conductance sums in distribution order (eden/Eden.cpp:5805-5898).
__device__ __forceinline__ void tree_channels(const CompConst &k, uint32_t tcm, const float *sF, float V, float &I_ch,
                                              float &G_ch) {
    // channel slot c uses the open fraction of its channel task (gated channels, in slot
    // order from the compartment's first task) or 1 (conductance_scaling alone)
    const int nch = tcm & 7;
    const uint32_t mask = (tcm >> 3) & 15;
    int idx = tcm >> 16;
    I_ch = 0.0f; G_ch = 0.0f;
#pragma unroll
    for (int c = 0; c < NCH; c++) {
        const bool gated = (mask >> c) & 1;
        const float f = gated ? sF[idx] : 1.0f;
        idx += gated ? 1 : 0;
        const float Gscaled = k.g[c] * f;
        const float iDensity = Gscaled * (k.e[c] - V) * 1e-5f;
        const float gDensity = (Gscaled) * 1e-5f;
        I_ch = (c < nch) ? I_ch + (iDensity * k.A) : I_ch;
        G_ch = (c < nch) ? G_ch + (gDensity * k.A) : G_ch;
    }
}

// Open fraction of a channel from its gate values, multiplied in gate order, each gate
// `instances` times, starting from conductance_scaling = 1 (eden/Eden.cpp:5679-5721).
__device__ __forceinline__ float channel_fopen(const DGate *gates, int type, const float *q) {
    const int ng = gates[type * GATES_PER_TYPE].forms >> 24;
    float f = 1.0f;
    for (int gi = 0; gi < ng; gi++) {
        const int inst = (gates[type * GATES_PER_TYPE + gi].forms >> 16) & 0xFF;
        const float qv = q[gi];
        for (int i = 0; i < inst; i++) f = f * qv;
    }
    return f;
}

// Channel phase of a cell, on the lanes `lane` + `stride` k of the channel group: every gate
// variable's next value from its compartment's voltage (one lane per gate), then, after
// `sync`, each gated channel's open fraction for the next step (one lane per channel).
template <class Sync>
__device__ __forceinline__ void channel_phase(const DGate *gates, const uint32_t *sCT, const uint32_t *sGT, float *sQ,
                                              float *sF_out, const float *sV_in, int NG, int NCT, int lane,
                                              int stride, bool init, float dt, Sync sync) {
    for (int g = lane; g < NG; g += stride) {
        const uint32_t gt = sGT[g];
        const uint32_t w = sCT[gt & 0xFFFFu];
        const int type = (w >> 11) & 63;
        sQ[g] = gate_next(gates[type * GATES_PER_TYPE + (gt >> 16)], sV_in[w & 0x7FF], sQ[g], init, dt);
    }
    sync();
    for (int i = lane; i < NCT; i += stride) {
        const uint32_t w = sCT[i];
        sF_out[i] = channel_fopen(gates, (w >> 11) & 63, sQ + (w >> 17));
    }
}

// One workgroup per multi-compartment cell, cells popped from a queue in decreasing size.
// The cable group (TPB threads) owns compartments t + TPB j (j < 4) in registers and runs,
// per step:
//   inputs     per-cell list of compartments with Poisson inputs -> sIn (LDS)
//   membrane   currents from the channels' open fractions (sF) and inputs, semi-implicit RHS
//   Hines      in LDS on breadth-first positions (tree_schedule.hpp): a 16-byte record per
//              node, (diagonal, rhs) then (Ui, Uj) of the edge to its parent, one task word per
//              thread per round
//   readback   new voltages, spike detection
// The channel phase (next gate values and open fractions from the present voltages) runs
// either in the cable group after the membrane phase, or, with SPLIT, in one extra wave at
// the same time as the cable group's step; voltages and open fractions are then
// double-buffered and the two meet at one workgroup barrier per step.
// NET: networks, launched one step at a time. A compartment with chemical synapses
// (TCM_SYN) runs them in its complex update, with the trigger flags of the step's parity as
// the point kernel does, and a compartment with spike recipients sends its spikes to the
// next step's flags; both are complex compartments, so the per-slot code of the other
// compartments is unchanged and the NET = false kernels carry none of it.
template <int TPB, bool SPLIT, bool NET>
__global__ void __launch_bounds__(TPB + (SPLIT ? 64 : 0)) __attribute__((amdgpu_waves_per_eu(4)))
tree_kernel(DevTables T, const DGate *__restrict__ gates_g, int n_gates, const DSynType *__restrict__ stypes_g,
            int n_stypes, const TreeCell *__restrict__ cells, const int32_t *__restrict__ task_order, int32_t n_tasks,
            int32_t *__restrict__ queue, const uint32_t *__restrict__ words_g, const int32_t *__restrict__ pos_g,
            const float2 *__restrict__ u_g, const uint32_t *__restrict__ tcm_g, const uint32_t *__restrict__ ctask_g,
            float *__restrict__ gq_g, const uint32_t *__restrict__ gt_g, const uint32_t *__restrict__ inl_g,
            const uint32_t *__restrict__ cflag_g, const TreeComplex *__restrict__ cx_g, int32_t cxmax_lds, int32_t nmax,
            int32_t nvmax, int32_t ngmax, int32_t nctmax, int32_t ninmax, int32_t words_max, int64_t s0, int64_t s1,
            int64_t row0, float dt) {
    constexpr int K = TREE_SLOTS;
    constexpr int NT = TPB + (SPLIT ? 64 : 0);
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ int32_t s_task;
    // dynamic LDS, laid out by tree_lds()
    const TreeLds L = tree_lds(nmax, nvmax, ngmax, nctmax, ninmax, words_max, cxmax_lds, n_gates, n_stypes, SPLIT);
    float *sm = reinterpret_cast<float *>(smem);
    DGate *gates = reinterpret_cast<DGate *>(sm + L.gtab);
    DSynType *stypes = reinterpret_cast<DSynType *>(sm + L.stab);
    TreeComplex *sCX = reinterpret_cast<TreeComplex *>(sm + L.cx);
    // Hines node records, 16 bytes at position p: sDX(p) = (diagonal, rhs), then sU(p) =
    // (Ui, Uj) of the edge from p to its parent; one record serves a task's reads of both
    char *const nodes = reinterpret_cast<char *>(sm + L.node);
    auto sDX = [nodes](int p) -> float2 & { return *reinterpret_cast<float2 *>(nodes + 16 * p); };
    auto sU = [nodes](int p) -> float2 & { return *reinterpret_cast<float2 *>(nodes + 16 * p + 8); };
    float2 *sIn = reinterpret_cast<float2 *>(sm + L.in);
    float *sVb = sm + L.v;
    float *sQ = sm + L.q;
    uint32_t *sGT = reinterpret_cast<uint32_t *>(sm + L.gt);
    float *sFb = sm + L.f;
    uint32_t *sCT = reinterpret_cast<uint32_t *>(sm + L.ct);
    uint32_t *sIL = reinterpret_cast<uint32_t *>(sm + L.il);
    uint32_t *sW = reinterpret_cast<uint32_t *>(sm + L.w);   // schedule block: round headers, task words
    const int tid = threadIdx.x;
    const bool cable = !SPLIT || tid < TPB;
    for (int i = tid; i < n_gates; i += NT) gates[i] = gates_g[i];
    for (int i = tid; i < n_stypes; i += NT) stypes[i] = stypes_g[i];

    for (;;) {
        __syncthreads();
        if (tid == 0) s_task = atomicAdd(&queue[0], 1);
        __syncthreads();
        const int32_t task = s_task;
        if (task >= n_tasks) break;
        const TreeCell cell = cells[task_order[task]];
        const int32_t N = cell.n;
        const int32_t NG = cell.n_gate;
        const int32_t NCT = cell.n_ctask;
        const int32_t NIN = cell.n_input;
        const int64_t c0 = cell.comp_begin;
        const int64_t gq0 = cell.gate_off;
        // const_tree: every non-root compartment has a constant diagonal (ungated channels,
        // no conductance input), so below the root the eliminated diagonals D' and the ratios
        // Uj / D' are the same every step. They are computed once per task below, with the
        // per-step operations, and the forward sweep then reads the ratio in place of Uj.
        const bool const_tree = cell.const_tree != 0;
        {
            const uint32_t *Wg = words_g + cell.words_off;
            for (int i = tid; i < cell.words_len; i += NT) sW[i] = Wg[i];
        }
        // Hines task lists (tree_schedule.hpp, plan_tree_waves): cable wave wv runs the lists
        // of its subtrees, wave 0 also those of the junction. run_list calls body(word) round
        // after round with no barrier (a wave's LDS accesses are performed in order), on the
        // lanes below the round's task count; lanes beyond it are masked off (their word is
        // the idle word, which points every access at the dummy slot N). Task words hold byte
        // offsets of node records, which start at a constant LDS address. Every lane reads a word
        // from the round (the block ends with padding for the lanes beyond a round's count).
        // The loop is unrolled twice, so that no register moves between rounds: after a
        // round's body is issued, the words of the round after next and the header of the
        // round two after that are read into the registers the round had used (five sentinel
        // headers end the block). Each value is then complete before its use, with no wait on
        // the loads just issued.
        constexpr int W = TPB / 64;
        const int wv = tid >> 6, lane = tid & 63;
        const uint32_t dummy_off = 16u * (uint32_t)N;
        const uint2 idle_fold = make_uint2(dummy_off | (dummy_off << 16), dummy_off);
        const uint2 idle_bwd = make_uint2(dummy_off | (dummy_off << 16), dummy_off);
        // constant trees take four-word fold tasks with merged children (plan_tree_waves
        // with merge): c0, c1 and the grandchildren refolded through each, dummy where absent
        const uint32_t dd = dummy_off | (dummy_off << 16);
        const uint4 idle_fold4 = make_uint4(dd, dd, dd, dummy_off);
        char *const dxb = nodes;       // task words: byte offsets 16 p of node records
        char *const ub = nodes + 8;
        // a list's directory entry and first four round headers: the same every step, so they
        // are read once per cell into scalar registers, and a list starts with the reads of its
        // first task words
        struct ListHead {
            uint32_t list, h0, h1, h2, h3;
        };
        auto list_head = [&](int dir) {
            ListHead L;
            L.list = __builtin_amdgcn_readfirstlane(sW[dir]);
            const int first = L.list & 0xFFFFu;
            L.h0 = __builtin_amdgcn_readfirstlane(sW[first]);
            L.h1 = __builtin_amdgcn_readfirstlane(sW[first + 1]);
            L.h2 = __builtin_amdgcn_readfirstlane(sW[first + 2]);
            L.h3 = __builtin_amdgcn_readfirstlane(sW[first + 3]);
            return L;
        };
        auto run_list = [&](const ListHead &L, auto idle, auto body) {
            using Word = decltype(idle);
            const int first = L.list & 0xFFFFu, cnt = L.list >> 16;
            auto word = [&](uint32_t h, bool &act) {
                const Word w = reinterpret_cast<const Word *>(sW + (h & 0xFFFFu))[lane];
                act = lane < (int)(h >> 16);
                return act ? w : idle;
            };
            bool aa, ab;
            uint32_t ha = L.h2, hb = L.h3;
            Word wa = word(L.h0, aa), wb = word(L.h1, ab);
            for (int r = 0; r < cnt; r += 2) {
                if (aa) body(wa);
                wa = word(ha, aa);
                ha = sW[first + r + 4];
                if (r + 1 < cnt) {
                    if (ab) body(wb);
                    wb = word(hb, ab);
                    hb = sW[first + r + 5];
                }
            }
        };
        for (int i = tid; i < N; i += NT) sU(i) = u_g[c0 + i];
        for (int i = tid; i < NG; i += NT) {
            sQ[i] = gq_g[gq0 + i];
            sGT[i] = gt_g[gq0 + i];
        }
        for (int i = tid; i < NCT; i += NT) sCT[i] = ctask_g[cell.ctask_off + i];
        for (int i = tid; i < NIN; i += NT) sIL[i] = inl_g[cell.input_off + i];
        // complex-compartment records: staged in LDS when the bucket's largest count fits,
        // else read from global memory
        const TreeComplex *CX = cx_g + cell.cx_off;
        if (cxmax_lds > 0) {
            const uint4 *src = reinterpret_cast<const uint4 *>(cx_g + cell.cx_off);
            uint4 *dst = reinterpret_cast<uint4 *>(sCX);
            for (int i = tid; i < cell.n_cx * 5; i += NT) dst[i] = src[i];
            CX = sCX;
        }
        // voltages read from LDS: the root's in a constant tree (no gated channel or input
        // below it), else every compartment's
        const int32_t NV = const_tree ? 1 : N;
        for (int i = tid; i < NV; i += NT) sVb[i] = T.V[c0 + i];
        if (tid == 0) {
            // dummy slot for threads without a task: finite values, never read back
            sU(N) = make_float2(0.0f, 0.0f);
            sDX(N) = make_float2(1.0f, 0.0f);
        }

        // Per slot registers: voltage, BFS position, flag word, trace column, and for a simple
        // compartment its channel's Gbase and Erev, area, capacitance, and the conductance sum
        // and Hines diagonal, which do not change from step to step (computed here with the
        // per-step operation order). A complex compartment's constants come from its record.
        // Slots beyond the cell behave as simple compartments with no channel conductance,
        // C = 1 and the dummy position N, so they compute exactly the dummy slot's (1, 0).
        float V[K], g0[K], e0[K], Ar[K], Cr[K], Gs[K], Ds[K];
        // fl: the compartment's flag word (TCX_COMPLEX | record index, or 0) | BFS position << 16
        uint32_t fl[K];
        auto posof = [&](int j) { return (int)((fl[j] >> 16) & 0x7FFu); };
        bool slot_any[K], slot_simple[K];   // per wave and slot: any valid lane, no complex lane
        const int wbase = __builtin_amdgcn_readfirstlane(tid & ~63);
#pragma unroll
        for (int j = 0; j < K; j++) {
            const int c = tid + TPB * j;
            const bool valid = cable && c < N;
            V[j] = 0.0f; g0[j] = 0.0f; e0[j] = 0.0f; Ar[j] = 0.0f; Cr[j] = 1.0f; Gs[j] = 0.0f; Ds[j] = 1.0f;
            fl[j] = (uint32_t)N << 16;
            if (valid) {
                const int64_t g = c0 + c;
                fl[j] = cflag_g[g] | ((uint32_t)pos_g[g] << 16);
                V[j] = T.V[g];
                if (!(fl[j] & TCX_COMPLEX)) {
                    g0[j] = T.chg[g * NCH];
                    e0[j] = T.che[g * NCH];
                    Ar[j] = T.A[g];
                    Cr[j] = T.C[g];
                    const float G_ch = 0.0f + (((g0[j] * 1.0f) * 1e-5f) * Ar[j]);
                    const float G_int = (G_ch + 0.0f) + 0.0f;
                    const float Dd = 1 * (G_int / Cr[j]) + T.invrc[g];
                    Gs[j] = G_int;
                    Ds[j] = 1 + Dd * dt;
                }
            }
            slot_any[j] = __builtin_amdgcn_readfirstlane(wbase + TPB * j) < N && cable;
            // the root (thread 0, slot 0) is updated separately, after wave 0's fold rounds
            slot_simple[j] = __ballot(valid && c != 0 && (fl[j] & TCX_COMPLEX)) == 0ull;
        }
        __syncthreads();
        // the Hines lists of this wave (for the channel wave, unused) and of the junction
        const ListHead Lf = list_head(2 * wv), Lb = list_head(2 * wv + 1);
        const ListHead Ljf = list_head(2 * W), Ljb = list_head(2 * W + 1);
        if (const_tree) {
            // eliminated diagonals of the constant subtrees: the forward sweep's diagonal part
            // on this step-independent matrix, with the per-step operations
#pragma unroll
            for (int j = 0; j < K; j++) {
                const int c = tid + TPB * j;
                if (cable && c < N) {
                    float Dr = Ds[j];
                    if (fl[j] & TCX_COMPLEX) {
                        const TreeComplex &cx = CX[fl[j] & 0xFFFFu];
                        const CompConst k = cx.k;
                        float I_ch, G_ch;
                        tree_channels(k, cx.tcm, sFb, V[j], I_ch, G_ch);
                        const float G_in = 0.0f, G_syn = 0.0f;
                        const float G_int = G_ch + G_in + G_syn;
                        const float Dd = 1 * (G_int / k.C) + k.invrc;
                        Dr = 1 + Dd * dt;
                    }
                    sDX(posof(j)) = make_float2(Dr, 0.0f);
                }
            }
            __syncthreads();
            // a fold task's diagonal part; a merged child's eliminated diagonal is formed again
            // from its children, as that child's own task forms it in the same round (the
            // dummy slot, D 1 and U 0, gives exactly zero terms for an unmerged child)
            auto diag_task = [&](uint4 w) {
                const uint32_t po = w.x & 0xFFFFu, ao = w.x >> 16, bo = w.y & 0xFFFFu;
                auto dprime = [&](uint32_t c, uint32_t g0, uint32_t g1) {
                    const float dc = reinterpret_cast<const float2 *>(dxb + c)->x;
                    const float d0 = reinterpret_cast<const float2 *>(dxb + g0)->x;
                    const float d1 = reinterpret_cast<const float2 *>(dxb + g1)->x;
                    const float2 u0 = *reinterpret_cast<const float2 *>(ub + g0);
                    const float2 u1 = *reinterpret_cast<const float2 *>(ub + g1);
                    return (dc - (u0.y / d0) * u0.x) - (u1.y / d1) * u1.x;
                };
                const float d = reinterpret_cast<const float2 *>(dxb + po)->x;
                const float da = dprime(ao, w.y >> 16, w.z & 0xFFFFu);
                const float db = dprime(bo, w.z >> 16, w.w);
                const float2 ua = *reinterpret_cast<const float2 *>(ub + ao);
                const float2 ubb = *reinterpret_cast<const float2 *>(ub + bo);
                reinterpret_cast<float2 *>(dxb + po)->x = (d - (ua.y / da) * ua.x) - (ubb.y / db) * ubb.x;
            };
            if (cable) run_list(Lf, idle_fold4, diag_task);
            __syncthreads();
            if (cable && wv == 0) run_list(Ljf, idle_fold4, diag_task);
            __syncthreads();
            // from here on, below the root: sDX[c] = (ratio Uj / D', right-hand side) and
            // sU[c] = (Ui, D'); the per-step diagonal register Ds holds the ratio
#pragma unroll
            for (int j = 0; j < K; j++) {
                const int c = tid + TPB * j;
                if (cable && c < N && posof(j) != 0) {
                    const float Dp = sDX(posof(j)).x;
                    Ds[j] = sU(posof(j)).y / Dp;
                    sU(posof(j)).y = Dp;
                }
            }
            // dummy slot: ratio 1 (times right-hand side 0), D' = 1 for the idle backward task
            if (tid == 0) sU(N) = make_float2(0.0f, 1.0f);
        }
        __syncthreads();
        // open fractions of the gated channels from the present gate values
        for (int i = tid; i < NCT; i += NT) {
            const uint32_t w = sCT[i];
            sFb[i] = channel_fopen(gates, (w >> 11) & 63, sQ + (w >> 17));
        }
        __syncthreads();

        // diagonal and right-hand side of a complex compartment (slot j, cell-local index c)
        auto complex_update = [&](int j, int c, const float *sF, float t, bool init, int64_t s, float &Dn,
                                  float &Vn) {
            const TreeComplex &cx = CX[fl[j] & 0xFFFFu];
            const CompConst k = cx.k;
            float I_ch, G_ch;
            tree_channels(k, cx.tcm, sF, V[j], I_ch, G_ch);
            if (k.chmeta == CHM_EXTERNAL) {
                const float2 e = T.chx_ext[c0 + c];   // channels of this step (ext_channels_kernel)
                I_ch = e.x;
                G_ch = e.y;
            }
            float I_in = pulses_current(T, k, c0 + c, t), G_in = 0.0f;
            if (cx.tcm & TCM_INPUT) {
                const float2 in = sIn[c];
                I_in += in.x;
                G_in += in.y;
            }
            float I_syn = 0.0f, G_syn = 0.0f;
            if constexpr (NET) {
                // synapses of this compartment, with the flags of this step's parity (the
                // trigger buffers alternate as in point_kernel)
                if (cx.tcm & TCM_SYN) {
                    const int64_t g = c0 + c;
                    if (T.syn_ext) {
                        const float2 e = T.syn_ext[g];   // generic synapses of this step (lems_syn_kernel)
                        I_syn = e.x;
                        G_syn = e.y;
                        if (T.inp_ext) {
                            const float2 e2 = T.inp_ext[g];   // LEMS inputs, after the native ones
                            I_in += e2.x;
                            G_in += e2.y;
                        }
                    } else {
                        synapses_step(T, stypes, T.syn_base[g], T.syn_cnt[g], V[j], init, t, dt, (int)((s + 3) & 1), s,
                                      I_syn, G_syn);
                    }
                }
            }
            const float I_int = I_ch + I_in + I_syn;
            const float G_int = G_ch + G_in + G_syn;
            const float Dd = 1 * (G_int / k.C) + k.invrc;
            Dn = (const_tree && posof(j) != 0) ? Ds[j] : 1 + Dd * dt;
            Vn = init ? V[j] : V[j] + (dt * (I_int + ((V[j] * G_int)))) / k.C;
        };
        int cur = 0, fcur = 0;
        if (SPLIT && !cable) {
            // channel wave, in a step loop of its own (so that none of the cable group's
            // per-slot registers are live across its gate updates): next open fractions from
            // this step's voltages, overlapping the cable group's step
            for (int64_t s = s0; s < s1; s++) {
                const bool init = s <= 0;
                const float *sV = sVb + cur * (nvmax + 1);
                float *sFn = sFb + (fcur ^ 1) * nctmax;
                // With several cable waves the channel wave takes part in their barriers
                // (the input and compartment barriers first, then, after its own work, the
                // three of the Hines solve), so that every wave passes the same barriers.
                if constexpr (TPB > 64) {
                    if (NIN > 0) lds_barrier<TPB>();
                    lds_barrier<TPB>();
                }
                channel_phase(gates, sCT, sGT, sQ, sFn, sV, NG, NCT, tid - TPB, 64, init, dt, [] { lds_barrier<64>(); });
                if constexpr (TPB > 64) {
                    if (!init) {
                        lds_barrier<TPB>();
                        lds_barrier<TPB>();
                        lds_barrier<TPB>();
                    }
                }
                // the end-of-step barrier of the cable group's loop
                asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
                cur ^= 1;
                fcur ^= 1;
            }
        } else {
            // time_f64 accumulates `time += dt` in double, as the reference's engine loop does
            double time = (s0 >= 1) ? T.t_row[s0 - 1] : 0.0;
            for (int64_t s = s0; s < s1; s++) {
                const bool init = s <= 0;
                const float t = (float)time;
                if (!init) time += (double)dt;
                const int nxt = SPLIT ? cur ^ 1 : cur;
                const float *sV = sVb + cur * (nvmax + 1);
                float *sVn = sVb + nxt * (nvmax + 1);
                // open fractions double-buffered in every mode: the root compartment's update, which
                // reads them, runs after the gate updates may have written the next ones
                const float *sF = sFb + fcur * nctmax;
                float *sFn = sFb + (fcur ^ 1) * nctmax;
                if (cable) {
                    // Poisson inputs of the compartments that have any
                    if (NIN > 0) {
                        for (int i = tid; i < NIN; i += TPB) {
                            const uint32_t e = sIL[i];
                            const int c = e & 0x7FF;
                            float Ii = 0.0f, Gi = 0.0f;
                            const PoiState ps{T.poi_tsince, T.poi_isi, T.poi_A, T.poi_B};
                            poisson_step(T, stypes, c0 + c, sV[c], init, s, dt, Ii, Gi, ps, ps);
                            sIn[c] = make_float2(Ii, Gi);
                        }
                        lds_barrier<TPB>();
                    }
                    // compartments: currents from the present open fractions, semi-implicit RHS.
                    // Slots without a valid lane in this wave are skipped; a wave whose valid
                    // lanes in a slot are all simple compartments takes the short form of the
                    // same operations (the constant conductance and diagonal computed above).
    #pragma unroll
                    for (int j = 0; j < K; j++) {
                        if (!slot_any[j]) continue;
                        const int c = tid + TPB * j;
                        // simple form: channel slot 0, Gbase * 1 (ungated) * (E - V) * 1e-5f * A,
                        // then the empty input and synapse terms, as in the general form below
                        const float I_ch_s = 0.0f + ((((g0[j] * 1.0f) * (e0[j] - V[j])) * 1e-5f) * Ar[j]);
                        const float I_int_s = (I_ch_s + 0.0f) + 0.0f;
                        float Dn = Ds[j];
                        float Vn = init ? V[j] : V[j] + (dt * (I_int_s + ((V[j] * Gs[j])))) / Cr[j];
                        if (!slot_simple[j] && c != 0 && (fl[j] & TCX_COMPLEX)) complex_update(j, c, sF, t, init, s, Dn, Vn);
                        sDX(posof(j)) = make_float2(Dn, Vn);
                    }
                    lds_barrier<TPB>();
                    if constexpr (!SPLIT) {
                        // gates on the last cable wave (lanes of that wave, wave-local sync), before
                        // its Hines rounds; the schedule gives that wave a lighter share of the tree.
                        // The open fractions are read again only after the step's last barrier.
                        if (wv == W - 1)
                            channel_phase(gates, sCT, sGT, sQ, sFn, sV, NG, NCT, lane, 64, init, dt, [] { lds_barrier<64>(); });
                    }
                    if (!init) {
                        // forward elimination: fold tasks of at most two children, applied in
                        // ascending child order (the dummy slot N, with right-hand side and Ui 0,
                        // stands in for a missing second child: its terms are exactly zero)
                        // constant tree: below the root sDX[c] = (constant ratio Uj / D', rhs), so
                        // only right-hand sides change; a merged child's right-hand side is folded
                        // again from its children, with the operations of its own task, which
                        // reads the same values and writes that child in the same round
                        auto fold_const = [&](uint4 w) {
                            const uint32_t po = w.x & 0xFFFFu, ao = w.x >> 16, bo = w.y & 0xFFFFu;
                            auto rhs = [&](float x, uint32_t g0, uint32_t g1) {
                                const float2 q0 = *reinterpret_cast<const float2 *>(dxb + g0);
                                const float2 q1 = *reinterpret_cast<const float2 *>(dxb + g1);
                                return (x - q0.x * q0.y) - q1.x * q1.y;
                            };
                            const float2 a = *reinterpret_cast<const float2 *>(dxb + ao);
                            const float2 b = *reinterpret_cast<const float2 *>(dxb + bo);
                            const float xa = rhs(a.y, w.y >> 16, w.z & 0xFFFFu);
                            const float xb = rhs(b.y, w.z >> 16, w.w);
                            float *xp = reinterpret_cast<float *>(dxb + po) + 1;
                            *xp = (*xp - a.x * xa) - b.x * xb;
                        };
                        // general tree: diagonal and right-hand side
                        auto fold_task = [&](uint2 w) {
                            const uint32_t po = w.x & 0xFFFFu, ao = w.x >> 16, bo = w.y;
                            float2 *dp = reinterpret_cast<float2 *>(dxb + po);
                            const float2 dx = *dp;
                            const float2 da = *reinterpret_cast<const float2 *>(dxb + ao);
                            const float2 ua = *reinterpret_cast<const float2 *>(ub + ao);
                            const float2 db = *reinterpret_cast<const float2 *>(dxb + bo);
                            const float2 ubb = *reinterpret_cast<const float2 *>(ub + bo);
                            const float r0 = ua.y / da.x;
                            const float r1 = ubb.y / db.x;
                            float2 o;
                            o.x = (dx.x - r0 * ua.x) - r1 * ubb.x;
                            o.y = (dx.y - r0 * da.y) - r1 * db.y;
                            *dp = o;
                        };
                        // back-substitution of one non-root node from its solved parent, or from its
                        // solved grandparent through the parent (solved again, the same operations,
                        // when the parent's own task shares the round); mid = dummy slot N for the
                        // first form, whose extra solve is discarded
                        auto bwd_task = [&](uint2 w) {
                            const uint32_t po = w.x & 0xFFFFu, so = w.x >> 16, mo = w.y;
                            const float xs = reinterpret_cast<const float2 *>(dxb + so)->y;
                            const float2 dm = *reinterpret_cast<const float2 *>(dxb + mo);
                            const float2 um = *reinterpret_cast<const float2 *>(ub + mo);
                            const float2 dp = *reinterpret_cast<const float2 *>(dxb + po);
                            const float2 up = *reinterpret_cast<const float2 *>(ub + po);
                            float xm, x;
                            if (const_tree) {
                                // sU[c] = (Ui, D')
                                xm = (dm.y - um.x * xs) / um.y;
                                xm = mo == dummy_off ? xs : xm;
                                x = (dp.y - up.x * xm) / up.y;
                            } else {
                                xm = (dm.y - um.x * xs) / dm.x;
                                xm = mo == dummy_off ? xs : xm;
                                x = (dp.y - up.x * xm) / dp.x;
                            }
                            reinterpret_cast<float2 *>(dxb + po)->y = x;
                        };
                        if (const_tree)
                            run_list(Lf, idle_fold4, fold_const);
                        else
                            run_list(Lf, idle_fold, fold_task);
                        if (tid == 0 && (fl[0] & TCX_COMPLEX)) {
                            // the root compartment's update, left out of the compartment phase: only
                            // the root fold reads the root's diagonal and right-hand side
                            float Dn = Ds[0], Vn = V[0];
                            complex_update(0, 0, sF, t, false, s, Dn, Vn);
                            sDX(0) = make_float2(Dn, Vn);
                        }
                        lds_barrier<TPB>();
                        if (wv == 0) {
                            // the junction is the serial part of the step; without a channel
                            // wave, wave 0 takes it at raised issue priority among the waves of
                            // its SIMD (config 4 about 1.5 % faster; with a channel wave, which
                            // it would hold back, config 2 was slower)
                            if constexpr (!SPLIT) __builtin_amdgcn_s_setprio(3);
                            if (const_tree)
                                run_list(Ljf, idle_fold4, fold_const);
                            else
                                run_list(Ljf, idle_fold, fold_task);
                            if (lane == 0) {
                                // the root: all its children (positions 1 .. nr) in order, loads
                                // issued four at a time, then its solve
                                const int nr = cell.root_children;
                                float2 r0 = sDX(0);
                                int c = 1;
                                for (; c + 3 <= nr; c += 4) {
                                    float2 a[4], u[4];
    #pragma unroll
                                    for (int k = 0; k < 4; k++) {
                                        a[k] = sDX(c + k);
                                        u[k] = sU(c + k);
                                    }
    #pragma unroll
                                    for (int k = 0; k < 4; k++) {
                                        const float r = const_tree ? a[k].x : u[k].y / a[k].x;
                                        r0.x = r0.x - r * u[k].x;
                                        r0.y = r0.y - r * a[k].y;
                                    }
                                }
                                for (; c <= nr; c++) {
                                    const float2 a = sDX(c), u = sU(c);
                                    const float r = const_tree ? a.x : u.y / a.x;
                                    r0.x = r0.x - r * u.x;
                                    r0.y = r0.y - r * a.y;
                                }
                                r0.y = r0.y / r0.x;
                                sDX(0) = r0;
                            }
                            run_list(Ljb, idle_bwd, bwd_task);
                            if constexpr (!SPLIT) __builtin_amdgcn_s_setprio(0);
                        }
                        lds_barrier<TPB>();
                        run_list(Lb, idle_bwd, bwd_task);
                        lds_barrier<TPB>();
                    }
    #pragma unroll
                    for (int j = 0; j < K; j++) {
                        const int c = tid + TPB * j;
                        if (slot_any[j] && c < N) {
                            const float Vn = init ? V[j] : sDX(posof(j)).y;
                            if (fl[j] & TCX_COMPLEX) {
                                const TreeComplex &cx = CX[fl[j] & 0xFFFFu];
                                // a recorded compartment is complex: its trace column is in its
                                // record; the trace row of step s holds the voltage the step
                                // started from
                                if (!init && cx.k.rec >= 0) trace_at(T, cx.k.rec, s - 1 - row0) = V[j];
                                const float vthr = cx.k.vthr;
                                if (V[j] < vthr && vthr < Vn) {
                                    log_event(T, cx.k.ev, s);
                                    if constexpr (NET) send_spike(T, c0 + c, (int)((s + 3) & 1) ^ 1);
                                }
                            }
                            V[j] = Vn;
                            if (c < NV) sVn[c] = Vn;
                        }
                    }
                }
                // the next step reads voltages and open fractions written above by other lanes
                if constexpr (SPLIT) {
                    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
                } else {
                    lds_barrier<TPB>();
                }
                cur = nxt;
                fcur ^= 1;
            }
        }
        if (cable) {
#pragma unroll
            for (int j = 0; j < K; j++) {
                const int c = tid + TPB * j;
                if (c < N) T.V[c0 + c] = V[j];
            }
        }
        for (int i = tid; i < NG; i += NT) gq_g[gq0 + i] = sQ[i];
    }
    // the last workgroup out resets the queue for the next launch
    if (tid == 0) {
        __threadfence();
        if (atomicAdd(&queue[1], 1) == (int)gridDim.x - 1) {
            queue[0] = 0;
            queue[1] = 0;
        }
    }
}

#define EDENHIP_TREE_INST(TPB, SPLIT, NET)                                                                         \
    template __global__ void tree_kernel<TPB, SPLIT, NET>(                                                         \
        DevTables, const DGate *, int, const DSynType *, int, const TreeCell *, const int32_t *, int32_t, int32_t *,  \
        const uint32_t *, const int32_t *, const float2 *, const uint32_t *, const uint32_t *, float *,             \
        const uint32_t *, const uint32_t *, const uint32_t *, const TreeComplex *, int32_t, int32_t, int32_t, int32_t, \
        int32_t, int32_t, int32_t, int64_t, int64_t, int64_t, float);
EDENHIP_TREE_INST(64, true, false)
EDENHIP_TREE_INST(64, false, false)
EDENHIP_TREE_INST(128, true, false)
EDENHIP_TREE_INST(128, false, false)
EDENHIP_TREE_INST(256, true, false)
EDENHIP_TREE_INST(256, false, false)
EDENHIP_TREE_INST(512, true, false)
EDENHIP_TREE_INST(512, false, false)
EDENHIP_TREE_INST(64, false, true)
EDENHIP_TREE_INST(128, false, true)
EDENHIP_TREE_INST(256, false, true)
EDENHIP_TREE_INST(512, false, true)

const void *tree_kernel_ptr(int tpb, bool split, bool net) {
    if (net) {
        if (split) return nullptr;
        switch (tpb) {
        case 64: return reinterpret_cast<const void *>(&tree_kernel<64, false, true>);
        case 128: return reinterpret_cast<const void *>(&tree_kernel<128, false, true>);
        case 256: return reinterpret_cast<const void *>(&tree_kernel<256, false, true>);
        case 512: return reinterpret_cast<const void *>(&tree_kernel<512, false, true>);
        default: return nullptr;
        }
    }
    if (split) {
        switch (tpb) {
        case 64: return reinterpret_cast<const void *>(&tree_kernel<64, true, false>);
        case 128: return reinterpret_cast<const void *>(&tree_kernel<128, true, false>);
        case 256: return reinterpret_cast<const void *>(&tree_kernel<256, true, false>);
        case 512: return reinterpret_cast<const void *>(&tree_kernel<512, true, false>);
        default: return nullptr;
        }
    }
    switch (tpb) {
    case 64: return reinterpret_cast<const void *>(&tree_kernel<64, false, false>);
    case 128: return reinterpret_cast<const void *>(&tree_kernel<128, false, false>);
    case 256: return reinterpret_cast<const void *>(&tree_kernel<256, false, false>);
    case 512: return reinterpret_cast<const void *>(&tree_kernel<512, false, false>);
    default: return nullptr;
    }
}

// Channels of compartments beyond the compartment kernels' slots (more than NCH channel
// distributions or NQ gate variables), one lane per compartment, before the step's
// compartment kernels: each channel in distribution order takes its open fraction from its
// gates' present values (each gate `instances` times, conductance_scaling = 1), the gates move
// to their next values, and the channel's current and conductance join the totals
// (eden/Eden.cpp:5679-5898), as channels_step does for the others.
__global__ void __launch_bounds__(64) ext_channels_kernel(DevTables T, const DGate *__restrict__ gates, ExtChanArgs X,
                                                          int64_t step, float dt) {
    ext_channels(T, gates, X, step, dt);
}

const void *ext_channels_kernel_ptr() { return reinterpret_cast<const void *>(&ext_channels_kernel); }
const void *point_kernel_ptr() { return reinterpret_cast<const void *>(&point_kernel); }
const void *point_quad_kernel_ptr() { return reinterpret_cast<const void *>(&point_quad_kernel); }

}  // namespace edenhip
