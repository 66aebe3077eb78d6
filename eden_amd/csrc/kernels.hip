// kernels.hip -- gfx950 kernels of the per-timestep loop.
//
//  point_kernel   one lane per single-compartment cell; time loop inside the launch, state in
//                 registers (a launch may cover one step, as in the network path).
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

__device__ __forceinline__ void log_event(const DevTables &T, int32_t col, int64_t step) {
    // spike detected while computing step `step` is logged at row `step` (the next row)
    if (col >= 0 && step < T.total_rows) {
        uint32_t slot = atomicAdd(T.n_events, 1u);
        if (slot < T.max_events) T.events[slot] = ((int64_t)step << 24) | (int64_t)col;
    }
}

__device__ __forceinline__ void send_spike(const DevTables &T, int64_t gidx, int nxt) {
    const int32_t b = T.spk_begin[gidx], e = T.spk_begin[gidx + 1];
    int32_t *trig = T.trig[nxt];
    for (int32_t k = b; k < e; k++) trig[T.spk_target[k]] = 1;   // idempotent flag (EDEN: atomic OR 1)
}

__global__ void __launch_bounds__(256)
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
    const bool has_syn = T.syn_begin[g + 1] > T.syn_begin[g];
    const bool has_poi = T.poi_begin[g + 1] > T.poi_begin[g];
    const bool has_spk = T.spk_begin[g + 1] > T.spk_begin[g];

    for (int64_t s = s0; s < s1; s++) {
        const bool init = s <= 0;
        const float t = init ? 0.0f : T.t_step[s - 1];
        const int now = (int)((s + 3) & 1), nxt = now ^ 1;
        if (!init && k.rec >= 0) T.trace[(s - 1 - row0) * T.n_trace_cols + k.rec] = V;
        float I_ch, G_ch;
        channels_step(gates, k, V, q, init, dt, I_ch, G_ch);
        float I_syn = 0.0f, G_syn = 0.0f;
        if (has_syn) synapses_step(T, stypes, g, V, init, t, dt, now, I_syn, G_syn);
        float I_in = pulses_current(T, k, g, t), G_in = 0.0f;
        if (has_poi) poisson_step(T, stypes, g, V, init, s, dt, I_in, G_in);
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
    }
    T.V[g] = V;
#pragma unroll
    for (int s = 0; s < NQ; s++) T.q[g * NQ + s] = q[s];
}

// LDS-only workgroup barrier: waits for this wave's LDS operations, then (for more than one
// wave) the hardware barrier. Outstanding global loads/stores (the trace write, the next
// step's time prefetch) are not drained, unlike __syncthreads().
template <int TPB>
__device__ __forceinline__ void lds_barrier() {
    if constexpr (TPB <= 64) {
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    } else {
        asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    }
}

// One workgroup per multi-compartment cell, cells popped from a queue in decreasing size.
// Thread t holds compartments t, t + TPB, t + 2 TPB, t + 3 TPB in registers. The Hines solve
// works in LDS on breadth-first positions (tree_schedule.hpp): sDX[p] = (diagonal, rhs) and
// sU[p] = (Ui, Uj) of the edge from p to its parent, driven by one task word per thread per
// round.
template <int TPB>
__global__ void __launch_bounds__(TPB)
tree_kernel(DevTables T, const DGate *__restrict__ gates_g, int n_gates, const DSynType *__restrict__ stypes_g,
            int n_stypes, const TreeCell *__restrict__ cells, const int32_t *__restrict__ task_order, int32_t n_tasks,
            int32_t *__restrict__ queue, const uint32_t *__restrict__ words_g, const int32_t *__restrict__ pos_g,
            const float2 *__restrict__ u_g, int32_t nmax, int64_t s0, int64_t s1, int64_t row0, float dt) {
    constexpr int K = 4;
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ DGate gates[MAX_CHAN_TYPES * GATES_PER_TYPE];
    __shared__ DSynType stypes[64];
    __shared__ int32_t s_task;
    float2 *sDX = reinterpret_cast<float2 *>(smem);
    float2 *sU = sDX + nmax;
    uint32_t *sW = reinterpret_cast<uint32_t *>(sU + nmax);
    const int tid = threadIdx.x;
    for (int i = tid; i < n_gates; i += TPB) gates[i] = gates_g[i];
    for (int i = tid; i < n_stypes; i += TPB) stypes[i] = stypes_g[i];

    for (;;) {
        __syncthreads();
        if (tid == 0) s_task = atomicAdd(&queue[0], 1);
        __syncthreads();
        const int32_t task = s_task;
        if (task >= n_tasks) break;
        const TreeCell cell = cells[task_order[task]];
        const int32_t N = cell.n;
        const int64_t c0 = cell.comp_begin;
        const int RF = cell.rounds_fwd, RB = cell.rounds_bwd;
        const int nw = (RF + RB) * TPB;
        for (int i = tid; i < nw; i += TPB) sW[i] = words_g[cell.words_off + i];
        for (int i = tid; i < N; i += TPB) sU[i] = u_g[c0 + i];

        CompConst kc[K];
        float V[K];
        float q[K][NQ];
        int pos[K];
        bool has_poi[K];
#pragma unroll
        for (int j = 0; j < K; j++) {
            const int c = tid + TPB * j;
            has_poi[j] = false;
            V[j] = 0.0f;
            pos[j] = 0;
#pragma unroll
            for (int s = 0; s < NQ; s++) q[j][s] = 0.0f;
            if (c < N) {
                load_comp_const(T, c0 + c, kc[j]);
                V[j] = T.V[c0 + c];
                pos[j] = pos_g[c0 + c];
#pragma unroll
                for (int s = 0; s < NQ; s++) q[j][s] = T.q[(c0 + c) * NQ + s];
                has_poi[j] = T.poi_begin[c0 + c + 1] > T.poi_begin[c0 + c];
            }
        }
        __syncthreads();

        float t_next = (s0 >= 1) ? T.t_step[s0 - 1] : 0.0f;
        for (int64_t s = s0; s < s1; s++) {
            const bool init = s <= 0;
            const float t = t_next;
            if (s + 1 < s1) t_next = (s + 1 >= 1) ? T.t_step[s] : 0.0f;   // prefetch the next step's time
#pragma unroll
            for (int j = 0; j < K; j++) {
                const int c = tid + TPB * j;
                if (c < N) {
                    if (!init && kc[j].rec >= 0) T.trace[(s - 1 - row0) * T.n_trace_cols + kc[j].rec] = V[j];
                    float I_ch, G_ch;
                    channels_step(gates, kc[j], V[j], q[j], init, dt, I_ch, G_ch);
                    float I_in = pulses_current(T, kc[j], c0 + c, t), G_in = 0.0f;
                    if (has_poi[j]) poisson_step(T, stypes, c0 + c, V[j], init, s, dt, I_in, G_in);
                    const float I_syn = 0.0f, G_syn = 0.0f;
                    const float I_int = I_ch + I_in + I_syn;
                    const float G_int = G_ch + G_in + G_syn;
                    const float Dd = 1 * (G_int / kc[j].C) + kc[j].invrc;
                    const float Vn = init ? V[j] : V[j] + (dt * (I_int + ((V[j] * G_int))) / kc[j].C);
                    sDX[pos[j]] = make_float2(1 + Dd * dt, Vn);
                }
            }
            if (!init) {
                lds_barrier<TPB>();
                // forward elimination, rounds of independent gathers
                uint32_t w = RF > 0 ? sW[tid] : TASK_IDLE;
                for (int r = 0; r < RF; r++) {
                    const uint32_t wn = (r + 1 < RF) ? sW[(r + 1) * TPB + tid] : TASK_IDLE;
                    if (w != TASK_IDLE) {
                        const int p = w & 0x7FF, fc = (w >> 11) & 0x7FF, nc = w >> 22;
                        float2 dx = sDX[p];
                        for (int i = 0; i < nc; i++) {
                            const float2 dxc = sDX[fc + i];
                            const float2 u = sU[fc + i];
                            const float ratio = u.y / dxc.x;
                            dx.x -= ratio * u.x;
                            dx.y -= ratio * dxc.y;
                        }
                        sDX[p] = dx;
                    }
                    lds_barrier<TPB>();
                    w = wn;
                }
                if (tid == 0) {
                    float2 r0 = sDX[0];
                    r0.y = r0.y / r0.x;
                    sDX[0] = r0;
                }
                lds_barrier<TPB>();
                // back-substitution, rounds by depth
                const uint32_t *sWb = sW + RF * TPB;
                w = RB > 0 ? sWb[tid] : TASK_IDLE;
                for (int r = 0; r < RB; r++) {
                    const uint32_t wn = (r + 1 < RB) ? sWb[(r + 1) * TPB + tid] : TASK_IDLE;
                    if (w != TASK_IDLE) {
                        const int p = w & 0x7FF, pp = (w >> 11) & 0x7FF;
                        float2 dx = sDX[p];
                        const float xp = sDX[pp].y;
                        dx.y = (dx.y - sU[p].x * xp) / dx.x;
                        sDX[p] = dx;
                    }
                    lds_barrier<TPB>();
                    w = wn;
                }
            }
#pragma unroll
            for (int j = 0; j < K; j++) {
                const int c = tid + TPB * j;
                if (c < N) {
                    const float Vn = init ? V[j] : sDX[pos[j]].y;
                    if (V[j] < kc[j].vthr && kc[j].vthr < Vn) log_event(T, kc[j].ev, s);
                    V[j] = Vn;
                }
            }
        }
#pragma unroll
        for (int j = 0; j < K; j++) {
            const int c = tid + TPB * j;
            if (c < N) {
                T.V[c0 + c] = V[j];
#pragma unroll
                for (int s = 0; s < NQ; s++) T.q[(c0 + c) * NQ + s] = q[j][s];
            }
        }
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

#define EDENHIP_TREE_INST(TPB)                                                                                     \
    template __global__ void tree_kernel<TPB>(DevTables, const DGate *, int, const DSynType *, int, const TreeCell *, \
                                              const int32_t *, int32_t, int32_t *, const uint32_t *, const int32_t *,  \
                                              const float2 *, int32_t, int64_t, int64_t, int64_t, float);
EDENHIP_TREE_INST(64)
EDENHIP_TREE_INST(128)
EDENHIP_TREE_INST(256)
EDENHIP_TREE_INST(512)

const void *tree_kernel_ptr(int tpb) {
    switch (tpb) {
    case 64: return reinterpret_cast<const void *>(&tree_kernel<64>);
    case 128: return reinterpret_cast<const void *>(&tree_kernel<128>);
    case 256: return reinterpret_cast<const void *>(&tree_kernel<256>);
    case 512: return reinterpret_cast<const void *>(&tree_kernel<512>);
    default: return nullptr;
    }
}

const void *point_kernel_ptr() { return reinterpret_cast<const void *>(&point_kernel); }

}  // namespace edenhip
