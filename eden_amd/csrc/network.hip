// network.hip -- windowed time loop for networks of one-compartment cells.
//
// The reference advances every synapse every step and carries spikes to the next step
// through trigger flags (eden/Eden.cpp:3636-3740, 3946-3995): a spike at step n sets the
// target's trigger, read at step n + 1, where the single-slot NextSpike rule either keeps a
// pending arrival or replaces it with t + delay; the synapse receives its event at the step
// whose window [t, t + dt) holds NextSpike. With every delay at least D steps, a spike
// cannot reach any cell within D - 1 steps, so cells can be advanced W < D steps at a time
// without seeing each other:
//
//   net_window_kernel   W steps of every cell (one lane per cell, synapses in the
//                       block-interleaved slots); a synapse's events come from a device
//                       ring buffer of per-step bit masks instead of its trigger flag,
//                       NextSpike and delay; the spikes of the window are kept as a bit mask
//                       per cell and the spiking cells are compacted into a list (wavefront
//                       ballot + prefix count, one atomic per wave)
//   net_deliver_kernel  for every spiking cell and target synapse, in step order: the
//                       trigger read at step n + 1 and the NextSpike rule with the
//                       reference's float comparisons; the steps whose windows hold the
//                       new NextSpike get the synapse's bit in the ring
//
// One case cannot be decided inside a window: a pending arrival that lies in two
// consecutive step windows (their float bounds overlap by an ulp) and is replaced at the
// first of them, so the reference drops the second delivery. The deliver kernel's check
// pass reports it (and any arrival inside the window just run); the host then runs that
// window again one step at a time, from the state copy the window did not write.
#include <hip/hip_runtime.h>

#include "device_model.hpp"
#include "kernels.hpp"

namespace edenhip {

namespace {

__device__ __forceinline__ void log_event_n(const DevTables &T, int32_t col, int64_t step) {
    if (col >= 0 && step < T.total_rows) {
        uint32_t slot = atomicAdd(T.n_events, 1u);
        if (slot < T.max_events) T.events[slot] = ((int64_t)step << 24) | (int64_t)col;
    }
}

// A cell's chemical synapses for one step: events from the ring words, expTwo dynamics,
// per-type aggregation in list order (eden/Eden.cpp:3636-3740): each type run summed from
// 0, the run sums added in order. State is read from (rA, rB) and written to (wA, wB).
constexpr int NET_CH = 8;

__device__ __forceinline__ void net_synapses(const DevTables &T, const DSynType *stypes, const uint32_t (&runs)[4],
                                             int32_t base, int32_t cnt, const uint32_t *wd, float V, bool init,
                                             float dt, int64_t step, const float *rA, const float *rB, float *wA,
                                             float *wB, uint32_t &n_ev, float &I_syn, float &G_syn) {
    I_syn = 0.0f;
    G_syn = 0.0f;
    int32_t k = 0;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int32_t e = (int32_t)(runs[r] & 0xFFFFu);
        if (e == 0) break;
        const DSynType st = stypes[runs[r] >> 16];
        float I_agg = 0.0f, G_agg = 0.0f;
        for (; k < e; k += NET_CH) {
            float a[NET_CH], b[NET_CH], w[NET_CH];
#pragma unroll
            for (int u = 0; u < NET_CH; u++) {
                const int32_t s = base + 64 * min(k + u, e - 1);
                a[u] = rA[s];
                b[u] = rB[s];
                w[u] = T.syn_weight[s];
            }
#pragma unroll
            for (int u = 0; u < NET_CH; u++) {
                if (k + u >= e) break;
                const int32_t kk = k + u;
                const int32_t s = base + 64 * kk;
                const bool ev = !init && ((wd[kk >> 5] >> (kk & 31)) & 1u);
                if (ev) {
                    n_ev++;
                    if (T.log_syn) {
                        const uint32_t slot = atomicAdd(T.n_syn_events, 1u);
                        if (slot < T.max_syn_events) T.syn_events[slot] = (step << 32) | (int64_t)T.syn_orig[s];
                    }
                }
                float A = a[u], B = b[u], i, g;
                exptwo_step(st, V, A, B, init, ev, dt, i, g);
                wA[s] = A;
                wB[s] = B;
                I_agg += i * w[u];
                G_agg += g * w[u];
            }
        }
        I_syn += I_agg;
        G_syn += G_agg;
        k = e;   // the chunked loop may have stepped past the run's end
    }
    (void)cnt;
}

}  // namespace

__global__ void __launch_bounds__(64)
net_window_kernel(DevTables T, NetTables Nt, const DGate *__restrict__ gates_g, int n_gates,
                  const DSynType *__restrict__ stypes_g, int n_stypes, const int32_t *__restrict__ comps, int32_t n,
                  int64_t s0, int64_t s1, int64_t row0, float dt, int cur) {
    extern __shared__ __align__(16) unsigned char net_smem[];
    DGate *gates = reinterpret_cast<DGate *>(net_smem);   // [n_gates], dynamic LDS
    __shared__ DSynType stypes[64];
    const int lane = threadIdx.x;
    for (int i = lane; i < n_gates; i += 64) gates[i] = gates_g[i];
    for (int i = lane; i < n_stypes; i += 64) stypes[i] = stypes_g[i];
    __syncthreads();
    const int nxt = cur ^ 1;
    const int32_t idx = blockIdx.x * 64 + lane;
    const bool valid = idx < n;
    uint64_t mask = 0;
    int64_t g = 0;
    if (valid) {
        g = comps[idx];
        CompConst k;
        load_comp_const(T, g, k);
        float V = Nt.V[cur][g];
        float q[NQ];
#pragma unroll
        for (int s = 0; s < NQ; s++) q[s] = Nt.q[cur][g * NQ + s];
        const int32_t syn_base = T.syn_base[g], syn_cnt = T.syn_cnt[g];
        uint32_t runs[4];
#pragma unroll
        for (int r = 0; r < 4; r++) runs[r] = Nt.runs[g * 4 + r];
        const int nwc = (syn_cnt + 31) >> 5;
        const int64_t rbase = ((int64_t)blockIdx.x * Nt.nw) * 64 + lane;
        const bool has_poi = T.poi_begin[g + 1] > T.poi_begin[g];
        uint32_t n_ev = 0;
        double time = (s0 >= 1) ? T.t_row[s0 - 1] : 0.0;
        for (int64_t s = s0; s < s1; s++) {
            const bool init = s <= 0;
            const float t = (float)time;
            const int rb = s == s0 ? cur : nxt;
            if (!init && k.rec >= 0) trace_at(T, k.rec, s - 1 - row0) = V;
            float I_ch, G_ch;
            channels_step(gates, k, V, q, init, dt, I_ch, G_ch);
            float I_syn = 0.0f, G_syn = 0.0f;
            if (syn_cnt > 0) {
                uint32_t wd[8] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u};
                if (!init) {
                    const uint32_t *rw = Nt.ring + (s & Nt.ring_mask) * Nt.ring_stride + rbase;
#pragma unroll
                    for (int w = 0; w < 8; w++)
                        if (w < nwc) wd[w] = rw[w * 64];
                }
                net_synapses(T, stypes, runs, syn_base, syn_cnt, wd, V, init, dt, s, Nt.gA[rb], Nt.gB[rb],
                             Nt.gA[nxt], Nt.gB[nxt], n_ev, I_syn, G_syn);
            }
            float I_in = pulses_current(T, k, g, t), G_in = 0.0f;
            if (has_poi) {
                const PoiState rd{Nt.poi_tsince[rb], Nt.poi_isi[rb], Nt.poi_A[rb], Nt.poi_B[rb]};
                const PoiState wr{Nt.poi_tsince[nxt], Nt.poi_isi[nxt], Nt.poi_A[nxt], Nt.poi_B[nxt]};
                poisson_step(T, stypes, g, V, init, s, dt, I_in, G_in, rd, wr);
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
                if (s - s0 < 64) mask |= 1ull << (s - s0);
                log_event_n(T, k.ev, s);
            }
            V = Vn;
            if (!init) time += (double)dt;
        }
        Nt.V[nxt][g] = V;
#pragma unroll
        for (int s = 0; s < NQ; s++) Nt.q[nxt][g * NQ + s] = q[s];
        if (n_ev) atomicAdd(Nt.n_deliv, (unsigned long long)n_ev);
        if (mask) Nt.spike_mask[g] = mask;
    }
    // spiking cells of this wave: ballot, prefix count, one atomic
    const bool spk = valid && mask != 0ull;
    const uint64_t bal = __ballot(spk);
    if (bal) {
        const int first = __ffsll((long long)bal) - 1;
        uint32_t base = 0;
        if (lane == first) base = atomicAdd(Nt.spk_count, (uint32_t)__popcll(bal));
        base = __shfl(base, first);
        const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(bal >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)bal, 0u));
        if (spk) Nt.spk_list[base + below] = (int32_t)g;
    }
}

namespace {

// steps j >= jmin (and <= last) whose window [t_j, t_j + dt) in float holds x
template <class F>
__device__ __forceinline__ void for_deliveries(const DevTables &T, float x, float dt, int64_t m, float tm, int64_t jmin,
                                               int64_t last, F f) {
    if (!(x > -3.0e38f && x < 3.0e38f)) return;
    const int64_t jc = m + (int64_t)floorf((x - tm) / dt);
    for (int64_t j = jc - 2; j <= jc + 2; j++) {
        if (j < jmin || j < 1 || j > last) continue;
        const float tj = (float)T.t_row[j - 1];
        if (tj <= x && x < tj + dt) f(j);
    }
}

}  // namespace

__global__ void __launch_bounds__(64)
net_deliver_kernel(DevTables T, NetTables Nt, int64_t s0, int64_t s1, int64_t last, float dt, int commit) {
    const int lane = threadIdx.x;
    const uint32_t count = *Nt.spk_count;
    const int64_t ring_stride = Nt.ring_stride;
    for (uint32_t i = blockIdx.x; i < count; i += gridDim.x) {
        const int32_t gc = Nt.spk_list[i];
        const uint64_t M = Nt.spike_mask[gc];
        for (int32_t kt = T.spk_begin[gc] + lane; kt < T.spk_begin[gc + 1]; kt += 64) {
            const int32_t slot = T.spk_target[kt];
            const uint32_t rl = Nt.tgt_rw[kt];
            const int64_t rword = (int64_t)(rl >> 5);
            const uint32_t bit = 1u << (rl & 31u);
            float nn = T.syn_next[slot];
            const float delay = T.syn_delay[slot];
            bool bad = false;
            for (uint64_t mm = M; mm; mm &= mm - 1ull) {
                const int b = __ffsll((long long)mm) - 1;
                const int64_t m = s0 + b + 1;     // the trigger is read one step after the spike
                if (m > last) break;
                const float tm = (float)T.t_row[m - 1];
                const float tpd = tm + dt;
                const float fresh = tm + delay;
                if (nn < tpd && nn < fresh) {
                    // the replaced value's delivery after step m (overlapping windows) is dropped
                    for_deliveries(T, nn, dt, m, tm, m + 1, last, [&](int64_t j) {
                        if (j < s1) bad = true;
                        else if (commit) atomicAnd(&Nt.ring[(j & Nt.ring_mask) * ring_stride + rword], ~bit);
                    });
                    nn = fresh;
                    for_deliveries(T, nn, dt, m, tm, m, last, [&](int64_t j) {
                        if (j < s1) bad = true;
                        else if (commit) atomicOr(&Nt.ring[(j & Nt.ring_mask) * ring_stride + rword], bit);
                    });
                }
            }
            if (commit) T.syn_next[slot] = nn;
            if (bad) atomicOr(Nt.conflict, 1);
        }
    }
    if (commit) {
        // the window's ring slots were consumed: clear them for their next use
        const int64_t a = s0 < 1 ? 1 : s0;
        const int64_t n = (s1 > a ? s1 - a : 0) * ring_stride;
        for (int64_t i = (int64_t)blockIdx.x * 64 + lane; i < n; i += (int64_t)gridDim.x * 64) {
            const int64_t j = a + i / ring_stride;
            Nt.ring[(j & Nt.ring_mask) * ring_stride + (i % ring_stride)] = 0u;
        }
    }
}

const void *net_window_kernel_ptr() { return reinterpret_cast<const void *>(&net_window_kernel); }
const void *net_deliver_kernel_ptr() { return reinterpret_cast<const void *>(&net_deliver_kernel); }

}  // namespace edenhip
