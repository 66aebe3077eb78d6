// lems_kernel.hip -- synapse kernel template for generic LEMS synapse types, compiled at
// context creation with hiprtc together with the generated functions (edenhip_lems_synapses:
// `lems_syn`, emitted by eden_amd.lems), the counterpart of the reference's per-cell-type C
// emission and gcc step (eden/Eden.cpp:2339-2995, 3561-3740, 6992-7216).
//
// One lane per compartment that has chemical synapses; per step, for each of its synapse
// instances in list order: the trigger read and clear and the single-slot delay rule
// (eden/Eden.cpp:3636-3707, the same code as synapses_step), then the instance's LEMS step
// (lems_syn, given the voltage at the start of the step of the instance's peer compartment
// for gap junctions), its exposures times the connection weight aggregated per synapse type,
// the types' aggregates summed in order (eden/Eden.cpp:3721-3735). Every compartment's V is
// the start-of-step value here: this kernel runs before the step's compartment kernels. The totals go to syn_ext,
// where the compartment kernels of the same step read them in place of their own synapse
// code. State is updated in place (the generated code reads an instance's whole present
// state before writing).
//
// lems_cell_kernel runs LEMS point neurons (the reference's artificial cells,
// eden/Eden.cpp:6603-6910): one lane per cell; per step, the external current (the cell's
// synapse currents from syn_ext, then its pulse inputs), the generated `lems_cell` step, the
// compartment's V set to the membrane-voltage state variable, and a spike sent to the
// recipients and logged as a physical cell's is.
//
// The generated source (`lems_syn` and `lems_cell`) is inserted where
// lems_synapse_functions.h is included.
#define EDENHIP_LEMS_RATES 1
#include "device_model.hpp"

#define EXP edenhip::exp_f
#define LOG edenhip::log_f
#define POW edenhip::pow_f
#define SIN(x) ((float)sin((double)(x)))
#define COS(x) ((float)cos((double)(x)))
#define TAN(x) ((float)tan((double)(x)))
#define SINH(x) ((float)sinh((double)(x)))
#define COSH(x) ((float)cosh((double)(x)))
#define TANH(x) ((float)tanh((double)(x)))
__device__ __forceinline__ float stepf(float x) { return x < 0.0f ? 0.0f : 1.0f; }
// random(x) of a point neuron's generated step: randof stamped with the cell's RNG seed, the
// step, the call's number and, as the instance, the reference's NOT_AN_INSTANCE marker of
// code outside a table loop (~0xFee1600dLL, eden/Eden.cpp:3255-3256)
__device__ __forceinline__ float lems_randof(float x, int seed, long long step, int n) {
    return edenhip::randof(x, seed, (int64_t)~0xFee1600dLL, (int64_t)step, n);
}

#include "lems_synapse_functions.h"

extern "C" __global__ void __launch_bounds__(64)
lems_syn_kernel(edenhip::DevTables T, edenhip::LemsArgs L, int64_t step, float dt) {
    using namespace edenhip;
    const int32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= L.n_comps) return;
    const int32_t g = L.comps[idx];
    const bool init = step <= 0;
    // time_f32 of the step, from the reference's double accumulation
    const float t = (float)(step >= 1 ? T.t_row[step - 1] : 0.0);
    const int now = (int)((step + 3) & 1);
    const float V = T.V[g];
    const int32_t base = T.syn_base[g], cnt = T.syn_cnt[g];
    int32_t *trig_now = T.trig[now];
    float I_syn = 0.0f, G_syn = 0.0f, I_agg = 0.0f, G_agg = 0.0f;
    float I_inp = 0.0f, G_inp = 0.0f;   // LEMS inputs: i * weight per instance (eden/Eden.cpp:3929)
    int32_t cur = -1;
    for (int32_t k = 0; k < cnt; k++) {
        const int32_t s = base + 64 * k;
        const int32_t ty = T.syn_type[s];
        const bool input = L.flags && (L.flags[ty] & 1);
        if (ty != cur && !input) {
            if (cur >= 0) { I_syn += I_agg; G_syn += G_agg; }
            I_agg = 0.0f; G_agg = 0.0f; cur = ty;
        }
        const float weight = T.syn_weight[s];
        bool spike_now = false;
        const int32_t tp = L.tpos ? L.tpos[s] : -1;
        if (tp >= 0) {
            // a firing-synapse input: its spike list's times before the end of the step fire it
            // in the same step (eden/Eden.cpp:3780-3788); the list ends with a sentinel
            // (events of the initial steps are ignored by the generated code, so not raised)
            int32_t pos = tp;
            bool fired = false;
            while (L.ttimes[pos] < t + dt) { fired = true; pos++; }
            if (!init) {
                L.tpos[s] = pos;
                spike_now = fired;
            }
        } else if (L.pspk && L.pspk[s] >= 0) {
            // a Poisson firing-synapse input's spiker (poissonFiringSynapse,
            // transientPoissonFiringSynapse): tsince' = 1; when tsince > isi, tsince = 0, a new
            // interval and a spike, the synapse's event of the same step
            const int32_t k = L.pspk[s];
            const float averageIsi = L.pk_isi[k];
            const int32_t seed = (int32_t)T.cell_seed[T.comp_cell[g]], inst = L.pk_inst[k], roff = L.pk_roff[k];
            const bool transient = L.pk_transient[k] != 0;
            const float2 win = L.pk_window[k];
            const float2 st = L.pk_state[k];
            float tn, in;
            if (init) {
                tn = 0.0f;
                const float nextIsi = (-averageIsi) * log_f(1.0f - randof(1.0f, seed, inst, step, roff + 0));
                in = transient ? nextIsi + win.x : nextIsi;
            } else {
                tn = st.x + dt * 1.0f;
                in = st.y;
                if (st.x > st.y) {
                    tn = 0.0f;
                    const float nextIsi = (-averageIsi) * log_f(1.0f - randof(1.0f, seed, inst, step, roff + 1));
                    // after the window, the next interval is pushed 1e9 hours (3.6e18 us) away
                    in = transient ? nextIsi + stepf(((t + nextIsi) - (win.x + win.y)) / win.y) * 3.6e18f : nextIsi;
                    spike_now = true;
                }
            }
            L.pk_state[k] = make_float2(tn, in);
        } else {
            bool flag = false;
            if (!init) {
                flag = trig_now[s] != 0;
                if (flag) trig_now[s] = 0;
            }
            const float delay = T.syn_delay[s];
            const float next_spike = T.syn_next[s];
            float nn = next_spike;
            if (!init) {
                if (t <= next_spike && next_spike < t + dt) spike_now = true;
                if (flag) {
                    const float fresh = t + delay;
                    if (t <= fresh && fresh < t + dt) spike_now = true;
                    if (nn < t + dt && nn < fresh) nn = fresh;
                }
            }
            if (nn != next_spike) T.syn_next[s] = nn;
        }
        if (spike_now && T.log_syn) {
            const uint32_t slot = atomicAdd(T.n_syn_events, 1u);
            if (slot < T.max_syn_events) T.syn_events[slot] = (step << 32) | (int64_t)T.syn_orig[s];
        }
        float i, gx;
        float *S = L.state + s;
        const int32_t pc = L.peer[s];
        const float vpeer = pc >= 0 ? T.V[pc] : V;
        const float *P = L.params + (L.poff ? L.poff[s] : L.param_begin[ty]);
        lems_syn(ty, P, S, S, (long)L.stride, V, vpeer, t, init ? 1 : 0, spike_now ? 1 : 0, dt, &i, &gx);
        if (input) {
            I_inp += i * weight;
            G_inp += gx * weight;
        } else {
            I_agg += i * weight;
            G_agg += gx * weight;
        }
    }
    I_syn += I_agg;
    G_syn += G_agg;
    L.syn_ext[g] = make_float2(I_syn, G_syn);
    if (L.inp_ext) L.inp_ext[g] = make_float2(I_inp, G_inp);
}

extern "C" __global__ void __launch_bounds__(64)
lems_cell_kernel(edenhip::DevTables T, edenhip::LemsCellArgs A, int64_t step, float dt) {
    using namespace edenhip;
    const int32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= A.n) return;
    const int32_t g = A.comps[idx], ty = A.type[idx];
    const bool init = step <= 0;
    const float t = (float)(step >= 1 ? T.t_row[step - 1] : 0.0);
    const int nxt = (int)((step + 3) & 1) ^ 1;
    // external current: synapses, then inputs (eden/Eden.cpp:6856-6880)
    float I_syn = 0.0f;
    if (T.syn_ext && T.syn_cnt[g] > 0) I_syn = T.syn_ext[g].x;
    float I_in = 0.0f;
    const int32_t pb = T.pulse_begin[g], pe = T.pulse_begin[g + 1];
    if (pe > pb) {
        float Ip = 0.0f;
        for (int32_t p = pb; p < pe; p++)
            if (T.pulse_start[p] <= t && t <= T.pulse_end[p]) Ip += T.pulse_amp[p];
        I_in += Ip;
        I_in += 0.0f;
    }
    if (T.inp_ext && T.syn_cnt[g] > 0) I_in += T.inp_ext[g].x;   // LEMS inputs, after the native ones
    const float ext = I_syn + I_in;
    float *S = A.state + idx;
    int spike = 0;
    const float *P = A.params + (A.poff ? A.poff[idx] : A.param_begin[ty]);
    lems_cell(ty, P, S, S, (long)A.n, t, dt, init ? 1 : 0, ext, (int)T.cell_seed[T.comp_cell[g]], (long long)step,
              &spike);
    const int32_t vs = A.v_state[ty];
    if (vs >= 0) T.V[g] = S[(int64_t)vs * A.n];
    if (spike) {
        send_spike(T, g, nxt);
        log_event(T, T.ev_col[g], step);
    }
}

// Channels of the compartments run outside the compartment kernels, here with the user rate
// ComponentTypes' generated lems_rate (gates of form RATE_LEMS); the statically compiled
// ext_channels_kernel runs the same code for models without them.
extern "C" __global__ void __launch_bounds__(64)
ext_channels_kernel_rtc(edenhip::DevTables T, const edenhip::DGate *__restrict__ gates, edenhip::ExtChanArgs X,
                        int64_t step, float dt) {
    edenhip::ext_channels(T, gates, X, step, dt);
}
