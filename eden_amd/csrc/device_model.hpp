// device_model.hpp -- device-side tables and the per-compartment mechanism step shared by
// the point-cell and tree-cell kernels. Built for gfx950 only.
//
// The per-compartment math restates the C that EDEN generates per cell type:
//   gate rate forms            eden/Eden.cpp:5126-5200
//   cnexp update + clipping     eden/Eden.cpp:5236-5259, 5336-5375
//   open fraction, density      eden/Eden.cpp:5679-5721, 5805-5898
//   pulse input                 eden/Eden.cpp:3798-3826
//   synapse trigger/delay rule  eden/Eden.cpp:3636-3707
//   expTwoSynapse dynamics      eden/Eden.cpp:2614-2994, LEMS_CoreComponents.inc.xml:145-180
//   Poisson firing synapse      eden/Eden.cpp:3859-3912, LEMS_CoreComponents.inc.xml:623-672
//   membrane RHS and diagonal   eden/Eden.cpp:6057-6107
// with the same operation order; kernels are built with -ffp-contract=off and exp/log are
// rounded once from double (exp_f, log_f), so results match the CPU path bit for bit
// except in rare exp/log rounding cases.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#ifdef EDENHIP_LEMS_RATES
// a user rate ComponentType's instance k (generated code compiled with hiprtc, defined at
// global scope); only the run-time compiled channel kernel sees gates of this form
__device__ float lems_rate(int k, float v);
#endif

namespace edenhip {

constexpr int NCH = 4;      // channel instances per compartment (register slots)
constexpr int NQ = 4;       // gate state variables per compartment (register slots)
constexpr int GATES_PER_TYPE = 4;
constexpr int MAX_CHAN_TYPES = 64;

enum RateForm : int32_t { RATE_EXP = 0, RATE_EXPLINEAR = 1, RATE_SIGMOID = 2, RATE_FIXED = 3, RATE_LEMS = 4 };

// One gate of a channel type, 32 bytes (two ds_read_b128 from LDS).
struct alignas(16) DGate {
    int32_t forms;   // alpha form | kind << 4 | beta form << 8 | instances << 16 | gates of the type << 24
    float q10;
    float ar, am;
    float as, br;
    float bm, bs;
};

// Per expTwoSynapse type, with the waveform factor precomputed on the host in fp32 with
// the reference's operation order (peakTime, waveformFactor are derived parameters).
struct alignas(16) DSynType {
    float gbase, erev, tau_rise, tau_decay;
    float weight_prop, wf, pad0, pad1;
};

// Flat pointers to every device table (one struct passed by value to kernels).
struct DevTables {
    // compartments
    float *V;               // [NC] membrane voltage (state)
    float *q;               // [NC*NQ] gate states (state)
    const float *C, *A, *invrc, *vthr;
    const uint32_t *chmeta; // [NC] 4 x 8-bit channel type (0xFF = empty)
    const uint32_t *gmeta;  // [NC] gate slot map: slot s: bits 4s..4s+3 = (chan slot | gate idx << 2);
                            //      bits 24..26 = gate count, bits 28..30 = channel count
    const float *chg, *che; // [NC*NCH] channel Gbase (mS/cm2), Erev (mV)
    const int32_t *pulse_begin;   // [NC+1]
    const float *pulse_start, *pulse_end, *pulse_amp;  // start, start+duration, Imax*weight
    const int32_t *rec_col; // [NC] device trace column or -1
    const int32_t *ev_col;  // [NC] event column or -1
    // chemical synapses of one-compartment cells, block-interleaved: the cells of a wave
    // (64 consecutive point cells) store their k-th synapses in 64 consecutive slots, so
    // synapse k of lane l sits at syn_base[comp] + 64 k
    const int32_t *syn_base;      // [NC] first slot, or 0
    const int32_t *syn_cnt;       // [NC] synapses of the compartment
    const int32_t *syn_orig;      // [slots] instance index in the descriptor (for the event log)
    const int32_t *syn_type;
    const float *syn_weight, *syn_delay;
    float *syn_next, *syn_A, *syn_B;   // state
    int32_t *trig[2];                  // trigger flags, double-buffered
    const int32_t *spk_begin;          // [NC+1]
    const int32_t *spk_target;
    // Poisson inputs (CSR over compartments)
    const int32_t *poi_begin;     // [NC+1]
    const float *poi_isi_mean;    // averageIsi = 1 / averageRate (host fp32)
    const float *poi_weight;
    const int32_t *poi_syn_type, *poi_inst, *poi_roff;
    float *poi_tsince, *poi_isi, *poi_A, *poi_B;   // state
    const uint32_t *cell_seed;    // [n_cells]
    const int32_t *comp_cell;     // [NC] owning cell
    // time
    const double *t_row;          // [rows] time (double, as accumulated) at each recorded row
    // outputs
    float *trace;                 // [n_trace_cols][trace_rows]: column-major, a column's rows
                                  // contiguous (see trace_at)
    int32_t n_trace_cols;
    int64_t trace_rows;           // rows of a chunk (the column stride)
    int64_t *events;              // [max_events] packed (row << 24 | col)
    uint32_t *n_events;
    uint32_t max_events;
    int64_t total_rows;
    // optional log of synaptic events: (step << 32 | synapse instance)
    int64_t *syn_events;
    uint32_t *n_syn_events;
    uint32_t max_syn_events;
    int32_t log_syn;
    // generic LEMS synapses: per compartment (I_syn, G_syn) of the present step, written by
    // lems_syn_kernel before the compartment kernels of the step run (null: synapses_step)
    const float2 *syn_ext;
    // LEMS inputs: per compartment (I, G) of the present step, summed after the native inputs
    // (null: none)
    const float2 *inp_ext;
    const float2 *chx_ext;  // [NC] (I, G) of the channels of a compartment whose channels run in
                            // ext_channels_kernel (chmeta == CHM_EXTERNAL), or null
};

// Arguments of lems_syn_kernel (the generic LEMS synapse kernel, compiled at run time from
// lems_kernel.hip): shared by that compilation and the host code that launches it.
struct LemsArgs {
    const int32_t *comps;         // compartments with synapses
    int32_t n_comps;
    const float *params;          // all types' parameter vectors
    const int32_t *param_begin;   // [n_types + 1]
    float *state;                 // [max state][stride]: variable k of slot s at k * stride + s
    int64_t stride;
    float2 *syn_ext;              // [NC] per-compartment (I_syn, G_syn) of the step
    const int32_t *peer;          // per synapse slot: the gap junction's peer compartment, or -1
    const int32_t *flags;         // per type: bit 0 an input type (null: none)
    const int32_t *poff;          // per slot: offset of its parameter vector in params (null: its type's)
    int32_t *tpos;                // per slot: a firing-synapse input's position in ttimes, or -1 (null: none)
    const float *ttimes;          // spike lists, each ended by a value >= FLT_MAX
    const int32_t *pspk;          // per slot: a Poisson spiker's index, or -1 (null: none)
    float2 *pk_state;             // per spiker: (tsince, isi), updated in place
    const float *pk_isi;          // per spiker: averageIsi = 1 / averageRate (us)
    const float2 *pk_window;      // per spiker: (delay, duration) of a transient spiker
    const int32_t *pk_transient;  // per spiker: 1 transient
    const int32_t *pk_inst;       // per spiker: randof instance
    const int32_t *pk_roff;       // per spiker: randof invocation offset
    float2 *inp_ext;              // [NC] per-compartment (I, G) of the LEMS inputs of the step, or null
};

// Arguments of lems_cell_kernel (LEMS point neurons, compiled at run time with the synapse
// kernel): one slot per point-neuron cell.
struct LemsCellArgs {
    const int32_t *comps;         // [n] the cell's compartment
    const int32_t *type;          // [n] its type
    int32_t n;
    const float *params;          // all types' parameter vectors
    const int32_t *param_begin;   // [n_types + 1]
    const int32_t *v_state;       // [n_types] membrane-voltage state variable, or -1
    float *state;                 // [max state][n]: variable k of slot i at k * n + i
    const int32_t *poff;          // [n] offset of the cell's parameter vector in params (null: its type's)
};

// Correctly rounded fp32 division whose expansion does not go through VCC.
//
// hipcc expands n / d into: v_div_scale of d, v_rcp, one Newton step on the reciprocal,
// v_div_scale of n, the quotient, two residual corrections (the last one by v_div_fmas, an
// fma that applies the scaling recorded in VCC) and v_div_fixup for special operands.
// v_div_scale leaves its operand unchanged and VCC clear unless d is zero, denormal or so
// large that 1/d is denormal, n is zero or has a biased exponent <= 23, the exponent of n
// exceeds that of d by 96 or more, or n/d is denormal. div_rn_fast requires, with margins,
// the opposite: d's biased exponent in [2, 250], and n either a zero (v_div_fixup then
// returns the signed zero whatever its first operand) or with biased exponent in [26, 252]
// and exponent difference n - d in [-120, 90]. Then the sequence below performs the same
// operations on the same values as the compiler's and gives the same correctly rounded
// quotient; `ok` reports whether the operands were in that range. It does not use VCC, so
// independent divisions can overlap instead of being serialised on that register.
__device__ __forceinline__ float div_rn_fast(float n, float d, bool &ok) {
    // range test without short-circuit operators, so it stays branch-free
    const uint32_t un = __float_as_uint(n), ud = __float_as_uint(d);
    const int32_t en = (int32_t)((un >> 23) & 0xFFu);
    const int32_t ed = (int32_t)((ud >> 23) & 0xFFu);
    const bool d_ok = (uint32_t)(ed - 2) <= 248u;
    const bool n_zero = (un << 1) == 0u;
    const bool n_ok = ((uint32_t)(en - 26) <= 226u) & ((uint32_t)(en - ed + 120) <= 210u);
    ok = d_ok & (n_zero | n_ok);
    float y = __builtin_amdgcn_rcpf(d);
    const float e = __builtin_fmaf(-d, y, 1.0f);
    y = __builtin_fmaf(e, y, y);
    float q = n * y;
    float r = __builtin_fmaf(-d, q, n);
    q = __builtin_fmaf(r, y, q);
    r = __builtin_fmaf(-d, q, n);
    q = __builtin_fmaf(r, y, q);
    return __builtin_amdgcn_div_fixupf(q, d, n);
}

__device__ __forceinline__ float div_rn(float n, float d) {
    bool ok;
    float q = div_rn_fast(n, d, ok);
    if (__builtin_expect(!ok, 0)) q = n / d;
    return q;
}

// Two independent divisions whose fast paths share one basic block (so they interleave),
// with one fallback branch for both.
__device__ __forceinline__ void div2_rn(float n0, float d0, float n1, float d1, float &q0, float &q1) {
    bool ok0, ok1;
    q0 = div_rn_fast(n0, d0, ok0);
    q1 = div_rn_fast(n1, d1, ok1);
    if (__builtin_expect(!(ok0 & ok1), 0)) {
        if (!ok0) q0 = n0 / d0;
        if (!ok1) q1 = n1 / d1;
    }
}

// fp32 exp and log rounded once from a double evaluation: correctly rounded except when the
// double result lies within one double ulp of a float rounding boundary (about 2^-28 of
// arguments). glibc's expf/logf, which the reference's generated C calls, are within
// 0.502 ulp, so both sides agree except in those rare cases. Building with
// -DEDENHIP_NATIVE_EXPF selects the device's own expf/logf instead.
//
// exp in double: Cody-Waite reduction x = k ln2 + r with |r| <= ln2/2 (ln2 split in a
// 53-bit head and a tail), the Taylor series of e^r to degree 13 (truncation error below
// 4e-18 relative), then scaling by 2^k. The polynomial is evaluated by Estrin's scheme
// (pairs, then powers r^2, r^4, r^8), four dependent fma levels instead of Horner's
// thirteen. Total relative error is a few double ulps, far below the 2^-24 spacing the
// final rounding to float resolves.
__device__ __forceinline__ double exp_taylor(double x) {
    const double LOG2E = 1.4426950408889634074;
    const double LN2_HI = 0x1.62e42ff000000p-1;    // ln2 rounded to 32 significant bits: k * LN2_HI is exact
    const double LN2_LO = -0x1.718432a1b0e26p-35;  // ln2 - LN2_HI
    const double k = rint(x * LOG2E);
    double r = fma(-k, LN2_HI, x);
    r = fma(-k, LN2_LO, r);
    const double r2 = r * r, r4 = r2 * r2, r8 = r4 * r4;
    const double p01 = r + 1.0;
    const double p23 = fma(r, 1.0 / 6.0, 0.5);
    const double p45 = fma(r, 1.0 / 120.0, 1.0 / 24.0);
    const double p67 = fma(r, 1.0 / 5040.0, 1.0 / 720.0);
    const double p89 = fma(r, 1.0 / 362880.0, 1.0 / 40320.0);
    const double pab = fma(r, 1.0 / 39916800.0, 1.0 / 3628800.0);
    const double pcd = fma(r, 1.0 / 6227020800.0, 1.0 / 479001600.0);
    const double q0 = fma(r2, p23, p01);
    const double q1 = fma(r2, p67, p45);
    const double q2 = fma(r2, pab, p89);
    const double s0 = fma(r4, q1, q0);
    const double s1 = fma(r4, pcd, q2);
    const double p = fma(r8, s1, s0);
    return ldexp(p, (int)k);
}

__device__ __forceinline__ float exp_f(float x) {
#ifdef EDENHIP_NATIVE_EXPF
    return expf(x);
#else
    if (!(x > -200.0f && x < 200.0f)) return expf(x);   // saturated, NaN or inf: the float result is exact
    return (float)exp_taylor((double)x);
#endif
}

__device__ __forceinline__ float log_f(float x) {
#ifdef EDENHIP_NATIVE_EXPF
    return logf(x);
#else
    return (float)log((double)x);
#endif
}

// x^y for generated LEMS code, rounded once from the double power (the oracle's C dialect
// does the same with the host's pow).
__device__ __forceinline__ float pow_f(float x, float y) { return (float)pow((double)x, (double)y); }

// A spike detected while computing step `step` is logged at row `step` (the next row).
__device__ __forceinline__ void log_event(const DevTables &T, int32_t col, int64_t step) {
    if (col >= 0 && step < T.total_rows) {
        uint32_t slot = atomicAdd(T.n_events, 1u);
        if (slot < T.max_events) T.events[slot] = ((int64_t)step << 24) | (int64_t)col;
    }
}

// A spike of compartment gidx sets its recipients' trigger flags for the next step.
__device__ __forceinline__ void send_spike(const DevTables &T, int64_t gidx, int nxt) {
    const int32_t b = T.spk_begin[gidx], e = T.spk_begin[gidx + 1];
    int32_t *trig = T.trig[nxt];
    for (int32_t k = b; k < e; k++) trig[T.spk_target[k]] = 1;   // idempotent flag (EDEN: atomic OR 1)
}

// Rate forms without divergent branches: one exp and at most one extra division per rate,
// each form's value picked by selects. Every form keeps the reference's operation order:
//   EXP        r * exp((V - mid) / scale)
//   EXPLINEAR  r * ((V == mid) ? 1 : ((V - mid) / scale) / (1 - exp(-(V - mid) / scale)))
//   SIGMOID    r / (1 + exp((mid - V) / scale))
//   FIXED      r
// Division policies for the gate math: DivFast takes div_rn_fast and remembers whether all
// operands were in its range; DivIeee is the compiler's division. Both give the same
// quotient wherever DivFast reports ok.
struct DivFast {
    bool ok = true;
    __device__ __forceinline__ float operator()(float n, float d) {
        bool o;
        const float q = div_rn_fast(n, d, o);
        ok = ok & o;
        return q;
    }
};
struct DivIeee {
    __device__ __forceinline__ float operator()(float n, float d) const { return n / d; }
};

template <class Div>
__device__ __forceinline__ float rate_eval_t(Div &dv, int form, float r, float mid, float scale, float V) {
#ifdef EDENHIP_LEMS_RATES
    if (form == RATE_LEMS) return ::lems_rate((int)r, V);
#endif
    const bool sig = form == RATE_SIGMOID, lin = form == RATE_EXPLINEAR;
    const float x = dv(sig ? (mid - V) : (V - mid), scale);
    const float e = exp_f(lin ? -x : x);
    const float num = lin ? x : r;
    const float den = lin ? (1.0f - e) : (1.0f + e);
    const float q = dv(num, den);
    const float v_exp = r * e;
    const float v_lin = r * ((V == mid) ? 1.0f : q);
    return form == RATE_EXP ? v_exp : (lin ? v_lin : (sig ? q : r));
}

// One gate's next value from the membrane voltage: rates, tau, inf, cnexp with clipping
// (eden/Eden.cpp:5236-5259, 5282-5383); initialisation steps set the gate to inf. A tau/inf
// gate (kind 1) holds its time course in the alpha slots and its steady state in the beta
// slots, and takes them as they are.
template <class Div>
__device__ __forceinline__ float gate_next_t(Div &dv, const DGate &G, float V, float q, bool init, float dt) {
    const float alpha = rate_eval_t(dv, G.forms & 0xF, G.ar, G.am, G.as, V);
    const float beta = rate_eval_t(dv, (G.forms >> 8) & 0xFF, G.br, G.bm, G.bs, V);
    const bool tau_inf = (G.forms & 0xF0) != 0;
    const float tau = tau_inf ? alpha : dv(1.0f, alpha + beta);
    const float inf = tau_inf ? beta : dv(alpha, alpha + beta);
    if (init) return inf;
    const float tau_factor = dv(dt * G.q10, tau);
    const float blend = exp_f(-tau_factor);
    float nv = (blend) * q + (1.0f - blend) * inf;
    if (!(nv > 1e-6f)) nv = 1e-6f;
    if (!(nv < (float)(1 - 1e-6))) nv = (float)(1 - 1e-6);
    return nv;
}

// The cnexp update from a given tau and inf (the tail of gate_next_t, IEEE divisions).
__device__ __forceinline__ float gate_cnexp(float q10, float tau, float inf, float q, bool init, float dt) {
    if (init) return inf;
    const float tau_factor = (dt * q10) / tau;
    const float blend = exp_f(-tau_factor);
    float nv = (blend) * q + (1.0f - blend) * inf;
    if (!(nv > 1e-6f)) nv = 1e-6f;
    if (!(nv < (float)(1 - 1e-6))) nv = (float)(1 - 1e-6);
    return nv;
}

__device__ __forceinline__ float gate_next(const DGate &G, float V, float q, bool init, float dt) {
    DivFast f;
    float nv = gate_next_t(f, G, V, q, init, dt);
    if (__builtin_expect(!f.ok, 0)) {
        DivIeee d;
        nv = gate_next_t(d, G, V, q, init, dt);
    }
    return nv;
}

// Thomas Wang's 64-bit integer hash (published 1997); EDEN's RNG mixer
// (eden/Eden.cpp:3200-3212).
__device__ __forceinline__ uint64_t hash64shift(uint64_t key) {
    key = (~key) + (key << 21);
    key = key ^ (key >> 24);
    key = (key + (key << 3)) + (key << 8);
    key = key ^ (key >> 14);
    key = (key + (key << 2)) + (key << 4);
    key = key ^ (key >> 28);
    key = key + (key << 31);
    return key;
}

// randof (eden/Eden.cpp:3213-3233): counter-based uniform in [0, x).
__device__ __forceinline__ float randof(float x, int32_t work_item, int64_t instance, int64_t step, int32_t invocation) {
    uint64_t hi = (uint64_t)(int64_t)work_item * (1ULL << 24) | ((uint64_t)instance % (1ULL << 24));
    uint64_t lo = (uint64_t)(int64_t)invocation * (1ULL << 40) | ((uint64_t)step % (1ULL << 40));
    uint64_t s = hash64shift(hash64shift(lo) ^ hi);
    float r = ((float)(s % (1ULL << 23))) / ((float)(1 << 23));
    return x * r;
}

// The trace element of device column `col` at chunk row `row`. Column-major: a recorded
// compartment's consecutive steps go to consecutive words, so L2 merges a cell's stores of
// successive steps into whole lines before they reach HBM (in a row-major layout every store
// of a step lands in a line of its own).
__device__ __forceinline__ float &trace_at(const DevTables &T, int32_t col, int64_t row) {
    return T.trace[(int64_t)col * T.trace_rows + row];
}

// chmeta of a compartment with more channel distributions (NCH) or gate variables (NQ) than
// the compartment kernels hold: ext_channels_kernel computes its channels each step, in
// distribution order, and the kernels take the totals from DevTables::chx_ext.
constexpr uint32_t CHM_EXTERNAL = 0xFFFFFFFEu;

// Arguments of ext_channels_kernel: one lane per such compartment.
struct ExtChanArgs {
    const int32_t *comps;      // [n] compartments
    const int32_t *ch_begin;   // [n + 1] into the channel arrays
    const int32_t *ch_type;    // per channel: its type
    const float *ch_g, *ch_e;  // per channel: Gbase, Erev
    const int32_t *q_begin;    // per channel: its first gate variable in q
    float *q;                  // gate variables, updated in place
    float2 *out;               // [NC] the compartment's (I, G) of the step
    int32_t n;
    const float *scaling;      // per channel type: conductance_scaling, or null (1)
    const float4 *gate_extra;  // per gate table entry: (form, rate, midpoint, scale) of the time
                               // course of a rates+tau gate or the steady state of a rates+inf
                               // gate, or null
};

// One lane of ext_channels_kernel: a compartment's channels in distribution order, each from
// its gates' present values (each gate `instances` times, conductance_scaling = 1), the gates
// moving to their next values, currents and conductances summed from zero
// (eden/Eden.cpp:5679-5898), as channels_step does for the compartments it holds.
__device__ __forceinline__ void ext_channels(const DevTables &T, const DGate *__restrict__ gates,
                                             const ExtChanArgs &X, int64_t step, float dt) {
    const int32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= X.n) return;
    const int32_t g = X.comps[idx];
    const bool init = step <= 0;
    const float V = T.V[g], A = T.A[g];
    float I_ch = 0.0f, G_ch = 0.0f;
    for (int32_t c = X.ch_begin[idx]; c < X.ch_begin[idx + 1]; c++) {
        const int32_t type = X.ch_type[c];
        const int ng = gates[type * GATES_PER_TYPE].forms >> 24;
        float fopen = X.scaling ? X.scaling[type] : 1.0f;   // conductance_scaling
        for (int gi = 0; gi < ng; gi++) {
            const DGate G = gates[type * GATES_PER_TYPE + gi];
            float *qp = X.q + X.q_begin[c] + gi;
            // an instantaneous gate (kind 2) is its steady state at this step's voltage
            const int kind = (G.forms >> 4) & 0xF;
            const bool instantaneous = kind == 2;
            DivIeee dv;
            const float qs = instantaneous ? rate_eval_t(dv, (G.forms >> 8) & 0xFF, G.br, G.bm, G.bs, V) : *qp;
            const int inst = (G.forms >> 16) & 0xFF;
            for (int i = 0; i < inst; i++) fopen = fopen * qs;
            if (kind == 3 || kind == 4) {
                // rates with a time course (3) or a steady state (4) of their own
                const float4 xr = X.gate_extra[type * GATES_PER_TYPE + gi];
                const float a = rate_eval_t(dv, G.forms & 0xF, G.ar, G.am, G.as, V);
                const float b = rate_eval_t(dv, (G.forms >> 8) & 0xFF, G.br, G.bm, G.bs, V);
                const float x = rate_eval_t(dv, (int)xr.x, xr.y, xr.z, xr.w, V);
                const float tau = kind == 3 ? x : 1.0f / (a + b);
                const float inf = kind == 4 ? x : a / (a + b);
                *qp = gate_cnexp(G.q10, tau, inf, qs, init, dt);
            } else {
                *qp = instantaneous ? qs : gate_next(G, V, qs, init, dt);
            }
        }
        const float Gscaled = X.ch_g[c] * fopen;
        const float iDensity = Gscaled * (X.ch_e[c] - V) * 1e-5f;
        const float gDensity = (Gscaled) * 1e-5f;
        I_ch += (iDensity * A);
        G_ch += (gDensity * A);
    }
    X.out[g] = make_float2(I_ch, G_ch);
}

// Per-compartment constants cached in registers for the whole time loop.
// gmeta bit 27: the compartment has pulse inputs (the first one is cached here); bit 31: it
// has more than one (the rest are read from global memory).
constexpr uint32_t GM_PULSE = 1u << 27;
constexpr uint32_t GM_MORE_PULSES = 1u << 31;

struct CompConst {
    float C, A, invrc, vthr;
    uint32_t chmeta, gmeta;
    float g[NCH], e[NCH];
    float p0s, p0e, p0a;   // first pulse: start, start + duration, Imax * weight
    int32_t rec, ev;
};

// A tree-cell compartment that is not simple: its constants and compartment word, staged
// per cell (80 bytes, read with 16-byte loads).
struct alignas(16) TreeComplex {
    CompConst k;
    uint32_t tcm;
};
static_assert(sizeof(TreeComplex) == 80, "TreeComplex layout");

__device__ __forceinline__ void load_comp_const(const DevTables &T, int64_t gidx, CompConst &k) {
    k.C = T.C[gidx]; k.A = T.A[gidx]; k.invrc = T.invrc[gidx]; k.vthr = T.vthr[gidx];
    k.chmeta = T.chmeta[gidx]; k.gmeta = T.gmeta[gidx];
#pragma unroll
    for (int c = 0; c < NCH; c++) { k.g[c] = T.chg[gidx * NCH + c]; k.e[c] = T.che[gidx * NCH + c]; }
    const int32_t pb = T.pulse_begin[gidx], pe = T.pulse_begin[gidx + 1];
    if (pe > pb) {
        k.p0s = T.pulse_start[pb]; k.p0e = T.pulse_end[pb]; k.p0a = T.pulse_amp[pb];
        k.gmeta |= GM_PULSE;
        if (pe > pb + 1) k.gmeta |= GM_MORE_PULSES;
    } else {
        k.p0s = 0.f; k.p0e = -1.f; k.p0a = 0.f;
    }
    k.rec = T.rec_col[gidx]; k.ev = T.ev_col[gidx];
}

// Channels: gate updates (in place, after each gate's current value has entered the open
// fraction, eden/Eden.cpp:5330) and channel current/conductance sums in distribution order.
__device__ __forceinline__ void channels_step(const DGate *__restrict__ gates, const CompConst &k, float V,
                                              float (&q)[NQ], bool init, float dt, float &I_ch, float &G_ch) {
    const int nq = (k.gmeta >> 24) & 7;
    const int nch = (k.gmeta >> 28) & 7;
    float fop[NCH];
#pragma unroll
    for (int c = 0; c < NCH; c++) fop[c] = 1.0f;   // conductance_scaling = 1
#pragma unroll
    for (int s = 0; s < NQ; s++) {
        if (s < nq) {
            const int cs = (k.gmeta >> (4 * s)) & 3;
            const int gi = (k.gmeta >> (4 * s + 2)) & 3;
            const int type = (k.chmeta >> (8 * cs)) & 0xFF;
            const DGate G = gates[type * GATES_PER_TYPE + gi];
            const float qs = q[s];
            const float nv = gate_next(G, V, qs, init, dt);
            const int inst = (G.forms >> 16) & 0xFF;
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                if (c == cs) {
                    float p = fop[c];
                    for (int i = 0; i < inst; i++) p = p * qs;
                    fop[c] = p;
                }
            }
            q[s] = nv;
        }
    }
    I_ch = 0.0f; G_ch = 0.0f;
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
}

__device__ __forceinline__ float pulses_current(const DevTables &T, const CompConst &k, int64_t gidx, float t) {
    // one pulse: 0 + (0 + amplitude) + 0 = amplitude exactly, or 0
    if (!(k.gmeta & GM_MORE_PULSES)) return (k.p0s <= t && t <= k.p0e) ? k.p0a : 0.0f;
    float I = 0.0f;
    {
        float Ip = 0.0f;
        if (k.p0s <= t && t <= k.p0e) Ip += k.p0a;
        {
            const int32_t pb = T.pulse_begin[gidx], pe = T.pulse_begin[gidx + 1];
            for (int p = pb + 1; p < pe; p++)
                if (T.pulse_start[p] <= t && t <= T.pulse_end[p]) Ip += T.pulse_amp[p];
        }
        I += Ip;
        I += 0.0f;
    }
    return I;
}

// expTwoSynapse LEMS dynamics for one instance: exposures from the current state, then
// Euler decay, or on an event the OnEvent assignment that overrides the decay.
__device__ __forceinline__ void exptwo_step(const DSynType &st, float V, float &A, float &B, bool init, bool ev,
                                            float dt, float &i_out, float &g_out) {
    const float g = st.gbase * (B - A);
    const float i = g * (st.erev - V);
    float An, Bn;
    if (init) {
        An = 0.0f; Bn = 0.0f;
    } else {
        const float dA = (-A) / st.tau_rise;
        An = A + dt * dA;
        const float dB = (-B) / st.tau_decay;
        Bn = B + dt * dB;
        if (ev) {
            An = A + (st.weight_prop * st.wf);
            Bn = B + (st.weight_prop * st.wf);
        }
    }
    A = An; B = Bn;
    i_out = i; g_out = g;
}

// All chemical synapses of a compartment: trigger read + clear, single-slot delay rule,
// expTwo dynamics, per-type aggregation in list order (eden/Eden.cpp:3636-3740). The fields
// of SYN_CHUNK synapses are loaded before any of them is computed (independent loads in
// flight); the accumulation stays sequential in synapse order.
constexpr int SYN_CHUNK = 8;

__device__ __forceinline__ void synapses_step(const DevTables &T, const DSynType *__restrict__ stypes, int32_t base,
                                              int32_t cnt, float V, bool init, float t, float dt, int now,
                                              int64_t step, float &I_syn, float &G_syn) {
    I_syn = 0.0f; G_syn = 0.0f;
    if (cnt <= 0) return;
    float I_agg = 0.0f, G_agg = 0.0f;
    int32_t cur = -1;
    int32_t *trig_now = T.trig[now];
    for (int32_t k0 = 0; k0 < cnt; k0 += SYN_CHUNK) {
        int32_t ty[SYN_CHUNK], tg[SYN_CHUNK];
        float w[SYN_CHUNK], dl[SYN_CHUNK], nx[SYN_CHUNK], a[SYN_CHUNK], b[SYN_CHUNK];
#pragma unroll
        for (int u = 0; u < SYN_CHUNK; u++) {
            const int32_t s = base + 64 * min(k0 + u, cnt - 1);
            ty[u] = T.syn_type[s];
            tg[u] = trig_now[s];
            w[u] = T.syn_weight[s];
            dl[u] = T.syn_delay[s];
            nx[u] = T.syn_next[s];
            a[u] = T.syn_A[s];
            b[u] = T.syn_B[s];
        }
#pragma unroll
        for (int u = 0; u < SYN_CHUNK; u++) {
            if (k0 + u >= cnt) break;
            const int32_t s = base + 64 * (k0 + u);
            if (ty[u] != cur) {
                if (cur >= 0) { I_syn += I_agg; G_syn += G_agg; }
                I_agg = 0.0f; G_agg = 0.0f; cur = ty[u];
            }
            const float weight = w[u];
            bool flag = false;
            if (!init) {
                flag = tg[u] != 0;
                if (flag) trig_now[s] = 0;
            }
            const float delay = dl[u];
            const float next_spike = nx[u];
            float nn = next_spike;
            bool spike_now = false;
            if (!init) {
                if (t <= next_spike && next_spike < t + dt) spike_now = true;
                if (flag) {
                    const float fresh = t + delay;
                    if (t <= fresh && fresh < t + dt) spike_now = true;
                    if (nn < t + dt && nn < fresh) nn = fresh;
                }
            }
            if (nn != next_spike) T.syn_next[s] = nn;
            if (spike_now && T.log_syn) {
                const uint32_t slot = atomicAdd(T.n_syn_events, 1u);
                if (slot < T.max_syn_events) T.syn_events[slot] = (step << 32) | (int64_t)T.syn_orig[s];
            }
            const DSynType st = stypes[ty[u]];
            float A = a[u], B = b[u], i, g;
            exptwo_step(st, V, A, B, init, spike_now, dt, i, g);
            T.syn_A[s] = A; T.syn_B[s] = B;
            I_agg += i * weight;
            G_agg += g * weight;
        }
    }
    I_syn += I_agg; G_syn += G_agg;
}

// Poisson firing synapse inputs of a compartment (LEMS OnStart / OnCondition with the hash
// RNG), each driving its own expTwo synapse in the same step.
// Poisson input state arrays (tsince, isi, expTwo A, B), so a caller can read one copy and
// write another.
struct PoiState {
    float *tsince, *isi, *A, *B;
};

__device__ __forceinline__ void poisson_step(const DevTables &T, const DSynType *__restrict__ stypes, int64_t gidx,
                                             float V, bool init, int64_t step, float dt, float &I_in, float &G_in,
                                             const PoiState &rd, const PoiState &wr) {
    const int32_t b = T.poi_begin[gidx], e = T.poi_begin[gidx + 1];
    if (e <= b) return;
    const int32_t seed = (int32_t)T.cell_seed[T.comp_cell[gidx]];
    float I_agg = 0.0f, G_agg = 0.0f;
    for (int32_t p = b; p < e; p++) {
        const float weight = T.poi_weight[p];
        const float tsince = rd.tsince[p], isi = rd.isi[p];
        const float averageIsi = T.poi_isi_mean[p];
        const int32_t inst = T.poi_inst[p], roff = T.poi_roff[p];
        bool fired = false;
        float tn, in;
        if (init) {
            tn = 0.0f;
            in = (-averageIsi) * log_f(1.0f - randof(1.0f, seed, inst, step, roff + 0));
        } else {
            tn = tsince + dt * 1.0f;
            in = isi;
            if (tsince > isi) {
                tn = 0.0f;
                in = (-averageIsi) * log_f(1.0f - randof(1.0f, seed, inst, step, roff + 1));
                fired = true;
            }
        }
        wr.tsince[p] = tn; wr.isi[p] = in;
        const DSynType st = stypes[T.poi_syn_type[p]];
        float A = rd.A[p], B = rd.B[p], i, g;
        exptwo_step(st, V, A, B, init, fired, dt, i, g);
        wr.A[p] = A; wr.B[p] = B;
        I_agg += i * weight;
        G_agg += g * weight;
    }
    I_in += I_agg;
    G_in += G_agg;
}

}  // namespace edenhip
