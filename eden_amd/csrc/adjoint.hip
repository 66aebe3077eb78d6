// adjoint.hip -- discrete adjoint of the fixed-step integrator (include/edenhip_adjoint.h).
//
// Everything here is double precision and follows the operation order of the reference's
// per-step update (as the FP64 restatement in oracle/ does):
//   gates          eden/Eden.cpp:5236-5259, 5336-5375 (rates, tau, inf, cnexp, clipping)
//   channels       eden/Eden.cpp:5679-5721, 5805-5898 (fopen, Gbase*fopen*(E-V)*1e-5f*A)
//   pulses         eden/Eden.cpp:3798-3826 (inclusive window, Imax * weight)
//   synapses       eden/Eden.cpp:2614-2994, 3640-3643, 3710 (expTwo, weight times i and g)
//   membrane       eden/Eden.cpp:6057-6107 (D = G/C + invRC, V' = V + dt(I + V G)/C)
//   Hines          eden/Eden.cpp:6114-6225 (D = 1 + D dt, elimination in DFS order)
// The Hines solve runs in the breadth-first rounds of tree_schedule.hpp, which apply each
// node's operations in the reference's order; one wave runs all rounds, and a node's
// eliminated diagonal is kept as its reciprocal (ratio = Uj * (1/D) rather than Uj / D).
//
// One workgroup per cell. A kernel first copies its cell's static records (compartments,
// channels, gates, pulses, synapse types, schedule) and parameters into LDS. adj_forward runs
// the whole time loop, writes the trace and a checkpoint of (V, q) at the start of every K-th
// recorded step; when the history fits in device memory it also writes a record of every
// step: its starting state and its Jacobian factors (the reciprocal eliminated diagonals and
// each gate's dq'/dq and dq'/dV), in the segment layout the backward pass reads.
// adj_backward handles one segment of K steps: it takes the segment's records from that
// history, or recomputes them from the segment's checkpoint, then walks the steps backwards:
//   lambda_b = M^-T lambda_V'   (Hines on the transposed matrix: Ui and Uj swap roles; the
//                                stored reciprocal diagonals serve both matrices)
//   lambda_D = -lambda_b V'
//   per channel c of compartment i:  dL/dGscaled = dt/C * 1e-5f * A * (lambda_b E_c + lambda_D)
//       dL/dGbase += dL/dGscaled * fopen,   dL/dfopen = dL/dGscaled * Gbase
//   per pulse p active at t:          dL/dAmplitude += lambda_b dt / C * weight
//   per synapse instance:             dL/dw += dt / C * (lambda_b erev + lambda_D) * g_s
//   per gate g (unclipped update):    dq'/dq = blend,  dq'/dV = (q - inf) dblend/dV + (1 - blend) dinf/dV
//   lambda_V = lambda_b + sum_g lambda_q'_g dq'_g/dV + dLoss/dV,  lambda_q = blend lambda_q' + dfopen/dq lambda_fopen
// (b = V + dt(I + V G)/C has db/dV = 1: the semi-implicit form's V terms cancel.)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/edenhip_adjoint.h"
#include "tree_schedule.hpp"

namespace edenhip {
namespace {

struct AdjGate {
    int32_t alpha_form, beta_form, instances, pad;
    double q10, ar, am, as, br, bm, bs;
};

struct AdjCell {
    int64_t comp_begin;
    int32_t n, rounds_fwd, rounds_bwd, words_off, words_len;
    int32_t chan_begin, n_chan, gate_begin, n_gate, pulse_begin, n_pulse;
    int32_t syn_begin, n_syn;
};

// expTwoSynapse type in double: parameters and the waveform factor (a derived parameter,
// evaluated with the reference's expression tree)
struct AdjSynType {
    double gbase, erev, tau_rise, tau_decay, weight_prop, wf;
};

// Static records of a cell, indexed locally within the cell. A kernel copies its cell's
// records into LDS when it starts, so the step loops read no table from global memory.
struct CompRec {
    float A, C, invrc;      // the descriptor's fp32 values, widened (exactly) where used
    int32_t rec_col;        // trace column, or -1
    uint16_t pos;           // breadth-first position in the Hines schedule
    uint16_t leaf;          // 1: no children (its reciprocal diagonal is formed in the compartment phase)
    uint16_t ch_b, ch_e;    // channel distributions [ch_b, ch_e)
    uint16_t p_b, p_e;      // pulse inputs
    uint16_t s_b, s_e;      // synapse instances
};
static_assert(sizeof(CompRec) == 32, "CompRec is four doubles");

struct ChanRec {
    float erev;
    uint16_t comp, q0;      // compartment, first gate variable
    uint32_t shape;         // gate count in bits 0-2, instances of gate k in bits 3 + 6k .. 8 + 6k
    uint32_t pad;
};
static_assert(sizeof(ChanRec) == 16, "ChanRec is two doubles");

struct GateRec {
    uint16_t comp, chan;    // compartment whose V drives the gate, the gate's channel
    uint16_t param, k;      // gate table entry (channel type * 4 + k), gate index within the channel
};
static_assert(sizeof(GateRec) == 8, "GateRec is one double");

struct PulseRec {
    float start, dur, weight, pad;
};

struct AdjTables {
    double dt;
    int64_t rows;
    int32_t n_record, n_gate_params, n_syn_types;
    const AdjCell *cells;
    const CompRec *comp;     // by compartment
    const ChanRec *chan;     // by channel distribution
    const GateRec *gate;     // by gate variable
    const PulseRec *pulse;   // by pulse input
    const AdjGate *gates;    // gate table, 4 entries per channel type
    const int32_t *s_type;   // by synapse instance
    const AdjSynType *s_types;
    // edge constants (Ui, Uj) by global position, schedule blocks, row times
    const double2 *u;
    const uint32_t *words;
    const double *t_row;
    // the synapses' fixed event schedule: per instance its event steps, ascending
    const int32_t *ev_begin;
    const int64_t *ev_step;
};

// One step's record of a cell (history or segment buffer): the state at the start of the
// step, V[n] q[ng] A[ns] B[ns], then the step's Jacobian factors: the reciprocal eliminated
// Hines diagonal by position inv[n + 1] (the last entry, the dummy node's, is 1), and per
// gate variable dq'/dq and dq'/dV. The backward pass copies a record into LDS as it is and
// reads its fields in place.
struct RecLayout {
    int32_t q, a, b, inv, dq, dv, per;
};

__host__ __device__ inline RecLayout rec_layout(int32_t n, int32_t ng, int32_t ns) {
    RecLayout R;
    R.q = n;
    R.a = R.q + ng;
    R.b = R.a + ns;
    R.inv = R.b + ns;
    R.dq = R.inv + n + 1;
    R.dv = R.dq + ng;
    R.per = R.dv + ng;
    return R;
}

constexpr double K1E5 = (double)1e-5f;           // the reference's 1e-5f literal, widened
constexpr double QLO = (double)(1e-6f);          // clipping bounds as the FP64 restatement writes them
constexpr double QHI = (double)(1 - 1e-6f);

template <int TPB>
__device__ __forceinline__ void bar() {
    if constexpr (TPB <= 64) {
        __builtin_amdgcn_wave_barrier();
        asm volatile("" ::: "memory");
    } else {
        asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
    }
}

// rounds of one wave: its LDS accesses execute in program order, so only the compiler has
// to be kept from moving them across a round boundary
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
}

// rate value and, when asked, its derivative in V. One exponential per rate whatever the
// form: exp(x) for EXP, exp(-x) for EXPLINEAR and SIGMOID, x = (V - m) / s ((m - V) / s and
// -(V - m) / s equal -x exactly), and the forms' remaining operations as selects, so the
// lanes of a wave holding gates of different forms do not run every form's branch. Each
// form's value is the same operations as its textbook expression.
__device__ __forceinline__ double rate64(int form, double r, double m, double s, double V, bool deriv, double &dv) {
    const bool is_exp = form == EDENHIP_RATE_EXP, is_lin = form == EDENHIP_RATE_EXPLINEAR;
    const bool is_sig = form == EDENHIP_RATE_SIGMOID, is_fix = !(is_exp || is_lin || is_sig);
    // EXPLINEAR at V == m is its limit r (derivative r / 2s); selected, not branched
    const bool at_mid = is_lin && V == m;
    const double x = is_fix ? 0.0 : (V - m) / s;
    const double e = is_fix ? 0.0 : exp(is_exp ? x : -x);
    const double den = is_lin ? 1.0 - e : is_sig ? 1.0 + e : 1.0;
    const double t = (is_exp ? e : is_lin ? x : r) / den;
    const double v = at_mid ? r : (is_exp || is_lin) ? r * t : t;   // r e, r (x / (1 - e^-x)), r / (1 + e), r
    if (deriv) {
        // d/dx x/(1 - e^-x) = ((1 - e^-x) - x e^-x) / (1 - e^-x)^2, with its series near 0
        const double gl = fabs(x) < 1e-5 ? 0.5 + x * (1.0 / 6.0) : ((den - x * e) / (den * den));
        const double dnum = is_exp ? v : is_lin ? r * gl : is_sig ? r * e : 0.0;
        const double dden = is_sig ? s * (1.0 + e) * (1.0 + e) : is_fix ? 1.0 : s;
        dv = at_mid ? r / (2.0 * s) : dnum / dden;
    } else {
        dv = 0.0;
    }
    return v;
}

// next gate value (the reference's order, widened), and, when asked, the partial
// derivatives of the update in q and V (zero when the result was clipped)
__device__ __forceinline__ double gate64(const AdjGate &G, double V, double q, bool init, double dt, double *dq,
                                         double *dV) {
    double da, db;
    const bool deriv = dq != nullptr;
    const double alpha = rate64(G.alpha_form, G.ar, G.am, G.as, V, deriv, da);
    const double beta = rate64(G.beta_form, G.br, G.bm, G.bs, V, deriv, db);
    const double tau = 1.0 / (alpha + beta);
    const double inf = alpha / (alpha + beta);
    if (init) {
        if (dq) *dq = 0.0;
        if (dV) *dV = 0.0;
        return inf;
    }
    const double tau_factor = ((dt * G.q10) / tau);
    const double blend = exp(-tau_factor);
    double qn = (blend) * q + (1.0 - blend) * inf;
    bool clipped = false;
    if (!(qn > QLO)) { qn = QLO; clipped = true; }
    if (!(qn < QHI)) { qn = QHI; clipped = true; }
    if (dq) {
        const double sum = alpha + beta;
        const double dblend = -blend * dt * G.q10 * (da + db);
        const double dinf = (da * beta - alpha * db) / (sum * sum);
        *dq = clipped ? 0.0 : blend;
        *dV = clipped ? 0.0 : (q - inf) * dblend + (1.0 - blend) * dinf;
    }
    return qn;
}

// LDS of one cell, in doubles
struct AdjLds {
    int32_t v, vn, q, f, lf, x, inv, u, lv, lq, gv, gg, ga, gw, sa, sb, sc, sn, gb, amp, wt, stage, comp, chan, gate,
        pulse, stype, gtab, sttab, w, total;
};

__host__ __device__ inline AdjLds adj_lds(int32_t nmax, int32_t nqmax, int32_t nchmax, int32_t npmax, int32_t wmax,
                                          int32_t nsmax, int32_t n_gate_params, int32_t n_syn_types) {
    AdjLds L;
    int32_t o = 0;
    L.u = o;     o += 2 * (nmax + 1);      // double2, first so it stays 16-byte aligned
    L.v = o;     o += nmax;
    L.vn = o;    o += nmax;                // V at the start of the next step (backward)
    L.q = o;     o += 2 * nqmax;           // gate values, two buffers
    L.f = o;     o += nchmax;              // open fractions
    L.lf = o;    o += nchmax;              // dLoss/dfopen
    L.x = o;     o += nmax + 1;            // Hines right-hand side / solution, by position
    L.inv = o;   o += nmax + 1;            // diagonal, then its reciprocal after elimination, by position
    L.lv = o;    o += nmax;
    L.lq = o;    o += 2 * nqmax;
    L.gv = o;    o += nqmax;
    L.gg = o;    o += nchmax;              // gradient accumulators: Gbase, amplitude, weight
    L.ga = o;    o += npmax;
    L.gw = o;    o += nsmax;
    L.sa = o;    o += nsmax;               // synapse state A, B
    L.sb = o;    o += nsmax;
    L.sc = o;    o += nsmax;               // event cursor and the step of the next event (int64)
    L.sn = o;    o += nsmax;
    L.gb = o;    o += nchmax;              // parameters: Gbase, amplitude, weight
    L.amp = o;   o += npmax;
    L.wt = o;    o += nsmax;
    L.stage = o; o += rec_layout(nmax, nqmax, nsmax).per;   // the step's record (backward)
    L.comp = o;  o += 4 * nmax;            // static records
    L.chan = o;  o += 2 * nchmax;
    L.gate = o;  o += nqmax;
    L.pulse = o; o += 2 * npmax;
    L.stype = o; o += (nsmax + 1) / 2;
    L.gtab = o;  o += 8 * n_gate_params;
    L.sttab = o; o += 6 * n_syn_types;
    L.w = o;     o += (wmax + 1) / 2;
    L.total = o;
    return L;
}

template <int TPB>
struct CellView {
    const AdjTables &T;
    const AdjCell &cell;
    int tid, nqmax;
    double *sV, *sVn, *sQ, *sF, *sLF, *sX, *sInv, *sLV, *sLQ, *sGV, *sGg, *sGa, *sGw, *sA, *sB, *sGb, *sAmp, *sWt,
        *sDQ, *sDV, *sStage;
    int64_t *sCur, *sNext;
    double2 *sU;
    CompRec *cr;
    ChanRec *chr;
    GateRec *gr;
    PulseRec *pr;
    int32_t *sty;
    AdjGate *gtab;
    AdjSynType *sttab;
    uint32_t *sW;

    // The schedule block (tree_schedule.hpp, rounds of at most 64 tasks) is read two rounds
    // ahead: while round r runs, a lane reads its task word of round r + 1 (from that round's
    // header, read during round r - 1) and the header of round r + 2. Lanes beyond a round's
    // tasks take the idle word, which points every access at the dummy node n.
    __device__ uint32_t task_word(uint32_t h, uint32_t idle) const {
        const uint32_t hu = __builtin_amdgcn_readfirstlane(h);
        return (uint32_t)tid < (hu >> 16) ? sW[(hu & 0xFFFFu) + tid] : idle;
    }

    // forward Hines solve (wave 0 runs every round; the other waves wait at the closing
    // barrier): each fold task eliminates its node's children in ascending order with the
    // children's reciprocal diagonals, then forms its own; back-substitution multiplies by it
    __device__ void hines_fwd() {
        if (tid < 64) {
            const int N = cell.n, RF = cell.rounds_fwd, RB = cell.rounds_bwd;
            const uint32_t idle_fwd = (uint32_t)N | ((uint32_t)N << 11) | (1u << 22);
            const uint32_t idle_bwd = (uint32_t)N | ((uint32_t)N << 11);
            uint32_t w = task_word(sW[0], idle_fwd), hn = sW[1];
            for (int r = 0; r < RF; r++) {
                const uint32_t wn = task_word(hn, idle_fwd);
                const uint32_t hnn = sW[r + 2];
                const int p = w & 0x7FF, fc = (w >> 11) & 0x7FF, nc = w >> 22;
                double dp = sInv[p], xp = sX[p];
                for (int i = 0; i < nc; i++) {
                    const int c = fc + i;
                    const double2 u = sU[c];
                    const double ratio = u.y * sInv[c];
                    dp -= ratio * u.x;
                    xp -= ratio * sX[c];
                }
                sX[p] = xp;
                sInv[p] = 1.0 / dp;
                wave_sync();
                w = wn;
                hn = hnn;
            }
            w = task_word(sW[RF + 1], idle_bwd);
            hn = RB > 0 ? sW[RF + 2] : 0u;
            if (tid == 0) sX[0] = sX[0] * sInv[0];
            wave_sync();
            for (int r = 0; r < RB; r++) {
                const uint32_t wn = task_word(hn, idle_bwd);
                const uint32_t hnn = r + 2 <= RB ? sW[RF + 3 + r] : 0u;
                const int p = w & 0x7FF, pp = (w >> 11) & 0x7FF;
                sX[p] = (sX[p] - sU[p].x * sX[pp]) * sInv[p];
                wave_sync();
                w = wn;
                hn = hnn;
            }
        }
        bar<TPB>();
    }

    // transposed solve with a step's stored reciprocal diagonals (Ui and Uj exchange roles;
    // the eliminated diagonals of the transposed matrix are the same)
    __device__ void hines_T() {
        if (tid < 64) {
            const int N = cell.n, RF = cell.rounds_fwd, RB = cell.rounds_bwd;
            const uint32_t idle_fwd = (uint32_t)N | ((uint32_t)N << 11) | (1u << 22);
            const uint32_t idle_bwd = (uint32_t)N | ((uint32_t)N << 11);
            uint32_t w = task_word(sW[0], idle_fwd), hn = sW[1];
            for (int r = 0; r < RF; r++) {
                const uint32_t wn = task_word(hn, idle_fwd);
                const uint32_t hnn = sW[r + 2];
                const int p = w & 0x7FF, fc = (w >> 11) & 0x7FF, nc = w >> 22;
                double xp = sX[p];
                for (int i = 0; i < nc; i++) {
                    const int c = fc + i;
                    xp -= (sU[c].x * sInv[c]) * sX[c];
                }
                sX[p] = xp;
                wave_sync();
                w = wn;
                hn = hnn;
            }
            w = task_word(sW[RF + 1], idle_bwd);
            hn = RB > 0 ? sW[RF + 2] : 0u;
            if (tid == 0) sX[0] = sX[0] * sInv[0];
            wave_sync();
            for (int r = 0; r < RB; r++) {
                const uint32_t wn = task_word(hn, idle_bwd);
                const uint32_t hnn = r + 2 <= RB ? sW[RF + 3 + r] : 0u;
                const int p = w & 0x7FF, pp = (w >> 11) & 0x7FF;
                sX[p] = (sX[p] - sU[p].y * sX[pp]) * sInv[p];
                wave_sync();
                w = wn;
                hn = hnn;
            }
        }
        bar<TPB>();
    }

    // open fractions of the cell's channels from the gate values in sQ[buf]
    __device__ void fopen(int buf) {
        const double *q = sQ + buf * nqmax;
        for (int c = tid; c < cell.n_chan; c += TPB) {
            const ChanRec ch = chr[c];
            const int ng = (int)(ch.shape & 7u);
            double f = 1.0;
            for (int k = 0; k < ng; k++) {
                const double qv = q[ch.q0 + k];
                const int inst = (int)((ch.shape >> (3 + 6 * k)) & 63u);
                for (int i = 0; i < inst; i++) f = f * qv;
            }
            sF[c] = f;
        }
    }

    // synaptic current and conductance of a compartment from the present synapse state:
    // per instance g = gbase (B - A), i = g (erev - V), times the connection weight, summed per
    // type in list order, the types' sums in order (as the oracle's FP64 flavour)
    __device__ void synapse_sums(const CompRec &c, double V, double &Is, double &Gs) const {
        Is = 0.0;
        Gs = 0.0;
        double Ia = 0.0, Ga = 0.0;
        int32_t cur = -1;
        for (int ls = c.s_b; ls < c.s_e; ls++) {
            const int32_t ty = sty[ls];
            if (ty != cur) {
                if (cur >= 0) { Is += Ia; Gs += Ga; }
                Ia = 0.0; Ga = 0.0; cur = ty;
            }
            const AdjSynType &st = sttab[ty];
            const double gs = st.gbase * (sB[ls] - sA[ls]);
            const double is = gs * (st.erev - V);
            Ia += is * sWt[ls];
            Ga += gs * sWt[ls];
        }
        if (cur >= 0) { Is += Ia; Gs += Ga; }
    }

    // next synapse state (in place) at step s: Euler decay, or on the instance's event the
    // OnEvent assignment; zero on initialisation steps
    __device__ void synapse_update(int64_t s, double dt) {
        for (int ls = tid; ls < cell.n_syn; ls += TPB) {
            const AdjSynType &st = sttab[sty[ls]];
            const double A = sA[ls], B = sB[ls];
            double An, Bn;
            if (s <= 0) {
                An = 0.0; Bn = 0.0;
            } else {
                An = A + dt * ((-A) / st.tau_rise);
                Bn = B + dt * ((-B) / st.tau_decay);
                if (sNext[ls] == s) {
                    An = A + (st.weight_prop * st.wf);
                    Bn = B + (st.weight_prop * st.wf);
                    const int64_t c = sCur[ls] + 1;
                    sCur[ls] = c;
                    sNext[ls] = c < T.ev_begin[cell.syn_begin + ls + 1] ? T.ev_step[c] : -1;
                }
            }
            sA[ls] = An;
            sB[ls] = Bn;
        }
    }

    // membrane currents of a compartment at time t: conductance and current sums
    __device__ void currents(const CompRec &c, double V, double t, double &I, double &G) const {
        const double A = (double)c.A;
        double Ich = 0.0, Gch = 0.0;
        for (int ch = c.ch_b; ch < c.ch_e; ch++) {
            const double Gs = sGb[ch] * sF[ch];
            const double iD = Gs * ((double)chr[ch].erev - V) * K1E5;
            const double gD = (Gs) * K1E5;
            Ich += (iD * A);
            Gch += (gD * A);
        }
        double Iin = 0.0;
        if (c.p_e > c.p_b) {
            double Ip = 0.0;
            for (int p = c.p_b; p < c.p_e; p++) {
                const double ps = (double)pr[p].start, pd = (double)pr[p].dur;
                if (ps <= t && t <= ps + pd) Ip += sAmp[p] * (double)pr[p].weight;
            }
            Iin += Ip;
            Iin += 0.0;
        }
        double Isyn = 0.0, Gsyn = 0.0;
        if (c.s_e > c.s_b) synapse_sums(c, V, Isyn, Gsyn);
        I = Ich + Iin + Isyn;
        G = Gch + 0.0 + Gsyn;
    }

    // the state at the start of a step into a record
    __device__ void put_state(double *rec, int cur) const {
        const int N = cell.n, NG = cell.n_gate, NS = cell.n_syn;
        for (int i = tid; i < N; i += TPB) rec[i] = sV[i];
        for (int k = tid; k < NG; k += TPB) rec[N + k] = sQ[cur * nqmax + k];
        for (int ls = tid; ls < NS; ls += TPB) {
            rec[N + NG + ls] = sA[ls];
            rec[N + NG + NS + ls] = sB[ls];
        }
    }

    // one forward step s from the state in sV / sQ[cur] into sV / sQ[1 - cur]; with rec, the
    // step's Jacobian factors go into the step's record
    __device__ void step(int64_t s, double t, int cur, double *rec, const RecLayout &R) {
        const bool init = s <= 0;
        const bool keep = rec != nullptr && !init;
        const double dt = T.dt;
        const int N = cell.n;
        fopen(cur);
        bar<TPB>();
        for (int i = tid; i < N; i += TPB) {
            const CompRec c = cr[i];
            const double V = sV[i];
            double I, G;
            currents(c, V, t, I, G);
            const double C = (double)c.C;
            const double D = 1 * (G / C) + (double)c.invrc;
            const double d = 1 + D * dt;
            sX[c.pos] = init ? V : V + (dt * (I + ((V * G))) / C);
            sInv[c.pos] = c.leaf ? 1.0 / d : d;
        }
        const double *q = sQ + cur * nqmax;
        double *qn = sQ + (1 - cur) * nqmax;
        // gates k and k + TPB side by side (independent chains of divisions and exponentials;
        // a lane without a second gate repeats its first and discards the result)
        for (int k = tid; k < cell.n_gate; k += 2 * TPB) {
            const int k2 = k + TPB < cell.n_gate ? k + TPB : k;
            const GateRec g = gr[k], g2 = gr[k2];
            if (keep) {
                double dq, dv, dq2, dv2;
                const double a = gate64(gtab[g.param], sV[g.comp], q[k], false, dt, &dq, &dv);
                const double b = gate64(gtab[g2.param], sV[g2.comp], q[k2], false, dt, &dq2, &dv2);
                qn[k] = a;
                rec[R.dq + k] = dq;
                rec[R.dv + k] = dv;
                if (k2 != k) {
                    qn[k2] = b;
                    rec[R.dq + k2] = dq2;
                    rec[R.dv + k2] = dv2;
                }
            } else {
                const double a = gate64(gtab[g.param], sV[g.comp], q[k], init, dt, nullptr, nullptr);
                const double b = gate64(gtab[g2.param], sV[g2.comp], q[k2], init, dt, nullptr, nullptr);
                qn[k] = a;
                if (k2 != k) qn[k2] = b;
            }
        }
        if (tid == 0) {
            sX[N] = 0.0;
            sInv[N] = 1.0;
        }
        bar<TPB>();
        // the synapses' next state (read above for this step's currents)
        if (cell.n_syn > 0) synapse_update(s, dt);
        if (!init) hines_fwd();
        for (int i = tid; i < N; i += TPB) sV[i] = sX[cr[i].pos];
        if (keep)
            for (int p = tid; p <= N; p += TPB) rec[R.inv + p] = p < N ? sInv[p] : 1.0;
        bar<TPB>();
    }
};

template <typename R>
__device__ __forceinline__ void copy_recs(R *dst, const R *src, int n, int tid, int tpb) {
    for (int i = tid; i < n; i += tpb) dst[i] = src[i];
}

// A cell's view of its LDS: static records, parameters and the synapses' initial state
template <int TPB>
__device__ CellView<TPB> make_view(const AdjTables &T, const AdjCell &cell, unsigned char *smem, const AdjLds &L,
                                   int nqmax, const double *gbase, const double *amp, const double *weight) {
    double *sm = reinterpret_cast<double *>(smem);
    const int tid = threadIdx.x;
    CellView<TPB> v{T, cell, tid, nqmax};
    v.sU = reinterpret_cast<double2 *>(sm + L.u);
    v.sV = sm + L.v;
    v.sVn = sm + L.vn;
    v.sQ = sm + L.q;
    v.sF = sm + L.f;
    v.sLF = sm + L.lf;
    v.sX = sm + L.x;
    v.sInv = sm + L.inv;
    v.sLV = sm + L.lv;
    v.sLQ = sm + L.lq;
    v.sGV = sm + L.gv;
    v.sGg = sm + L.gg;
    v.sGa = sm + L.ga;
    v.sGw = sm + L.gw;
    v.sA = sm + L.sa;
    v.sB = sm + L.sb;
    v.sCur = reinterpret_cast<int64_t *>(sm + L.sc);
    v.sNext = reinterpret_cast<int64_t *>(sm + L.sn);
    v.sGb = sm + L.gb;
    v.sAmp = sm + L.amp;
    v.sWt = sm + L.wt;
    v.sStage = sm + L.stage;
    v.sDQ = nullptr;   // the backward pass points these into its record copy
    v.sDV = nullptr;
    v.cr = reinterpret_cast<CompRec *>(sm + L.comp);
    v.chr = reinterpret_cast<ChanRec *>(sm + L.chan);
    v.gr = reinterpret_cast<GateRec *>(sm + L.gate);
    v.pr = reinterpret_cast<PulseRec *>(sm + L.pulse);
    v.sty = reinterpret_cast<int32_t *>(sm + L.stype);
    v.gtab = reinterpret_cast<AdjGate *>(sm + L.gtab);
    v.sttab = reinterpret_cast<AdjSynType *>(sm + L.sttab);
    v.sW = reinterpret_cast<uint32_t *>(sm + L.w);
    copy_recs(v.cr, T.comp + cell.comp_begin, cell.n, tid, TPB);
    copy_recs(v.chr, T.chan + cell.chan_begin, cell.n_chan, tid, TPB);
    copy_recs(v.gr, T.gate + cell.gate_begin, cell.n_gate, tid, TPB);
    copy_recs(v.pr, T.pulse + cell.pulse_begin, cell.n_pulse, tid, TPB);
    copy_recs(v.sty, T.s_type + cell.syn_begin, cell.n_syn, tid, TPB);
    copy_recs(v.gtab, T.gates, T.n_gate_params, tid, TPB);
    copy_recs(v.sttab, T.s_types, T.n_syn_types, tid, TPB);
    copy_recs(v.sW, T.words + cell.words_off, cell.words_len, tid, TPB);
    copy_recs(v.sU, T.u + cell.comp_begin, cell.n, tid, TPB);
    if (tid == 0) v.sU[cell.n] = make_double2(0.0, 0.0);
    copy_recs(v.sGb, gbase + cell.chan_begin, cell.n_chan, tid, TPB);
    copy_recs(v.sAmp, amp + cell.pulse_begin, cell.n_pulse, tid, TPB);
    copy_recs(v.sWt, weight + cell.syn_begin, cell.n_syn, tid, TPB);
    for (int ls = tid; ls < cell.n_syn; ls += TPB) {
        v.sA[ls] = 0.0;
        v.sB[ls] = 0.0;
        const int64_t c = T.ev_begin[cell.syn_begin + ls];
        v.sCur[ls] = c;
        v.sNext[ls] = c < T.ev_begin[cell.syn_begin + ls + 1] ? T.ev_step[c] : -1;
    }
    return v;
}

// Whole forward pass: initial state from v0/q0, trace rows, checkpoints of the state at the
// start of steps 1, 1 + K, 1 + 2K, ... (ck[k] = V of all compartments then q of all gates).
// hist (may be null): a record of every step s >= 1 (its starting state and its Jacobian
// factors), in the backward pass's segment layout: segment k = (s - 1) / K holds steps
// 1 + k K .. 1 + (k + 1) K (K + 1 records; the last one is also the first of segment k + 1,
// of which the backward pass reads only the state), hist_stride doubles apart, cell c at
// seg_off[c], step s's record at (s - 1 - k K) * per after it.
template <int TPB>
__global__ void __launch_bounds__(TPB) adj_forward(AdjTables T, const double *__restrict__ v0,
                                                   const double *__restrict__ q0, const double *__restrict__ gbase,
                                                   const double *__restrict__ amp, const double *__restrict__ weight,
                                                   double *__restrict__ trace, double *__restrict__ ck, int64_t ck_stride,
                                                   int64_t nc_total, int32_t K, AdjLds L, int32_t nqmax,
                                                   double *__restrict__ hist, int64_t hist_stride, int64_t n_seg,
                                                   const int64_t *__restrict__ seg_off) {
    extern __shared__ __align__(16) unsigned char smem[];
    const AdjCell cell = T.cells[blockIdx.x];
    auto v = make_view<TPB>(T, cell, smem, L, nqmax, gbase, amp, weight);
    const int tid = threadIdx.x;
    const RecLayout R = rec_layout(cell.n, cell.n_gate, cell.n_syn);
    for (int i = tid; i < cell.n; i += TPB) v.sV[i] = v0[cell.comp_begin + i];
    for (int k = tid; k < cell.n_gate; k += TPB) v.sQ[k] = q0[cell.gate_begin + k];
    const int64_t soff = hist ? seg_off[blockIdx.x] : 0;
    double t_next = T.t_row[0];
    bar<TPB>();
    int cur = 0;
    for (int64_t s = -3; s <= T.rows; s++) {
        const double t = s >= 1 ? t_next : 0.0;
        if (s >= 1) t_next = T.t_row[s];    // read a step ahead (t_row has rows + 1 entries)
        double *rec = nullptr;
        if (s >= 1) {
            if ((s - 1) % K == 0) {
                double *c = ck + ((s - 1) / K) * ck_stride;
                for (int i = tid; i < cell.n; i += TPB) c[cell.comp_begin + i] = v.sV[i];
                for (int k = tid; k < cell.n_gate; k += TPB) c[nc_total + cell.gate_begin + k] = v.sQ[cur * nqmax + k];
            }
            for (int i = tid; i < cell.n; i += TPB) {
                const int rc = v.cr[i].rec_col;
                if (rc >= 0) trace[(s - 1) * T.n_record + rc] = v.sV[i];
            }
            if (hist) {
                const int64_t k = (s - 1) / K, j = (s - 1) - k * K;
                double *base = hist + soff;
                // the state of the last step opens no segment of its own when (rows - 1) is a
                // multiple of K: it is only the closing state of the segment before
                if (k < n_seg) {
                    rec = base + k * hist_stride + j * R.per;
                    v.put_state(rec, cur);
                }
                if (j == 0 && k > 0) v.put_state(base + (k - 1) * hist_stride + (int64_t)K * R.per, cur);
            }
        }
        if (s == T.rows) break;
        v.step(s, t, cur, rec, R);
        cur ^= 1;
    }
}

// One backward segment: steps [a, b) with a = 1 + k K. seg holds the records of steps a..b
// (the forward pass's history, or recomputed here from the segment's checkpoint). lam
// (global) carries (lambda_V, lambda_q) at the start of step b in and at the start of step a
// out; grad_g, grad_a and grad_w accumulate.
template <int TPB>
__global__ void __launch_bounds__(TPB) adj_backward(AdjTables T, const double *__restrict__ gbase,
                                                    const double *__restrict__ amp, const double *__restrict__ weight,
                                                    const double *__restrict__ gtrace, const double *__restrict__ ck,
                                                    int64_t nc_total, int64_t a, int64_t b, double *__restrict__ seg,
                                                    const int64_t *__restrict__ seg_off, double *__restrict__ lam,
                                                    double *__restrict__ grad_g, double *__restrict__ grad_a,
                                                    double *__restrict__ grad_w, AdjLds L, int32_t nqmax, int32_t stored) {
    extern __shared__ __align__(16) unsigned char smem[];
    const AdjCell cell = T.cells[blockIdx.x];
    auto v = make_view<TPB>(T, cell, smem, L, nqmax, gbase, amp, weight);
    const int tid = threadIdx.x;
    const int N = cell.n, NG = cell.n_gate, NS = cell.n_syn;
    const RecLayout R = rec_layout(N, NG, NS);
    double *S = seg + seg_off[blockIdx.x];
    const double dt = T.dt;
    for (int c = tid; c < cell.n_chan; c += TPB) v.sGg[c] = 0.0;
    for (int p = tid; p < cell.n_pulse; p += TPB) v.sGa[p] = 0.0;
    for (int ls = tid; ls < NS; ls += TPB) v.sGw[ls] = 0.0;
    if (!stored) {
        // recompute the segment's records from its checkpoint (each thread later reads back
        // the entries it wrote here)
        for (int i = tid; i < N; i += TPB) v.sV[i] = ck[cell.comp_begin + i];
        for (int k = tid; k < NG; k += TPB) v.sQ[k] = ck[nc_total + cell.gate_begin + k];
        bar<TPB>();
        int cur = 0;
        for (int64_t s = a; s <= b; s++) {
            double *Ss = S + (s - a) * R.per;
            v.put_state(Ss, cur);
            if (s == b) break;
            v.step(s, T.t_row[s - 1], cur, Ss, R);
            cur ^= 1;
        }
        // the records are read below by other threads of the workgroup
        __threadfence_block();
        bar<TPB>();
        __threadfence_block();
    }
    // the record of the step being differentiated, in LDS: state, reciprocal diagonals, gate
    // derivatives read in place
    v.sV = v.sStage;
    v.sQ = v.sStage + R.q;
    v.sA = v.sStage + R.a;
    v.sB = v.sStage + R.b;
    v.sInv = v.sStage + R.inv;
    v.sDQ = v.sStage + R.dq;
    v.sDV = v.sStage + R.dv;
    // adjoint state at the start of step b, and V there
    for (int i = tid; i < N; i += TPB) v.sLV[i] = lam[cell.comp_begin + i];
    for (int k = tid; k < NG; k += TPB) v.sLQ[k] = lam[nc_total + cell.gate_begin + k];
    for (int i = tid; i < N; i += TPB) v.sVn[i] = S[(b - a) * R.per + i];
    double t_next = T.t_row[b - 2];
    bar<TPB>();
    for (int64_t s = b - 1; s >= a; s--) {
        const double *Ss = S + (s - a) * R.per;        // record of step s
        const double t = t_next;
        if (s >= 2) t_next = T.t_row[s - 2];          // read a step ahead
        // step s's record (reads issued in batches)
#pragma unroll 4
        for (int j = tid; j < R.per; j += TPB) v.sStage[j] = Ss[j];
        // dLoss/dV at the start of step s + 1 (trace row s)
        if (s < T.rows) {
            const double *gr = gtrace + s * T.n_record;
            for (int i = tid; i < N; i += TPB) {
                const int rc = v.cr[i].rec_col;
                if (rc >= 0) v.sLV[i] += gr[rc];
            }
        }
        bar<TPB>();
        v.fopen(0);
        // right-hand side lambda_V' of the transposed solve
        for (int i = tid; i < N; i += TPB) v.sX[v.cr[i].pos] = v.sLV[i];
        if (tid == 0) v.sX[N] = 0.0;
        bar<TPB>();
        v.hines_T();
        // membrane, channel, pulse and synapse adjoints
        for (int i = tid; i < N; i += TPB) {
            const CompRec c = v.cr[i];
            const double lb = v.sX[c.pos];
            const double lD = -lb * v.sVn[i];
            const double C = (double)c.C, A = (double)c.A;
            const double k = dt / C * K1E5 * A;
            for (int ch = c.ch_b; ch < c.ch_e; ch++) {
                const double lGs = k * (lb * (double)v.chr[ch].erev + lD);
                v.sGg[ch] += lGs * v.sF[ch];
                v.sLF[ch] = lGs * v.sGb[ch];
            }
            for (int p = c.p_b; p < c.p_e; p++) {
                const double ps = (double)v.pr[p].start, pd = (double)v.pr[p].dur;
                if (ps <= t && t <= ps + pd) v.sGa[p] += lb * dt / C * (double)v.pr[p].weight;
            }
            // synapse instances: a conductance w g_s(t) into reversal erev, like a channel with
            // no area or density factor; g_s does not depend on V or on the parameters
            for (int ls = c.s_b; ls < c.s_e; ls++) {
                const AdjSynType &st = v.sttab[v.sty[ls]];
                const double gs = st.gbase * (v.sB[ls] - v.sA[ls]);
                v.sGw[ls] += dt / C * (lb * st.erev + lD) * gs;
            }
            v.sLV[i] = lb;
        }
        bar<TPB>();
        // gate adjoints: lambda_q of step s, and each gate's term of lambda_V
        double *lqn = v.sLQ + nqmax;
        for (int kq = tid; kq < NG; kq += TPB) {
            const GateRec g = v.gr[kq];
            const ChanRec ch = v.chr[g.chan];
            const int ng = (int)(ch.shape & 7u);
            // dfopen/dq_k: product over the channel's gates with one factor q_k removed
            double df = 1.0;
            for (int m = 0; m < ng; m++) {
                const double qv = v.sQ[ch.q0 + m];
                const int inst = (int)((ch.shape >> (3 + 6 * m)) & 63u);
                if (m == g.k) {
                    double pw = (double)inst;
                    for (int r = 0; r < inst - 1; r++) pw = pw * qv;
                    df = df * (inst > 0 ? pw : 0.0);
                } else {
                    for (int r = 0; r < inst; r++) df = df * qv;
                }
            }
            const double lq1 = v.sLQ[kq];
            lqn[kq] = lq1 * v.sDQ[kq] + v.sLF[g.chan] * df;
            v.sGV[kq] = lq1 * v.sDV[kq];
        }
        bar<TPB>();
        for (int i = tid; i < N; i += TPB) {
            const CompRec c = v.cr[i];
            double lv = v.sLV[i];
            for (int ch = c.ch_b; ch < c.ch_e; ch++) {
                const ChanRec r = v.chr[ch];
                const int ng = (int)(r.shape & 7u);
                for (int m = 0; m < ng; m++) lv += v.sGV[r.q0 + m];
            }
            v.sLV[i] = lv;
            v.sVn[i] = v.sV[i];
        }
        for (int kq = tid; kq < NG; kq += TPB) v.sLQ[kq] = lqn[kq];
        bar<TPB>();
    }
    for (int i = tid; i < N; i += TPB) lam[cell.comp_begin + i] = v.sLV[i];
    for (int k = tid; k < NG; k += TPB) lam[nc_total + cell.gate_begin + k] = v.sLQ[k];
    for (int c = tid; c < cell.n_chan; c += TPB) grad_g[cell.chan_begin + c] += v.sGg[c];
    for (int p = tid; p < cell.n_pulse; p += TPB) grad_a[cell.pulse_begin + p] += v.sGa[p];
    for (int ls = tid; ls < NS; ls += TPB) grad_w[cell.syn_begin + ls] += v.sGw[ls];
}

template <int TPB>
const void *fwd_ptr() { return reinterpret_cast<const void *>(&adj_forward<TPB>); }
template <int TPB>
const void *bwd_ptr() { return reinterpret_cast<const void *>(&adj_backward<TPB>); }

}  // namespace
}  // namespace edenhip

using namespace edenhip;

struct edenhip_adj {
    std::string err;
    int device = 0;
    hipStream_t stream = nullptr;
    int tpb = 64;
    int32_t K = 500;
    int64_t rows = 0, n_record = 0, NC = 0, NQ = 0, NCH = 0, NP = 0, NSYN = 0;
    int32_t n_cells = 0, nqmax = 0;
    AdjTables T{};
    AdjLds L{};
    size_t lds = 0;
    std::vector<void *> allocs;
    double *d_v0 = nullptr, *d_q0 = nullptr, *d_gbase0 = nullptr, *d_amp0 = nullptr, *d_w0 = nullptr;
    int32_t *d_ev_begin = nullptr;
    int64_t *d_ev_step = nullptr;
    std::vector<void *> ev_allocs;        // the event schedule's tables (replaced by set_events)
    const double *last_w = nullptr;
    double *d_ck = nullptr, *d_seg = nullptr, *d_lam = nullptr;
    double *d_hist = nullptr;          // state history (store mode), n_seg segments
    int64_t hist_stride = 0, n_seg = 0;
    int32_t store = 0;                 // options.store_states
    int64_t *d_seg_off = nullptr;
    int64_t n_ck = 0, ck_stride = 0;
    const double *last_gbase = nullptr, *last_amp = nullptr;
    bool have_forward = false;
    double fwd_ms = 0, bwd_ms = 0;
};

namespace {

int afail(edenhip_adj *a, int code, const std::string &msg) {
    if (a) a->err = msg;
    return code;
}

#define AHIP(a, call)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess) return afail((a), EDENHIP_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
int aupload(edenhip_adj *a, T **dptr, const std::vector<T> &v) {
    const size_t bytes = std::max<size_t>(v.size(), 1) * sizeof(T);
    void *p = nullptr;
    AHIP(a, hipMalloc(&p, bytes));
    a->allocs.push_back(p);
    if (!v.empty()) AHIP(a, hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    else AHIP(a, hipMemset(p, 0, bytes));
    *dptr = static_cast<T *>(p);
    return 0;
}

int aalloc(edenhip_adj *a, void **dptr, size_t bytes) {
    void *p = nullptr;
    AHIP(a, hipMalloc(&p, std::max<size_t>(bytes, 8)));
    a->allocs.push_back(p);
    *dptr = p;
    return 0;
}

#define ATRY(x)                 \
    do {                        \
        int r_ = (x);           \
        if (r_) return r_;      \
    } while (0)

const void *fwd_kernel(int tpb) {
    switch (tpb) {
    case 64: return fwd_ptr<64>();
    case 128: return fwd_ptr<128>();
    case 256: return fwd_ptr<256>();
    default: return nullptr;
    }
}

const void *bwd_kernel(int tpb) {
    switch (tpb) {
    case 64: return bwd_ptr<64>();
    case 128: return bwd_ptr<128>();
    case 256: return bwd_ptr<256>();
    default: return nullptr;
    }
}

int build_adj(edenhip_adj *a, const edenhip_model_desc *d) {
    if (d->abi_version < 1 || d->abi_version > EDENHIP_ABI_VERSION)
        return afail(a, EDENHIP_E_INVALID, "descriptor ABI version mismatch");
    if (d->n_cells <= 0 || d->n_comps <= 0 || !(d->dt > 0.0f)) return afail(a, EDENHIP_E_INVALID, "empty model or dt <= 0");
    const int64_t NC = d->n_comps;
    if (d->comp_poisson_begin[NC] != 0)
        return afail(a, EDENHIP_E_INVALID, "the adjoint supports channels, pulse inputs and expTwo synapses (no Poisson inputs)");
    if (d->abi_version >= 2 && d->lems && d->lems->n_types > 0)
        return afail(a, EDENHIP_E_INVALID, "the adjoint supports expTwo synapses only (no generic LEMS synapse types)");
    a->NSYN = d->comp_syn_begin[NC];
    if (a->NSYN >= ((int64_t)1 << 31)) return afail(a, EDENHIP_E_LIMIT, "more than 2^31 synapses");
    for (int64_t sy = 0; sy < a->NSYN; sy++)
        if (d->syn_type[sy] < 0 || d->syn_type[sy] >= d->n_syn_types) return afail(a, EDENHIP_E_INVALID, "synapse type index out of range");
    a->NC = NC;
    a->NCH = d->comp_chan_begin[NC];
    a->NP = d->comp_pulse_begin[NC];
    a->n_cells = d->n_cells;
    const double dt = (double)d->dt;
    a->T.dt = dt;
    // rows and row times, accumulated in double as the engine loop does
    std::vector<double> trow;
    {
        double time = 0.0;
        while (time <= d->t_final) {
            trow.push_back(time);
            time += (double)d->dt;
        }
    }
    a->rows = (int64_t)trow.size();
    trow.push_back(trow.empty() ? 0.0 : trow.back() + dt);
    a->T.rows = a->rows;
    // channel types: the gate table (4 entries per type) and each type's gate shape word
    const int n_types = std::max(1, d->n_channel_types);
    if ((int64_t)n_types * 4 > 65535) return afail(a, EDENHIP_E_LIMIT, "more than 16,383 channel types");
    std::vector<AdjGate> gates((size_t)n_types * 4);
    std::vector<uint32_t> type_shape((size_t)n_types, 0u);
    for (int ty = 0; ty < d->n_channel_types; ty++) {
        const edenhip_channel_type &ct = d->channel_types[ty];
        if (ct.n_gates < 0 || ct.n_gates > 4) return afail(a, EDENHIP_E_LIMIT, "channel type with more than 4 gates");
        uint32_t shape = (uint32_t)ct.n_gates;
        for (int k = 0; k < ct.n_gates; k++) {
            const edenhip_gate &g = ct.gates[k];
            if (g.instances < 0 || g.instances > 63) return afail(a, EDENHIP_E_LIMIT, "gate with more than 63 instances");
            AdjGate x{};
            x.alpha_form = g.alpha.form; x.beta_form = g.beta.form; x.instances = g.instances;
            x.q10 = (double)g.q10;
            x.ar = (double)g.alpha.rate; x.am = (double)g.alpha.midpoint; x.as = (double)g.alpha.scale;
            x.br = (double)g.beta.rate; x.bm = (double)g.beta.midpoint; x.bs = (double)g.beta.scale;
            gates[(size_t)ty * 4 + k] = x;
            shape |= (uint32_t)g.instances << (3 + 6 * k);
        }
        type_shape[(size_t)ty] = shape;
    }
    // recorded compartments
    if (d->abi_version >= 4 && d->gate_kind)
        for (int64_t k = 0; k < (int64_t)d->n_channel_types * EDENHIP_MAX_GATES; k++)
            if (d->gate_kind[k] != EDENHIP_GATE_RATES)
                return afail(a, EDENHIP_E_INVALID, "the adjoint model takes alpha/beta rate gates only");
    a->n_record = d->n_record;
    a->T.n_record = (int32_t)d->n_record;
    std::vector<int32_t> rec(NC, -1);
    for (int64_t r = 0; r < d->n_record; r++) {
        const int64_t g = d->record_comp[r];
        if (g < 0 || g >= NC) return afail(a, EDENHIP_E_INVALID, "record_comp out of range");
        if (rec[g] >= 0) return afail(a, EDENHIP_E_INVALID, "a compartment may be recorded once in the adjoint model");
        if (d->abi_version >= 3 && d->record_gate && d->record_gate[r] != -1)
            return afail(a, EDENHIP_E_INVALID, "the adjoint model records membrane voltages only (record_gate -1)");
        rec[g] = (int32_t)r;
    }
    // synapses: types in double with the waveform factor evaluated in double with the
    // reference's expression tree (as the oracle's FP64 flavour per step)
    std::vector<int32_t> sty(std::max<int64_t>(a->NSYN, 1), 0);
    for (int64_t sy = 0; sy < a->NSYN; sy++) sty[sy] = d->syn_type[sy];
    const int n_stypes = std::max(1, d->n_syn_types);
    std::vector<AdjSynType> stypes((size_t)n_stypes);
    for (int t = 0; t < d->n_syn_types; t++) {
        const edenhip_exptwo_type &e = d->syn_types[t];
        const double tr = (double)e.tau_rise, td = (double)e.tau_decay;
        const double peak = (std::log(td / tr) * (tr * td)) / (td - tr);
        const double wf = (double)1 / ((-std::exp((-peak) / tr)) + std::exp((-peak) / td));
        stypes[t] = AdjSynType{(double)e.gbase, (double)e.erev, tr, td, (double)e.weight_prop, wf};
    }
    std::vector<double> w0(std::max<int64_t>(a->NSYN, 1), 0.0);
    for (int64_t sy = 0; sy < a->NSYN; sy++) w0[sy] = (double)d->syn_weight[sy];
    // per compartment / channel / gate / pulse records and initial values
    std::vector<CompRec> comp(NC);
    std::vector<ChanRec> chan(std::max<int64_t>(a->NCH, 1));
    std::vector<GateRec> gate;
    std::vector<PulseRec> pulse(std::max<int64_t>(a->NP, 1));
    std::vector<double> v0(NC), gb0(std::max<int64_t>(a->NCH, 1)), amp0(std::max<int64_t>(a->NP, 1)), q0;
    std::vector<double2> u(NC, make_double2(0.0, 0.0));
    for (int64_t g = 0; g < NC; g++) v0[g] = (double)d->v0[g];
    for (int64_t p = 0; p < a->NP; p++) {
        pulse[p] = PulseRec{d->pulse_start[p], d->pulse_duration[p], d->pulse_weight[p], 0.0f};
        amp0[p] = (double)d->pulse_amplitude[p];
    }
    for (int64_t ch = 0; ch < a->NCH; ch++) gb0[ch] = (double)d->chan_gbase[ch];
    int32_t nsmax = 1;
    std::vector<AdjCell> cells(d->n_cells);
    std::vector<uint32_t> words;
    std::vector<int64_t> seg_off(d->n_cells);
    int64_t seg_total = 0;
    int32_t nmax = 1, nqmax = 1, nchmax = 1, npmax = 1, wmax = 1;
    int tpb = 64;
    for (int c = 0; c < d->n_cells; c++) {
        const int64_t c0 = d->cell_comp_begin[c], c1 = d->cell_comp_begin[c + 1];
        const int32_t n = (int32_t)(c1 - c0);
        if (n > MAX_TREE_N) return afail(a, EDENHIP_E_LIMIT, "cell with more than 2047 compartments");
        nmax = std::max(nmax, n);
    }
    // workgroup size: enough lanes for a cell's compartments and gates
    for (int c = 0; c < d->n_cells; c++) {
        const int64_t c0 = d->cell_comp_begin[c], c1 = d->cell_comp_begin[c + 1];
        int64_t ngate = 0;
        for (int64_t ch = d->comp_chan_begin[c0]; ch < d->comp_chan_begin[c1]; ch++)
            ngate += d->channel_types[d->chan_type[ch]].n_gates;
        const int64_t want = std::max<int64_t>(c1 - c0, ngate / 2);
        tpb = std::max(tpb, want > 128 ? 256 : want > 64 ? 128 : 64);
    }
    tpb = std::min(tpb, 256);
    // opt overrides applied by the caller
    a->tpb = a->tpb > 0 ? a->tpb : tpb;
    for (int c = 0; c < d->n_cells; c++) {
        const int64_t c0 = d->cell_comp_begin[c], c1 = d->cell_comp_begin[c + 1];
        const int32_t n = (int32_t)(c1 - c0);
        if (n <= 0) return afail(a, EDENHIP_E_INVALID, "cell without compartments");
        if (d->parent[c0] != -1) return afail(a, EDENHIP_E_INVALID, "compartment 0 of a cell must be its root");
        for (int32_t i = 1; i < n; i++)
            if (d->parent[c0 + i] < 0 || d->parent[c0 + i] >= i) return afail(a, EDENHIP_E_INVALID, "parents must precede children");
        AdjCell cell{};
        cell.comp_begin = c0;
        cell.n = n;
        // rounds of at most 64 tasks: one wave runs the whole solve
        const TreePlan P = plan_tree(n, d->parent + c0, 64);
        if (!P.ok) return afail(a, EDENHIP_E_LIMIT, "the adjoint supports at most 1,023 children per compartment");
        cell.rounds_fwd = n > 1 ? P.rounds_fwd : 0;
        cell.rounds_bwd = n > 1 ? P.rounds_bwd : 0;
        cell.words_off = (int32_t)words.size();
        cell.words_len = (int32_t)P.block.size();
        words.insert(words.end(), P.block.begin(), P.block.end());
        wmax = std::max(wmax, cell.words_len);
        std::vector<char> has_child((size_t)n, 0);
        for (int32_t i = 1; i < n; i++) {
            has_child[(size_t)d->parent[c0 + i]] = 1;
            // Ui = -dt/(R C_i), Uj = -dt/(R C_parent) with R = R_axial of the child (eden/Eden.cpp:6136-6150)
            const double R = (double)d->r_axial[c0 + i];
            const int32_t j = d->parent[c0 + i];
            u[c0 + P.pos[i]] = make_double2(-dt / (R * (double)d->capacitance[c0 + i]), -dt / (R * (double)d->capacitance[c0 + j]));
        }
        cell.chan_begin = (int32_t)d->comp_chan_begin[c0];
        cell.n_chan = (int32_t)(d->comp_chan_begin[c1] - d->comp_chan_begin[c0]);
        cell.pulse_begin = (int32_t)d->comp_pulse_begin[c0];
        cell.n_pulse = (int32_t)(d->comp_pulse_begin[c1] - d->comp_pulse_begin[c0]);
        cell.syn_begin = (int32_t)d->comp_syn_begin[c0];
        cell.n_syn = (int32_t)(d->comp_syn_begin[c1] - d->comp_syn_begin[c0]);
        if (cell.n_chan > 65535 || cell.n_pulse > 65535 || cell.n_syn > 65535)
            return afail(a, EDENHIP_E_LIMIT, "more than 65,535 channel distributions, pulse inputs or synapses on one cell");
        cell.gate_begin = (int32_t)q0.size();
        for (int64_t g = c0; g < c1; g++) {
            CompRec r{};
            r.A = d->area[g];
            r.C = d->capacitance[g];
            r.invrc = d->inv_rc_axial[g];
            r.rec_col = rec[g];
            r.pos = (uint16_t)P.pos[g - c0];
            r.leaf = has_child[(size_t)(g - c0)] ? 0 : 1;
            r.ch_b = (uint16_t)(d->comp_chan_begin[g] - cell.chan_begin);
            r.ch_e = (uint16_t)(d->comp_chan_begin[g + 1] - cell.chan_begin);
            r.p_b = (uint16_t)(d->comp_pulse_begin[g] - cell.pulse_begin);
            r.p_e = (uint16_t)(d->comp_pulse_begin[g + 1] - cell.pulse_begin);
            r.s_b = (uint16_t)(d->comp_syn_begin[g] - cell.syn_begin);
            r.s_e = (uint16_t)(d->comp_syn_begin[g + 1] - cell.syn_begin);
            comp[g] = r;
            for (int64_t ch = d->comp_chan_begin[g]; ch < d->comp_chan_begin[g + 1]; ch++) {
                const int ty = d->chan_type[ch];
                if (ty < 0 || ty >= d->n_channel_types) return afail(a, EDENHIP_E_INVALID, "channel type index out of range");
                ChanRec x{};
                x.erev = d->chan_erev[ch];
                x.comp = (uint16_t)(g - c0);
                x.q0 = (uint16_t)(q0.size() - (size_t)cell.gate_begin);
                x.shape = type_shape[(size_t)ty];
                chan[ch] = x;
                for (int k = 0; k < d->channel_types[ty].n_gates; k++) {
                    const edenhip_gate &gt = d->channel_types[ty].gates[k];
                    const double al = edenhip_rate_host(gt.alpha, d->v0[g]), be = edenhip_rate_host(gt.beta, d->v0[g]);
                    q0.push_back(al / (al + be));
                    gate.push_back(GateRec{(uint16_t)(g - c0), (uint16_t)(ch - cell.chan_begin), (uint16_t)(ty * 4 + k), (uint16_t)k});
                }
            }
        }
        cell.n_gate = (int32_t)(q0.size() - (size_t)cell.gate_begin);
        if (cell.n_gate > 65535) return afail(a, EDENHIP_E_LIMIT, "more than 65,535 gate variables on one cell");
        nsmax = std::max(nsmax, cell.n_syn);
        nqmax = std::max(nqmax, cell.n_gate);
        nchmax = std::max(nchmax, cell.n_chan);
        npmax = std::max(npmax, cell.n_pulse);
        cells[c] = cell;
        seg_off[c] = seg_total;
        seg_total += rec_layout(n, cell.n_gate, cell.n_syn).per;
    }
    a->NQ = (int64_t)q0.size();
    a->nqmax = nqmax;
    a->L = adj_lds(nmax, nqmax, nchmax, npmax, wmax, nsmax, n_types * 4, n_stypes);
    a->lds = (size_t)a->L.total * sizeof(double);
    if (a->lds > 160 * 1024) {
        char msg[160];
        std::snprintf(msg, sizeof msg, "cell too large for the adjoint kernels' LDS (%zu bytes needed, 163,840 available)", a->lds);
        return afail(a, EDENHIP_E_LIMIT, msg);
    }
    a->K = std::max(1, a->K);
    const int64_t steps = a->rows;   // steps 1..rows-1 are differentiated
    a->n_ck = (steps + a->K - 1) / a->K + 1;
    a->ck_stride = NC + a->NQ;
    // segment buffer: (K + 1) records per cell
    for (auto &o : seg_off) o *= (int64_t)(a->K + 1);
    AdjTables &T = a->T;
    T.n_gate_params = n_types * 4;
    T.n_syn_types = n_stypes;
    ATRY(aupload(a, (AdjCell **)&T.cells, cells));
    ATRY(aupload(a, (CompRec **)&T.comp, comp));
    ATRY(aupload(a, (ChanRec **)&T.chan, chan));
    ATRY(aupload(a, (GateRec **)&T.gate, gate));
    ATRY(aupload(a, (PulseRec **)&T.pulse, pulse));
    ATRY(aupload(a, (AdjGate **)&T.gates, gates));
    ATRY(aupload(a, (double2 **)&T.u, u));
    ATRY(aupload(a, (uint32_t **)&T.words, words));
    ATRY(aupload(a, (double **)&T.t_row, trow));
    ATRY(aupload(a, (int32_t **)&T.s_type, sty));
    ATRY(aupload(a, (AdjSynType **)&T.s_types, stypes));
    ATRY(aupload(a, &a->d_w0, w0));
    // no events until edenhip_adj_set_events
    ATRY(aupload(a, &a->d_ev_begin, std::vector<int32_t>((size_t)a->NSYN + 1, 0)));
    ATRY(aupload(a, &a->d_ev_step, std::vector<int64_t>(1, 0)));
    T.ev_begin = a->d_ev_begin;
    T.ev_step = a->d_ev_step;
    ATRY(aupload(a, &a->d_v0, v0));
    ATRY(aupload(a, &a->d_q0, q0));
    ATRY(aupload(a, &a->d_gbase0, gb0));
    ATRY(aupload(a, &a->d_amp0, amp0));
    ATRY(aupload(a, &a->d_seg_off, seg_off));
    ATRY(aalloc(a, (void **)&a->d_ck, (size_t)(a->n_ck * a->ck_stride) * sizeof(double)));
    ATRY(aalloc(a, (void **)&a->d_seg, (size_t)(seg_total * (a->K + 1)) * sizeof(double)));
    // history of every step (store mode): n_seg segments of seg_total (K + 1) doubles
    a->hist_stride = seg_total * (a->K + 1);
    a->n_seg = a->rows >= 2 ? (a->rows - 2) / a->K + 1 : 0;
    {
        const size_t bytes = (size_t)(a->hist_stride * std::max<int64_t>(a->n_seg, 1)) * sizeof(double);
        size_t free_b = 0, total_b = 0;
        AHIP(a, hipMemGetInfo(&free_b, &total_b));
        const bool want = a->store > 0 || (a->store == 0 && bytes <= free_b / 2);
        if (a->store > 0 && bytes > free_b) return afail(a, EDENHIP_E_LIMIT, "state history does not fit in device memory");
        if (want && a->n_seg > 0) ATRY(aalloc(a, (void **)&a->d_hist, bytes));
    }
    ATRY(aalloc(a, (void **)&a->d_lam, (size_t)a->ck_stride * sizeof(double)));
    if (!fwd_kernel(a->tpb) || !bwd_kernel(a->tpb)) return afail(a, EDENHIP_E_LIMIT, "unsupported threads_per_cell");
    return 0;
}

}  // namespace

extern "C" {

int edenhip_adj_create(const edenhip_model_desc *desc, const edenhip_adj_options *opt, edenhip_adj **out) {
    if (!out) return EDENHIP_E_INVALID;
    *out = nullptr;
    edenhip_adj *a = new (std::nothrow) edenhip_adj();
    if (!a) return EDENHIP_E_INVALID;
    *out = a;
    if (!desc) return afail(a, EDENHIP_E_INVALID, "null descriptor");
    if (opt) {
        a->K = opt->checkpoint_steps > 0 ? opt->checkpoint_steps : 500;
        a->device = opt->device;
        a->store = opt->store_states;
        if (opt->threads_per_cell) {
            if (opt->threads_per_cell != 64 && opt->threads_per_cell != 128 && opt->threads_per_cell != 256)
                return afail(a, EDENHIP_E_INVALID, "threads_per_cell must be 0, 64, 128 or 256");
            a->tpb = opt->threads_per_cell;
        } else {
            a->tpb = 0;
        }
    } else {
        a->tpb = 0;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return afail(a, EDENHIP_E_DEVICE, "no HIP device visible");
    AHIP(a, hipSetDevice(a->device));
    AHIP(a, hipStreamCreateWithFlags(&a->stream, hipStreamNonBlocking));
    return build_adj(a, desc);
}

int edenhip_adj_set_events(edenhip_adj *a, const int64_t *step, const int64_t *syn, int64_t n) {
    if (!a || n < 0 || (n > 0 && (!step || !syn))) return EDENHIP_E_INVALID;
    if (!a->stream) return afail(a, EDENHIP_E_STATE, "adjoint context was not created successfully");
    if (n >= ((int64_t)1 << 31)) return afail(a, EDENHIP_E_LIMIT, "more than 2^31 - 1 events");
    AHIP(a, hipSetDevice(a->device));
    // per synapse, its event steps in ascending order
    std::vector<int32_t> begin((size_t)a->NSYN + 1, 0);
    for (int64_t k = 0; k < n; k++) {
        if (syn[k] < 0 || syn[k] >= a->NSYN) return afail(a, EDENHIP_E_INVALID, "event synapse index out of range");
        if (step[k] < 1) return afail(a, EDENHIP_E_INVALID, "event steps count from 1");
        begin[(size_t)syn[k] + 1]++;
    }
    for (int64_t sy = 0; sy < a->NSYN; sy++) begin[(size_t)sy + 1] += begin[(size_t)sy];
    std::vector<int64_t> steps((size_t)std::max<int64_t>(n, 1), 0);
    std::vector<int32_t> fill(begin.begin(), begin.end() - 1);
    for (int64_t k = 0; k < n; k++) steps[(size_t)fill[(size_t)syn[k]]++] = step[k];
    for (int64_t sy = 0; sy < a->NSYN; sy++) {
        std::sort(steps.begin() + begin[(size_t)sy], steps.begin() + begin[(size_t)sy + 1]);
        for (int32_t k = begin[(size_t)sy] + 1; k < begin[(size_t)sy + 1]; k++)
            if (steps[(size_t)k] == steps[(size_t)k - 1]) return afail(a, EDENHIP_E_INVALID, "two events of one synapse at one step");
    }
    for (void *p : a->ev_allocs) (void)hipFree(p);
    a->ev_allocs.clear();
    void *pb = nullptr, *ps = nullptr;
    AHIP(a, hipMalloc(&pb, begin.size() * sizeof(int32_t)));
    a->ev_allocs.push_back(pb);
    AHIP(a, hipMalloc(&ps, steps.size() * sizeof(int64_t)));
    a->ev_allocs.push_back(ps);
    AHIP(a, hipMemcpy(pb, begin.data(), begin.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    AHIP(a, hipMemcpy(ps, steps.data(), steps.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    a->T.ev_begin = (const int32_t *)pb;
    a->T.ev_step = (const int64_t *)ps;
    a->have_forward = false;
    return 0;
}

int edenhip_adj_n_synapses(const edenhip_adj *a, int64_t *n_syn) {
    if (!a || !n_syn) return EDENHIP_E_INVALID;
    *n_syn = a->NSYN;
    return 0;
}

int edenhip_adj_sizes(const edenhip_adj *a, int64_t *rows, int64_t *n_record, int64_t *n_chan, int64_t *n_pulse) {
    if (!a) return EDENHIP_E_INVALID;
    if (rows) *rows = a->rows;
    if (n_record) *n_record = a->n_record;
    if (n_chan) *n_chan = a->NCH;
    if (n_pulse) *n_pulse = a->NP;
    return 0;
}

int edenhip_adj_forward_w(edenhip_adj *a, const double *gbase, const double *amp, const double *weight, double *trace) {
    if (!a || !trace) return EDENHIP_E_INVALID;
    if (!a->stream) return afail(a, EDENHIP_E_STATE, "adjoint context was not created successfully");
    AHIP(a, hipSetDevice(a->device));
    const double *gb = gbase ? gbase : a->d_gbase0;
    const double *am = amp ? amp : a->d_amp0;
    const double *wt = weight ? weight : a->d_w0;
    if (a->NSYN > 0 && !a->d_hist)
        return afail(a, EDENHIP_E_STATE, "synapses need the state history in device memory (store_states)");
    hipEvent_t e0, e1;
    AHIP(a, hipEventCreate(&e0));
    AHIP(a, hipEventCreate(&e1));
    AHIP(a, hipEventRecord(e0, a->stream));
    AdjTables T = a->T;
    int32_t K = a->K, nq = a->nqmax;
    int64_t cks = a->ck_stride, nct = a->NC;
    AdjLds L = a->L;
    int64_t hs = a->hist_stride, ns = a->n_seg;
    void *args[] = {&T, &a->d_v0, &a->d_q0, (void *)&gb, (void *)&am, (void *)&wt, &trace, &a->d_ck, &cks, &nct, &K, &L, &nq,
                    &a->d_hist, &hs, &ns, &a->d_seg_off};
    AHIP(a, hipLaunchKernel(fwd_kernel(a->tpb), dim3(a->n_cells), dim3(a->tpb), args, a->lds, a->stream));
    AHIP(a, hipEventRecord(e1, a->stream));
    AHIP(a, hipEventSynchronize(e1));
    float ms = 0;
    AHIP(a, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    a->fwd_ms = ms;
    a->last_gbase = gb;
    a->last_amp = am;
    a->last_w = wt;
    a->have_forward = true;
    return 0;
}

int edenhip_adj_forward(edenhip_adj *a, const double *gbase, const double *amp, double *trace) {
    return edenhip_adj_forward_w(a, gbase, amp, nullptr, trace);
}

int edenhip_adj_backward_w(edenhip_adj *a, const double *grad_trace, double *grad_gbase, double *grad_amp,
                           double *grad_weight) {
    if (!a || !grad_trace) return EDENHIP_E_INVALID;
    if (!a->stream) return afail(a, EDENHIP_E_STATE, "adjoint context was not created successfully");
    if (!a->have_forward) return afail(a, EDENHIP_E_STATE, "backward before forward");
    AHIP(a, hipSetDevice(a->device));
    double *gg = grad_gbase, *ga = grad_amp, *gw = grad_weight;
    void *tmp_g = nullptr, *tmp_a = nullptr, *tmp_w = nullptr;
    if (!gw) {
        AHIP(a, hipMalloc(&tmp_w, (size_t)std::max<int64_t>(a->NSYN, 1) * sizeof(double)));
        gw = (double *)tmp_w;
    }
    if (!gg) {
        AHIP(a, hipMalloc(&tmp_g, (size_t)std::max<int64_t>(a->NCH, 1) * sizeof(double)));
        gg = (double *)tmp_g;
    }
    if (!ga) {
        AHIP(a, hipMalloc(&tmp_a, (size_t)std::max<int64_t>(a->NP, 1) * sizeof(double)));
        ga = (double *)tmp_a;
    }
    hipEvent_t e0, e1;
    AHIP(a, hipEventCreate(&e0));
    AHIP(a, hipEventCreate(&e1));
    AHIP(a, hipEventRecord(e0, a->stream));
    AHIP(a, hipMemsetAsync(gg, 0, (size_t)std::max<int64_t>(a->NCH, 1) * sizeof(double), a->stream));
    AHIP(a, hipMemsetAsync(ga, 0, (size_t)std::max<int64_t>(a->NP, 1) * sizeof(double), a->stream));
    AHIP(a, hipMemsetAsync(gw, 0, (size_t)std::max<int64_t>(a->NSYN, 1) * sizeof(double), a->stream));
    AHIP(a, hipMemsetAsync(a->d_lam, 0, (size_t)a->ck_stride * sizeof(double), a->stream));
    AdjTables T = a->T;
    int32_t nq = a->nqmax;
    int64_t nct = a->NC;
    AdjLds L = a->L;
    // differentiated steps: 1 .. rows-1 (step rows only produces the state after the last row)
    const int64_t last = a->rows;   // segments cover [1, rows)
    for (int64_t k = (last - 2) / a->K; k >= 0; k--) {
        int64_t sa = 1 + k * a->K, sb = std::min<int64_t>(1 + (k + 1) * a->K, last);
        if (sa >= sb) continue;
        const double *ckp = a->d_ck + k * a->ck_stride;
        int32_t stored = a->d_hist ? 1 : 0;
        double *segp = a->d_hist ? a->d_hist + k * a->hist_stride : a->d_seg;
        void *args[] = {&T, (void *)&a->last_gbase, (void *)&a->last_amp, (void *)&a->last_w, (void *)&grad_trace,
                        (void *)&ckp, &nct, &sa, &sb, &segp, &a->d_seg_off, &a->d_lam, &gg, &ga, &gw, &L, &nq, &stored};
        AHIP(a, hipLaunchKernel(bwd_kernel(a->tpb), dim3(a->n_cells), dim3(a->tpb), args, a->lds, a->stream));
    }
    AHIP(a, hipEventRecord(e1, a->stream));
    AHIP(a, hipEventSynchronize(e1));
    float ms = 0;
    AHIP(a, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    a->bwd_ms = ms;
    if (tmp_g) (void)hipFree(tmp_g);
    if (tmp_a) (void)hipFree(tmp_a);
    if (tmp_w) (void)hipFree(tmp_w);
    return 0;
}

int edenhip_adj_backward(edenhip_adj *a, const double *grad_trace, double *grad_gbase, double *grad_amp) {
    return edenhip_adj_backward_w(a, grad_trace, grad_gbase, grad_amp, nullptr);
}

int edenhip_adj_timing(const edenhip_adj *a, double *forward_ms, double *backward_ms) {
    if (!a) return EDENHIP_E_INVALID;
    if (forward_ms) *forward_ms = a->fwd_ms;
    if (backward_ms) *backward_ms = a->bwd_ms;
    return 0;
}

const char *edenhip_adj_last_error(const edenhip_adj *a) {
    if (!a) return "null adjoint context";
    return a->err.c_str();
}

void edenhip_adj_destroy(edenhip_adj *a) {
    if (!a) return;
    if (a->stream) {
        (void)hipSetDevice(a->device);
        (void)hipStreamSynchronize(a->stream);
        for (void *p : a->allocs) (void)hipFree(p);
        for (void *p : a->ev_allocs) (void)hipFree(p);
        (void)hipStreamDestroy(a->stream);
    }
    delete a;
}

}  // extern "C"
