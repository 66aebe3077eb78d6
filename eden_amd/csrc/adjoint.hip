// adjoint.hip -- discrete adjoint of the fixed-step integrator (include/edenhip_adjoint.h).
//
// Everything here is double precision and follows the operation order of the reference's
// per-step update (as the FP64 restatement in oracle/ does):
//   gates          eden/Eden.cpp:5236-5259, 5336-5375 (rates, tau, inf, cnexp, clipping)
//   channels       eden/Eden.cpp:5679-5721, 5805-5898 (fopen, Gbase*fopen*(E-V)*1e-5f*A)
//   pulses         eden/Eden.cpp:3798-3826 (inclusive window, Imax * weight)
//   membrane       eden/Eden.cpp:6057-6107 (D = G/C + invRC, V' = V + dt(I + V G)/C)
//   Hines          eden/Eden.cpp:6114-6225 (D = 1 + D dt, elimination in DFS order)
// The Hines solve runs in the breadth-first rounds of tree_schedule.hpp, which apply each
// node's operations in the reference's order.
//
// One workgroup per cell. adj_forward runs the whole time loop, writes the trace and a
// checkpoint of (V, q) at the start of every K-th recorded step; when the state history fits
// in device memory it also writes (V, q) of every step, in the segment layout the backward
// pass reads. adj_backward handles one segment of K steps: it takes the segment's states from
// that history, or recomputes them from the segment's checkpoint into a segment buffer, then
// walks the steps backwards:
//   lambda_b = M^-T lambda_V'   (Hines on the transposed matrix: Ui and Uj swap roles)
//   lambda_D = -lambda_b V'
//   per channel c of compartment i:  dL/dGscaled = dt/C * 1e-5f * A * (lambda_b E_c + lambda_D)
//       dL/dGbase += dL/dGscaled * fopen,   dL/dfopen = dL/dGscaled * Gbase
//   per pulse p active at t:          dL/dAmplitude += lambda_b dt / C * weight
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

struct AdjTables {
    double dt;
    int64_t rows;
    int32_t n_record;
    const AdjCell *cells;
    // per compartment
    const double *C, *A, *invrc;
    const int32_t *pos, *ch_begin, *p_begin, *rec_col;
    // per channel distribution
    const int32_t *ch_type, *ch_gate, *ch_comp;
    const double *ch_erev;
    // per gate variable
    const int32_t *g_chan, *g_k;
    // channel types: gate count, gate table (4 per type)
    const int32_t *type_ngates;
    const AdjGate *gates;
    // pulses
    const double *p_start, *p_dur, *p_weight;
    // edge constants (Ui, Uj) by global position, schedule blocks, row times
    const double2 *u;
    const uint32_t *words;
    const double *t_row;
    // chemical synapses (expTwo) with a fixed event schedule: per compartment CSR of
    // synapse instances, per instance its type, per instance the sorted steps at which it
    // receives its event
    const int32_t *s_begin;
    const int32_t *s_type;
    const AdjSynType *s_types;
    const int32_t *ev_begin;
    const int64_t *ev_step;
};

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

// rate value and its derivative in V
__device__ __forceinline__ double rate64(int form, double r, double m, double s, double V, double &dv) {
    switch (form) {
    case EDENHIP_RATE_EXP: {
        const double v = r * exp((V - m) / s);
        dv = v / s;
        return v;
    }
    case EDENHIP_RATE_EXPLINEAR: {
        const double x = (V - m) / s;
        if (V == m) {
            dv = r / (2.0 * s);
            return r;
        }
        const double em = exp(-(V - m) / s);
        const double v = r * (x / (1.0 - em));
        // d/dx x/(1 - e^-x) = ((1 - e^-x) - x e^-x) / (1 - e^-x)^2, with its series near 0
        const double den = 1.0 - em;
        const double gx = fabs(x) < 1e-5 ? 0.5 + x / 6.0 : ((den - x * em) / (den * den));
        dv = r * gx / s;
        return v;
    }
    case EDENHIP_RATE_SIGMOID: {
        const double e = exp((m - V) / s);
        const double v = r / (1.0 + e);
        dv = r * e / (s * (1.0 + e) * (1.0 + e));
        return v;
    }
    default:
        dv = 0.0;
        return r;
    }
}

// next gate value (the reference's order, widened), and, when asked, the partial
// derivatives of the update in q and V (zero when the result was clipped)
__device__ __forceinline__ double gate64(const AdjGate &G, double V, double q, bool init, double dt, double *dq,
                                         double *dV) {
    double da, db;
    const double alpha = rate64(G.alpha_form, G.ar, G.am, G.as, V, da);
    const double beta = rate64(G.beta_form, G.br, G.bm, G.bs, V, db);
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
    int32_t v, q, f, lf, d, x, u, lv, lq, gv, gg, ga, sa, sb, sc, gw, w, total;
};

__host__ __device__ inline AdjLds adj_lds(int32_t nmax, int32_t nqmax, int32_t nchmax, int32_t npmax, int32_t wmax,
                                          int32_t nsmax) {
    AdjLds L;
    L.v = 0;
    L.q = L.v + nmax;
    L.f = L.q + 2 * nqmax;
    L.lf = L.f + nchmax;
    L.d = L.lf + nchmax;
    L.x = L.d + (nmax + 1);
    L.u = L.x + (nmax + 1);
    L.lv = L.u + 2 * (nmax + 1);
    L.lq = L.lv + nmax;
    L.gv = L.lq + 2 * nqmax;
    L.gg = L.gv + nqmax;
    L.ga = L.gg + nchmax;
    L.sa = L.ga + npmax;           // synapse state A, B, event cursor (as an index), weight gradient
    L.sb = L.sa + nsmax;
    L.sc = L.sb + nsmax;
    L.gw = L.sc + nsmax;
    L.w = L.gw + nsmax;
    L.total = L.w + (wmax + 1) / 2;
    return L;
}

template <int TPB>
struct CellView {
    const AdjTables &T;
    const AdjCell &cell;
    int tid, wbase;
    double *sV, *sQ, *sF, *sLF, *sD, *sX, *sLV, *sLQ, *sGV, *sGg, *sGa, *sA, *sB, *sGw;
    int64_t *sCur;
    double2 *sU;
    const uint32_t *sW;

    // Hines solve of (D, X) in place; transposed: the matrix with Ui and Uj exchanged
    __device__ void hines(bool transposed) {
        const int N = cell.n;
        const int RF = cell.rounds_fwd, RB = cell.rounds_bwd;
        const uint32_t idle_fwd = (uint32_t)N | ((uint32_t)N << 11) | (1u << 22);
        const uint32_t idle_bwd = (uint32_t)N | ((uint32_t)N << 11);
        for (int r = 0; r < RF; r++) {
            const uint32_t h = sW[r];
            const int cnt = (int)(__builtin_amdgcn_readfirstlane(h) >> 16);
            if (wbase < cnt) {
                const uint32_t w = tid < cnt ? sW[(h & 0xFFFFu) + tid] : idle_fwd;
                const int p = w & 0x7FF, fc = (w >> 11) & 0x7FF, nc = w >> 22;
                double dp = sD[p], xp = sX[p];
                for (int i = 0; i < nc; i++) {
                    const int c = fc + i;
                    const double2 u = sU[c];
                    const double ui = transposed ? u.y : u.x, uj = transposed ? u.x : u.y;
                    const double ratio = uj / sD[c];
                    dp -= ratio * ui;
                    xp -= ratio * sX[c];
                }
                sD[p] = dp;
                sX[p] = xp;
            }
            bar<TPB>();
        }
        if (tid == 0) sX[0] = sX[0] / sD[0];
        bar<TPB>();
        for (int r = 0; r < RB; r++) {
            const uint32_t h = sW[RF + 1 + r];
            const int cnt = (int)(__builtin_amdgcn_readfirstlane(h) >> 16);
            if (wbase < cnt) {
                const uint32_t w = tid < cnt ? sW[(h & 0xFFFFu) + tid] : idle_bwd;
                const int p = w & 0x7FF, pp = (w >> 11) & 0x7FF;
                const double2 u = sU[p];
                const double ui = transposed ? u.y : u.x;
                sX[p] = (sX[p] - ui * sX[pp]) / sD[p];
            }
            bar<TPB>();
        }
    }

    // open fractions of the cell's channels from the gate values in sQ[buf]
    __device__ void fopen(int buf, int nqmax) {
        const double *q = sQ + buf * nqmax;
        for (int c = tid; c < cell.n_chan; c += TPB) {
            const int ch = cell.chan_begin + c;
            const int ty = T.ch_type[ch];
            const int q0 = T.ch_gate[ch] - cell.gate_begin;
            double f = 1.0;
            for (int k = 0; k < T.type_ngates[ty]; k++) {
                const double qv = q[q0 + k];
                for (int i = 0; i < T.gates[ty * 4 + k].instances; i++) f = f * qv;
            }
            sF[c] = f;
        }
    }

    // synaptic current and conductance of compartment i from the present synapse state:
    // per instance g = gbase (B - A), i = g (erev - V), times the connection weight, summed per
    // type in list order, the types' sums in order (as the oracle's FP64 flavour)
    __device__ void synapse_sums(int i, double V, const double *weight, double &Is, double &Gs) const {
        const int64_t g = cell.comp_begin + i;
        Is = 0.0;
        Gs = 0.0;
        double Ia = 0.0, Ga = 0.0;
        int32_t cur = -1;
        for (int sy = T.s_begin[g]; sy < T.s_begin[g + 1]; sy++) {
            const int32_t ty = T.s_type[sy];
            if (ty != cur) {
                if (cur >= 0) { Is += Ia; Gs += Ga; }
                Ia = 0.0; Ga = 0.0; cur = ty;
            }
            const AdjSynType st = T.s_types[ty];
            const int ls = sy - cell.syn_begin;
            const double gs = st.gbase * (sB[ls] - sA[ls]);
            const double is = gs * (st.erev - V);
            Ia += is * weight[sy];
            Ga += gs * weight[sy];
        }
        if (cur >= 0) { Is += Ia; Gs += Ga; }
    }

    // next synapse state (in place) at step s: Euler decay, or on the instance's event the
    // OnEvent assignment; zero on initialisation steps
    __device__ void synapse_update(int64_t s, double dt) {
        for (int ls = tid; ls < cell.n_syn; ls += TPB) {
            const int sy = cell.syn_begin + ls;
            const AdjSynType st = T.s_types[T.s_type[sy]];
            const double A = sA[ls], B = sB[ls];
            double An, Bn;
            if (s <= 0) {
                An = 0.0; Bn = 0.0;
            } else {
                An = A + dt * ((-A) / st.tau_rise);
                Bn = B + dt * ((-B) / st.tau_decay);
                int64_t c = sCur[ls];
                if (c < T.ev_begin[sy + 1] && T.ev_step[c] == s) {
                    An = A + (st.weight_prop * st.wf);
                    Bn = B + (st.weight_prop * st.wf);
                    sCur[ls] = c + 1;
                }
            }
            sA[ls] = An;
            sB[ls] = Bn;
        }
    }

    // membrane currents of compartment i (local) at time t: conductance and current sums
    __device__ void currents(int i, double V, double t, const double *gbase, const double *amp, double &I,
                             double &G, const double *weight) const {
        const int64_t g = cell.comp_begin + i;
        const double A = T.A[g];
        double Ich = 0.0, Gch = 0.0;
        for (int ch = T.ch_begin[g]; ch < T.ch_begin[g + 1]; ch++) {
            const double Gs = gbase[ch] * sF[ch - cell.chan_begin];
            const double iD = Gs * (T.ch_erev[ch] - V) * K1E5;
            const double gD = (Gs) * K1E5;
            Ich += (iD * A);
            Gch += (gD * A);
        }
        double Iin = 0.0;
        if (T.p_begin[g + 1] > T.p_begin[g]) {
            double Ip = 0.0;
            for (int p = T.p_begin[g]; p < T.p_begin[g + 1]; p++)
                if (T.p_start[p] <= t && t <= T.p_start[p] + T.p_dur[p]) Ip += amp[p] * T.p_weight[p];
            Iin += Ip;
            Iin += 0.0;
        }
        double Isyn = 0.0, Gsyn = 0.0;
        if (cell.n_syn > 0) synapse_sums(i, V, weight, Isyn, Gsyn);
        I = Ich + Iin + Isyn;
        G = Gch + 0.0 + Gsyn;
    }

    // one forward step s from the state in sV / sQ[cur] into sV / sQ[1 - cur]
    __device__ void step(int64_t s, double t, int cur, int nqmax, const double *gbase, const double *amp,
                         const double *weight) {
        const bool init = s <= 0;
        const double dt = T.dt;
        const int N = cell.n;
        fopen(cur, nqmax);
        bar<TPB>();
        for (int i = tid; i < N; i += TPB) {
            const int64_t g = cell.comp_begin + i;
            const double V = sV[i];
            double I, G;
            currents(i, V, t, gbase, amp, I, G, weight);
            const double C = T.C[g];
            const double D = 1 * (G / C) + T.invrc[g];
            const int p = T.pos[g];
            sD[p] = 1 + D * dt;
            sX[p] = init ? V : V + (dt * (I + ((V * G))) / C);
        }
        const double *q = sQ + cur * nqmax;
        double *qn = sQ + (1 - cur) * nqmax;
        for (int k = tid; k < cell.n_gate; k += TPB) {
            const int gq = cell.gate_begin + k;
            const int ch = T.g_chan[gq];
            const int ty = T.ch_type[ch];
            qn[k] = gate64(T.gates[ty * 4 + T.g_k[gq]], sV[T.ch_comp[ch]], q[k], init, dt, nullptr, nullptr);
        }
        if (tid == 0) {
            sD[N] = 1.0;
            sX[N] = 0.0;
        }
        bar<TPB>();
        // the synapses' next state (read above for this step's currents)
        if (cell.n_syn > 0) synapse_update(s, dt);
        if (!init) hines(false);
        for (int i = tid; i < N; i += TPB) sV[i] = sX[T.pos[cell.comp_begin + i]];
        bar<TPB>();
    }
};

template <int TPB>
__device__ CellView<TPB> make_view(const AdjTables &T, const AdjCell &cell, unsigned char *smem, const AdjLds &L,
                                   int nqmax) {
    double *sm = reinterpret_cast<double *>(smem);
    const int tid = threadIdx.x;
    CellView<TPB> v{T, cell, tid, __builtin_amdgcn_readfirstlane(tid & ~63)};
    v.sV = sm + L.v;
    v.sQ = sm + L.q;
    v.sF = sm + L.f;
    v.sLF = sm + L.lf;
    v.sD = sm + L.d;
    v.sX = sm + L.x;
    v.sU = reinterpret_cast<double2 *>(sm + L.u);
    v.sLV = sm + L.lv;
    v.sLQ = sm + L.lq;
    v.sGV = sm + L.gv;
    v.sGg = sm + L.gg;
    v.sGa = sm + L.ga;
    v.sA = sm + L.sa;
    v.sB = sm + L.sb;
    v.sCur = reinterpret_cast<int64_t *>(sm + L.sc);
    v.sGw = sm + L.gw;
    for (int ls = tid; ls < cell.n_syn; ls += TPB) {
        v.sA[ls] = 0.0;
        v.sB[ls] = 0.0;
        v.sCur[ls] = T.ev_begin[cell.syn_begin + ls];
        v.sGw[ls] = 0.0;
    }
    uint32_t *w = reinterpret_cast<uint32_t *>(sm + L.w);
    for (int i = tid; i < cell.words_len; i += TPB) w[i] = T.words[cell.words_off + i];
    for (int i = tid; i < cell.n; i += TPB) v.sU[i] = T.u[cell.comp_begin + i];
    if (tid == 0) v.sU[cell.n] = make_double2(0.0, 0.0);
    v.sW = w;
    (void)nqmax;
    return v;
}

// Whole forward pass: initial state from v0/q0, trace rows, checkpoints of the state at the
// start of steps 1, 1 + K, 1 + 2K, ... (ck[k] = V of all compartments then q of all gates).
// hist (may be null): the state at the start of every step s >= 1, in the backward pass's
// segment layout: segment k = (s - 1) / K holds steps 1 + k K .. 1 + (k + 1) K (K + 1
// states; the last one is also the first of segment k + 1), hist_stride doubles apart, cell
// c at seg_off[c], the state of step s at (s - 1 - k K) * (N + NG) after it.
template <int TPB>
__global__ void __launch_bounds__(TPB) adj_forward(AdjTables T, int32_t n_cells, const double *__restrict__ v0,
                                                   const double *__restrict__ q0, const double *__restrict__ gbase,
                                                   const double *__restrict__ amp, double *__restrict__ trace,
                                                   double *__restrict__ ck, int64_t ck_stride, int64_t nc_total,
                                                   int32_t K, AdjLds L, int32_t nqmax, double *__restrict__ hist,
                                                   int64_t hist_stride, int64_t n_seg,
                                                   const int64_t *__restrict__ seg_off,
                                                   const double *__restrict__ weight) {
    extern __shared__ __align__(16) unsigned char smem[];
    const AdjCell cell = T.cells[blockIdx.x];
    auto v = make_view<TPB>(T, cell, smem, L, nqmax);
    const int tid = threadIdx.x;
    for (int i = tid; i < cell.n; i += TPB) v.sV[i] = v0[cell.comp_begin + i];
    for (int k = tid; k < cell.n_gate; k += TPB) v.sQ[k] = q0[cell.gate_begin + k];
    bar<TPB>();
    int cur = 0;
    for (int64_t s = -3; s <= T.rows; s++) {
        const double t = s >= 1 ? T.t_row[s - 1] : 0.0;
        if (s >= 1) {
            if ((s - 1) % K == 0) {
                double *c = ck + ((s - 1) / K) * ck_stride;
                for (int i = tid; i < cell.n; i += TPB) c[cell.comp_begin + i] = v.sV[i];
                for (int k = tid; k < cell.n_gate; k += TPB) c[nc_total + cell.gate_begin + k] = v.sQ[cur * nqmax + k];
            }
            for (int i = tid; i < cell.n; i += TPB) {
                const int rc = T.rec_col[cell.comp_begin + i];
                if (rc >= 0) trace[(s - 1) * T.n_record + rc] = v.sV[i];
            }
            if (hist) {
                const int64_t per = (int64_t)cell.n + cell.n_gate + 2 * cell.n_syn;
                const int64_t k = (s - 1) / K, j = (s - 1) - k * K;
                auto put = [&](int64_t seg, int64_t idx) {
                    double *h = hist + seg * hist_stride + seg_off[blockIdx.x] + idx * per;
                    for (int i = tid; i < cell.n; i += TPB) h[i] = v.sV[i];
                    for (int q = tid; q < cell.n_gate; q += TPB) h[cell.n + q] = v.sQ[cur * nqmax + q];
                    double *hs = h + cell.n + cell.n_gate;
                    for (int ls = tid; ls < cell.n_syn; ls += TPB) {
                        hs[ls] = v.sA[ls];
                        hs[cell.n_syn + ls] = v.sB[ls];
                    }
                };
                // the state of the last step opens no segment of its own when (rows - 1) is a
                // multiple of K: it is only the closing state of the segment before
                if (k < n_seg) put(k, j);
                if (j == 0 && k > 0) put(k - 1, K);
            }
        }
        if (s == T.rows) break;
        v.step(s, t, cur, nqmax, gbase, amp, weight);
        cur ^= 1;
    }
}

// One backward segment: steps [a, b) with a = 1 + k K. seg holds the states at the start of
// steps a..b (per step: V of the cell's compartments then its q). lam (global) carries
// (lambda_V, lambda_q) at the start of step b in and at the start of step a out; grad_g and
// grad_a accumulate.
template <int TPB>
__global__ void __launch_bounds__(TPB) adj_backward(AdjTables T, const double *__restrict__ gbase,
                                                    const double *__restrict__ amp, const double *__restrict__ gtrace,
                                                    const double *__restrict__ ck, int64_t nc_total, int64_t a,
                                                    int64_t b, double *__restrict__ seg, const int64_t *__restrict__ seg_off,
                                                    double *__restrict__ lam, double *__restrict__ grad_g,
                                                    double *__restrict__ grad_a, AdjLds L, int32_t nqmax, int32_t stored,
                                                    const double *__restrict__ weight, double *__restrict__ grad_w) {
    extern __shared__ __align__(16) unsigned char smem[];
    const AdjCell cell = T.cells[blockIdx.x];
    auto v = make_view<TPB>(T, cell, smem, L, nqmax);
    const int tid = threadIdx.x;
    const int N = cell.n, NG = cell.n_gate, NS = cell.n_syn;
    const int64_t per = (int64_t)N + NG + 2 * NS;
    double *S = seg + seg_off[blockIdx.x];
    const double dt = T.dt;
    for (int c = tid; c < cell.n_chan; c += TPB) v.sGg[c] = 0.0;
    for (int p = tid; p < cell.n_pulse; p += TPB) v.sGa[p] = 0.0;
    if (!stored) {
        // recompute the segment from its checkpoint (else seg is the forward pass's history)
        for (int i = tid; i < N; i += TPB) v.sV[i] = ck[cell.comp_begin + i];
        for (int k = tid; k < NG; k += TPB) v.sQ[k] = ck[nc_total + cell.gate_begin + k];
        bar<TPB>();
        int cur = 0;
        for (int64_t s = a; s <= b; s++) {
            double *Ss = S + (s - a) * per;
            for (int i = tid; i < N; i += TPB) Ss[i] = v.sV[i];
            for (int k = tid; k < NG; k += TPB) Ss[N + k] = v.sQ[cur * nqmax + k];
            if (s == b) break;
            v.step(s, T.t_row[s - 1], cur, nqmax, gbase, amp, weight);
            cur ^= 1;
        }
    }
    // adjoint state at the start of step b
    for (int i = tid; i < N; i += TPB) v.sLV[i] = lam[cell.comp_begin + i];
    for (int k = tid; k < NG; k += TPB) v.sLQ[k] = lam[nc_total + cell.gate_begin + k];
    bar<TPB>();
    for (int64_t s = b - 1; s >= a; s--) {
        const double *Ss = S + (s - a) * per;        // state at the start of step s
        const double *Sn = Ss + per;                   // state at the start of step s + 1
        const double t = T.t_row[s - 1];
        // dLoss/dV at the start of step s + 1 (trace row s)
        if (s < T.rows) {
            const double *gr = gtrace + s * T.n_record;
            for (int i = tid; i < N; i += TPB) {
                const int rc = T.rec_col[cell.comp_begin + i];
                if (rc >= 0) v.sLV[i] += gr[rc];
            }
        }
        // state of step s into LDS; open fractions
        for (int i = tid; i < N; i += TPB) v.sV[i] = Ss[i];
        for (int k = tid; k < NG; k += TPB) v.sQ[k] = Ss[N + k];
        for (int ls = tid; ls < NS; ls += TPB) {
            v.sA[ls] = Ss[N + NG + ls];
            v.sB[ls] = Ss[N + NG + NS + ls];
        }
        bar<TPB>();
        v.fopen(0, nqmax);
        bar<TPB>();
        // diagonal of step s and right-hand side lambda_V' for the transposed solve
        for (int i = tid; i < N; i += TPB) {
            const int64_t g = cell.comp_begin + i;
            double I, G;
            v.currents(i, v.sV[i], t, gbase, amp, I, G, weight);
            const double D = 1 * (G / T.C[g]) + T.invrc[g];
            const int p = T.pos[g];
            v.sD[p] = 1 + D * dt;
            v.sX[p] = v.sLV[i];
        }
        if (tid == 0) {
            v.sD[N] = 1.0;
            v.sX[N] = 0.0;
        }
        bar<TPB>();
        v.hines(true);
        // membrane and channel adjoints
        for (int i = tid; i < N; i += TPB) {
            const int64_t g = cell.comp_begin + i;
            const double lb = v.sX[T.pos[g]];
            const double Vn = Sn[i];
            const double lD = -lb * Vn;
            const double C = T.C[g], A = T.A[g];
            const double k = dt / C * K1E5 * A;
            for (int ch = T.ch_begin[g]; ch < T.ch_begin[g + 1]; ch++) {
                const int c = ch - cell.chan_begin;
                const double lGs = k * (lb * T.ch_erev[ch] + lD);
                v.sGg[c] += lGs * v.sF[c];
                v.sLF[c] = lGs * gbase[ch];
            }
            for (int p = T.p_begin[g]; p < T.p_begin[g + 1]; p++)
                if (T.p_start[p] <= t && t <= T.p_start[p] + T.p_dur[p])
                    v.sGa[p - cell.pulse_begin] += lb * dt / C * T.p_weight[p];
            // synapse instances: a conductance w g_s(t) into reversal erev, like a channel with
            // no area or density factor; g_s does not depend on V or on the parameters
            for (int sy = T.s_begin[g]; sy < T.s_begin[g + 1]; sy++) {
                const AdjSynType st = T.s_types[T.s_type[sy]];
                const int ls = sy - cell.syn_begin;
                const double gs = st.gbase * (v.sB[ls] - v.sA[ls]);
                v.sGw[ls] += dt / C * (lb * st.erev + lD) * gs;
            }
            v.sLV[i] = lb;
        }
        bar<TPB>();
        // gate adjoints: lambda_q of step s, and each gate's term of lambda_V
        double *lqn = v.sLQ + nqmax;
        for (int kq = tid; kq < NG; kq += TPB) {
            const int gq = cell.gate_begin + kq;
            const int ch = T.g_chan[gq];
            const int ty = T.ch_type[ch];
            const int kk = T.g_k[gq];
            const int q0 = T.ch_gate[ch] - cell.gate_begin;
            const double V = v.sV[T.ch_comp[ch]];
            double dq, dV;
            (void)gate64(T.gates[ty * 4 + kk], V, v.sQ[kq], false, dt, &dq, &dV);
            // dfopen/dq_kk: product over the channel's gates with one factor q_kk removed
            double df = 1.0;
            for (int m = 0; m < T.type_ngates[ty]; m++) {
                const double qv = v.sQ[q0 + m];
                const int inst = T.gates[ty * 4 + m].instances;
                if (m == kk) {
                    double pw = (double)inst;
                    for (int r = 0; r < inst - 1; r++) pw = pw * qv;
                    df = df * (inst > 0 ? pw : 0.0);
                } else {
                    for (int r = 0; r < inst; r++) df = df * qv;
                }
            }
            const double lq1 = v.sLQ[kq];
            lqn[kq] = lq1 * dq + v.sLF[ch - cell.chan_begin] * df;
            v.sGV[kq] = lq1 * dV;
        }
        bar<TPB>();
        for (int i = tid; i < N; i += TPB) {
            const int64_t g = cell.comp_begin + i;
            double lv = v.sLV[i];
            for (int ch = T.ch_begin[g]; ch < T.ch_begin[g + 1]; ch++) {
                const int q0 = T.ch_gate[ch] - cell.gate_begin;
                for (int m = 0; m < T.type_ngates[T.ch_type[ch]]; m++) lv += v.sGV[q0 + m];
            }
            v.sLV[i] = lv;
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
    if (d->abi_version != 1 && d->abi_version != EDENHIP_ABI_VERSION)
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
    // channel types
    std::vector<int32_t> tng(std::max(1, d->n_channel_types));
    std::vector<AdjGate> gates((size_t)std::max(1, d->n_channel_types) * 4);
    for (int ty = 0; ty < d->n_channel_types; ty++) {
        const edenhip_channel_type &ct = d->channel_types[ty];
        if (ct.n_gates < 0 || ct.n_gates > 4) return afail(a, EDENHIP_E_LIMIT, "channel type with more than 4 gates");
        tng[ty] = ct.n_gates;
        for (int k = 0; k < ct.n_gates; k++) {
            const edenhip_gate &g = ct.gates[k];
            AdjGate x{};
            x.alpha_form = g.alpha.form; x.beta_form = g.beta.form; x.instances = g.instances;
            x.q10 = (double)g.q10;
            x.ar = (double)g.alpha.rate; x.am = (double)g.alpha.midpoint; x.as = (double)g.alpha.scale;
            x.br = (double)g.beta.rate; x.bm = (double)g.beta.midpoint; x.bs = (double)g.beta.scale;
            gates[(size_t)ty * 4 + k] = x;
        }
    }
    // per compartment / channel / gate / pulse tables
    std::vector<double> C(NC), A(NC), invrc(NC), v0(NC), erev(a->NCH), gb0(a->NCH), amp0(a->NP), ps(a->NP), pd(a->NP),
        pw(a->NP);
    std::vector<int32_t> pos(NC), chb(NC + 1), pb(NC + 1), rec(NC, -1), cht(a->NCH), chg(a->NCH), chc(a->NCH);
    std::vector<int32_t> gch, gk;
    std::vector<double> q0;
    std::vector<double2> u(NC, make_double2(0.0, 0.0));
    for (int64_t g = 0; g < NC; g++) {
        C[g] = (double)d->capacitance[g]; A[g] = (double)d->area[g]; invrc[g] = (double)d->inv_rc_axial[g];
        v0[g] = (double)d->v0[g];
        chb[g] = (int32_t)d->comp_chan_begin[g]; pb[g] = (int32_t)d->comp_pulse_begin[g];
    }
    chb[NC] = (int32_t)a->NCH; pb[NC] = (int32_t)a->NP;
    for (int64_t p = 0; p < a->NP; p++) {
        ps[p] = (double)d->pulse_start[p]; pd[p] = (double)d->pulse_duration[p];
        pw[p] = (double)d->pulse_weight[p]; amp0[p] = (double)d->pulse_amplitude[p];
    }
    a->n_record = d->n_record;
    a->T.n_record = (int32_t)d->n_record;
    for (int64_t r = 0; r < d->n_record; r++) {
        const int64_t g = d->record_comp[r];
        if (g < 0 || g >= NC) return afail(a, EDENHIP_E_INVALID, "record_comp out of range");
        if (rec[g] >= 0) return afail(a, EDENHIP_E_INVALID, "a compartment may be recorded once in the adjoint model");
        rec[g] = (int32_t)r;
    }
    // synapses: per compartment CSR, types in double with the waveform factor evaluated in
    // double with the reference's expression tree (as the oracle's FP64 flavour per step)
    std::vector<int32_t> sb(NC + 1), sty(std::max<int64_t>(a->NSYN, 1), 0);
    for (int64_t g = 0; g <= NC; g++) sb[g] = (int32_t)d->comp_syn_begin[g];
    for (int64_t sy = 0; sy < a->NSYN; sy++) sty[sy] = d->syn_type[sy];
    std::vector<AdjSynType> stypes(std::max(1, d->n_syn_types));
    for (int t = 0; t < d->n_syn_types; t++) {
        const edenhip_exptwo_type &e = d->syn_types[t];
        const double tr = (double)e.tau_rise, td = (double)e.tau_decay;
        const double peak = (std::log(td / tr) * (tr * td)) / (td - tr);
        const double wf = (double)1 / ((-std::exp((-peak) / tr)) + std::exp((-peak) / td));
        stypes[t] = AdjSynType{(double)e.gbase, (double)e.erev, tr, td, (double)e.weight_prop, wf};
    }
    std::vector<double> w0(std::max<int64_t>(a->NSYN, 1), 0.0);
    for (int64_t sy = 0; sy < a->NSYN; sy++) w0[sy] = (double)d->syn_weight[sy];
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
        const TreePlan P = plan_tree(n, d->parent + c0, a->tpb);
        if (!P.ok) return afail(a, EDENHIP_E_LIMIT, "the adjoint supports at most 1,023 children per compartment");
        cell.rounds_fwd = n > 1 ? P.rounds_fwd : 0;
        cell.rounds_bwd = n > 1 ? P.rounds_bwd : 0;
        cell.words_off = (int32_t)words.size();
        cell.words_len = (int32_t)P.block.size();
        words.insert(words.end(), P.block.begin(), P.block.end());
        wmax = std::max(wmax, cell.words_len);
        for (int32_t i = 0; i < n; i++) pos[c0 + i] = P.pos[i];
        for (int32_t i = 1; i < n; i++) {
            // Ui = -dt/(R C_i), Uj = -dt/(R C_parent) with R = R_axial of the child (eden/Eden.cpp:6136-6150)
            const double R = (double)d->r_axial[c0 + i];
            const int32_t j = d->parent[c0 + i];
            u[c0 + P.pos[i]] = make_double2(-dt / (R * (double)d->capacitance[c0 + i]), -dt / (R * (double)d->capacitance[c0 + j]));
        }
        cell.chan_begin = (int32_t)d->comp_chan_begin[c0];
        cell.n_chan = (int32_t)(d->comp_chan_begin[c1] - d->comp_chan_begin[c0]);
        cell.gate_begin = (int32_t)q0.size();
        for (int64_t g = c0; g < c1; g++)
            for (int64_t ch = d->comp_chan_begin[g]; ch < d->comp_chan_begin[g + 1]; ch++) {
                const int ty = d->chan_type[ch];
                if (ty < 0 || ty >= d->n_channel_types) return afail(a, EDENHIP_E_INVALID, "channel type index out of range");
                cht[ch] = ty;
                erev[ch] = (double)d->chan_erev[ch];
                gb0[ch] = (double)d->chan_gbase[ch];
                chg[ch] = (int32_t)q0.size();
                chc[ch] = (int32_t)(g - c0);
                for (int k = 0; k < d->channel_types[ty].n_gates; k++) {
                    const edenhip_gate &gt = d->channel_types[ty].gates[k];
                    const double al = edenhip_rate_host(gt.alpha, d->v0[g]), be = edenhip_rate_host(gt.beta, d->v0[g]);
                    q0.push_back(al / (al + be));
                    gch.push_back((int32_t)ch);
                    gk.push_back(k);
                }
            }
        cell.n_gate = (int32_t)(q0.size() - (size_t)cell.gate_begin);
        cell.pulse_begin = (int32_t)d->comp_pulse_begin[c0];
        cell.n_pulse = (int32_t)(d->comp_pulse_begin[c1] - d->comp_pulse_begin[c0]);
        cell.syn_begin = (int32_t)d->comp_syn_begin[c0];
        cell.n_syn = (int32_t)(d->comp_syn_begin[c1] - d->comp_syn_begin[c0]);
        nsmax = std::max(nsmax, cell.n_syn);
        nqmax = std::max(nqmax, cell.n_gate);
        nchmax = std::max(nchmax, cell.n_chan);
        npmax = std::max(npmax, cell.n_pulse);
        cells[c] = cell;
        seg_off[c] = seg_total;
        seg_total += ((int64_t)n + cell.n_gate + 2 * (int64_t)cell.n_syn);
    }
    a->NQ = (int64_t)q0.size();
    a->nqmax = nqmax;
    a->L = adj_lds(nmax, nqmax, nchmax, npmax, wmax, nsmax);
    a->lds = (size_t)a->L.total * sizeof(double);
    if (a->lds > 160 * 1024) return afail(a, EDENHIP_E_LIMIT, "cell too large for the adjoint kernels' LDS");
    a->K = std::max(1, a->K);
    const int64_t steps = a->rows;   // steps 1..rows-1 are differentiated
    a->n_ck = (steps + a->K - 1) / a->K + 1;
    a->ck_stride = NC + a->NQ;
    // segment buffer: (K + 1) states per cell
    for (auto &o : seg_off) o *= (int64_t)(a->K + 1);
    AdjTables &T = a->T;
    ATRY(aupload(a, (AdjCell **)&T.cells, cells));
    ATRY(aupload(a, (double **)&T.C, C));
    ATRY(aupload(a, (double **)&T.A, A));
    ATRY(aupload(a, (double **)&T.invrc, invrc));
    ATRY(aupload(a, (int32_t **)&T.pos, pos));
    ATRY(aupload(a, (int32_t **)&T.ch_begin, chb));
    ATRY(aupload(a, (int32_t **)&T.p_begin, pb));
    ATRY(aupload(a, (int32_t **)&T.rec_col, rec));
    ATRY(aupload(a, (int32_t **)&T.ch_type, cht));
    ATRY(aupload(a, (int32_t **)&T.ch_gate, chg));
    ATRY(aupload(a, (int32_t **)&T.ch_comp, chc));
    ATRY(aupload(a, (double **)&T.ch_erev, erev));
    ATRY(aupload(a, (int32_t **)&T.g_chan, gch));
    ATRY(aupload(a, (int32_t **)&T.g_k, gk));
    ATRY(aupload(a, (int32_t **)&T.type_ngates, tng));
    ATRY(aupload(a, (AdjGate **)&T.gates, gates));
    ATRY(aupload(a, (double **)&T.p_start, ps));
    ATRY(aupload(a, (double **)&T.p_dur, pd));
    ATRY(aupload(a, (double **)&T.p_weight, pw));
    ATRY(aupload(a, (double2 **)&T.u, u));
    ATRY(aupload(a, (uint32_t **)&T.words, words));
    ATRY(aupload(a, (double **)&T.t_row, trow));
    ATRY(aupload(a, (int32_t **)&T.s_begin, sb));
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
    // state history of every step (store mode): n_seg segments of seg_total (K + 1) doubles
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
    int32_t ncell = a->n_cells, K = a->K, nq = a->nqmax;
    int64_t cks = a->ck_stride, nct = a->NC;
    AdjLds L = a->L;
    int64_t hs = a->hist_stride, ns = a->n_seg;
    void *args[] = {&T, &ncell, &a->d_v0, &a->d_q0, (void *)&gb, (void *)&am, &trace, &a->d_ck, &cks, &nct, &K, &L, &nq,
                    &a->d_hist, &hs, &ns, &a->d_seg_off, (void *)&wt};
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
        void *args[] = {&T, (void *)&a->last_gbase, (void *)&a->last_amp, (void *)&grad_trace, (void *)&ckp, &nct, &sa, &sb,
                        &segp, &a->d_seg_off, &a->d_lam, &gg, &ga, &L, &nq, &stored, (void *)&a->last_w, &gw};
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
