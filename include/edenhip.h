/*
 * edenhip.h -- C-ABI of the MI355X (gfx950) per-timestep simulation loop for EDEN models.
 *
 * The library replaces the body of EDEN's engine loop: the per-work-item callback
 *   typedef void (*IterationCallback)(double time, float dt, ...)      eden/Eden.cpp:53-76
 * exported by every generated `<Cell>.gen.so` as `doit` (eden/Eden.cpp:3239-3246), the
 *   #pragma omp parallel for schedule(runtime) ... tabs.callbacks[item](...)  eden/Eden.cpp:11848-11869
 * dispatch around it, and the step / swap logic of eden/Eden.cpp:11486-11938.
 * A per-item GPU callback at that call site would mean 100k host calls per step, so the
 * boundary sits one level up: the caller hands over the whole instantiated model once, in
 * structure-of-arrays form, and asks for N steps at a time.
 *
 * Conventions
 *  - Engine units are EDEN's native units (eden/NeuroML.cpp:305-425): time us, voltage mV,
 *    current nA, conductance uS, capacitance pF, resistance MOhm, conductivity mS/cm^2,
 *    length um, rates 1/us.
 *  - Every entry point returns 0 on success and a negative EDENHIP_E* code on failure;
 *    edenhip_last_error() then describes the failure. No C++ exception crosses the ABI.
 *  - Descriptor arrays are borrowed only for the duration of edenhip_create(); the context
 *    owns device copies. One host thread per context.
 *  - Compartment indices inside a cell are EDEN's: parents before children, root = 0.
 */
#ifndef EDENHIP_H
#define EDENHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EDENHIP_ABI_VERSION 4   /* 2 adds edenhip_model_desc.lems, 3 adds edenhip_options.channel_wave,
                                   edenhip_model_desc.record_gate and edenhip_lems_synapses.peer,
                                   4 adds edenhip_model_desc.lems_cells; descriptors of versions 1-3
                                   are accepted (options of versions 1 and 2 are read without
                                   channel_wave) */

enum {
    EDENHIP_OK = 0,
    EDENHIP_E_INVALID = -1,   /* malformed descriptor or argument */
    EDENHIP_E_DEVICE = -2,    /* HIP runtime error */
    EDENHIP_E_LIMIT = -3,     /* size beyond a compiled-in limit */
    EDENHIP_E_STATE = -4      /* call not valid in the context's current state */
};

/* Gate rate forms of NeuroML's HHExpRate / HHExpLinearRate / HHSigmoidRate and fixed rates,
 * emitted at eden/Eden.cpp:5126-5200. */
enum {
    EDENHIP_RATE_EXP = 0,        /* r * exp((V - mid) / scale) */
    EDENHIP_RATE_EXPLINEAR = 1,  /* r * (V == mid ? 1 : x / (1 - exp(-x))), x = (V - mid) / scale */
    EDENHIP_RATE_SIGMOID = 2,    /* r / (1 + exp((mid - V) / scale)) */
    EDENHIP_RATE_FIXED = 3,      /* r */
    EDENHIP_RATE_LEMS = 4        /* (abi_version >= 4) lems_rate((int)r, V): instance r of a user
                                    rate, time-course or steady-state ComponentType, from
                                    edenhip_model_desc.lems_rates_source */
};

#define EDENHIP_MAX_GATES 4

/* How a gate's tau and inf arise (eden/Eden.cpp:5282-5383). EDENHIP_GATE_RATES: from the
 * forward and reverse rates alpha, beta (gateHHrates): tau = 1 / (alpha + beta), inf =
 * alpha / (alpha + beta). EDENHIP_GATE_TAU_INF: given directly (gateHHtauInf, and
 * gateHHratesTauInf, whose rates then play no part): the gate's `alpha` holds the time course
 * (EDENHIP_RATE_FIXED with rate = tau in us: fixedTimeCourse) and its `beta` the steady state
 * (EXP / EXPLINEAR / SIGMOID forms with a dimensionless rate: HHExpVariable,
 * HHExpLinearVariable, HHSigmoidVariable). Both then take the same cnexp update. */
enum {
    EDENHIP_GATE_RATES = 0,
    EDENHIP_GATE_TAU_INF = 1,
    /* (gateHHInstantaneous) no state: the gate's value in a step is its steady state, held in
     * its `beta`, at the step's voltage (eden/Eden.cpp:5269-5281); compartments with such
     * gates run their channels in the channel pre-pass (one launch per step) */
    EDENHIP_GATE_INSTANTANEOUS = 2,
    /* (gateHHratesTau, gateHHratesInf) the rates give inf (resp. tau) and the third rate, in
     * edenhip_model_desc.gate_extra_rate, gives tau (resp. inf); channels with such gates run
     * in the channel pre-pass */
    EDENHIP_GATE_RATES_TAU = 3,
    EDENHIP_GATE_RATES_INF = 4
};

typedef struct edenhip_rate {
    int32_t form;        /* EDENHIP_RATE_* */
    float rate;          /* 1/us */
    float midpoint;      /* mV */
    float scale;         /* mV */
} edenhip_rate;

/* One gateHHrates gate: alpha/beta form, tau = 1/(alpha+beta), inf = alpha/(alpha+beta),
 * cnexp update with clipping (eden/Eden.cpp:5236-5259, 5336-5375). */
typedef struct edenhip_gate {
    int32_t instances;   /* gate power in the open fraction (eden/Eden.cpp:5679-5694) */
    float q10;           /* fixed rate multiplier (1 when the gate has no q10Settings) */
    edenhip_rate alpha;  /* forward rate */
    edenhip_rate beta;   /* reverse rate */
} edenhip_gate;

/* An ion channel type. n_gates == 0 is a passive (leak) channel: open fraction 1. */
typedef struct edenhip_channel_type {
    int32_t n_gates;
    edenhip_gate gates[EDENHIP_MAX_GATES];
} edenhip_channel_type;

/* expTwoSynapse (eden/neuroml/LEMS_CoreComponents.inc.xml:145-180), run by EDEN as a
 * lemsified component (eden/Eden.cpp:2614-2994). */
typedef struct edenhip_exptwo_type {
    float gbase;         /* uS */
    float erev;          /* mV */
    float tau_rise;      /* us */
    float tau_decay;     /* us */
    float weight_prop;   /* the LEMS `weight` property, default 1 */
} edenhip_exptwo_type;

/* Generic LEMS synapse types (alphaSynapse, expThreeSynapse, the PyNN synapses, gap
 * junctions, ...) and LEMS current inputs (sine and ramp generators, voltage clamps, ...). The
 * reference emits C for each lemsified ComponentType and compiles it with gcc
 * (eden/Eden.cpp:2339-2995, 6992-7216); here the caller supplies HIP source defining
 *   __device__ void lems_syn(int type, const float *P, const float *S, float *Sn, long stride,
 *                            float v, float vpeer, float t, int init, int event, float dt,
 *                            float *i, float *g)
 * (one instance's step: P its type's parameters, S / Sn its present / next state with
 * variable k at k * stride, v the compartment voltage, vpeer the voltage at the start of the
 * step of the instance's peer compartment (gap junctions; v when it has none), t the step's
 * time, event whether it receives its event this step; i, g its current and conductance
 * exposures), which edenhip_create() compiles with hiprtc into the library's synapse kernel
 * template (eden_amd.lems emits it). When present, every synapse instance's syn_type indexes
 * these types, and syn_types serves only the Poisson inputs. An instance of an input type
 * (type_flags bit 0) sits in its compartment's synapse list after the synapse types (it never
 * receives spikes); its current times its weight joins the compartment's input total, after
 * the native pulse and Poisson inputs (eden/Eden.cpp:3914-3931). */
typedef struct edenhip_lems_synapses {
    int32_t n_types;
    const char *source;           /* HIP C++ source, NUL-terminated */
    const int32_t *param_begin;   /* [n_types + 1] into params */
    const float *params;          /* per type: its parameter vector */
    const int32_t *n_state;       /* [n_types] state variables per instance (at most 8) */
    const int64_t *peer;          /* (abi_version >= 3) [n_syn] the global compartment whose voltage
                                     is the instance's vpeer (the other end of a gap junction,
                                     eden/Eden.cpp:9225-9248), or -1; NULL = none */
    const int32_t *type_flags;    /* (abi_version >= 4) [n_types] bit 0: an input type; NULL = none */
    const int64_t *instance_param;/* (abi_version >= 4) [n_syn] offset into params of the instance's own
                                     parameter vector (per-instance settings: the reference's
                                     EdenCustomSetup, eden/NeuroML.cpp:7362-7980), or -1 for its
                                     type's; NULL = every instance its type's */
    int64_t n_params;             /* (abi_version >= 4) floats in params: the types' vectors, then any
                                     per-instance vectors; 0 = param_begin[n_types] */
    const int64_t *spike_list;    /* (abi_version >= 4) [n_syn] for a firing-synapse input (the
                                     reference's timedSynapticInput, eden/Eden.cpp:3759-3791,
                                     3859-3912, 8840-8852) the offset into spike_times of its spike
                                     list, else -1; NULL = none. Such an instance takes no incoming
                                     connections: on each step it receives an event when its next
                                     listed times fall before t + dt (all of them consumed, the
                                     position advancing after the initial steps), and its type is an
                                     input type (type_flags bit 0) */
    const float *spike_times;     /* spike lists (us, in listed order), each ended by a value >= FLT_MAX */
    int64_t n_spike_times;
    /* (abi_version >= 4) Poisson firing-synapse inputs on the generic path (poissonFiringSynapse
     * with a synapse other than expTwoSynapse, transientPoissonFiringSynapse; their spiker's
     * LEMS dynamics at eden/neuroml/LEMS_CoreComponents.inc.xml:623-735): per instance an index
     * into the spiker tables, or -1 (NULL = none). The spiker draws its intervals with randof
     * (work item = the cell's RNG seed, the given instance, the step, invocation rng_offset + 0
     * at the initial steps and + 1 on each spike); its spike is the instance's event in the
     * same step. Such an instance takes no incoming connections and its type is an input type. */
    const int64_t *poisson_spiker;
    int64_t n_poisson_spikers;
    const float *spiker_rate;       /* [n] averageRate, 1/us */
    const int32_t *spiker_transient;/* [n] 1: transientPoissonFiringSynapse (delay and duration apply) */
    const float *spiker_delay;      /* [n] us */
    const float *spiker_duration;   /* [n] us */
    const int32_t *spiker_instance; /* [n] randof instance (earlier instances of the same input on the cell) */
    const int32_t *spiker_rng_offset; /* [n] */
} edenhip_lems_synapses;

/* LEMS point neurons (the reference's artificial cells: iafCell, izhikevich2007Cell,
 * adExIaFCell, the PyNN cells, user-defined cell ComponentTypes, ...), whose generated C the
 * reference compiles per cell type (eden/Eden.cpp:6603-6910). The caller supplies HIP source
 * defining
 *   __device__ void lems_cell(int type, const float *P, const float *S, float *Sn, long stride,
 *                             float t, float dt, int init, float isyn, int rng_seed,
 *                             long long rng_step, int *spike)
 * (one cell's step: P its type's parameters, S / Sn its present / next state with variable k
 * at k * stride, t the step's time, isyn its external current (its synapses' currents, then
 * its inputs'), rng_seed and rng_step the cell's RNG seed and the step (the stamp of the
 * type's random() calls, eden/Eden.cpp:2450-2459, 6628-6637), *spike set when it emits a
 * spike), which edenhip_create() compiles with hiprtc
 * beside lems_syn (eden_amd.lems_cells emits it). A point-neuron cell has one compartment,
 * without channels or Poisson inputs; its synapses take the generic path (lems), and the
 * compartment's V mirrors the type's membrane-voltage state variable (its synapses' and gap
 * junction peers' voltage). Models with point neurons run one launch per step. */
typedef struct edenhip_lems_cells {
    int32_t n_types;
    const char *source;           /* HIP C++ source, NUL-terminated */
    const int32_t *param_begin;   /* [n_types + 1] into params */
    const float *params;          /* per type: its parameter vector */
    const int32_t *n_state;       /* [n_types] state variables per cell (at most 16) */
    const int32_t *v_state;       /* [n_types] the state variable exposed as the membrane voltage, or -1 */
    const int32_t *cell_type;     /* [n_cells] each cell's type, -1 for a physical cell */
    const int64_t *cell_param;    /* [n_cells] offset into params of the cell's own parameter vector
                                     (per-instance settings), or -1 for its type's; NULL = types' */
    int64_t n_params;             /* floats in params: the types' vectors, then any per-instance
                                     vectors; 0 = param_begin[n_types] */
} edenhip_lems_cells;

/*
 * The instantiated model. Global compartment index g runs over all cells in cell order;
 * cell c owns compartments [cell_comp_begin[c], cell_comp_begin[c+1]).
 * CSR arrays "x_begin" have n_comps + 1 entries and index per-compartment lists.
 * Lists are in EDEN's instance order (eden/Eden.cpp:8791-9281): summation order of the
 * per-compartment currents follows it.
 */
typedef struct edenhip_model_desc {
    int32_t abi_version;          /* EDENHIP_ABI_VERSION */
    float dt;                     /* us, EngineConfig::dt */
    double t_final;               /* us, EngineConfig::t_final (= <Simulation length>) */

    int32_t n_cells;
    int64_t n_comps;
    const int64_t *cell_comp_begin;   /* [n_cells + 1] */
    const uint32_t *cell_rng_seed;    /* [n_cells] ReverseBits(sim_seed) ^ gid (eden/Eden.cpp:7679-7691) */

    /* per compartment [n_comps] */
    const float *v0;              /* initial V, mV */
    const float *capacitance;     /* C, pF */
    const float *area;            /* membrane area, um^2 */
    const float *inv_rc_axial;    /* sum over neighbours of 1/(R C), 1/us (eden/Eden.cpp:4488-4501) */
    const float *r_axial;         /* resistance to the parent, MOhm; unused for a root */
    const float *v_threshold;     /* spike threshold, mV; NaN = no spike output */
    const int32_t *parent;        /* cell-local parent index, -1 for the root */

    /* ion channel distributions (channelDensity): CSR over compartments */
    int32_t n_channel_types;
    const edenhip_channel_type *channel_types;   /* [n_channel_types] */
    const int64_t *comp_chan_begin;              /* [n_comps + 1] */
    const int32_t *chan_type;                    /* [n_chan] */
    const float *chan_gbase;                     /* [n_chan] mS/cm^2 */
    const float *chan_erev;                      /* [n_chan] mV */

    /* pulseGenerator inputs: CSR over compartments (eden/Eden.cpp:3798-3826) */
    const int64_t *comp_pulse_begin;             /* [n_comps + 1] */
    const float *pulse_start;                    /* us */
    const float *pulse_duration;                 /* us */
    const float *pulse_amplitude;                /* nA */
    const float *pulse_weight;                   /* input weight */

    /* synapse types */
    int32_t n_syn_types;
    const edenhip_exptwo_type *syn_types;        /* [n_syn_types] */

    /* chemical synapse instances (spike-triggered, with delay): CSR over compartments
     * (eden/Eden.cpp:3561-3740) */
    const int64_t *comp_syn_begin;               /* [n_comps + 1] */
    const int32_t *syn_type;                     /* [n_syn] */
    const float *syn_weight;                     /* [n_syn] connection weight */
    const float *syn_delay;                      /* [n_syn] us */

    /* spike recipients: for every compartment, the synapse instances (global index into the
     * syn_* arrays) that a spike of it triggers (eden/Eden.cpp:3946-3995, 8029-8046) */
    const int64_t *comp_spike_begin;             /* [n_comps + 1] */
    const int64_t *spike_target;                 /* [n_targets] */

    /* poissonFiringSynapse inputs: CSR over compartments, each with its own synapse
     * (eden/Eden.cpp:3859-3912; LEMS_CoreComponents.inc.xml:623-672) */
    const int64_t *comp_poisson_begin;           /* [n_comps + 1] */
    const float *poisson_rate;                   /* 1/us */
    const int32_t *poisson_syn_type;             /* synapse type driven by the generator */
    const float *poisson_weight;                 /* input weight */
    const int32_t *poisson_instance;             /* instance index for the RNG stamp */
    const int32_t *poisson_rng_offset;           /* rng_offset + random_call_counter base */

    /* recording */
    int64_t n_record;                            /* trajectory columns */
    const int64_t *record_comp;                  /* [n_record] compartment whose V is logged */
    int64_t n_event_cols;                        /* spike-event logger columns */
    const int64_t *event_comp;                   /* [n_event_cols] compartment spiking */

    /* (abi_version >= 2) generic LEMS synapse types, or NULL */
    const edenhip_lems_synapses *lems;

    /* (abi_version >= 3) what each recorded column holds: -1 = the compartment's V; j >= 0 =
     * its j-th gate variable (its channel distributions in order, each type's gates in order),
     * as an OutputColumn of a channel gate's q records it; for a point neuron (abi_version >= 4)
     * j >= 0 is its j-th state variable and -1 its membrane-voltage state variable. NULL = V
     * everywhere. A model that records gate variables runs one launch per step. */
    const int32_t *record_gate;

    /* (abi_version >= 4) LEMS point neurons, or NULL */
    const edenhip_lems_cells *lems_cells;

    /* (abi_version >= 4) [n_channel_types * EDENHIP_MAX_GATES] each gate's EDENHIP_GATE_* kind
     * (gate k of type t at t * EDENHIP_MAX_GATES + k), or NULL: every gate EDENHIP_GATE_RATES.
     * The adjoint takes EDENHIP_GATE_RATES gates only. */
    const int32_t *gate_kind;

    /* (abi_version >= 4) HIP source defining
     *   __device__ float lems_rate(int k, float v)
     * (instance k of a user gate rate / time course / steady state ComponentType at membrane
     * voltage v; eden_amd.lems_rates emits it), for gates with EDENHIP_RATE_LEMS forms, or
     * NULL. edenhip_create() compiles it with hiprtc into the channel pre-pass; compartments
     * with such gates run their channels there (one launch per step). */
    const char *lems_rates_source;

    /* (abi_version >= 4) [n_channel_types * EDENHIP_MAX_GATES] the time course of an
     * EDENHIP_GATE_RATES_TAU gate or the steady state of an EDENHIP_GATE_RATES_INF gate, or
     * NULL when there are none */
    const edenhip_rate *gate_extra_rate;

    /* (abi_version >= 4) [n_channel_types] each channel type's conductance_scaling, the start
     * of its open fraction (q10ConductanceScaling, eden/Eden.cpp:5700-5721), or NULL: 1
     * everywhere. Channels whose scaling is not 1 run in the channel pre-pass. */
    const float *chan_scaling;
} edenhip_model_desc;

/* Launch-shape choices. Zero-initialise for defaults. */
typedef struct edenhip_options {
    int32_t threads_per_cell;     /* 0 = by size (size_limits); else 64, 128, 256 or 512 for
                                     every tree cell */
    int32_t persistent;           /* 1 (default when 0) = time loop inside one launch per chunk;
                                     -1 = one launch per step (the plain per-step launch) */
    int32_t chunk_steps;          /* steps per trace chunk; 0 = all steps */
    int32_t use_graph;            /* 1 = with one launch per step (persistent = -1, or networks
                                     of multi-compartment cells): capture the launches in
                                     hipGraphs of 256 steps, kept for later runs over the same
                                     steps; the profile then has whole-chunk times only */
    int32_t profile;              /* 1 = time each launch with HIP events */
    int32_t device;               /* HIP device ordinal */
    int32_t log_synapse_events;   /* > 0: log up to this many synaptic events (see below) */
    int32_t network_window;       /* networks of one-compartment cells: 0 = advance cells several
                                     steps per launch (up to the minimum delay) with ring-buffer
                                     delivery; -1 = one launch per step; -2 = windows, each run
                                     again step by step (exercises the exact fallback) */
    int32_t size_limits[4];       /* largest tree cell (compartments) given 64, 128, 256 and
                                     512 threads; all 0 = 256, 512, 1024, 2047. A cell takes
                                     the first size whose limit holds it (each limit at most 4
                                     compartments per thread; the 512 limit must be 2047).
                                     eden_amd.autotune picks them by timing short runs. */
    int32_t channel_wave[4];      /* (abi_version >= 3) tree cells with 64, 128, 256, 512 threads:
                                     0 = by the rule (512: on; 64: on when the bucket has at most
                                     4 cells per CU; 128 and 256: off), 1 = on, -1 = off. On:
                                     one extra wave per workgroup updates the gates and open
                                     fractions for the next step while the cable waves run the
                                     step. Applies to persistent launches of models without
                                     synapses or spike recipients on tree cells.
                                     eden_amd.autotune times both settings per size. */
} edenhip_options;

typedef struct edenhip_ctx edenhip_ctx;

/* Build device tables from the descriptor (replaces RawTables instantiation +
 * the initial state copy, eden/Eden.cpp:7643-9281, 11093-11150). */
int edenhip_create(const edenhip_model_desc *desc, const edenhip_options *opt, edenhip_ctx **out);

/* Rows the full run records: floor(t_final / dt) + 1 (eden/Eden.cpp:11488). */
int64_t edenhip_total_rows(const edenhip_ctx *ctx);

/* Run the 4 initialisation steps (if not yet run) and then up to `n_rows` recorded steps
 * (each: record the state at `time`, compute, time += dt; eden/Eden.cpp:11549-11938).
 * Recorded V values go to the context's device trace buffer for this chunk; spike events
 * go to its device event list. Returns the number of rows done in *rows_done. */
int edenhip_run(edenhip_ctx *ctx, int64_t n_rows, int64_t *rows_done);

/* No counterpart in the reference: for callers that cannot allocate device memory while they
 * run (code under stream capture, tight serving loops, harnesses that forbid allocation).
 * Same as edenhip_run(), except that the tree-cell kernels' work-queue counters (2 ints per
 * workgroup-size bucket, scratch that edenhip_run() takes from the context) come from a
 * static device array of 32 buckets' counters that each GPU holds until the process ends;
 * the call allocates nothing. The counters are cleared by a kernel at the start of each call.
 * Every call on one GPU shares that GPU's array: calls that overlap in time on one GPU (from
 * several host threads) race. Rejected with EDENHIP_E_LIMIT: models on the windowed network
 * path (their scratch grows with the network) and more than 32 buckets; with
 * EDENHIP_E_STATE: contexts created with use_graph or profile (graph instantiation and
 * profiling events allocate). */
int edenhip_run_static(edenhip_ctx *ctx, int64_t n_rows, int64_t *rows_done);

/* Rewind to the state edenhip_create() built (before the initialisation steps), clearing the
 * event log; device tables stay resident. No counterpart in EDEN, which runs a model once. */
int edenhip_reset(edenhip_ctx *ctx);

/* Checkpoint and restore of the simulation state between edenhip_run() calls. EDEN keeps its
 * step state in the double-buffered tables main() owns (eden/Eden.cpp:11093-11150) and has
 * no save/restore; here the state is an opaque blob: every live state table of the context
 * (compartment voltages, gate variables, synapse trigger flags, pending arrival times and
 * LEMS state, Poisson input state, the network path's delivery ring) and the step counters.
 * Outputs already logged (events) are not part of it. edenhip_set_state() accepts a blob of a
 * context built from the same descriptor and options; running on then continues exactly as
 * the context the blob came from would. */
int edenhip_state_size(const edenhip_ctx *ctx, int64_t *bytes);
int edenhip_get_state(edenhip_ctx *ctx, void *dst, int64_t bytes);
int edenhip_set_state(edenhip_ctx *ctx, const void *src, int64_t bytes);

/* Copy the trace rows of the last edenhip_run() call: [rows x n_record] floats (mV),
 * row r holding the state at time (row0 + r) * dt. dst is host memory. */
int edenhip_read_trace(edenhip_ctx *ctx, float *dst, int64_t max_rows, int64_t *rows, int64_t *row0);

/* Device pointer of the trace buffer of the last run. Layout: column-major, one column per
 * distinct recorded compartment (in order of first appearance in record_comp), each column
 * a chunk's rows long (chunk_steps, or all rows): the value of that column at row r of the
 * chunk is at column * chunk_rows + r. */
int edenhip_trace_device_ptr(edenhip_ctx *ctx, const float **dptr);

/* Spike events logged so far, as (event column, row) pairs in ascending (row, column)
 * order: the spike detected while computing row r - 1 is logged at row r with timestamp
 * r * dt (eden/Eden.cpp:11598-11635). Copies up to max_events pairs; *n gets the total. */
int edenhip_read_events(edenhip_ctx *ctx, int64_t *col, int64_t *row, int64_t max_events, int64_t *n);

/* Synaptic events logged so far (options.log_synapse_events > 0): every (synapse instance,
 * step) at which a spike-triggered synapse receives its event, i.e. the step whose time
 * window [t, t + dt) holds the pending arrival time (eden/Eden.cpp:3649-3685). Sorted by
 * (step, synapse); steps count from 1 for the first recorded row, as in the event log. */
int edenhip_read_synapse_events(edenhip_ctx *ctx, int64_t *syn, int64_t *step, int64_t max_events, int64_t *n);

/* Copy the membrane voltage of every compartment (n_comps floats) to host memory. */
int edenhip_get_voltages(edenhip_ctx *ctx, float *dst);

/* Launch statistics of the hot kernel (only with options.profile = 1): the summed duration
 * of the kernel launches in ms, measured with HIP events on the stream they ran on, and
 * the launch count. */
int edenhip_kernel_stats(const edenhip_ctx *ctx, double *total_ms, int64_t *launches, char *name, int32_t name_len);

/* Per-kernel launch statistics (options.profile = 1). Kernel 0 is the point-cell kernel when
 * the model has one-compartment cells; then one entry per tree-cell bucket (cells grouped
 * by workgroup size, each bucket one kernel on a stream of its own). total_ms sums, over
 * the chunks run since create or reset, the time from the event recorded before the
 * kernel's first launch in a chunk to the event after its last, on the kernel's stream. */
typedef struct edenhip_kernel_info {
    char name[64];               /* kernel name as a profiler lists it, without arguments */
    int32_t threads_per_block;
    int32_t grid;                /* workgroups per launch */
    int32_t n_cells;             /* cells this kernel advances */
    int32_t lds_bytes;           /* dynamic LDS per workgroup */
    int64_t n_comps;             /* compartments of those cells */
    int64_t launches;
    double total_ms;
} edenhip_kernel_info;

int32_t edenhip_kernel_count(const edenhip_ctx *ctx);
int edenhip_get_kernel_info(const edenhip_ctx *ctx, int32_t i, edenhip_kernel_info *out);
/* Descriptor cell indices advanced by kernel i (n_cells of them), in work-queue order. */
int edenhip_get_kernel_cells(const edenhip_ctx *ctx, int32_t i, int32_t *cells, int32_t max_cells);

/* Network path figures since create or reset: steps per window (1 = one launch per step),
 * windows run, windows run again step by step (see edenhip_options.network_window), and
 * spike-triggered synaptic events received (-1 when not counted: per-step path). */
int edenhip_network_stats(edenhip_ctx *ctx, int32_t *window_steps, int64_t *windows, int64_t *windows_rerun,
                          int64_t *synaptic_events);

/* Describe the launch shape chosen for the tree-cell kernel (autotuned or given). */
int edenhip_launch_shape(const edenhip_ctx *ctx, int32_t *threads_per_cell, int32_t *slots_per_thread);

const char *edenhip_last_error(const edenhip_ctx *ctx);
void edenhip_destroy(edenhip_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif /* EDENHIP_H */
