// edenhip.hip -- host side of the C-ABI (include/edenhip.h): builds the device tables from the
// instantiated model, plans launches, runs chunks of steps and returns recorded outputs.
//
// Replaces, for the GPU path: RawTables instantiation and the initial-state copy
// (eden/Eden.cpp:7643-9281, 11093-11150), the OpenMP dispatch of generated callbacks
// (eden/Eden.cpp:11848-11869) and the step/swap logic of the engine loop
// (eden/Eden.cpp:11486-11938).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <atomic>
#include <mutex>
#include <cmath>
#include <thread>
#include <unordered_map>
#include <cstdio>
#include <cstddef>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/edenhip.h"
#include "device_model.hpp"
#include "kernels.hpp"
#include "rtc.hpp"
#include "tree_schedule.hpp"

using namespace edenhip;

namespace {

// A tree whose non-root compartments have no gated channel, no Poisson input and no chemical
// synapse: their Hines diagonals are the same every step (the tree kernel's constant-tree
// path).
bool const_below_root(const edenhip_model_desc *d, int64_t c0, int32_t n) {
    for (int32_t i = 1; i < n; i++) {
        const int64_t g = c0 + i;
        if (d->comp_poisson_begin[g + 1] != d->comp_poisson_begin[g]) return false;
        if (d->comp_syn_begin[g + 1] != d->comp_syn_begin[g]) return false;
        for (int64_t ch = d->comp_chan_begin[g]; ch < d->comp_chan_begin[g + 1]; ch++)
            if (d->channel_types[d->chan_type[ch]].n_gates != 0) return false;
        // (more channel distributions than the kernels' slots: channels of their own pass)
        if (d->comp_chan_begin[g + 1] - d->comp_chan_begin[g] > NCH) return false;
    }
    return true;
}

struct Bucket {
    int tpb = 64;
    bool split = false;   // one extra wave runs the channel phase concurrently
    std::vector<int32_t> order;    // tree-cell indices, largest first
    int32_t nmax = 0;
    int32_t nvmax = 0;        // voltage slots in LDS (1 for a constant tree, else its size)
    int32_t ngmax = 0;
    int32_t ninmax = 0;
    int32_t nctmax = 0;
    int32_t words_max = 0;
    int32_t cxmax = 0;        // largest complex-compartment count of the bucket's cells
    int32_t cxmax_lds = 0;    // cxmax when the records are staged in LDS, else 0
    int64_t comps = 0;
    size_t lds = 0;
    int grid = 0;
    int32_t *d_order = nullptr;
    int32_t *d_queue = nullptr;
    TreeCell *d_cells = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    // profiling (options.profile): events around the bucket's launches of a chunk
    hipEvent_t p0 = nullptr, p1 = nullptr;
    double ms = 0;
    int64_t launches = 0;
};

}  // namespace

struct edenhip_ctx {
    std::string err;
    edenhip_options opt{};
    int device = 0;
    hipStream_t stream = nullptr;
    float dt = 0;
    int64_t NC = 0, n_poi = 0, n_slots = 0, total_rows = 0, rows_done = 0, next_step = -3;
    int n_cells = 0;
    bool network = false;      // spike recipients: one launch per step
    bool tree_net = false;     // tree kernels of the NET form (network, or synapses on tree cells)
    bool point_quad = false;   // point cells run four lanes per cell (no synapses, inputs, spikes)
    // host mirrors for output mapping
    std::vector<int32_t> desc_to_dev_col;   // per desc trace column
    int32_t n_dev_cols = 0;
    // recorded gate variables: device column and source (>= 0: index into T.q of a point
    // cell; < 0: ~index into the tree cells' gate table), written by record_gates_kernel
    int32_t *d_grec_col = nullptr;
    int64_t *d_grec_src = nullptr;
    int32_t n_grec = 0;
    std::vector<std::vector<int64_t>> ev_expand;   // dev event column -> desc columns
    int64_t chunk_rows = 0, last_row0 = 0, last_rows = 0;
    DevTables T{};
    DGate *d_gates = nullptr;
    int n_gates = 0;
    DSynType *d_stypes = nullptr;
    int n_stypes = 0;
    // point cells
    int32_t *d_point = nullptr;
    int32_t n_point = 0;
    // tree cells
    uint32_t *d_words = nullptr;
    int32_t *d_pos = nullptr;
    float2 *d_u = nullptr;
    uint32_t *d_tcm = nullptr, *d_ctask = nullptr, *d_inl = nullptr, *d_gt = nullptr, *d_cflag = nullptr;
    TreeComplex *d_cx = nullptr;
    float *d_gq = nullptr;
    std::vector<Bucket> buckets;
    hipEvent_t ev_start = nullptr;
    // use_graph: executable graphs of per-step launch ranges, by (first step, end step, row0)
    struct GraphEntry {
        int64_t a, b, row0;
        hipGraphExec_t exec;
    };
    std::vector<GraphEntry> graphs;
    int n_cu = 256;
    // initial copies of the state tables (for edenhip_reset)
    struct Snap { void *dst; void *src; size_t bytes; };
    std::vector<Snap> snaps;
    // profiling
    double kernel_ms = 0;
    int64_t launches = 0;
    // per-kernel profile of the point kernel, and descriptor cell index of every point / tree cell
    hipEvent_t pp0 = nullptr, pp1 = nullptr;
    hipEvent_t nw0 = nullptr, nw1 = nullptr;   // around each network window launch (profile)
    double point_ms = 0;
    int64_t point_launches = 0;
    std::vector<int32_t> point_desc, tree_desc;
    // windowed network path (networks of one-compartment cells)
    bool netwin = false;
    int32_t net_W = 0;
    int net_cur = 0;
    NetTables N{};
    uint32_t *d_ev_save = nullptr;   // event-log counters saved before a window
    unsigned long long *d_deliv_save = nullptr;
    int32_t *h_conflict = nullptr;   // pinned host copy of the conflict flag
    int64_t net_windows = 0, net_redone = 0;
    std::string kernel_name;
    std::vector<void *> allocs;
    // generic LEMS synapses: the run-time compiled synapse kernel and its tables
    hipFunction_t lems_fn = nullptr;
    int32_t *d_lems_comps = nullptr;
    int32_t n_lems_comps = 0;
    float *d_lems_params = nullptr, *d_lems_state = nullptr;
    int32_t *d_lems_pbegin = nullptr;
    int32_t *d_lems_peer = nullptr;      // per synapse slot: gap junction peer compartment, or -1
    float2 *d_syn_ext = nullptr;
    int32_t *d_lems_flags = nullptr;     // per generic type: bit 0 an input type
    int32_t *d_lems_poff = nullptr;      // per synapse slot: offset of its parameter vector (per-instance settings)
    int32_t *d_lems_tpos = nullptr;      // per synapse slot: a firing-synapse input's spike-list position, or -1
    float *d_lems_ttimes = nullptr;      // firing-synapse inputs' spike lists
    int32_t *d_pk_slot = nullptr;        // per synapse slot: its Poisson spiker, or -1
    // compartments whose channels run in ext_channels_kernel
    int32_t n_xcomp = 0;
    int64_t n_xq = 0;
    int32_t *d_x_comp = nullptr, *d_x_chb = nullptr, *d_x_type = nullptr, *d_x_qb = nullptr;
    float *d_x_g = nullptr, *d_x_e = nullptr, *d_x_q = nullptr;
    float2 *d_chx_ext = nullptr;
    hipFunction_t ext_rtc_fn = nullptr;  // the pre-pass compiled with the model's user rate types
    float4 *d_gate_extra = nullptr;      // per gate table entry: a rates+tau/inf gate's third rate
    float *d_chan_scaling = nullptr;     // per channel type: conductance_scaling
    float2 *d_pk_state = nullptr;        // per Poisson spiker: (tsince, isi)
    float *d_pk_isi = nullptr;
    float2 *d_pk_window = nullptr;
    int32_t *d_pk_transient = nullptr, *d_pk_inst = nullptr, *d_pk_roff = nullptr;
    int64_t n_pk = 0;
    int32_t *d_acell_poff = nullptr;     // per point neuron: offset of its parameter vector (per-instance settings)
    float2 *d_inp_ext = nullptr;         // per compartment: the LEMS inputs' (I, G) of the step
    // LEMS point neurons: the run-time compiled cell kernel and its tables (one slot per cell)
    hipFunction_t lems_cell_fn = nullptr;
    int32_t n_acells = 0;
    int32_t *d_acell_comps = nullptr, *d_acell_type = nullptr, *d_acell_pbegin = nullptr, *d_acell_vstate = nullptr;
    float *d_acell_params = nullptr, *d_acell_state = nullptr;
    // edenhip_run_static: the tree buckets' work-queue counters come from the device's static
    // scratch array for the duration of that call (nullptr: each bucket's own d_queue)
    int32_t *queue_base = nullptr;
};

#define HIPCHECK(ctx, call)                                                                       \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) {                                                                   \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);                       \
            return EDENHIP_E_DEVICE;                                                              \
        }                                                                                         \
    } while (0)

template <typename T>
static int upload(edenhip_ctx *ctx, T **dptr, const std::vector<T> &v) {
    size_t bytes = std::max<size_t>(v.size(), 1) * sizeof(T);
    void *p = nullptr;
    HIPCHECK(ctx, hipMalloc(&p, bytes));
    ctx->allocs.push_back(p);
    if (!v.empty()) HIPCHECK(ctx, hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    else HIPCHECK(ctx, hipMemset(p, 0, bytes));
    *dptr = static_cast<T *>(p);
    return 0;
}

template <typename T>
static int alloc_zero(edenhip_ctx *ctx, T **dptr, size_t n) {
    size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    void *p = nullptr;
    HIPCHECK(ctx, hipMalloc(&p, bytes));
    ctx->allocs.push_back(p);
    HIPCHECK(ctx, hipMemset(p, 0, bytes));
    *dptr = static_cast<T *>(p);
    return 0;
}

#define TRY(x)                         \
    do {                               \
        int rc_ = (x);                 \
        if (rc_ != 0) return rc_;      \
    } while (0)

static int fail(edenhip_ctx *ctx, int code, const std::string &msg) {
    ctx->err = msg;
    return code;
}

// time of every recorded row: the reference accumulates `time += dt` in double from
// t_initial = 0 while time <= t_final, and casts to float per step (eden/Eden.cpp:11488,
// 11930-11932, 3252). Kernels continue the same double accumulation from a chunk's first row.
static std::vector<double> row_times(float dt, double t_final) {
    std::vector<double> ts;
    double time = 0;
    while (time <= t_final) {
        ts.push_back(time);
        time += (double)dt;
    }
    return ts;
}

static int build_netwin(edenhip_ctx *ctx, const edenhip_model_desc *d, const std::vector<int32_t> &point);

// Scratch of edenhip_run_static: the tree kernels' work-queue counters (next task, finished
// workgroups) of up to STATIC_QUEUE_BUCKETS buckets, in a file-scope device array that each
// GPU holds for the life of the process. Host code never passes the array's name to a kernel
// (in host code it is a stand-in, not a device address): it uses the address
// hipGetSymbolAddress returns for the current device, looked up once per device when the
// library first creates a context there.
constexpr int STATIC_QUEUE_BUCKETS = 32;
static __device__ __align__(256) int32_t scratch[2 * STATIC_QUEUE_BUCKETS];

static int static_scratch_ptr(edenhip_ctx *ctx, int32_t **out) {
    static std::mutex mu;
    static std::vector<int32_t *> ptrs;   // per device ordinal
    std::lock_guard<std::mutex> lk(mu);
    if ((int)ptrs.size() <= ctx->device) ptrs.resize((size_t)ctx->device + 1, nullptr);
    if (!ptrs[(size_t)ctx->device]) {
        int32_t *p = nullptr;
        HIPCHECK(ctx, hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(ctx, EDENHIP_E_DEVICE, "hipGetSymbolAddress(scratch) returned no address");
        ptrs[(size_t)ctx->device] = p;
    }
    *out = ptrs[(size_t)ctx->device];
    return 0;
}

// Zero n ints with a kernel (not a memset node, so that a caller may capture the call).
__global__ void zero_i32_kernel(int32_t *p, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// Sources of recorded gate and state variables: the kind in the top bits, the index below.
constexpr int64_t REC_TREE_GATE = (int64_t)1 << 60;   // the tree cells' gate table
constexpr int64_t REC_CELL_STATE = (int64_t)2 << 60;  // the point neurons' state table
constexpr int64_t REC_EXT_GATE = (int64_t)3 << 60;    // the channel pre-pass's gate variables
constexpr int64_t REC_INDEX = ((int64_t)1 << 60) - 1; // (kind 0: the point cells' gates, T.q)

// Recorded gate and state variables into trace row `row`: column col[i] takes the value src[i]
// names.
__global__ void record_gates_kernel(DevTables T, const float *__restrict__ gq, const float *__restrict__ cs,
                                    const float *__restrict__ xq, const int32_t *__restrict__ col,
                                    const int64_t *__restrict__ src, int32_t n, int64_t row) {
    const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const int64_t k = src[i] & REC_INDEX, kind = src[i] >> 60;
    trace_at(T, col[i], row) = kind == 0 ? T.q[k] : kind == 1 ? gq[k] : kind == 2 ? cs[k] : xq[k];
}

// Streams come from a per-device pool created once per process and kept until it ends:
// slot 0 is a context's main stream, which also carries its first (largest) tree bucket, and
// slots 1-3 the other buckets' streams. The runtime binds each stream to one of a few
// hardware queues (4 by default) when the stream is created, so streams made afresh for
// every context land on different queues from one context to the next (measured: the same
// C4 run took 656 ms in a process's first context, 725 ms in later ones), and two streams
// that share a queue run their kernels one after the other (measured with five streams: the
// 64-thread bucket waited for the 128-thread one, 496 ms per C4 run). Four pool streams keep
// one binding each for every context. Contexts of one device share the pool's streams, so
// their work is ordered on them (one host thread per context, as documented).
static int pool_stream(edenhip_ctx *ctx, int slot, hipStream_t *out) {
    static std::mutex mu;
    static std::vector<std::vector<hipStream_t>> pools;
    std::lock_guard<std::mutex> lk(mu);
    if ((int)pools.size() <= ctx->device) pools.resize((size_t)ctx->device + 1);
    auto &p = pools[(size_t)ctx->device];
    if (p.empty()) {
        p.assign(4, nullptr);
        for (auto &st : p) HIPCHECK(ctx, hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    }
    *out = p[(size_t)slot];
    return 0;
}

// floats in a LEMS parameter table: the types' vectors, then any per-instance vectors
static int64_t lems_param_count(const int32_t *param_begin, int32_t n_types, int64_t n_params) {
    return std::max<int64_t>(param_begin[n_types], n_params);
}

static int build(edenhip_ctx *ctx, const edenhip_model_desc *d) {
    if (d->abi_version < 1 || d->abi_version > EDENHIP_ABI_VERSION) return fail(ctx, EDENHIP_E_INVALID, "abi_version mismatch");
    const edenhip_lems_synapses *LS = d->abi_version >= 2 ? d->lems : nullptr;
    const bool lems = LS && LS->n_types > 0;
    if (lems) {
        if (!LS->source || !LS->param_begin || !LS->n_state) return fail(ctx, EDENHIP_E_INVALID, "lems: null table");
        if (LS->n_types > 4096) return fail(ctx, EDENHIP_E_LIMIT, "lems: more than 4096 synapse types");
        for (int t = 0; t < LS->n_types; t++) {
            if (LS->n_state[t] < 0 || LS->n_state[t] > 8) return fail(ctx, EDENHIP_E_LIMIT, "lems: more than 8 state variables");
            if (LS->param_begin[t + 1] < LS->param_begin[t]) return fail(ctx, EDENHIP_E_INVALID, "lems: param_begin decreasing");
        }
        if (d->abi_version >= 4 && LS->instance_param) {
            const int64_t nsyn = d->comp_syn_begin[d->n_comps], np = lems_param_count(LS->param_begin, LS->n_types,
                                                                                          LS->n_params);
            for (int64_t k = 0; k < nsyn; k++) {
                const int64_t o = LS->instance_param[k];
                const int32_t ty = d->syn_type[k];
                if (ty < 0 || ty >= LS->n_types) return fail(ctx, EDENHIP_E_INVALID, "synapse type index out of range");
                if (o < -1 || (o >= 0 && o + (LS->param_begin[ty + 1] - LS->param_begin[ty]) > np))
                    return fail(ctx, EDENHIP_E_INVALID, "lems: instance_param out of range");
            }
        }
        if (d->abi_version >= 4 && LS->spike_list) {
            // firing-synapse inputs: lists inside spike_times, which ends with a sentinel, on
            // instances of input types that take no incoming connections
            const int64_t nsyn = d->comp_syn_begin[d->n_comps], nt = LS->n_spike_times;
            if (!LS->spike_times || nt < 1 || !(LS->spike_times[nt - 1] >= FLT_MAX))
                return fail(ctx, EDENHIP_E_INVALID, "lems: spike_times must end with a value >= FLT_MAX");
            if (nt >= (int64_t)1 << 31) return fail(ctx, EDENHIP_E_LIMIT, "lems: more than 2^31 spike times");
            for (int64_t k = 0; k < nsyn; k++) {
                const int64_t o = LS->spike_list[k];
                if (o < -1 || o >= nt) return fail(ctx, EDENHIP_E_INVALID, "lems: spike_list out of range");
                const int32_t ty = d->syn_type[k];
                if (o >= 0 && (ty < 0 || ty >= LS->n_types || !LS->type_flags || !(LS->type_flags[ty] & 1)))
                    return fail(ctx, EDENHIP_E_INVALID, "lems: a firing-synapse input's type must be an input type");
            }
            const int64_t nspk = d->comp_spike_begin[d->n_comps];
            for (int64_t k = 0; k < nspk; k++) {
                const int64_t tgt = d->spike_target[k];
                if (tgt >= 0 && tgt < nsyn && LS->spike_list[tgt] >= 0)
                    return fail(ctx, EDENHIP_E_INVALID, "lems: a firing-synapse input takes no incoming connections");
            }
        }
        if (d->abi_version >= 4 && LS->poisson_spiker) {
            // Poisson spikers: tables present, indices in range, on instances of input types
            // without a spike list that take no incoming connections
            const int64_t nsyn = d->comp_syn_begin[d->n_comps], n = LS->n_poisson_spikers;
            if (n < 0 || n >= (int64_t)1 << 31) return fail(ctx, EDENHIP_E_LIMIT, "lems: poisson spiker count");
            if (n > 0 && (!LS->spiker_rate || !LS->spiker_transient || !LS->spiker_delay || !LS->spiker_duration ||
                          !LS->spiker_instance || !LS->spiker_rng_offset))
                return fail(ctx, EDENHIP_E_INVALID, "lems: poisson spiker table missing");
            for (int64_t k = 0; k < nsyn; k++) {
                const int64_t o = LS->poisson_spiker[k];
                if (o < -1 || o >= n) return fail(ctx, EDENHIP_E_INVALID, "lems: poisson_spiker out of range");
                if (o < 0) continue;
                const int32_t ty = d->syn_type[k];
                if (ty < 0 || ty >= LS->n_types || !LS->type_flags || !(LS->type_flags[ty] & 1))
                    return fail(ctx, EDENHIP_E_INVALID, "lems: a Poisson spiker's instance must be of an input type");
                if (LS->spike_list && LS->spike_list[k] >= 0)
                    return fail(ctx, EDENHIP_E_INVALID, "lems: an instance has both a spike list and a Poisson spiker");
            }
            const int64_t nspk = d->comp_spike_begin[d->n_comps];
            for (int64_t k = 0; k < nspk; k++) {
                const int64_t tgt = d->spike_target[k];
                if (tgt >= 0 && tgt < nsyn && LS->poisson_spiker[tgt] >= 0)
                    return fail(ctx, EDENHIP_E_INVALID, "lems: a Poisson spiker's instance takes no incoming connections");
            }
        }
        if (d->abi_version >= 3 && LS->peer) {
            const int64_t nsyn = d->comp_syn_begin[d->n_comps];
            for (int64_t k = 0; k < nsyn; k++)
                if (LS->peer[k] < -1 || LS->peer[k] >= d->n_comps)
                    return fail(ctx, EDENHIP_E_INVALID, "lems: peer compartment out of range");
        }
    }
    const edenhip_lems_cells *LC = d->abi_version >= 4 ? d->lems_cells : nullptr;
    if (LC && LC->n_types > 0) {
        if (!LC->source || !LC->param_begin || !LC->params || !LC->n_state || !LC->v_state || !LC->cell_type)
            return fail(ctx, EDENHIP_E_INVALID, "lems_cells: null table");
        if (LC->n_types > 4096) return fail(ctx, EDENHIP_E_LIMIT, "lems_cells: more than 4096 cell types");
        for (int t = 0; t < LC->n_types; t++) {
            if (LC->n_state[t] < 0 || LC->n_state[t] > 16)
                return fail(ctx, EDENHIP_E_LIMIT, "lems_cells: more than 16 state variables");
            if (LC->v_state[t] < -1 || LC->v_state[t] >= LC->n_state[t])
                return fail(ctx, EDENHIP_E_INVALID, "lems_cells: v_state out of range");
            if (LC->param_begin[t + 1] < LC->param_begin[t])
                return fail(ctx, EDENHIP_E_INVALID, "lems_cells: param_begin decreasing");
        }
        if (LC->cell_param) {
            const int64_t np = lems_param_count(LC->param_begin, LC->n_types, LC->n_params);
            for (int c = 0; c < d->n_cells; c++) {
                const int64_t o = LC->cell_param[c];
                const int32_t ty = LC->cell_type[c];
                if (o == -1) continue;
                if (ty < 0 || ty >= LC->n_types || o < 0 || o + (LC->param_begin[ty + 1] - LC->param_begin[ty]) > np)
                    return fail(ctx, EDENHIP_E_INVALID, "lems_cells: cell_param out of range");
            }
        }
    } else {
        LC = nullptr;
    }
    if (!(d->dt > 0) || !(d->t_final >= 0)) return fail(ctx, EDENHIP_E_INVALID, "dt must be > 0 and t_final >= 0");
    if (d->n_cells < 0 || d->n_comps < 0) return fail(ctx, EDENHIP_E_INVALID, "negative sizes");
    const int64_t NC = d->n_comps;
    if (NC >= (int64_t)1 << 31) return fail(ctx, EDENHIP_E_LIMIT, "more than 2^31 compartments");
    if (d->cell_comp_begin[0] != 0 || d->cell_comp_begin[d->n_cells] != NC)
        return fail(ctx, EDENHIP_E_INVALID, "cell_comp_begin must start at 0 and end at n_comps");
    if (d->n_channel_types > MAX_CHAN_TYPES) return fail(ctx, EDENHIP_E_LIMIT, "more than 64 channel types");
    if (d->n_syn_types > 64) return fail(ctx, EDENHIP_E_LIMIT, "more than 64 synapse types");
    ctx->NC = NC;
    ctx->n_cells = d->n_cells;
    ctx->dt = d->dt;

    // LEMS point neurons: one compartment each, without channels or Poisson inputs; their
    // synapses take the generic path
    std::vector<int32_t> acell_slot(NC, -1), acell_comp, acell_type;
    for (int c = 0; LC && c < d->n_cells; c++) {
        const int32_t ty = LC->cell_type[c];
        if (ty < 0) continue;
        if (ty >= LC->n_types) return fail(ctx, EDENHIP_E_INVALID, "lems_cells: cell type out of range");
        const int64_t g = d->cell_comp_begin[c];
        if (d->cell_comp_begin[c + 1] != g + 1) return fail(ctx, EDENHIP_E_INVALID, "a point neuron has one compartment");
        if (d->comp_chan_begin[g + 1] != d->comp_chan_begin[g] || d->comp_poisson_begin[g + 1] != d->comp_poisson_begin[g])
            return fail(ctx, EDENHIP_E_INVALID, "a point neuron takes no channel distributions or Poisson inputs");
        if (d->comp_syn_begin[g + 1] != d->comp_syn_begin[g] && !lems)
            return fail(ctx, EDENHIP_E_INVALID, "synapses on a point neuron need the generic (lems) synapse path");
        acell_slot[g] = (int32_t)acell_comp.size();
        acell_comp.push_back((int32_t)g);
        acell_type.push_back(ty);
    }
    ctx->n_acells = (int32_t)acell_comp.size();

    // channel type table (types with user rate ComponentTypes run in the run-time compiled
    // channel pre-pass)
    const char *rates_src = d->abi_version >= 4 ? d->lems_rates_source : nullptr;
    std::vector<char> lems_rate_type((size_t)std::max(d->n_channel_types, 1), 0);
    std::vector<float4> gate_extra((size_t)std::max(d->n_channel_types, 1) * GATES_PER_TYPE, make_float4(3.f, 0.f, 0.f, 0.f));
    std::vector<DGate> gates((size_t)std::max(d->n_channel_types, 1) * GATES_PER_TYPE);
    std::memset(gates.data(), 0, gates.size() * sizeof(DGate));
    for (int t = 0; t < d->n_channel_types; t++) {
        const edenhip_channel_type &ct = d->channel_types[t];
        if (ct.n_gates < 0 || ct.n_gates > GATES_PER_TYPE) return fail(ctx, EDENHIP_E_LIMIT, "channel type with more than 4 gates");
        // a conductance scaling other than 1: the channel runs in the pre-pass
        if (d->abi_version >= 4 && d->chan_scaling && d->chan_scaling[t] != 1.0f) lems_rate_type[(size_t)t] = 1;
        for (int gi = 0; gi < ct.n_gates; gi++) {
            const edenhip_gate &g = ct.gates[gi];
            if (g.instances < 0 || g.instances > 255) return fail(ctx, EDENHIP_E_INVALID, "gate instances out of range");
            for (const edenhip_rate *r : {&g.alpha, &g.beta}) {
                if (r->form < 0 || r->form > EDENHIP_RATE_LEMS) return fail(ctx, EDENHIP_E_INVALID, "unknown rate form");
                if (r->form == EDENHIP_RATE_LEMS && !rates_src)
                    return fail(ctx, EDENHIP_E_INVALID, "a LEMS rate form needs lems_rates_source");
                if (r->form == EDENHIP_RATE_LEMS) lems_rate_type[(size_t)t] = 1;
            }
            const int32_t kind = (d->abi_version >= 4 && d->gate_kind) ? d->gate_kind[t * EDENHIP_MAX_GATES + gi]
                                                                       : EDENHIP_GATE_RATES;
            if (kind < EDENHIP_GATE_RATES || kind > EDENHIP_GATE_RATES_INF)
                return fail(ctx, EDENHIP_E_INVALID, "unknown gate kind");
            // instantaneous gates take the step's voltage, rates with a third rate take it from
            // the extra table: their channels run in the pre-pass
            if (kind >= EDENHIP_GATE_INSTANTANEOUS) lems_rate_type[(size_t)t] = 1;
            if (kind == EDENHIP_GATE_RATES_TAU || kind == EDENHIP_GATE_RATES_INF) {
                if (!d->gate_extra_rate) return fail(ctx, EDENHIP_E_INVALID, "a rates+tau/inf gate needs gate_extra_rate");
                const edenhip_rate &x = d->gate_extra_rate[t * EDENHIP_MAX_GATES + gi];
                if (x.form < 0 || x.form > EDENHIP_RATE_LEMS || (x.form == EDENHIP_RATE_LEMS && !rates_src))
                    return fail(ctx, EDENHIP_E_INVALID, "unknown extra rate form");
                gate_extra[(size_t)t * GATES_PER_TYPE + gi] = make_float4((float)x.form, x.rate, x.midpoint, x.scale);
            }
            DGate &G = gates[(size_t)t * GATES_PER_TYPE + gi];
            G.forms = g.alpha.form | (kind << 4) | (g.beta.form << 8) | (g.instances << 16) | (ct.n_gates << 24);
            G.q10 = g.q10;
            G.ar = g.alpha.rate; G.am = g.alpha.midpoint; G.as = g.alpha.scale;
            G.br = g.beta.rate; G.bm = g.beta.midpoint; G.bs = g.beta.scale;
        }
    }
    ctx->n_gates = (int)gates.size();
    TRY(upload(ctx, &ctx->d_gates, gates));

    // synapse types; waveformFactor from the derived parameters with the reference's
    // expression tree: peakTime = (log(td/tr) * (tr*td)) / (td - tr)
    std::vector<DSynType> st((size_t)std::max(d->n_syn_types, 1));
    std::memset(st.data(), 0, st.size() * sizeof(DSynType));
    for (int t = 0; t < d->n_syn_types; t++) {
        const edenhip_exptwo_type &s = d->syn_types[t];
        const float tr = s.tau_rise, td = s.tau_decay;
        // logf/expf rounded once from double, as the kernels' log_f/exp_f
        const float peak = ((float)std::log((double)(td / tr)) * (tr * td)) / (td - tr);
        const float wf = 1.0f / ((-(float)std::exp((double)((-peak) / tr))) + (float)std::exp((double)((-peak) / td)));
        st[t] = DSynType{s.gbase, s.erev, s.tau_rise, s.tau_decay, s.weight_prop, wf, 0.f, 0.f};
    }
    ctx->n_stypes = (int)st.size();
    TRY(upload(ctx, &ctx->d_stypes, st));

    // per-compartment packing
    std::vector<float> V(NC), C(NC), A(NC), invrc(NC), vthr(NC), chg(NC * NCH, 0.f), che(NC * NCH, 0.f), q(NC * NQ, 0.f);
    std::vector<uint32_t> chmeta(NC), gmeta(NC);
    std::vector<int32_t> pulse_begin(NC + 1), spk_begin(NC + 1), poi_begin(NC + 1), comp_cell(NC);
    for (int c = 0; c < d->n_cells; c++)
        for (int64_t g = d->cell_comp_begin[c]; g < d->cell_comp_begin[c + 1]; g++) comp_cell[g] = c;
    const int64_t n_syn = d->comp_syn_begin[NC], n_spk = d->comp_spike_begin[NC], n_poi = d->comp_poisson_begin[NC];
    ctx->n_poi = n_poi;
    const int64_t n_pulse = d->comp_pulse_begin[NC];
    if (n_syn >= (int64_t)1 << 31 || n_spk >= (int64_t)1 << 31 || n_poi >= (int64_t)1 << 31)
        return fail(ctx, EDENHIP_E_LIMIT, "more than 2^31 synapses or spike targets");
    // compartments with more channel distributions or gate variables than the compartment
    // kernels hold run their channels in ext_channels_kernel
    std::vector<int32_t> x_comp, x_chb{0}, x_type, x_qb;
    std::vector<float> x_g, x_e, x_q;
    std::vector<char> ext_comp(NC, 0);
    std::vector<int64_t> x_first_q(NC, -1), x_nq(NC, 0);   // per such compartment: its gate variables
    for (int64_t g = 0; g < NC; g++) {
        V[g] = d->v0[g]; C[g] = d->capacitance[g]; A[g] = d->area[g]; invrc[g] = d->inv_rc_axial[g];
        vthr[g] = d->v_threshold ? d->v_threshold[g] : NAN;
        const int64_t cb = d->comp_chan_begin[g], ce = d->comp_chan_begin[g + 1];
        int64_t n_gvars = 0;
        bool lems_rates_here = false;
        for (int64_t ch = cb; ch < ce; ch++) {
            const int ty = d->chan_type[ch];
            if (ty < 0 || ty >= d->n_channel_types) return fail(ctx, EDENHIP_E_INVALID, "channel type index out of range");
            n_gvars += d->channel_types[ty].n_gates;
            lems_rates_here |= lems_rate_type[(size_t)ty] != 0;
        }
        if (ce - cb > NCH || n_gvars > NQ || lems_rates_here) {
            if (acell_slot[g] >= 0) return fail(ctx, EDENHIP_E_INVALID, "a point neuron takes no channel distributions");
            ext_comp[g] = 1;
            x_comp.push_back((int32_t)g);
            x_first_q[g] = (int64_t)x_q.size();
            x_nq[g] = n_gvars;
            for (int64_t ch = cb; ch < ce; ch++) {
                const int ty = d->chan_type[ch];
                const edenhip_channel_type &ct = d->channel_types[ty];
                x_type.push_back(ty);
                x_g.push_back(d->chan_gbase[ch]);
                x_e.push_back(d->chan_erev[ch]);
                x_qb.push_back((int32_t)x_q.size());
                for (int gi = 0; gi < ct.n_gates; gi++) {
                    // setup value alpha/(alpha+beta) at V0; the initialisation steps overwrite it
                    const double a = edenhip_rate_host(ct.gates[gi].alpha, d->v0[g]);
                    const double b = edenhip_rate_host(ct.gates[gi].beta, d->v0[g]);
                    x_q.push_back((float)(a / (a + b)));
                }
            }
            if (x_q.size() >= ((size_t)1 << 31)) return fail(ctx, EDENHIP_E_LIMIT, "more than 2^31 gate variables");
            x_chb.push_back((int32_t)x_type.size());
            chmeta[g] = CHM_EXTERNAL; gmeta[g] = 0;
            pulse_begin[g + 1] = (int32_t)d->comp_pulse_begin[g + 1];
            spk_begin[g + 1] = (int32_t)d->comp_spike_begin[g + 1];
            poi_begin[g + 1] = (int32_t)d->comp_poisson_begin[g + 1];
            continue;
        }
        uint32_t cm = 0xFFFFFFFFu, gm = 0;
        int nq = 0;
        for (int64_t ch = cb; ch < ce; ch++) {
            const int slot = (int)(ch - cb);
            const int ty = d->chan_type[ch];
            if (ty < 0 || ty >= d->n_channel_types) return fail(ctx, EDENHIP_E_INVALID, "channel type index out of range");
            cm = (cm & ~(0xFFu << (8 * slot))) | ((uint32_t)ty << (8 * slot));
            chg[g * NCH + slot] = d->chan_gbase[ch];
            che[g * NCH + slot] = d->chan_erev[ch];
            const edenhip_channel_type &ct = d->channel_types[ty];
            for (int gi = 0; gi < ct.n_gates; gi++) {
                if (nq >= NQ) return fail(ctx, EDENHIP_E_LIMIT, "compartment with more than 4 gate variables");
                gm |= (uint32_t)(slot | (gi << 2)) << (4 * nq);
                // setup value alpha/(alpha+beta) at V0 (eden/Eden.cpp:5313-5321); the
                // initialisation steps overwrite it with inf(V0)
                const double a = edenhip_rate_host(ct.gates[gi].alpha, d->v0[g]);
                const double b = edenhip_rate_host(ct.gates[gi].beta, d->v0[g]);
                q[g * NQ + nq] = (float)(a / (a + b));
                nq++;
            }
        }
        gm |= (uint32_t)nq << 24;
        gm |= (uint32_t)(ce - cb) << 28;
        chmeta[g] = cm; gmeta[g] = gm;
        pulse_begin[g + 1] = (int32_t)d->comp_pulse_begin[g + 1];
        spk_begin[g + 1] = (int32_t)d->comp_spike_begin[g + 1];
        poi_begin[g + 1] = (int32_t)d->comp_poisson_begin[g + 1];
    }
    std::vector<float> pstart(n_pulse), pend(n_pulse), pamp(n_pulse);
    for (int64_t p = 0; p < n_pulse; p++) {
        pstart[p] = d->pulse_start[p];
        pend[p] = d->pulse_start[p] + d->pulse_duration[p];       // Start + Duration (fp32)
        pamp[p] = d->pulse_amplitude[p] * d->pulse_weight[p];     // Imax * Weight (fp32)
    }
    // synapse types must be grouped in ascending type order per compartment
    for (int64_t g = 0; g < NC; g++)
        for (int64_t s = d->comp_syn_begin[g] + 1; s < d->comp_syn_begin[g + 1]; s++)
            if (d->syn_type[s] < d->syn_type[s - 1]) return fail(ctx, EDENHIP_E_INVALID, "synapses of a compartment must be sorted by type");
    const int32_t n_inst_types = lems ? LS->n_types : d->n_syn_types;
    for (int64_t s = 0; s < n_syn; s++)
        if (d->syn_type[s] < 0 || d->syn_type[s] >= n_inst_types) return fail(ctx, EDENHIP_E_INVALID, "synapse type index out of range");
    for (int64_t k = 0; k < n_spk; k++)
        if (d->spike_target[k] < 0 || d->spike_target[k] >= n_syn) return fail(ctx, EDENHIP_E_INVALID, "spike target out of range");
    std::vector<float> poi_mean(n_poi), poi_w(n_poi);
    std::vector<int32_t> poi_st(n_poi), poi_inst(n_poi), poi_roff(n_poi);
    for (int64_t p = 0; p < n_poi; p++) {
        poi_mean[p] = 1.0f / d->poisson_rate[p];
        poi_w[p] = d->poisson_weight[p];
        poi_st[p] = d->poisson_syn_type[p];
        if (poi_st[p] < 0 || poi_st[p] >= d->n_syn_types) return fail(ctx, EDENHIP_E_INVALID, "poisson synapse type out of range");
        poi_inst[p] = d->poisson_instance[p];
        poi_roff[p] = d->poisson_rng_offset[p];
    }
    std::vector<uint32_t> seeds(std::max(d->n_cells, 1));
    for (int c = 0; c < d->n_cells; c++) seeds[c] = d->cell_rng_seed ? d->cell_rng_seed[c] : 0u;

    // recording maps
    std::vector<int32_t> rec_col(NC, -1), ev_col(NC, -1);
    ctx->desc_to_dev_col.assign(d->n_record, -1);
    ctx->n_dev_cols = 0;
    // V columns first, then one column per distinct recorded gate variable
    // (a point neuron's columns all read its state variables: V is its membrane-voltage one)
    std::vector<std::pair<int64_t, int32_t>> grec;   // (compartment, gate or state variable) of these columns
    std::unordered_map<int64_t, int32_t> grec_at;    // compartment * 32 + variable -> index in grec
    const bool gate_rec = d->abi_version >= 3 && d->record_gate;
    std::vector<int32_t> rec_var(d->n_record, -1);
    for (int64_t r = 0; r < d->n_record; r++) {
        int64_t g = d->record_comp[r];
        if (g < 0 || g >= NC) return fail(ctx, EDENHIP_E_INVALID, "record_comp out of range");
        const int32_t j = gate_rec ? d->record_gate[r] : -1;
        if (acell_slot[g] >= 0) {
            const int32_t ty = acell_type[(size_t)acell_slot[g]];
            const int32_t v = j < 0 ? LC->v_state[ty] : j;
            if (j < -1 || v < 0 || v >= LC->n_state[ty])
                return fail(ctx, EDENHIP_E_INVALID, "record_gate: a point neuron's state variable (-1: its membrane voltage)");
            rec_var[r] = v;
            continue;
        }
        const int64_t nq_here = ext_comp[g] ? x_nq[g] : (int64_t)((gmeta[g] >> 24) & 7u);
        if (j < -1 || j >= nq_here)
            return fail(ctx, EDENHIP_E_INVALID, "record_gate: -1 or a gate variable of the compartment");
        rec_var[r] = j;
        if (j >= 0) continue;
        if (rec_col[g] < 0) rec_col[g] = ctx->n_dev_cols++;
        ctx->desc_to_dev_col[r] = rec_col[g];
    }
    for (int64_t r = 0; r < d->n_record; r++) {
        const int64_t g = d->record_comp[r];
        const int32_t j = rec_var[r];
        if (j < 0) continue;
        auto it = grec_at.emplace(g * 32 + j, (int32_t)grec.size()).first;
        if (it->second == (int32_t)grec.size()) grec.push_back({g, j});
        ctx->desc_to_dev_col[r] = ctx->n_dev_cols + it->second;
    }
    ctx->n_dev_cols += (int32_t)grec.size();
    // state read back between launches: recorded gate variables need one launch per step
    if (!grec.empty()) {
        ctx->opt.persistent = -1;
        ctx->opt.network_window = -1;
    }
    if (d->n_event_cols >= (int64_t)1 << 24) return fail(ctx, EDENHIP_E_LIMIT, "more than 2^24 event columns");
    ctx->ev_expand.assign((size_t)d->n_event_cols, {});
    for (int64_t e = 0; e < d->n_event_cols; e++) {
        int64_t g = d->event_comp[e];
        if (g < 0 || g >= NC) return fail(ctx, EDENHIP_E_INVALID, "event_comp out of range");
        if (ev_col[g] < 0) ev_col[g] = (int32_t)e;
        ctx->ev_expand[ev_col[g]].push_back(e);
    }

    // time table
    std::vector<double> ts = row_times(d->dt, d->t_final);
    ctx->total_rows = (int64_t)ts.size();

    // cells: point (1 compartment) vs tree
    std::vector<int32_t> point;
    std::vector<TreeCell> cells;
    std::vector<int32_t> cell_tpb;
    std::vector<uint32_t> words;
    std::vector<int32_t> posv(NC, 0);
    std::vector<float2> uv(NC, make_float2(0.f, 0.f));
    // tree cells: compartment word (channel count | gated-slot mask << 3 | first channel task
    // << 16), channel tasks (compartment | channel type << 11 | first gate << 17) and the gate
    // state, in compartment / distribution / gate order
    std::vector<uint32_t> tcm(NC, 0), ctask, inl;
    std::vector<float> gqv;
    std::vector<uint32_t> cflag(NC, 0);
    std::vector<TreeComplex> cxv;
    std::vector<uint32_t> gtv;   // per gate variable: its channel task (local to the cell) | gate index << 16
    std::vector<int64_t> gq_first(NC, -1);   // tree compartments: index of the first gate variable in gqv
    // Hines schedules of the tree cells, planned in parallel (independent per cell); cells
    // with malformed parents are left unplanned here and rejected in the loop below
    std::vector<WavePlan> plans(d->n_cells);
    {
        auto plan_one = [&](int c) {
            const int64_t c0 = d->cell_comp_begin[c];
            const int32_t n = (int32_t)(d->cell_comp_begin[c + 1] - c0);
            if (n <= 1 || n > MAX_TREE_N || d->parent[c0] != -1) return;
            for (int32_t i = 1; i < n; i++)
                if (d->parent[c0 + i] < 0 || d->parent[c0 + i] >= i) return;
            const int32_t tpb = choose_tpb(n, ctx->opt.threads_per_cell, ctx->opt.size_limits);
            // wave 0 updates the root compartment during the first Hines phase; 2- and 4-wave
            // cells also update their gates on the last cable wave (they have no channel wave)
            const bool gates_on_last = tpb == 128 || tpb == 256;
            // constant trees take merged fold tasks (their kernel path forms right-hand sides only)
            plans[c] = plan_tree_waves(n, d->parent + c0, tpb / 64, 0.5, gates_on_last ? 2.0 : 0.0,
                                       const_below_root(d, c0, n));
        };
        const int nthr = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::atomic<int> next{0};
        auto worker = [&]() {
            for (int c; (c = next.fetch_add(1)) < d->n_cells;) plan_one(c);
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < nthr; t++) pool.emplace_back(worker);
        worker();
        for (auto &th : pool) th.join();
    }
    for (int c = 0; c < d->n_cells; c++) {
        const int64_t c0 = d->cell_comp_begin[c], c1 = d->cell_comp_begin[c + 1];
        const int32_t n = (int32_t)(c1 - c0);
        if (n <= 0) return fail(ctx, EDENHIP_E_INVALID, "cell without compartments");
        if (d->parent[c0] != -1) return fail(ctx, EDENHIP_E_INVALID, "compartment 0 of a cell must be its root");
        for (int32_t i = 1; i < n; i++)
            if (d->parent[c0 + i] < 0 || d->parent[c0 + i] >= i) return fail(ctx, EDENHIP_E_INVALID, "parents must precede children");
        if (acell_slot[c0] >= 0) continue;   // a point neuron: lems_cell_kernel
        if (n == 1) {
            point.push_back((int32_t)c0);
            ctx->point_desc.push_back(c);
            continue;
        }
        if (n > MAX_TREE_N) return fail(ctx, EDENHIP_E_LIMIT, "cell with more than 2047 compartments");
        const int32_t tpb = choose_tpb(n, ctx->opt.threads_per_cell, ctx->opt.size_limits);
        const WavePlan &P = plans[c];
        TreeCell tc{};
        tc.gate_off = (int64_t)gqv.size();
        tc.ctask_off = (int32_t)ctask.size();
        for (int32_t i = 0; i < n; i++) {
            const int64_t g = c0 + i;
            gq_first[g] = (int64_t)gqv.size();
            const int32_t first_task = (int32_t)(ctask.size() - tc.ctask_off);
            uint32_t mask = 0;
            for (int64_t ch = d->comp_chan_begin[g]; !ext_comp[g] && ch < d->comp_chan_begin[g + 1]; ch++) {
                const int slot = (int)(ch - d->comp_chan_begin[g]);
                const edenhip_channel_type &ct = d->channel_types[d->chan_type[ch]];
                if (ct.n_gates == 0) continue;
                const int64_t first_gate = (int64_t)gqv.size() - tc.gate_off;
                if (first_gate >= (1 << 15)) return fail(ctx, EDENHIP_E_LIMIT, "more than 32767 gate variables in a cell");
                mask |= 1u << slot;
                ctask.push_back((uint32_t)i | ((uint32_t)d->chan_type[ch] << 11) | ((uint32_t)first_gate << 17));
                for (int gi = 0; gi < ct.n_gates; gi++) {
                    gtv.push_back((uint32_t)(ctask.size() - 1 - (size_t)tc.ctask_off) | ((uint32_t)gi << 16));
                    const double a = edenhip_rate_host(ct.gates[gi].alpha, d->v0[g]);
                    const double b = edenhip_rate_host(ct.gates[gi].beta, d->v0[g]);
                    gqv.push_back((float)(a / (a + b)));
                }
            }
            const int nch = ext_comp[g] ? 0 : (int)(d->comp_chan_begin[g + 1] - d->comp_chan_begin[g]);
            if (first_task > 0xFFFF) return fail(ctx, EDENHIP_E_LIMIT, "more than 65535 gated channels in a cell");
            tcm[g] = (uint32_t)nch | (mask << 3) | ((uint32_t)first_task << 16) | (ext_comp[g] ? TCM_XCH : 0u);
        }
        tc.n_ctask = (int32_t)(ctask.size() - tc.ctask_off);
        tc.n_gate = (int32_t)(gqv.size() - tc.gate_off);
        // input list: compartments with Poisson synapse inputs
        tc.input_off = (int32_t)inl.size();
        for (int32_t i = 0; i < n; i++) {
            const int64_t g = c0 + i;
            if (d->comp_poisson_begin[g + 1] == d->comp_poisson_begin[g]) continue;
            inl.push_back((uint32_t)i);
            tcm[g] |= TCM_INPUT;
        }
        // compartments with chemical synapses run them in their complex update
        for (int32_t i = 0; i < n; i++) {
            const int64_t g = c0 + i;
            if (d->comp_syn_begin[g + 1] != d->comp_syn_begin[g]) tcm[g] |= TCM_SYN;
        }
        tc.n_input = (int32_t)(inl.size() - tc.input_off);
        // complex compartments: anything but one ungated channel with no pulse, no Poisson
        // input, no synapse, no spike event column, no spike recipients and no trace column;
        // their constants go to a per-cell record table
        tc.cx_off = (int32_t)cxv.size();
        for (int32_t i = 0; i < n; i++) {
            const int64_t g = c0 + i;
            const bool simple = (tcm[g] & 0xFFFFu) == 1u && pulse_begin[g + 1] == pulse_begin[g] && ev_col[g] < 0 &&
                                rec_col[g] < 0 && spk_begin[g + 1] == spk_begin[g];
            if (simple) continue;
            TreeComplex x{};
            x.k.C = C[g]; x.k.A = A[g]; x.k.invrc = invrc[g]; x.k.vthr = vthr[g];
            x.k.chmeta = chmeta[g]; x.k.gmeta = gmeta[g];
            for (int k = 0; k < NCH; k++) { x.k.g[k] = chg[g * NCH + k]; x.k.e[k] = che[g * NCH + k]; }
            const int32_t pb = pulse_begin[g], pe = pulse_begin[g + 1];
            if (pe > pb) {
                x.k.p0s = pstart[pb]; x.k.p0e = pend[pb]; x.k.p0a = pamp[pb];
                x.k.gmeta |= GM_PULSE;
                if (pe > pb + 1) x.k.gmeta |= GM_MORE_PULSES;
            } else {
                x.k.p0s = 0.f; x.k.p0e = -1.f; x.k.p0a = 0.f;
            }
            x.k.rec = rec_col[g]; x.k.ev = ev_col[g];
            x.tcm = tcm[g];
            const int32_t idx = (int32_t)(cxv.size() - (size_t)tc.cx_off);
            if (idx > 0xFFFF) return fail(ctx, EDENHIP_E_LIMIT, "more than 65536 complex compartments in a cell");
            cflag[g] = TCX_COMPLEX | (uint32_t)idx;
            cxv.push_back(x);
        }
        tc.n_cx = (int32_t)(cxv.size() - (size_t)tc.cx_off);

        tc.comp_begin = c0;
        tc.n = n;
        tc.longest = P.longest;
        tc.n_junction = P.n_junction;
        for (int32_t i = 1; i < n; i++) tc.root_children += d->parent[c0 + i] == 0 ? 1 : 0;
        tc.words_off = (int32_t)words.size();
        tc.words_len = (int32_t)P.block.size();
        words.insert(words.end(), P.block.begin(), P.block.end());
        // edge constants by position, fp32 as computed per step by the reference:
        // Ui = -dt/(R*C[i]), Uj = -dt/(R*C[j]) with R = R_Axial[max(i, j)] = R_Axial[i]
        for (int32_t i = 0; i < n; i++) posv[c0 + i] = P.pos[i];
        for (int32_t i = 1; i < n; i++) {
            const float R = d->r_axial[c0 + i];
            const int32_t j = d->parent[c0 + i];
            const float ui = -d->dt / (R * d->capacitance[c0 + i]);
            const float uj = -d->dt / (R * d->capacitance[c0 + j]);
            uv[c0 + P.pos[i]] = make_float2(ui, uj);
        }
        // constant tree below the root: no gated channel, Poisson input or synapse on any other
        // compartment, so their Hines diagonals are the same every step
        {
            bool ok = true;
            for (int32_t i = 1; i < n && ok; i++) ok = (tcm[c0 + i] & (0xF8u | TCM_SYN | TCM_XCH)) == 0u;
            tc.const_tree = ok ? 1 : 0;
            if (ok != P.merged) return fail(ctx, EDENHIP_E_INVALID, "internal: schedule format of a constant tree");
        }
        cells.push_back(tc);
        cell_tpb.push_back(tpb);
        ctx->tree_desc.push_back(c);
    }
    ctx->n_point = (int32_t)point.size();
    // block-interleaved synapse slots for the point cells (64 per wave, in point order)
    std::vector<int32_t> syn_base(NC, 0), syn_cnt(NC, 0);
    std::vector<int64_t> blocked_of(n_syn, -1);
    int64_t n_slots = 0;
    for (size_t b0 = 0; b0 < point.size(); b0 += 64) {
        int64_t kmax = 0;
        for (size_t l = 0; l < 64 && b0 + l < point.size(); l++) {
            const int64_t g = point[b0 + l];
            kmax = std::max<int64_t>(kmax, d->comp_syn_begin[g + 1] - d->comp_syn_begin[g]);
        }
        for (size_t l = 0; l < 64 && b0 + l < point.size(); l++) {
            const int64_t g = point[b0 + l];
            const int64_t cnt = d->comp_syn_begin[g + 1] - d->comp_syn_begin[g];
            syn_base[g] = (int32_t)(n_slots + (int64_t)l);
            syn_cnt[g] = (int32_t)cnt;
            for (int64_t k = 0; k < cnt; k++) blocked_of[d->comp_syn_begin[g] + k] = n_slots + 64 * k + (int64_t)l;
        }
        n_slots += 64 * kmax;
        if (n_slots >= ((int64_t)1 << 31)) return fail(ctx, EDENHIP_E_LIMIT, "more than 2^31 synapse slots");
    }
    // the same layout for the tree compartments that have synapses, 64 compartments a block
    {
        std::vector<int64_t> tsyn;
        for (int32_t c : ctx->tree_desc)
            for (int64_t g = d->cell_comp_begin[c]; g < d->cell_comp_begin[c + 1]; g++)
                if (d->comp_syn_begin[g + 1] > d->comp_syn_begin[g]) tsyn.push_back(g);
        for (size_t b0 = 0; b0 < tsyn.size(); b0 += 64) {
            int64_t kmax = 0;
            for (size_t l = 0; l < 64 && b0 + l < tsyn.size(); l++)
                kmax = std::max<int64_t>(kmax, d->comp_syn_begin[tsyn[b0 + l] + 1] - d->comp_syn_begin[tsyn[b0 + l]]);
            for (size_t l = 0; l < 64 && b0 + l < tsyn.size(); l++) {
                const int64_t g = tsyn[b0 + l];
                const int64_t cnt = d->comp_syn_begin[g + 1] - d->comp_syn_begin[g];
                syn_base[g] = (int32_t)(n_slots + (int64_t)l);
                syn_cnt[g] = (int32_t)cnt;
                for (int64_t k = 0; k < cnt; k++) blocked_of[d->comp_syn_begin[g] + k] = n_slots + 64 * k + (int64_t)l;
            }
            n_slots += 64 * kmax;
            if (n_slots >= ((int64_t)1 << 31)) return fail(ctx, EDENHIP_E_LIMIT, "more than 2^31 synapse slots");
        }
        ctx->tree_net = !tsyn.empty();
    }
    // and for the point neurons that have synapses
    for (size_t b0 = 0; b0 < acell_comp.size(); b0 += 64) {
        int64_t kmax = 0;
        for (size_t l = 0; l < 64 && b0 + l < acell_comp.size(); l++) {
            const int64_t g = acell_comp[b0 + l];
            kmax = std::max<int64_t>(kmax, d->comp_syn_begin[g + 1] - d->comp_syn_begin[g]);
        }
        for (size_t l = 0; l < 64 && b0 + l < acell_comp.size(); l++) {
            const int64_t g = acell_comp[b0 + l];
            const int64_t cnt = d->comp_syn_begin[g + 1] - d->comp_syn_begin[g];
            syn_base[g] = (int32_t)(n_slots + (int64_t)l);
            syn_cnt[g] = (int32_t)cnt;
            for (int64_t k = 0; k < cnt; k++) blocked_of[d->comp_syn_begin[g] + k] = n_slots + 64 * k + (int64_t)l;
        }
        n_slots += 64 * kmax;
        if (n_slots >= ((int64_t)1 << 31)) return fail(ctx, EDENHIP_E_LIMIT, "more than 2^31 synapse slots");
    }
    for (int64_t s = 0; s < n_syn; s++)
        if (blocked_of[s] < 0) return fail(ctx, EDENHIP_E_INVALID, "internal: synapse without a slot");
    std::vector<int32_t> syn_type(n_slots, 0), syn_orig(n_slots, -1), spk_target(n_spk);
    std::vector<float> syn_weight(n_slots, 0.f), syn_delay(n_slots, 0.f), syn_next(n_slots, -INFINITY);
    for (int64_t s = 0; s < n_syn; s++) {
        const int64_t k = blocked_of[s];
        syn_type[k] = d->syn_type[s];
        syn_weight[k] = d->syn_weight[s];
        syn_delay[k] = d->syn_delay[s];
        syn_orig[k] = (int32_t)s;
    }
    for (int64_t k = 0; k < n_spk; k++) spk_target[k] = (int32_t)blocked_of[d->spike_target[k]];
    bool any_targets = n_spk > 0;
    ctx->n_slots = n_slots;
    // generic synapses run in their own kernel before the compartments of each step: one
    // launch per step, like a network
    ctx->network = any_targets || (lems && n_syn > 0) || ctx->n_acells > 0 || !x_comp.empty();
    ctx->tree_net = ctx->tree_net || ctx->network;
    ctx->point_quad = n_syn == 0 && n_poi == 0 && n_spk == 0 && x_comp.empty();
    if (!x_comp.empty()) {
        // channels beyond the compartment kernels' slots: their own pre-pass each step
        ctx->n_xcomp = (int32_t)x_comp.size();
        ctx->n_xq = (int64_t)x_q.size();
        TRY(upload(ctx, &ctx->d_x_comp, x_comp));
        TRY(upload(ctx, &ctx->d_x_chb, x_chb));
        TRY(upload(ctx, &ctx->d_x_type, x_type));
        TRY(upload(ctx, &ctx->d_x_g, x_g));
        TRY(upload(ctx, &ctx->d_x_e, x_e));
        TRY(upload(ctx, &ctx->d_x_qb, x_qb));
        TRY(upload(ctx, &ctx->d_x_q, x_q.empty() ? std::vector<float>(1, 0.0f) : x_q));
        TRY(alloc_zero(ctx, &ctx->d_chx_ext, (size_t)NC));
        ctx->T.chx_ext = ctx->d_chx_ext;
        TRY(upload(ctx, &ctx->d_gate_extra, gate_extra));
        if (d->abi_version >= 4 && d->chan_scaling)
            TRY(upload(ctx, &ctx->d_chan_scaling,
                       std::vector<float>(d->chan_scaling, d->chan_scaling + std::max(d->n_channel_types, 1))));
    }

    // recorded gate variables: point cells keep theirs in T.q (NQ slots per compartment in
    // distribution / gate order), tree cells in the gate table gqv
    std::vector<int32_t> grec_col(grec.size());
    std::vector<int64_t> grec_src(grec.size());
    for (size_t k = 0; k < grec.size(); k++) {
        const int64_t g = grec[k].first;
        const int64_t j = grec[k].second;
        grec_col[k] = ctx->n_dev_cols - (int32_t)grec.size() + (int32_t)k;
        if (acell_slot[g] >= 0)
            grec_src[k] = REC_CELL_STATE | (j * ctx->n_acells + acell_slot[g]);
        else if (ext_comp[g])
            grec_src[k] = REC_EXT_GATE | (x_first_q[g] + j);
        else if (gq_first[g] >= 0)
            grec_src[k] = REC_TREE_GATE | (gq_first[g] + j);
        else
            grec_src[k] = g * NQ + j;
    }
    ctx->n_grec = (int32_t)grec.size();
    if (ctx->n_grec > 0) {
        TRY(upload(ctx, &ctx->d_grec_col, grec_col));
        TRY(upload(ctx, &ctx->d_grec_src, grec_src));
    }

    // device uploads
    DevTables &T = ctx->T;
    TRY(upload(ctx, &T.V, V));
    TRY(upload(ctx, &T.q, q));
    TRY(upload(ctx, (float **)&T.C, C));
    TRY(upload(ctx, (float **)&T.A, A));
    TRY(upload(ctx, (float **)&T.invrc, invrc));
    TRY(upload(ctx, (float **)&T.vthr, vthr));
    TRY(upload(ctx, (uint32_t **)&T.chmeta, chmeta));
    TRY(upload(ctx, (uint32_t **)&T.gmeta, gmeta));
    TRY(upload(ctx, (float **)&T.chg, chg));
    TRY(upload(ctx, (float **)&T.che, che));
    TRY(upload(ctx, (int32_t **)&T.pulse_begin, pulse_begin));
    TRY(upload(ctx, (float **)&T.pulse_start, pstart));
    TRY(upload(ctx, (float **)&T.pulse_end, pend));
    TRY(upload(ctx, (float **)&T.pulse_amp, pamp));
    TRY(upload(ctx, (int32_t **)&T.rec_col, rec_col));
    TRY(upload(ctx, (int32_t **)&T.ev_col, ev_col));
    TRY(upload(ctx, (int32_t **)&T.syn_base, syn_base));
    TRY(upload(ctx, (int32_t **)&T.syn_cnt, syn_cnt));
    TRY(upload(ctx, (int32_t **)&T.syn_orig, syn_orig));
    TRY(upload(ctx, (int32_t **)&T.syn_type, syn_type));
    TRY(upload(ctx, (float **)&T.syn_weight, syn_weight));
    TRY(upload(ctx, (float **)&T.syn_delay, syn_delay));
    TRY(upload(ctx, &T.syn_next, syn_next));
    TRY(alloc_zero(ctx, &T.syn_A, (size_t)n_slots));
    TRY(alloc_zero(ctx, &T.syn_B, (size_t)n_slots));
    TRY(alloc_zero(ctx, &T.trig[0], (size_t)n_slots));
    TRY(alloc_zero(ctx, &T.trig[1], (size_t)n_slots));
    TRY(upload(ctx, (int32_t **)&T.spk_begin, spk_begin));
    TRY(upload(ctx, (int32_t **)&T.spk_target, spk_target));
    TRY(upload(ctx, (int32_t **)&T.poi_begin, poi_begin));
    TRY(upload(ctx, (float **)&T.poi_isi_mean, poi_mean));
    TRY(upload(ctx, (float **)&T.poi_weight, poi_w));
    TRY(upload(ctx, (int32_t **)&T.poi_syn_type, poi_st));
    TRY(upload(ctx, (int32_t **)&T.poi_inst, poi_inst));
    TRY(upload(ctx, (int32_t **)&T.poi_roff, poi_roff));
    TRY(alloc_zero(ctx, &T.poi_tsince, (size_t)n_poi));
    TRY(alloc_zero(ctx, &T.poi_isi, (size_t)n_poi));
    TRY(alloc_zero(ctx, &T.poi_A, (size_t)n_poi));
    TRY(alloc_zero(ctx, &T.poi_B, (size_t)n_poi));
    if (lems && n_syn > 0) {
        // generic synapse tables: compartments with synapses, parameters, state [k][slot]
        std::vector<int32_t> lc;
        for (int64_t g = 0; g < NC; g++)
            if (syn_cnt[g] > 0) lc.push_back((int32_t)g);
        ctx->n_lems_comps = (int32_t)lc.size();
        int32_t smax = 1;
        for (int t = 0; t < LS->n_types; t++) smax = std::max(smax, LS->n_state[t]);
        const int64_t np = lems_param_count(LS->param_begin, LS->n_types, d->abi_version >= 4 ? LS->n_params : 0);
        if (np >= (int64_t)1 << 31) return fail(ctx, EDENHIP_E_LIMIT, "lems: more than 2^31 parameters");
        std::vector<float> prm(LS->params, LS->params + np);
        std::vector<int32_t> pb(LS->param_begin, LS->param_begin + LS->n_types + 1);
        if (d->abi_version >= 4 && LS->instance_param) {
            // each slot's parameter vector: its instance's own, else its type's
            std::vector<int32_t> poff((size_t)n_slots, 0);
            for (int64_t k = 0; k < n_slots; k++) {
                const int32_t s = syn_orig[(size_t)k];
                if (s < 0) continue;
                const int64_t o = LS->instance_param[s];
                poff[(size_t)k] = o >= 0 ? (int32_t)o : pb[(size_t)d->syn_type[s]];
            }
            TRY(upload(ctx, &ctx->d_lems_poff, poff));
        }
        if (d->abi_version >= 4 && LS->spike_list) {
            std::vector<int32_t> tpos((size_t)n_slots, -1);
            bool any = false;
            for (int64_t k = 0; k < n_slots; k++) {
                const int32_t s = syn_orig[(size_t)k];
                if (s >= 0 && LS->spike_list[s] >= 0) {
                    tpos[(size_t)k] = (int32_t)LS->spike_list[s];
                    any = true;
                }
            }
            if (any) {
                std::vector<float> tt(LS->spike_times, LS->spike_times + LS->n_spike_times);
                TRY(upload(ctx, &ctx->d_lems_tpos, tpos));
                TRY(upload(ctx, &ctx->d_lems_ttimes, tt));
            }
        }
        if (d->abi_version >= 4 && LS->poisson_spiker && LS->n_poisson_spikers > 0) {
            const int64_t n = LS->n_poisson_spikers;
            std::vector<int32_t> slot_pk((size_t)n_slots, -1);
            for (int64_t k = 0; k < n_slots; k++) {
                const int32_t s = syn_orig[(size_t)k];
                if (s >= 0 && LS->poisson_spiker[s] >= 0) slot_pk[(size_t)k] = (int32_t)LS->poisson_spiker[s];
            }
            std::vector<float> isi((size_t)n);
            std::vector<float2> win((size_t)n);
            for (int64_t k = 0; k < n; k++) {
                isi[(size_t)k] = 1.0f / LS->spiker_rate[k];   // averageIsi = 1 / averageRate
                win[(size_t)k] = make_float2(LS->spiker_delay[k], LS->spiker_duration[k]);
            }
            TRY(upload(ctx, &ctx->d_pk_slot, slot_pk));
            TRY(upload(ctx, &ctx->d_pk_isi, isi));
            TRY(upload(ctx, &ctx->d_pk_window, win));
            TRY(upload(ctx, &ctx->d_pk_transient, std::vector<int32_t>(LS->spiker_transient, LS->spiker_transient + n)));
            TRY(upload(ctx, &ctx->d_pk_inst, std::vector<int32_t>(LS->spiker_instance, LS->spiker_instance + n)));
            TRY(upload(ctx, &ctx->d_pk_roff, std::vector<int32_t>(LS->spiker_rng_offset, LS->spiker_rng_offset + n)));
            TRY(alloc_zero(ctx, &ctx->d_pk_state, (size_t)n));
            ctx->n_pk = n;
        }
        TRY(upload(ctx, &ctx->d_lems_comps, lc));
        TRY(upload(ctx, &ctx->d_lems_params, prm));
        TRY(upload(ctx, &ctx->d_lems_pbegin, pb));
        TRY(alloc_zero(ctx, &ctx->d_lems_state, (size_t)(smax * n_slots)));
        std::vector<int32_t> peer((size_t)n_slots, -1);
        if (d->abi_version >= 3 && LS->peer)
            for (int64_t k = 0; k < n_slots; k++)
                if (syn_orig[(size_t)k] >= 0) peer[(size_t)k] = (int32_t)LS->peer[syn_orig[(size_t)k]];
        TRY(upload(ctx, &ctx->d_lems_peer, peer));
        TRY(alloc_zero(ctx, &ctx->d_syn_ext, (size_t)NC));
        T.syn_ext = ctx->d_syn_ext;
        if (d->abi_version >= 4 && LS->type_flags) {
            std::vector<int32_t> fl(LS->type_flags, LS->type_flags + LS->n_types);
            bool any_input = false;
            for (int64_t k = 0; k < n_syn; k++) any_input |= (fl[(size_t)d->syn_type[k]] & 1) != 0;
            if (any_input) {
                TRY(upload(ctx, &ctx->d_lems_flags, fl));
                TRY(alloc_zero(ctx, &ctx->d_inp_ext, (size_t)NC));
                T.inp_ext = ctx->d_inp_ext;
            }
        }
    }
    const bool rates_rtc = rates_src && ctx->n_xcomp > 0;
    if ((lems && n_syn > 0) || ctx->n_acells > 0 || rates_rtc) {
        // the generated synapse, point-neuron and rate functions, each stubbed when absent
        std::string gen = lems ? std::string(LS->source)
                               : std::string("__device__ __forceinline__ void lems_syn(int, const float *, const float *, "
                                             "float *, long, float, float, float, int, int, float, float *i, float *g) "
                                             "{ *i = 0.0f; *g = 0.0f; }\n");
        gen += "\n";
        gen += LC ? std::string(LC->source)
                  : std::string("__device__ __forceinline__ void lems_cell(int, const float *, const float *, float *, "
                                "long, float, float, int, float, int, long long, int *spike) { *spike = 0; }\n");
        gen += "\n";
        gen += rates_src ? std::string(rates_src)
                         : std::string("__device__ float lems_rate(int, float) { return 0.0f; }\n");
        std::string err;
        hipFunction_t syn_fn = nullptr, cell_fn = nullptr, ext_fn = nullptr;
        if (rtc_lems_kernel(gen, &syn_fn, &cell_fn, &ext_fn, err) != 0) return fail(ctx, EDENHIP_E_INVALID, err);
        if (lems && n_syn > 0) ctx->lems_fn = syn_fn;
        if (ctx->n_acells > 0) ctx->lems_cell_fn = cell_fn;
        if (rates_rtc) ctx->ext_rtc_fn = ext_fn;
    }
    if (ctx->n_acells > 0) {
        // point-neuron tables: compartments, types, parameters, state [variable][slot]
        int32_t smax = 1;
        for (int t = 0; t < LC->n_types; t++) smax = std::max(smax, LC->n_state[t]);
        const int64_t np = lems_param_count(LC->param_begin, LC->n_types, LC->n_params);
        if (np >= (int64_t)1 << 31) return fail(ctx, EDENHIP_E_LIMIT, "lems_cells: more than 2^31 parameters");
        std::vector<float> prm(LC->params, LC->params + np);
        std::vector<int32_t> pb(LC->param_begin, LC->param_begin + LC->n_types + 1);
        std::vector<int32_t> vs(LC->v_state, LC->v_state + LC->n_types);
        if (LC->cell_param) {
            // each point neuron's parameter vector: its own, else its type's
            std::vector<int32_t> poff;
            for (int c = 0; c < d->n_cells; c++) {
                const int32_t ty = LC->cell_type[c];
                if (ty < 0) continue;
                poff.push_back(LC->cell_param[c] >= 0 ? (int32_t)LC->cell_param[c] : pb[(size_t)ty]);
            }
            TRY(upload(ctx, &ctx->d_acell_poff, poff));
        }
        TRY(upload(ctx, &ctx->d_acell_comps, acell_comp));
        TRY(upload(ctx, &ctx->d_acell_type, acell_type));
        TRY(upload(ctx, &ctx->d_acell_params, prm));
        TRY(upload(ctx, &ctx->d_acell_pbegin, pb));
        TRY(upload(ctx, &ctx->d_acell_vstate, vs));
        TRY(alloc_zero(ctx, &ctx->d_acell_state, (size_t)smax * (size_t)ctx->n_acells));
    }
    TRY(upload(ctx, (uint32_t **)&T.cell_seed, seeds));
    TRY(upload(ctx, (int32_t **)&T.comp_cell, comp_cell));
    TRY(upload(ctx, (double **)&T.t_row, ts));
    T.n_trace_cols = ctx->n_dev_cols;
    T.total_rows = ctx->total_rows;

    // trace chunk: all rows unless that exceeds 1 GiB or the caller asked for less
    int64_t chunk = ctx->opt.chunk_steps > 0 ? ctx->opt.chunk_steps : ctx->total_rows;
    const int64_t row_bytes = std::max<int64_t>(1, ctx->n_dev_cols) * 4;
    chunk = std::max<int64_t>(1, std::min<int64_t>(chunk, ((int64_t)1 << 30) / row_bytes));
    ctx->chunk_rows = chunk;
    T.trace_rows = chunk;
    TRY(alloc_zero(ctx, &T.trace, (size_t)(chunk * std::max<int64_t>(1, ctx->n_dev_cols))));
    T.max_events = 1u << 24;
    TRY(alloc_zero(ctx, &T.events, (size_t)T.max_events));
    TRY(alloc_zero(ctx, &T.n_events, 1));
    T.log_syn = ctx->opt.log_synapse_events > 0;
    T.max_syn_events = T.log_syn ? (uint32_t)std::min<int64_t>(ctx->opt.log_synapse_events, (int64_t)1 << 30) : 1u;
    TRY(alloc_zero(ctx, &T.syn_events, (size_t)T.max_syn_events));
    TRY(alloc_zero(ctx, &T.n_syn_events, 1));

    // snapshot every state table so edenhip_reset can rewind without re-uploading
    auto snap = [&](void *dst, size_t bytes) -> int {
        void *p = nullptr;
        HIPCHECK(ctx, hipMalloc(&p, std::max<size_t>(bytes, 4)));
        ctx->allocs.push_back(p);
        HIPCHECK(ctx, hipMemcpy(p, dst, std::max<size_t>(bytes, 4), hipMemcpyDeviceToDevice));
        ctx->snaps.push_back({dst, p, std::max<size_t>(bytes, 4)});
        return 0;
    };
    TRY(snap(T.V, NC * sizeof(float)));
    TRY(snap(T.q, NC * NQ * sizeof(float)));
    TRY(snap(T.syn_next, n_slots * sizeof(float)));
    TRY(snap(T.syn_A, n_slots * sizeof(float)));
    TRY(snap(T.syn_B, n_slots * sizeof(float)));
    TRY(snap(T.trig[0], n_slots * sizeof(int32_t)));
    TRY(snap(T.trig[1], n_slots * sizeof(int32_t)));
    TRY(snap(T.poi_tsince, n_poi * sizeof(float)));
    TRY(snap(T.poi_isi, n_poi * sizeof(float)));
    TRY(snap(T.poi_A, n_poi * sizeof(float)));
    TRY(snap(T.poi_B, n_poi * sizeof(float)));
    if (ctx->d_lems_state) {
        int32_t smax = 1;
        for (int t = 0; t < LS->n_types; t++) smax = std::max(smax, LS->n_state[t]);
        TRY(snap(ctx->d_lems_state, (size_t)(smax * n_slots) * sizeof(float)));
    }
    if (ctx->d_lems_tpos) TRY(snap(ctx->d_lems_tpos, (size_t)n_slots * sizeof(int32_t)));
    if (ctx->d_x_q) TRY(snap(ctx->d_x_q, (size_t)std::max<int64_t>(ctx->n_xq, 1) * sizeof(float)));
    if (ctx->d_pk_state) TRY(snap(ctx->d_pk_state, (size_t)ctx->n_pk * sizeof(float2)));
    if (ctx->d_acell_state) {
        int32_t smax = 1;
        for (int t = 0; t < LC->n_types; t++) smax = std::max(smax, LC->n_state[t]);
        TRY(snap(ctx->d_acell_state, (size_t)smax * (size_t)ctx->n_acells * sizeof(float)));
    }

    TRY(upload(ctx, &ctx->d_point, point));
    TRY(upload(ctx, &ctx->d_words, words));
    TRY(upload(ctx, &ctx->d_pos, posv));
    TRY(upload(ctx, &ctx->d_u, uv));
    TRY(upload(ctx, &ctx->d_tcm, tcm));
    TRY(upload(ctx, &ctx->d_ctask, ctask));
    TRY(upload(ctx, &ctx->d_gq, gqv));
    TRY(upload(ctx, &ctx->d_gt, gtv));
    TRY(upload(ctx, &ctx->d_cflag, cflag));
    TRY(upload(ctx, &ctx->d_cx, cxv));
    TRY(upload(ctx, &ctx->d_inl, inl));
    TRY(snap(ctx->d_gq, gqv.size() * sizeof(float)));

    // launch plan for tree cells: bucket by launch shape
    hipDeviceProp_t prop;
    HIPCHECK(ctx, hipGetDeviceProperties(&prop, ctx->device));
    ctx->n_cu = prop.multiProcessorCount;
    for (size_t i = 0; i < cells.size(); i++) {
        Bucket *b = nullptr;
        for (auto &bb : ctx->buckets)
            if (bb.tpb == cell_tpb[i]) b = &bb;
        if (!b) {
            ctx->buckets.emplace_back();
            b = &ctx->buckets.back();
            b->tpb = cell_tpb[i];
        }
        b->order.push_back((int32_t)i);
        b->comps += cells[i].n;
        b->nmax = std::max(b->nmax, cells[i].n);
        b->nvmax = std::max(b->nvmax, cells[i].const_tree ? 1 : cells[i].n);
        b->ngmax = std::max(b->ngmax, cells[i].n_gate);
        b->ninmax = std::max(b->ninmax, cells[i].n_input);
        b->nctmax = std::max(b->nctmax, cells[i].n_ctask);
        b->words_max = std::max(b->words_max, cells[i].words_len);
        b->cxmax = std::max(b->cxmax, cells[i].n_cx);
    }
    // launch order: workgroups with the most waves first, so that they find whole CUs
    std::stable_sort(ctx->buckets.begin(), ctx->buckets.end(), [](const Bucket &x, const Bucket &y) { return x.tpb > y.tpb; });
    std::vector<TreeCell> dcells;
    for (auto &b : ctx->buckets) {
        // work queue order: largest cells first
        std::stable_sort(b.order.begin(), b.order.end(), [&](int32_t x, int32_t y) { return cells[x].n > cells[y].n; });
        std::vector<TreeCell> bc;
        for (int32_t i : b.order) bc.push_back(cells[i]);
        std::vector<int32_t> seq(bc.size());
        for (size_t i = 0; i < seq.size(); i++) seq[i] = (int32_t)i;
        // sDX and sU with a dummy slot, Poisson input currents, sV, gate values, channel open
        // fractions and tasks,
        // input list, and the task words for workgroups of up to 128 threads
        b.nctmax = std::max(b.nctmax, 1);   // ungated slots read sF[first task], kept in bounds
        // a separate channel wave (gates overlapping the cable step) pays off for 8-wave
        // workgroups, and for one-wave workgroups when the bucket is latency-bound (at most
        // 4 cells per CU: every cell runs at once). With more one-wave cells the device runs
        // out of wave slots first and the extra wave per cell costs more than it hides. For
        // 2 and 4 cable waves the extra wave costs a workgroup per CU, which those sizes
        // need more.
        // Networks run one step per launch with the NET kernels, which have no channel wave.
        // channel_wave overrides the rule per workgroup size
        const int cw = ctx->opt.channel_wave[b.tpb == 64 ? 0 : b.tpb == 128 ? 1 : b.tpb == 256 ? 2 : 3];
        const bool rule = b.tpb == 512 || (b.tpb == 64 && (int64_t)bc.size() <= 4 * (int64_t)ctx->n_cu);
        b.split = (cw > 0 || (cw == 0 && rule)) && ctx->opt.persistent >= 0 && !ctx->tree_net;
        // complex-compartment records go to LDS when they take at most 20 KiB
        b.cxmax_lds = b.cxmax <= 256 ? b.cxmax : 0;
        b.lds = (size_t)tree_lds(b.nmax, b.nvmax, b.ngmax, b.nctmax, b.ninmax, b.words_max, b.cxmax_lds, ctx->n_gates,
                                 ctx->n_stypes, b.split).total *
                sizeof(float);
        const void *kp = tree_kernel_ptr(b.tpb, b.split, ctx->tree_net);
        if (!kp) return fail(ctx, EDENHIP_E_LIMIT, "no tree kernel for the chosen workgroup size");
        int per_cu = 0;
        HIPCHECK(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kp, b.tpb + (b.split ? 64 : 0), b.lds));
        if (per_cu <= 0) return fail(ctx, EDENHIP_E_LIMIT, "tree kernel does not fit on a CU (LDS)");
        b.grid = (int)std::min<int64_t>((int64_t)bc.size(), (int64_t)per_cu * ctx->n_cu);
        TRY(upload(ctx, &b.d_cells, bc));
        TRY(upload(ctx, &b.d_order, seq));
        TRY(alloc_zero(ctx, &b.d_queue, 2));
        // the first bucket on the main stream (the fork and join events below are then
        // recorded and waited on the same stream, which orders nothing more)
        TRY(pool_stream(ctx, (int)(&b - ctx->buckets.data()), &b.stream));
        HIPCHECK(ctx, hipEventCreateWithFlags(&b.done, hipEventDisableTiming));
        if (ctx->opt.profile) {
            HIPCHECK(ctx, hipEventCreate(&b.p0));
            HIPCHECK(ctx, hipEventCreate(&b.p1));
        }
    }
    if (ctx->opt.profile) {
        HIPCHECK(ctx, hipEventCreate(&ctx->pp0));
        HIPCHECK(ctx, hipEventCreate(&ctx->pp1));
        HIPCHECK(ctx, hipEventCreate(&ctx->nw0));
        HIPCHECK(ctx, hipEventCreate(&ctx->nw1));
    }
    HIPCHECK(ctx, hipEventCreateWithFlags(&ctx->ev_start, hipEventDisableTiming));
    TRY(build_netwin(ctx, d, point));
    HIPCHECK(ctx, hipDeviceSynchronize());
    return 0;
}

// Windowed network path: eligibility, per-compartment type runs, spike-target ring
// addresses, second state copies. Models it does not cover keep the per-step path.
static int build_netwin(edenhip_ctx *ctx, const edenhip_model_desc *d, const std::vector<int32_t> &point) {
    if (!ctx->network || !ctx->buckets.empty() || ctx->opt.network_window == -1 || ctx->lems_fn || ctx->n_xcomp > 0)
        return 0;
    const int64_t NC = d->n_comps;
    if (d->n_syn_types > 64 || (int64_t)point.size() != NC) return 0;
    const double dt = (double)d->dt;
    // window length: every synapse's delay spans at least W + 2 steps
    double dmin = INFINITY, dmax = 0.0;
    for (int64_t s = 0; s < d->comp_syn_begin[NC]; s++) {
        dmin = std::min(dmin, (double)d->syn_delay[s]);
        dmax = std::max(dmax, (double)d->syn_delay[s]);
    }
    if (!(dmin < INFINITY)) return 0;
    const int64_t Dmin = (int64_t)std::floor(dmin / dt);
    const int32_t W = (int32_t)std::min<int64_t>(64, Dmin - 2);
    if (W < 2) return 0;
    int64_t R = 1;
    while (R < (int64_t)std::ceil(dmax / dt) + W + 8) R <<= 1;
    // per compartment: up to four type runs (synapses are sorted by type), at most 256 synapses
    std::vector<uint32_t> runs(NC * 4, 0u);
    int64_t maxcnt = 0;
    for (int64_t g = 0; g < NC; g++) {
        const int64_t sb = d->comp_syn_begin[g], cnt = d->comp_syn_begin[g + 1] - sb;
        if (cnt > 256) return 0;
        maxcnt = std::max(maxcnt, cnt);
        int nr = 0;
        for (int64_t kq = 1; kq <= cnt; kq++)
            if (kq == cnt || d->syn_type[sb + kq] != d->syn_type[sb + kq - 1]) {
                if (nr >= 4) return 0;
                runs[g * 4 + nr++] = (uint32_t)kq | ((uint32_t)d->syn_type[sb + kq - 1] << 16);
            }
    }
    const int32_t nw = (int32_t)std::max<int64_t>(1, (maxcnt + 31) / 32);
    const int32_t n_blocks = (int32_t)((point.size() + 63) / 64);
    const int64_t stride = (int64_t)n_blocks * nw * 64;
    if (stride >= ((int64_t)1 << 27)) return 0;
    // ring address of every spike target: the target synapse's cell (block, lane) and index
    std::vector<int64_t> point_of(NC, -1);
    for (size_t p = 0; p < point.size(); p++) point_of[point[p]] = (int64_t)p;
    std::vector<int64_t> comp_of_syn(d->comp_syn_begin[NC]);
    for (int64_t g = 0; g < NC; g++)
        for (int64_t sq = d->comp_syn_begin[g]; sq < d->comp_syn_begin[g + 1]; sq++) comp_of_syn[sq] = g;
    const int64_t n_spk = d->comp_spike_begin[NC];
    std::vector<uint32_t> tgt_rw(std::max<int64_t>(n_spk, 1), 0u);
    for (int64_t kt = 0; kt < n_spk; kt++) {
        const int64_t sd = d->spike_target[kt], g = comp_of_syn[sd], p = point_of[g];
        const int64_t kk = sd - d->comp_syn_begin[g];
        const int64_t word = ((p / 64) * nw + kk / 32) * 64 + (p % 64);
        tgt_rw[kt] = ((uint32_t)word << 5) | (uint32_t)(kk % 32);
    }
    const int64_t n_slots = ctx->n_slots;
    NetTables &N = ctx->N;
    N.n_blocks = n_blocks;
    N.nw = nw;
    N.ring_mask = (int32_t)(R - 1);
    N.ring_stride = stride;
    TRY(upload(ctx, (uint32_t **)&N.runs, runs));
    TRY(upload(ctx, (uint32_t **)&N.tgt_rw, tgt_rw));
    N.gA[0] = ctx->T.syn_A;
    N.gB[0] = ctx->T.syn_B;
    TRY(alloc_zero(ctx, &N.gA[1], (size_t)n_slots));
    TRY(alloc_zero(ctx, &N.gB[1], (size_t)n_slots));
    TRY(alloc_zero(ctx, &N.ring, (size_t)(R * stride)));
    const int64_t n_poi = d->comp_poisson_begin[NC];
    N.V[0] = ctx->T.V; N.q[0] = ctx->T.q;
    N.poi_tsince[0] = ctx->T.poi_tsince; N.poi_isi[0] = ctx->T.poi_isi;
    N.poi_A[0] = ctx->T.poi_A; N.poi_B[0] = ctx->T.poi_B;
    TRY(alloc_zero(ctx, &N.V[1], (size_t)NC));
    TRY(alloc_zero(ctx, &N.q[1], (size_t)(NC * NQ)));
    TRY(alloc_zero(ctx, &N.poi_tsince[1], (size_t)n_poi));
    TRY(alloc_zero(ctx, &N.poi_isi[1], (size_t)n_poi));
    TRY(alloc_zero(ctx, &N.poi_A[1], (size_t)n_poi));
    TRY(alloc_zero(ctx, &N.poi_B[1], (size_t)n_poi));
    TRY(alloc_zero(ctx, &N.spike_mask, (size_t)NC));
    TRY(alloc_zero(ctx, &N.spk_list, (size_t)NC));
    TRY(alloc_zero(ctx, &N.spk_count, 1));
    TRY(alloc_zero(ctx, &N.conflict, 1));
    TRY(alloc_zero(ctx, &N.n_deliv, 1));
    TRY(alloc_zero(ctx, &ctx->d_deliv_save, 1));
    TRY(alloc_zero(ctx, &ctx->d_ev_save, 2));
    HIPCHECK(ctx, hipHostMalloc((void **)&ctx->h_conflict, sizeof(int32_t), hipHostMallocDefault));
    *ctx->h_conflict = 0;
    // reset rewinds the ring too (NextSpike and copy 0 of the state are already snapshotted)
    {
        void *p = nullptr;
        const size_t bytes = (size_t)(R * stride) * sizeof(uint32_t);
        HIPCHECK(ctx, hipMalloc(&p, bytes));
        ctx->allocs.push_back(p);
        HIPCHECK(ctx, hipMemset(p, 0, bytes));
        ctx->snaps.push_back({N.ring, p, bytes});
    }
    ctx->net_W = W;
    ctx->net_cur = 0;
    ctx->netwin = true;
    return 0;
}

// Copy the state copy `from` over copy `to` (compartments, gates, Poisson inputs, synapses).
static int net_copy_state(edenhip_ctx *ctx, int from, int to) {
    NetTables &N = ctx->N;
    const int64_t NC = ctx->NC, n_poi = ctx->n_poi;
    const int64_t n_slots = ctx->n_slots;
    const hipStream_t st = ctx->stream;
    HIPCHECK(ctx, hipMemcpyAsync(N.V[to], N.V[from], NC * sizeof(float), hipMemcpyDeviceToDevice, st));
    HIPCHECK(ctx, hipMemcpyAsync(N.q[to], N.q[from], NC * NQ * sizeof(float), hipMemcpyDeviceToDevice, st));
    if (n_poi > 0) {
        HIPCHECK(ctx, hipMemcpyAsync(N.poi_tsince[to], N.poi_tsince[from], n_poi * sizeof(float), hipMemcpyDeviceToDevice, st));
        HIPCHECK(ctx, hipMemcpyAsync(N.poi_isi[to], N.poi_isi[from], n_poi * sizeof(float), hipMemcpyDeviceToDevice, st));
        HIPCHECK(ctx, hipMemcpyAsync(N.poi_A[to], N.poi_A[from], n_poi * sizeof(float), hipMemcpyDeviceToDevice, st));
        HIPCHECK(ctx, hipMemcpyAsync(N.poi_B[to], N.poi_B[from], n_poi * sizeof(float), hipMemcpyDeviceToDevice, st));
    }
    HIPCHECK(ctx, hipMemcpyAsync(N.gA[to], N.gA[from], n_slots * sizeof(float), hipMemcpyDeviceToDevice, st));
    HIPCHECK(ctx, hipMemcpyAsync(N.gB[to], N.gB[from], n_slots * sizeof(float), hipMemcpyDeviceToDevice, st));
    return 0;
}

// One window [a, b) of the network path from state copy ctx->net_cur; with check, the
// deliver pass first only checks that the window matches step-by-step order. Returns 1 in
// *conflict when it does not (nothing committed then).
static int net_window(edenhip_ctx *ctx, int64_t a, int64_t b, int64_t row0, bool check, bool force, int *conflict,
                      int64_t *nl) {
    NetTables &N = ctx->N;
    DevTables T = ctx->T;
    const hipStream_t st = ctx->stream;
    float dt = ctx->dt;
    int cur = ctx->net_cur;
    int64_t last = ctx->total_rows;
    HIPCHECK(ctx, hipMemsetAsync(N.spk_count, 0, sizeof(uint32_t), st));
    int32_t np = ctx->n_point;
    void *wargs[] = {&T, &N, &ctx->d_gates, &ctx->n_gates, &ctx->d_stypes, &ctx->n_stypes, &ctx->d_point, &np,
                     &a, &b, &row0, &dt, &cur};
    const bool prof = ctx->opt.profile != 0;
    if (prof) HIPCHECK(ctx, hipEventRecord(ctx->nw0, st));
    HIPCHECK(ctx, hipLaunchKernel(net_window_kernel_ptr(), dim3(N.n_blocks), dim3(64), wargs,
                                  (size_t)std::max(1, ctx->n_gates) * sizeof(DGate), st));
    if (prof) {
        // the window kernel's own time (the delivery kernels are not included)
        HIPCHECK(ctx, hipEventRecord(ctx->nw1, st));
        HIPCHECK(ctx, hipEventSynchronize(ctx->nw1));
        float ms = 0;
        HIPCHECK(ctx, hipEventElapsedTime(&ms, ctx->nw0, ctx->nw1));
        ctx->point_ms += ms;
        ctx->point_launches++;
    }
    (*nl)++;
    const int grid = 1024;
    int commit = 0;
    *conflict = 0;
    if (check) {
        HIPCHECK(ctx, hipMemsetAsync(N.conflict, 0, sizeof(int32_t), st));
        void *dargs[] = {&T, &N, &a, &b, &last, &dt, &commit};
        HIPCHECK(ctx, hipLaunchKernel(net_deliver_kernel_ptr(), dim3(grid), dim3(64), dargs, 0, st));
        HIPCHECK(ctx, hipMemcpyAsync(ctx->h_conflict, N.conflict, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        HIPCHECK(ctx, hipStreamSynchronize(st));
        if (*ctx->h_conflict || force) {
            *conflict = 1;
            return 0;
        }
    }
    commit = 1;
    void *dargs[] = {&T, &N, &a, &b, &last, &dt, &commit};
    HIPCHECK(ctx, hipLaunchKernel(net_deliver_kernel_ptr(), dim3(grid), dim3(64), dargs, 0, st));
    ctx->net_cur ^= 1;
    return 0;
}

// Steps [s0, s1) of a network through windows of up to net_W steps, each checked against
// step-by-step order and, where it differs, run again one step at a time.
static int net_chunk(edenhip_ctx *ctx, int64_t s0, int64_t s1, int64_t row0, int64_t *nl) {
    const hipStream_t st = ctx->stream;
    for (int64_t a = s0; a < s1;) {
        const int64_t b = std::min<int64_t>(s1, a + ctx->net_W);
        // event-log counters, restored if the window is run again
        HIPCHECK(ctx, hipMemcpyAsync(ctx->d_ev_save, ctx->T.n_events, sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
        HIPCHECK(ctx, hipMemcpyAsync(ctx->d_ev_save + 1, ctx->T.n_syn_events, sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
        HIPCHECK(ctx, hipMemcpyAsync(ctx->d_deliv_save, ctx->N.n_deliv, sizeof(unsigned long long), hipMemcpyDeviceToDevice, st));
        int conflict = 0;
        TRY(net_window(ctx, a, b, row0, true, ctx->opt.network_window == -2, &conflict, nl));
        ctx->net_windows++;
        if (conflict) {
            HIPCHECK(ctx, hipMemcpyAsync(ctx->T.n_events, ctx->d_ev_save, sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
            HIPCHECK(ctx, hipMemcpyAsync(ctx->T.n_syn_events, ctx->d_ev_save + 1, sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
            HIPCHECK(ctx, hipMemcpyAsync(ctx->N.n_deliv, ctx->d_deliv_save, sizeof(unsigned long long), hipMemcpyDeviceToDevice, st));
            for (int64_t s = a; s < b; s++) {
                int c2 = 0;
                TRY(net_window(ctx, s, s + 1, row0, false, false, &c2, nl));
            }
            ctx->net_redone++;
        }
        a = b;
    }
    // leave the state in copy 0, where the rest of the library reads it
    if (ctx->net_cur != 0) {
        TRY(net_copy_state(ctx, 1, 0));
        ctx->net_cur = 0;
    }
    return 0;
}

static int launch_chunk(edenhip_ctx *ctx, int64_t s0, int64_t s1, int64_t row0) {
    DevTables T = ctx->T;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ctx->opt.profile) {
        HIPCHECK(ctx, hipEventCreate(&e0));
        HIPCHECK(ctx, hipEventCreate(&e1));
        HIPCHECK(ctx, hipEventRecord(e0, ctx->stream));
    }
    const bool per_step = ctx->network || ctx->opt.persistent < 0;
    const int64_t stride = per_step ? 1 : (s1 - s0);
    int64_t nl = 0;
    if (ctx->netwin) {
        if (ctx->opt.profile) HIPCHECK(ctx, hipEventRecord(ctx->pp0, ctx->stream));
        TRY(net_chunk(ctx, s0, s1, row0, &nl));
        if (ctx->opt.profile) {
            HIPCHECK(ctx, hipEventRecord(ctx->pp1, ctx->stream));
            HIPCHECK(ctx, hipEventRecord(e1, ctx->stream));
            HIPCHECK(ctx, hipEventSynchronize(e1));
            float ms = 0;
            HIPCHECK(ctx, hipEventElapsedTime(&ms, e0, e1));
            ctx->kernel_ms += ms;
            ctx->launches += nl;
            (void)nl;   // window launches and their time: counted in net_window
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        }
        return 0;
    }
    const bool graph = ctx->opt.use_graph != 0 && per_step;
    const bool prof = ctx->opt.profile != 0 && !graph;   // per-kernel events stay out of graphs
    // work-queue counters of edenhip_run_static: bucket b uses queue_base + 2 b
    std::vector<int32_t *> qs;
    if (ctx->queue_base)
        for (size_t b = 0; b < ctx->buckets.size(); b++) qs.push_back(ctx->queue_base + 2 * b);
    // the launches of steps [a0, a1), one range per launch
    auto issue = [&](int64_t a0, int64_t a1, bool pf) -> int {
        for (int64_t a = a0; a < a1; a += stride) {
            const int64_t b = std::min(a1, a + stride);
            const bool first = a == s0, last = b == s1;
            if (ctx->n_grec > 0 && a >= 1) {
                // gate variables after step a - 1 into row a - 1, before step a changes them
                // (one step per launch whenever gates are recorded)
                const int64_t row = a - 1 - row0;
                hipLaunchKernelGGL(record_gates_kernel, dim3((ctx->n_grec + 255) / 256), dim3(256), 0, ctx->stream, T,
                                   (const float *)ctx->d_gq, (const float *)ctx->d_acell_state,
                                   (const float *)ctx->d_x_q, (const int32_t *)ctx->d_grec_col, (const int64_t *)ctx->d_grec_src, ctx->n_grec,
                                   row);
                HIPCHECK(ctx, hipGetLastError());
                nl++;
            }
            if (ctx->lems_fn && ctx->n_lems_comps > 0) {
                // generic synapses of step a: per-compartment currents for the kernels below
                LemsArgs la{ctx->d_lems_comps, ctx->n_lems_comps, ctx->d_lems_params, ctx->d_lems_pbegin, ctx->d_lems_state,
                     ctx->n_slots, ctx->d_syn_ext, ctx->d_lems_peer, ctx->d_lems_flags, ctx->d_lems_poff,
                     ctx->d_lems_tpos, ctx->d_lems_ttimes, ctx->d_pk_slot, ctx->d_pk_state, ctx->d_pk_isi,
                     ctx->d_pk_window, ctx->d_pk_transient, ctx->d_pk_inst, ctx->d_pk_roff, ctx->d_inp_ext};
                int64_t st = a;
                float dt = ctx->dt;
                void *args[] = {&T, &la, &st, &dt};
                HIPCHECK(ctx, hipModuleLaunchKernel(ctx->lems_fn, (unsigned)((ctx->n_lems_comps + 63) / 64), 1, 1, 64, 1, 1,
                                                    0, ctx->stream, args, nullptr));
                nl++;
            }
            if (ctx->lems_cell_fn && ctx->n_acells > 0) {
                // point neurons of step a, after the synapse currents of the step
                LemsCellArgs ca{ctx->d_acell_comps, ctx->d_acell_type, ctx->n_acells, ctx->d_acell_params,
                                ctx->d_acell_pbegin, ctx->d_acell_vstate, ctx->d_acell_state, ctx->d_acell_poff};
                int64_t st = a;
                float dt = ctx->dt;
                void *args[] = {&T, &ca, &st, &dt};
                HIPCHECK(ctx, hipModuleLaunchKernel(ctx->lems_cell_fn, (unsigned)((ctx->n_acells + 63) / 64), 1, 1, 64, 1,
                                                    1, 0, ctx->stream, args, nullptr));
                nl++;
            }
            if (ctx->n_xcomp > 0) {
                // channels of the compartments beyond the kernels' slots, from this step's V
                ExtChanArgs xa{ctx->d_x_comp, ctx->d_x_chb, ctx->d_x_type, ctx->d_x_g, ctx->d_x_e, ctx->d_x_qb,
                               ctx->d_x_q, ctx->d_chx_ext, ctx->n_xcomp, ctx->d_chan_scaling, ctx->d_gate_extra};
                int64_t st = a;
                float dt = ctx->dt;
                void *args[] = {&T, &ctx->d_gates, &xa, &st, &dt};
                if (ctx->ext_rtc_fn)
                    HIPCHECK(ctx, hipModuleLaunchKernel(ctx->ext_rtc_fn, (unsigned)((ctx->n_xcomp + 63) / 64), 1, 1, 64, 1, 1,
                                                        0, ctx->stream, args, nullptr));
                else
                    HIPCHECK(ctx, hipLaunchKernel(ext_channels_kernel_ptr(), dim3((ctx->n_xcomp + 63) / 64), dim3(64), args,
                                                  0, ctx->stream));
                nl++;
            }
            if (ctx->n_point > 0) {
                const int tpb = 64;
                float dt = ctx->dt;
                int32_t n = ctx->n_point;
                int64_t aa = a, bb = b;
                if (pf && first) HIPCHECK(ctx, hipEventRecord(ctx->pp0, ctx->stream));
                if (ctx->point_quad) {
                    const int grid = (int)(((int64_t)ctx->n_point * 4 + tpb - 1) / tpb);
                    void *args[] = {&T, &ctx->d_gates, &ctx->n_gates, &ctx->d_point, &n, (void *)&aa, (void *)&bb, &row0, &dt};
                    HIPCHECK(ctx, hipLaunchKernel(point_quad_kernel_ptr(), dim3(grid), dim3(tpb), args, 0, ctx->stream));
                } else {
                    const int grid = (ctx->n_point + tpb - 1) / tpb;
                    void *args[] = {&T, &ctx->d_gates, &ctx->n_gates, &ctx->d_stypes, &ctx->n_stypes, &ctx->d_point, &n,
                                    (void *)&aa, (void *)&bb, &row0, &dt};
                    HIPCHECK(ctx, hipLaunchKernel(point_kernel_ptr(), dim3(grid), dim3(tpb), args, 0, ctx->stream));
                }
                if (pf && last) HIPCHECK(ctx, hipEventRecord(ctx->pp1, ctx->stream));
                nl++;
            }
            // tree buckets run concurrently, each on its own stream, joined back to the main stream
            if (!ctx->buckets.empty()) HIPCHECK(ctx, hipEventRecord(ctx->ev_start, ctx->stream));
            for (auto &bk : ctx->buckets) {
                float dt = ctx->dt;
                int32_t ntask = (int32_t)bk.order.size();
                int64_t aa = a, bb = b;
                void *args[] = {&T, &ctx->d_gates, &ctx->n_gates, &ctx->d_stypes, &ctx->n_stypes, &bk.d_cells,
                                &bk.d_order, &ntask, &bk.d_queue, &ctx->d_words, &ctx->d_pos, &ctx->d_u, &ctx->d_tcm,
                                &ctx->d_ctask, &ctx->d_gq, &ctx->d_gt, &ctx->d_inl, &ctx->d_cflag,
                                &ctx->d_cx, &bk.cxmax_lds, &bk.nmax, &bk.nvmax, &bk.ngmax, &bk.nctmax, &bk.ninmax,
                                &bk.words_max, (void *)&aa, (void *)&bb, &row0, &dt};
                if (ctx->queue_base) args[8] = &qs[&bk - ctx->buckets.data()];
                HIPCHECK(ctx, hipStreamWaitEvent(bk.stream, ctx->ev_start, 0));
                if (pf && first) HIPCHECK(ctx, hipEventRecord(bk.p0, bk.stream));
                HIPCHECK(ctx, hipLaunchKernel(tree_kernel_ptr(bk.tpb, bk.split, ctx->tree_net), dim3(bk.grid),
                                              dim3(bk.tpb + (bk.split ? 64 : 0)), args, bk.lds, bk.stream));
                if (pf && last) HIPCHECK(ctx, hipEventRecord(bk.p1, bk.stream));
                HIPCHECK(ctx, hipEventRecord(bk.done, bk.stream));
                HIPCHECK(ctx, hipStreamWaitEvent(ctx->stream, bk.done, 0));
                nl++;
                if (pf) bk.launches++;
            }
        }
        return 0;
    };
    if (!graph) {
        TRY(issue(s0, s1, prof));
    } else {
        // per-step launches captured in graphs of up to 256 steps (kernels only: the kernels
        // reset their own work queues, nothing is cleared with a memset node), kept for the
        // next run over the same steps (after edenhip_reset)
        constexpr int64_t G = 256;
        for (int64_t a = s0; a < s1; a += G) {
            const int64_t b = std::min(s1, a + G);
            hipGraphExec_t exec = nullptr;
            for (auto &g : ctx->graphs)
                if (g.a == a && g.b == b && g.row0 == row0) exec = g.exec;
            if (!exec) {
                hipGraph_t gr = nullptr;
                HIPCHECK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
                const int rc = issue(a, b, false);
                const hipError_t ec = hipStreamEndCapture(ctx->stream, &gr);
                if (rc != 0) {
                    if (gr) (void)hipGraphDestroy(gr);
                    return rc;
                }
                if (ec != hipSuccess) return fail(ctx, EDENHIP_E_DEVICE, hipGetErrorString(ec));
                const hipError_t ei = hipGraphInstantiate(&exec, gr, nullptr, nullptr, 0);
                (void)hipGraphDestroy(gr);
                if (ei != hipSuccess) return fail(ctx, EDENHIP_E_DEVICE, hipGetErrorString(ei));
                ctx->graphs.push_back({a, b, row0, exec});
            } else {
                nl += (ctx->n_point > 0 ? 1 : 0) * (b - a) + (int64_t)ctx->buckets.size() * (b - a);
            }
            HIPCHECK(ctx, hipGraphLaunch(exec, ctx->stream));
        }
    }
    if (graph && ctx->opt.profile) {
        // whole-chunk time only: the graphs hold no per-kernel events
        HIPCHECK(ctx, hipEventRecord(e1, ctx->stream));
        HIPCHECK(ctx, hipEventSynchronize(e1));
        float ms = 0;
        HIPCHECK(ctx, hipEventElapsedTime(&ms, e0, e1));
        ctx->kernel_ms += ms;
        ctx->launches += nl;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    if (prof) {
        HIPCHECK(ctx, hipEventRecord(e1, ctx->stream));
        HIPCHECK(ctx, hipEventSynchronize(e1));
        float ms = 0;
        HIPCHECK(ctx, hipEventElapsedTime(&ms, e0, e1));
        ctx->kernel_ms += ms;
        ctx->launches += nl;
        if (ctx->n_point > 0) {
            HIPCHECK(ctx, hipEventElapsedTime(&ms, ctx->pp0, ctx->pp1));
            ctx->point_ms += ms;
            ctx->point_launches += per_step ? (s1 - s0) : 1;
        }
        for (auto &bk : ctx->buckets) {
            HIPCHECK(ctx, hipEventElapsedTime(&ms, bk.p0, bk.p1));
            bk.ms += ms;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    return 0;
}

extern "C" {

int edenhip_create(const edenhip_model_desc *desc, const edenhip_options *opt, edenhip_ctx **out) {
    if (!out) return EDENHIP_E_INVALID;
    *out = nullptr;
    edenhip_ctx *ctx = new (std::nothrow) edenhip_ctx();
    if (!ctx) return EDENHIP_E_INVALID;
    // options of version 1 and 2 callers end before channel_wave (versions 3 and 4 share them)
    if (opt) std::memcpy(&ctx->opt, opt, desc && desc->abi_version >= 3 ? sizeof(edenhip_options)
                                                                         : offsetof(edenhip_options, channel_wave));
    if (ctx->opt.persistent == 0) ctx->opt.persistent = 1;
    ctx->device = ctx->opt.device;
    *out = ctx;
    if (!desc) return fail(ctx, EDENHIP_E_INVALID, "null descriptor");
    for (int k = 0; k < 4; k++) {
        const int32_t L = ctx->opt.size_limits[k];
        if (L < 0 || L > 4 * (64 << k) || (k > 0 && L != 0 && L < ctx->opt.size_limits[k - 1]))
            return fail(ctx, EDENHIP_E_INVALID, "size_limits: each in [0, 4 x threads], nondecreasing");
    }
    for (int k = 0; k < 4; k++)
        if (ctx->opt.channel_wave[k] < -1 || ctx->opt.channel_wave[k] > 1)
            return fail(ctx, EDENHIP_E_INVALID, "channel_wave: each -1, 0 or 1");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(ctx, EDENHIP_E_DEVICE, "no HIP device visible");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    TRY(pool_stream(ctx, 0, &ctx->stream));
    // the device's static scratch address, looked up once per device (edenhip_run_static)
    int32_t *sp = nullptr;
    TRY(static_scratch_ptr(ctx, &sp));
    return build(ctx, desc);
}

int64_t edenhip_total_rows(const edenhip_ctx *ctx) { return ctx ? ctx->total_rows : -1; }

int edenhip_run(edenhip_ctx *ctx, int64_t n_rows, int64_t *rows_done) {
    if (!ctx) return EDENHIP_E_INVALID;
    if (!ctx->stream) return fail(ctx, EDENHIP_E_STATE, "context was not created successfully");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    int64_t want = std::min(n_rows < 0 ? ctx->total_rows : n_rows, ctx->total_rows - ctx->rows_done);
    want = std::min(want, ctx->chunk_rows);
    if (want < 0) want = 0;
    const int64_t row0 = ctx->rows_done;
    int64_t s0 = ctx->next_step;           // -3 on the first call: the init steps go first
    const int64_t s1 = row0 + want + 1;    // step s records row s - 1
    if (want > 0 || s0 <= 0) {
        if (want == 0) return fail(ctx, EDENHIP_E_STATE, "no rows left to run");
        TRY(launch_chunk(ctx, s0, s1, row0));
        ctx->next_step = s1;
    }
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->rows_done += want;
    ctx->last_row0 = row0;
    ctx->last_rows = want;
    if (rows_done) *rows_done = want;
    return 0;
}

int edenhip_run_static(edenhip_ctx *ctx, int64_t n_rows, int64_t *rows_done) {
    if (!ctx) return EDENHIP_E_INVALID;
    if (!ctx->stream) return fail(ctx, EDENHIP_E_STATE, "context was not created successfully");
    if (ctx->netwin)
        return fail(ctx, EDENHIP_E_LIMIT,
                    "edenhip_run_static: the windowed network path keeps scratch that grows with the network; "
                    "use edenhip_run, or create the context with network_window = -1");
    if (ctx->opt.use_graph || ctx->opt.profile)
        return fail(ctx, EDENHIP_E_STATE,
                    "edenhip_run_static: graph capture and profiling events allocate; create the context without "
                    "use_graph and profile");
    if ((int64_t)ctx->buckets.size() > STATIC_QUEUE_BUCKETS)
        return fail(ctx, EDENHIP_E_LIMIT, "edenhip_run_static: more tree-cell buckets than the static scratch holds (32)");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    int32_t *sp = nullptr;
    TRY(static_scratch_ptr(ctx, &sp));
    // the array keeps the last call's contents (a call that failed midway may leave counters
    // behind): the counters start from zero
    if (!ctx->buckets.empty()) {
        const int n = 2 * (int)ctx->buckets.size();
        hipLaunchKernelGGL(zero_i32_kernel, dim3(1), dim3(64), 0, ctx->stream, sp, n);
        HIPCHECK(ctx, hipGetLastError());
    }
    ctx->queue_base = sp;
    const int rc = edenhip_run(ctx, n_rows, rows_done);
    ctx->queue_base = nullptr;
    return rc;
}

int edenhip_reset(edenhip_ctx *ctx) {
    if (!ctx) return EDENHIP_E_INVALID;
    if (!ctx->stream) return fail(ctx, EDENHIP_E_STATE, "context was not created successfully");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    for (auto &s : ctx->snaps) HIPCHECK(ctx, hipMemcpyAsync(s.dst, s.src, s.bytes, hipMemcpyDeviceToDevice, ctx->stream));
    HIPCHECK(ctx, hipMemsetAsync(ctx->T.n_events, 0, sizeof(uint32_t), ctx->stream));
    HIPCHECK(ctx, hipMemsetAsync(ctx->T.n_syn_events, 0, sizeof(uint32_t), ctx->stream));
    if (ctx->netwin) HIPCHECK(ctx, hipMemsetAsync(ctx->N.n_deliv, 0, sizeof(unsigned long long), ctx->stream));
    HIPCHECK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->rows_done = 0;
    ctx->next_step = -3;
    ctx->last_rows = 0;
    ctx->last_row0 = 0;
    ctx->kernel_ms = 0;
    ctx->launches = 0;
    ctx->point_ms = 0;
    ctx->point_launches = 0;
    ctx->net_cur = 0;
    ctx->net_windows = 0;
    ctx->net_redone = 0;
    for (auto &bk : ctx->buckets) {
        bk.ms = 0;
        bk.launches = 0;
    }
    return 0;
}

// State blob: a header, then every live state table in the order edenhip_reset restores them.
struct StateHeader {
    uint64_t magic;          // "EDENHIPS"
    int64_t n_tables, total_bytes, next_step, rows_done;
};
constexpr uint64_t STATE_MAGIC = 0x5350494845444545ull;

static int64_t state_bytes(const edenhip_ctx *ctx) {
    int64_t b = (int64_t)sizeof(StateHeader) + (int64_t)ctx->snaps.size() * 8;
    for (const auto &sn : ctx->snaps) b += (int64_t)sn.bytes;
    return b;
}

int edenhip_state_size(const edenhip_ctx *ctx, int64_t *bytes) {
    if (!ctx || !bytes) return EDENHIP_E_INVALID;
    *bytes = state_bytes(ctx);
    return 0;
}

int edenhip_get_state(edenhip_ctx *ctx, void *dst, int64_t bytes) {
    if (!ctx || !dst) return EDENHIP_E_INVALID;
    if (!ctx->stream) return fail(ctx, EDENHIP_E_STATE, "context was not created successfully");
    if (bytes < state_bytes(ctx)) return fail(ctx, EDENHIP_E_INVALID, "state buffer smaller than edenhip_state_size");
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    char *p = static_cast<char *>(dst);
    StateHeader h{STATE_MAGIC, (int64_t)ctx->snaps.size(), state_bytes(ctx), ctx->next_step, ctx->rows_done};
    std::memcpy(p, &h, sizeof(h));
    p += sizeof(h);
    for (const auto &sn : ctx->snaps) {
        const int64_t nb = (int64_t)sn.bytes;
        std::memcpy(p, &nb, 8);
        p += 8;
        HIPCHECK(ctx, hipMemcpy(p, sn.dst, sn.bytes, hipMemcpyDeviceToHost));
        p += sn.bytes;
    }
    return 0;
}

int edenhip_set_state(edenhip_ctx *ctx, const void *src, int64_t bytes) {
    if (!ctx || !src) return EDENHIP_E_INVALID;
    if (!ctx->stream) return fail(ctx, EDENHIP_E_STATE, "context was not created successfully");
    if (bytes < (int64_t)sizeof(StateHeader)) return fail(ctx, EDENHIP_E_INVALID, "state buffer too small");
    StateHeader h;
    std::memcpy(&h, src, sizeof(h));
    if (h.magic != STATE_MAGIC || h.n_tables != (int64_t)ctx->snaps.size() || h.total_bytes != state_bytes(ctx) ||
        bytes < h.total_bytes)
        return fail(ctx, EDENHIP_E_INVALID, "state blob does not belong to a context of this model");
    if (h.next_step < -3 || h.rows_done < 0 || h.rows_done > ctx->total_rows)
        return fail(ctx, EDENHIP_E_INVALID, "state blob with invalid step counters");
    const char *p = static_cast<const char *>(src) + sizeof(h);
    for (const auto &sn : ctx->snaps) {
        int64_t nb = 0;
        std::memcpy(&nb, p, 8);
        if (nb != (int64_t)sn.bytes) return fail(ctx, EDENHIP_E_INVALID, "state blob table size mismatch");
        p += 8 + sn.bytes;
    }
    HIPCHECK(ctx, hipSetDevice(ctx->device));
    p = static_cast<const char *>(src) + sizeof(h);
    for (const auto &sn : ctx->snaps) {
        p += 8;
        HIPCHECK(ctx, hipMemcpy(sn.dst, p, sn.bytes, hipMemcpyHostToDevice));
        p += sn.bytes;
    }
    ctx->next_step = h.next_step;
    ctx->rows_done = h.rows_done;
    ctx->last_rows = 0;
    ctx->last_row0 = h.rows_done;
    ctx->net_cur = 0;
    return 0;
}

int edenhip_read_trace(edenhip_ctx *ctx, float *dst, int64_t max_rows, int64_t *rows, int64_t *row0) {
    if (!ctx) return EDENHIP_E_INVALID;
    const int64_t nr = std::min(max_rows, ctx->last_rows);
    const int64_t ncol = (int64_t)ctx->desc_to_dev_col.size();
    if (rows) *rows = nr;
    if (row0) *row0 = ctx->last_row0;
    if (nr <= 0 || ncol == 0 || !dst) return 0;
    // device layout [column][chunk row]; the first nr rows of each column
    const int64_t stride = ctx->chunk_rows;
    std::vector<float> buf((size_t)(stride * ctx->n_dev_cols));
    HIPCHECK(ctx, hipMemcpy(buf.data(), ctx->T.trace, buf.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (int64_t c = 0; c < ncol; c++) {
        const float *src = buf.data() + (size_t)(ctx->desc_to_dev_col[c] * stride);
        for (int64_t r = 0; r < nr; r++) dst[r * ncol + c] = src[r];
    }
    return 0;
}

int edenhip_trace_device_ptr(edenhip_ctx *ctx, const float **dptr) {
    if (!ctx || !dptr) return EDENHIP_E_INVALID;
    *dptr = ctx->T.trace;
    return 0;
}

int edenhip_read_events(edenhip_ctx *ctx, int64_t *col, int64_t *row, int64_t max_events, int64_t *n) {
    if (!ctx) return EDENHIP_E_INVALID;
    uint32_t cnt = 0;
    HIPCHECK(ctx, hipMemcpy(&cnt, ctx->T.n_events, sizeof(cnt), hipMemcpyDeviceToHost));
    if (cnt > ctx->T.max_events) return fail(ctx, EDENHIP_E_LIMIT, "event buffer overflow");
    std::vector<int64_t> ev(cnt);
    if (cnt) HIPCHECK(ctx, hipMemcpy(ev.data(), ctx->T.events, cnt * sizeof(int64_t), hipMemcpyDeviceToHost));
    std::sort(ev.begin(), ev.end());
    int64_t k = 0;
    for (int64_t x : ev) {
        const int64_t r = x >> 24, c = x & 0xFFFFFF;
        for (int64_t dc : ctx->ev_expand[c]) {
            if (k < max_events && col && row) { col[k] = dc; row[k] = r; }
            k++;
        }
    }
    // columns that share a compartment interleave: restore ascending column order per row
    if (col && row) {
        const int64_t m = std::min(k, max_events);
        std::vector<std::pair<int64_t, int64_t>> pr((size_t)m);
        for (int64_t i = 0; i < m; i++) pr[i] = {row[i], col[i]};
        std::sort(pr.begin(), pr.end());
        for (int64_t i = 0; i < m; i++) { row[i] = pr[i].first; col[i] = pr[i].second; }
    }
    if (n) *n = k;
    return 0;
}

int edenhip_read_synapse_events(edenhip_ctx *ctx, int64_t *syn, int64_t *step, int64_t max_events, int64_t *n) {
    if (!ctx) return EDENHIP_E_INVALID;
    uint32_t cnt = 0;
    HIPCHECK(ctx, hipMemcpy(&cnt, ctx->T.n_syn_events, sizeof(cnt), hipMemcpyDeviceToHost));
    if (ctx->T.log_syn && cnt > ctx->T.max_syn_events) return fail(ctx, EDENHIP_E_LIMIT, "synaptic event log overflow");
    std::vector<int64_t> ev(cnt);
    if (cnt) HIPCHECK(ctx, hipMemcpy(ev.data(), ctx->T.syn_events, cnt * sizeof(int64_t), hipMemcpyDeviceToHost));
    std::sort(ev.begin(), ev.end());
    for (int64_t k = 0; k < (int64_t)cnt && k < max_events; k++) {
        if (syn) syn[k] = ev[k] & 0xFFFFFFFF;
        if (step) step[k] = ev[k] >> 32;
    }
    if (n) *n = cnt;
    return 0;
}

int edenhip_get_voltages(edenhip_ctx *ctx, float *dst) {
    if (!ctx || !dst) return EDENHIP_E_INVALID;
    HIPCHECK(ctx, hipMemcpy(dst, ctx->T.V, ctx->NC * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int edenhip_kernel_stats(const edenhip_ctx *ctx, double *total_ms, int64_t *launches, char *name, int32_t name_len) {
    if (!ctx) return EDENHIP_E_INVALID;
    if (total_ms) *total_ms = ctx->kernel_ms;
    if (launches) *launches = ctx->launches;
    if (name && name_len > 0) {
        std::string nm = !ctx->buckets.empty() ? "tree_kernel" : ctx->point_quad ? "point_quad_kernel" : "point_kernel";
        std::strncpy(name, nm.c_str(), (size_t)name_len - 1);
        name[name_len - 1] = 0;
    }
    return 0;
}

int32_t edenhip_kernel_count(const edenhip_ctx *ctx) {
    if (!ctx) return EDENHIP_E_INVALID;
    return (ctx->n_point > 0 ? 1 : 0) + (int32_t)ctx->buckets.size();
}

int edenhip_get_kernel_info(const edenhip_ctx *ctx, int32_t i, edenhip_kernel_info *out) {
    if (!ctx || !out || i < 0 || i >= edenhip_kernel_count(ctx)) return EDENHIP_E_INVALID;
    std::memset(out, 0, sizeof(*out));
    if (ctx->n_point > 0 && i == 0) {
        std::snprintf(out->name, sizeof(out->name), ctx->netwin ? "net_window_kernel" : ctx->point_quad ? "point_quad_kernel" : "point_kernel");
        out->threads_per_block = 64;
        out->grid = ctx->point_quad ? (int32_t)(((int64_t)ctx->n_point * 4 + 63) / 64) : (ctx->n_point + 63) / 64;
        out->n_cells = ctx->n_point;
        out->n_comps = ctx->n_point;
        out->launches = ctx->point_launches;
        out->total_ms = ctx->point_ms;
        return 0;
    }
    const Bucket &b = ctx->buckets[(size_t)(i - (ctx->n_point > 0 ? 1 : 0))];
    std::snprintf(out->name, sizeof(out->name), "tree_kernel<%d, %s, %s>", b.tpb, b.split ? "true" : "false",
                  ctx->tree_net ? "true" : "false");
    out->threads_per_block = b.tpb + (b.split ? 64 : 0);
    out->grid = b.grid;
    out->n_cells = (int32_t)b.order.size();
    out->lds_bytes = (int32_t)b.lds;
    out->n_comps = b.comps;
    out->launches = b.launches;
    out->total_ms = b.ms;
    return 0;
}

int edenhip_get_kernel_cells(const edenhip_ctx *ctx, int32_t i, int32_t *cells, int32_t max_cells) {
    if (!ctx || !cells || i < 0 || i >= edenhip_kernel_count(ctx)) return EDENHIP_E_INVALID;
    if (ctx->n_point > 0 && i == 0) {
        if (max_cells < ctx->n_point) return EDENHIP_E_INVALID;
        std::copy(ctx->point_desc.begin(), ctx->point_desc.end(), cells);
        return 0;
    }
    const Bucket &b = ctx->buckets[(size_t)(i - (ctx->n_point > 0 ? 1 : 0))];
    if (max_cells < (int32_t)b.order.size()) return EDENHIP_E_INVALID;
    for (size_t k = 0; k < b.order.size(); k++) cells[k] = ctx->tree_desc[(size_t)b.order[k]];
    return 0;
}

int edenhip_network_stats(edenhip_ctx *ctx, int32_t *window_steps, int64_t *windows, int64_t *windows_rerun,
                          int64_t *synaptic_events) {
    if (!ctx) return EDENHIP_E_INVALID;
    if (window_steps) *window_steps = ctx->netwin ? ctx->net_W : 1;
    if (windows) *windows = ctx->net_windows;
    if (windows_rerun) *windows_rerun = ctx->net_redone;
    if (synaptic_events) {
        unsigned long long n = 0;
        if (ctx->netwin) HIPCHECK(ctx, hipMemcpy(&n, ctx->N.n_deliv, sizeof(n), hipMemcpyDeviceToHost));
        *synaptic_events = ctx->netwin ? (int64_t)n : -1;
    }
    return 0;
}

int edenhip_launch_shape(const edenhip_ctx *ctx, int32_t *threads_per_cell, int32_t *slots_per_thread) {
    if (!ctx) return EDENHIP_E_INVALID;
    if (ctx->buckets.empty()) {
        if (threads_per_cell) *threads_per_cell = 0;
        if (slots_per_thread) *slots_per_thread = 0;
        return 0;
    }
    if (threads_per_cell) *threads_per_cell = ctx->buckets[0].tpb;
    if (slots_per_thread) *slots_per_thread = 4;
    return 0;
}

const char *edenhip_last_error(const edenhip_ctx *ctx) {
    if (!ctx) return "null context";
    return ctx->err.c_str();
}

void edenhip_destroy(edenhip_ctx *ctx) {
    if (!ctx) return;
    if (ctx->stream) {
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        for (void *p : ctx->allocs) (void)hipFree(p);
        for (auto &b : ctx->buckets) {
            (void)b;   // pool streams stay for the process
            if (b.done) (void)hipEventDestroy(b.done);
            if (b.p0) (void)hipEventDestroy(b.p0);
            if (b.p1) (void)hipEventDestroy(b.p1);
        }
        if (ctx->ev_start) (void)hipEventDestroy(ctx->ev_start);
        for (auto &g : ctx->graphs) (void)hipGraphExecDestroy(g.exec);
        if (ctx->h_conflict) (void)hipHostFree(ctx->h_conflict);
        if (ctx->pp0) (void)hipEventDestroy(ctx->pp0);
        if (ctx->pp1) (void)hipEventDestroy(ctx->pp1);
        if (ctx->nw0) (void)hipEventDestroy(ctx->nw0);
        if (ctx->nw1) (void)hipEventDestroy(ctx->nw1);
        // ctx->stream is a pool stream: kept for the process
    }
    delete ctx;
}

}  // extern "C"
