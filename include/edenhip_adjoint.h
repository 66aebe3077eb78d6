/* edenhip_adjoint.h -- discrete adjoint of the fixed-step integrator: HH-style channels,
 * passive leak, pulse inputs, expTwo chemical synapses with a fixed event schedule, tree
 * cables.
 *
 * No counterpart in the reference (EDEN has no gradients). The differentiated map is the
 * per-step update of eden/Eden.cpp:5236-5259 (gates), 5805-5898 (channel currents), 3798-3826
 * (pulses), 6057-6107 (membrane) and 6114-6225 (Hines solve), evaluated in double precision
 * with the reference's operations (as the FP64 restatement in oracle/ does), except that the
 * Hines solve multiplies by each node's reciprocal eliminated diagonal where the reference
 * divides by the diagonal (results differ in the last bits). Spike threshold crossings only
 * feed event logs here, so they carry no gradient.
 *
 * The forward pass records the voltage trace of the recorded compartments (double), a
 * checkpoint of the full state every `checkpoint_steps` steps and, when it fits (see
 * store_states), a record of every step: its starting state and its Jacobian factors (the
 * reciprocal eliminated Hines diagonals and each gate's dq'/dq and dq'/dV). The backward pass
 * takes dLoss/dtrace and, segment by segment from the last, takes the segment's records from
 * that history or recomputes them from the segment's checkpoint, and runs the adjoint
 * recursion backwards, accumulating dLoss/dGbase per channel distribution, dLoss/dAmplitude
 * per pulse input and dLoss/dWeight per synapse instance. Gates clipped to [1e-6, 1 - 1e-6]
 * in a step pass no gradient through that step's update.
 *
 * Sizes: one cell's tables, state and step record live in LDS (at most 160 KiB per cell; an
 * HH cell of 100 compartments takes about 50 KiB); larger cells are refused with
 * EDENHIP_E_LIMIT. The history takes (2 n + 1 + 3 gates + 2 synapses) doubles per cell and
 * step.
 *
 * Synapses: threshold crossings are fixed events (no gradient through spike timing). The
 * steps at which each synapse instance receives its event are given by the caller
 * (edenhip_adj_set_events, e.g. the schedule edenhip_read_synapse_events logs for a run of
 * the same model); each cell then evolves independently. An instance's conductance
 * g(t) = gbase (B - A) follows its expTwo dynamics (eden/Eden.cpp:2614-2994) and does not
 * depend on V or on the weights; the connection weight w multiplies its current and
 * conductance (eden/Eden.cpp:3710), so dLoss/dw = sum over steps of g dt / C
 * (lambda_b erev + lambda_D), as for a channel. Models with synapses need the state history
 * kept in device memory (store_states).
 *
 * Pointers marked (device) are device memory of the context's GPU; calls synchronise their
 * work before returning. */
#ifndef EDENHIP_ADJOINT_H
#define EDENHIP_ADJOINT_H
#include <stdint.h>

#include "edenhip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct edenhip_adj edenhip_adj;

typedef struct edenhip_adj_options {
    int32_t checkpoint_steps;    /* steps between checkpoints (0 = 500) */
    int32_t threads_per_cell;    /* 0 = by cell size, or 64/128/256 */
    int32_t device;
    int32_t store_states;        /* 0 = keep every step's state in device memory when it takes at
                                    most half of the free memory, else checkpoints and recompute;
                                    1 = always keep it (EDENHIP_E_LIMIT when it does not fit);
                                    -1 = never (checkpoints every checkpoint_steps, recompute) */
} edenhip_adj_options;

/* Build the differentiable model from the same descriptor edenhip_create() takes. The model
 * must have no Poisson inputs or generic LEMS synapse types (EDENHIP_E_INVALID otherwise);
 * spike recipients are ignored (the synapses' events come from edenhip_adj_set_events). The
 * descriptor's chan_gbase, pulse_amplitude and syn_weight are the default parameter values. */
int edenhip_adj_create(const edenhip_model_desc *desc, const edenhip_adj_options *opt, edenhip_adj **out);

/* The fixed event schedule of the synapses: n pairs (step, synapse instance), steps counting
 * from 1 as edenhip_read_synapse_events reports them (the step whose time window holds the
 * arrival). Replaces the previous schedule; none until called. */
int edenhip_adj_set_events(edenhip_adj *adj, const int64_t *step, const int64_t *syn, int64_t n);

/* Synapse instances of the model (the length of the weight vectors). */
int edenhip_adj_n_synapses(const edenhip_adj *adj, int64_t *n_syn);

/* Sizes: rows of the trace, recorded columns, channel distributions, pulse inputs. */
int edenhip_adj_sizes(const edenhip_adj *adj, int64_t *rows, int64_t *n_record, int64_t *n_chan, int64_t *n_pulse);

/* Forward pass with parameters gbase[n_chan] and amplitude[n_pulse] (device, double; NULL =
 * the descriptor's values): writes trace[rows x n_record] (device, double, row r = voltage
 * at time r * dt in mV) and keeps the checkpoints for edenhip_adj_backward. */
int edenhip_adj_forward(edenhip_adj *adj, const double *gbase, const double *amplitude, double *trace);

/* Forward pass with connection weights weight[n_syn] too (device, double; NULL = the
 * descriptor's syn_weight). */
int edenhip_adj_forward_w(edenhip_adj *adj, const double *gbase, const double *amplitude, const double *weight,
                          double *trace);

/* Backward pass for the parameters of the last forward pass: grad_trace[rows x n_record]
 * (device, double) is dLoss/dtrace; writes grad_gbase[n_chan] and grad_amplitude[n_pulse]
 * (device, double; either may be NULL). */
int edenhip_adj_backward(edenhip_adj *adj, const double *grad_trace, double *grad_gbase, double *grad_amplitude);

/* Backward pass with dLoss/dWeight too: grad_weight[n_syn] (device, double; may be NULL). */
int edenhip_adj_backward_w(edenhip_adj *adj, const double *grad_trace, double *grad_gbase, double *grad_amplitude,
                           double *grad_weight);

/* Milliseconds of device time of the last forward and backward passes (HIP events). */
int edenhip_adj_timing(const edenhip_adj *adj, double *forward_ms, double *backward_ms);

const char *edenhip_adj_last_error(const edenhip_adj *adj);
void edenhip_adj_destroy(edenhip_adj *adj);

#ifdef __cplusplus
}
#endif
#endif /* EDENHIP_ADJOINT_H */
