"""Gradients of voltage-trace losses through the discrete adjoint (include/edenhip_adjoint.h).

`Adjoint` wraps one differentiable model on the GPU; `trace(adj, gbase, amplitude, weight)`
is the differentiable simulation for PyTorch: it returns the recorded voltage trace (rows x
recorded compartments, float64, mV) and its backward pass runs the hand-written adjoint
kernels, giving dLoss/dGbase per channel distribution, dLoss/dAmplitude per pulse input and
dLoss/dWeight per synapse instance (with the synaptic event schedule fixed:
`Adjoint.set_events`).

The differentiated map is the reference's fixed-step update evaluated in double precision
(as the FP64 restatement in oracle/ computes it); the fp32 path of eden_amd.Simulation is
the bit-exact forward. Threshold crossings only feed event logs, so they carry no gradient.
Tensors live on the CUDA (HIP) device; calls synchronise before and after.
"""
import ctypes as C

import torch

from . import _native as N
from .model import EdenHipError


class Adjoint:
    """Differentiable model: `model` is an eden_amd.Model without synapses, spike recipients
    or Poisson inputs. Parameters default to the model's chan_gbase and pulse_amplitude."""

    def __init__(self, model, checkpoint_steps=500, threads_per_cell=0, device=0, store_states=0):
        """store_states: 0 = keep every step's state in device memory when it fits (no
        recomputation in the backward pass), 1 = always, -1 = checkpoints and recompute."""
        self.lib = N.load_library()
        self.model = model
        self._desc = model.desc()
        opt = N.AdjOptions(checkpoint_steps, threads_per_cell, device, store_states)
        self.h = C.c_void_p()
        rc = self.lib.edenhip_adj_create(C.byref(self._desc), C.byref(opt), C.byref(self.h))
        if rc != 0:
            msg = self.lib.edenhip_adj_last_error(self.h).decode()
            self.lib.edenhip_adj_destroy(self.h)
            self.h = None
            raise EdenHipError(f"adjoint create failed ({rc}): {msg}")
        rows, nrec, nch, npu, nsy = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        self.lib.edenhip_adj_sizes(self.h, C.byref(rows), C.byref(nrec), C.byref(nch), C.byref(npu))
        self.lib.edenhip_adj_n_synapses(self.h, C.byref(nsy))
        self.rows, self.n_record, self.n_chan, self.n_pulse = rows.value, nrec.value, nch.value, npu.value
        self.n_syn = nsy.value
        self.device = torch.device("cuda", device)

    def _check(self, rc, what):
        if rc != 0:
            raise EdenHipError(f"{what} failed ({rc}): {self.lib.edenhip_adj_last_error(self.h).decode()}")

    def default_parameters(self):
        g = torch.tensor(self.model.chan_gbase, dtype=torch.float64, device=self.device)
        a = torch.tensor(self.model.pulse_amplitude, dtype=torch.float64, device=self.device)
        return g, a

    def default_weights(self):
        return torch.tensor(self.model.syn_weight, dtype=torch.float64, device=self.device)

    def set_events(self, steps, synapses):
        """The synapses' fixed event schedule: (step, synapse instance) pairs, steps from 1,
        as Simulation.read_synapse_events returns them."""
        import numpy as np
        st = np.ascontiguousarray(steps, np.int64)
        sy = np.ascontiguousarray(synapses, np.int64)
        self._check(self.lib.edenhip_adj_set_events(self.h, N.ptr(st), N.ptr(sy), len(st)), "adjoint set_events")

    @staticmethod
    def _ptr(t):
        return None if t is None else C.c_void_p(t.data_ptr())

    def _param(self, t, n, name):
        if t is None:
            return None
        if t.dtype != torch.float64 or t.device != self.device or t.numel() != n:
            raise ValueError(f"{name}: expected {n} float64 values on {self.device}")
        return t.contiguous()

    def forward(self, gbase=None, amplitude=None, weight=None):
        gbase = self._param(gbase, self.n_chan, "gbase")
        amplitude = self._param(amplitude, self.n_pulse, "amplitude")
        weight = self._param(weight, self.n_syn, "weight")
        out = torch.empty((self.rows, self.n_record), dtype=torch.float64, device=self.device)
        torch.cuda.synchronize(self.device)
        self._check(self.lib.edenhip_adj_forward_w(self.h, self._ptr(gbase), self._ptr(amplitude), self._ptr(weight),
                                                   self._ptr(out)), "adjoint forward")
        self._last = (gbase, amplitude, weight)   # keep the parameters alive for the backward pass
        return out

    def backward(self, grad_trace):
        g = grad_trace.to(device=self.device, dtype=torch.float64).contiguous()
        if g.shape != (self.rows, self.n_record):
            raise ValueError(f"grad_trace: expected shape {(self.rows, self.n_record)}")
        gg = torch.empty(max(1, self.n_chan), dtype=torch.float64, device=self.device)
        ga = torch.empty(max(1, self.n_pulse), dtype=torch.float64, device=self.device)
        gw = torch.empty(max(1, self.n_syn), dtype=torch.float64, device=self.device)
        torch.cuda.synchronize(self.device)
        self._check(self.lib.edenhip_adj_backward_w(self.h, self._ptr(g), self._ptr(gg), self._ptr(ga), self._ptr(gw)),
                    "adjoint backward")
        return gg[: self.n_chan], ga[: self.n_pulse], gw[: self.n_syn]

    def timing(self):
        f, b = C.c_double(), C.c_double()
        self.lib.edenhip_adj_timing(self.h, C.byref(f), C.byref(b))
        return f.value, b.value

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.edenhip_adj_destroy(self.h)
        self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class TraceFunction(torch.autograd.Function):
    """trace = simulate(gbase, amplitude, weight); backward through the adjoint kernels."""

    @staticmethod
    def forward(ctx, gbase, amplitude, weight, adj):
        ctx.adj = adj
        return adj.forward(gbase.detach(), amplitude.detach(), weight.detach())

    @staticmethod
    def backward(ctx, grad_trace):
        gg, ga, gw = ctx.adj.backward(grad_trace)
        return gg, ga, gw, None


def trace(adj, gbase, amplitude, weight=None):
    """Differentiable voltage trace of `adj`'s model for parameters gbase[n_chan],
    amplitude[n_pulse] and (when the model has synapses) weight[n_syn] (float64 tensors on
    the device; weight None = the model's connection weights, not differentiated)."""
    if weight is None:
        weight = adj.default_weights()
    return TraceFunction.apply(gbase, amplitude, weight, adj)
