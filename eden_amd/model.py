"""Instantiated models (the SoA form the C-ABI consumes) and the GPU simulation handle.

``Model`` holds every array of ``edenhip_model_desc`` as numpy arrays, in EDEN's instance
order; ``Model.desc()`` returns the ctypes descriptor (the arrays stay owned by the Model).
``Simulation`` drives ``libedenhip.so``: create, run chunks of rows, read traces/events.
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N

# engine units (eden/NeuroML.cpp:305-425): us, mV, nA, uS, pF, MOhm, mS/cm^2, um, 1/us
PER_MS = 1e-3   # 1 per_ms in 1/us


def rate(form, r_per_ms, midpoint_mV, scale_mV):
    return N.Rate(form, np.float32(r_per_ms * PER_MS), midpoint_mV, scale_mV)


def channel_type(gates):
    """gates: list of (instances, alpha Rate, beta Rate[, q10])."""
    ct = N.ChannelType()
    ct.n_gates = len(gates)
    for i, g in enumerate(gates):
        inst, a, b = g[:3]
        q10 = g[3] if len(g) > 3 else 1.0
        ct.gates[i] = N.Gate(inst, q10, a, b)
    return ct


def hh_squid_types():
    """Leak, Na (m^3 h) and K (n^4) of the HH squid axon as in NeuroML's
    NML2_SingleCompHHCell example (examples/NML2_SingleCompHHCell.nml:14-43)."""
    leak = channel_type([])
    na = channel_type([
        (3, rate(N.RATE_EXPLINEAR, 1.0, -40.0, 10.0), rate(N.RATE_EXP, 4.0, -65.0, -18.0)),
        (1, rate(N.RATE_EXP, 0.07, -65.0, -20.0), rate(N.RATE_SIGMOID, 1.0, -35.0, 10.0)),
    ])
    k = channel_type([
        (4, rate(N.RATE_EXPLINEAR, 0.1, -55.0, 10.0), rate(N.RATE_EXP, 0.125, -65.0, -80.0)),
    ])
    return [leak, na, k]


def reverse_bits32(x):
    x = int(x) & 0xFFFFFFFF
    x = ((x & 0xFFFF0000) >> 16) | ((x & 0x0000FFFF) << 16)
    x = ((x & 0xFF00FF00) >> 8) | ((x & 0x00FF00FF) << 8)
    x = ((x & 0xF0F0F0F0) >> 4) | ((x & 0x0F0F0F0F) << 4)
    x = ((x & 0xCCCCCCCC) >> 2) | ((x & 0x33333333) << 2)
    x = ((x & 0xAAAAAAAA) >> 1) | ((x & 0x55555555) << 1)
    return x & 0xFFFFFFFF


def cell_rng_seeds(sim_seed, n_cells):
    """ReverseBits32(sim_seed) ^ cell_gid (eden/Eden.cpp:7679-7691)."""
    base = reverse_bits32(sim_seed)
    return (np.uint32(base) ^ np.arange(n_cells, dtype=np.uint32)).astype(np.uint32)


def _csr(counts):
    b = np.zeros(len(counts) + 1, np.int64)
    np.cumsum(counts, out=b[1:])
    return b


@dataclass
class Model:
    dt: float                          # us
    t_final: float                     # us
    cell_comp_begin: np.ndarray        # int64 [n_cells+1]
    v0: np.ndarray
    capacitance: np.ndarray
    area: np.ndarray
    inv_rc_axial: np.ndarray
    r_axial: np.ndarray
    v_threshold: np.ndarray
    parent: np.ndarray                 # int32 cell-local
    channel_types: list
    comp_chan_begin: np.ndarray
    chan_type: np.ndarray
    chan_gbase: np.ndarray
    chan_erev: np.ndarray
    cell_rng_seed: np.ndarray = None
    comp_pulse_begin: np.ndarray = None
    pulse_start: np.ndarray = None
    pulse_duration: np.ndarray = None
    pulse_amplitude: np.ndarray = None
    pulse_weight: np.ndarray = None
    syn_types: list = field(default_factory=list)
    comp_syn_begin: np.ndarray = None
    syn_type: np.ndarray = None
    syn_weight: np.ndarray = None
    syn_delay: np.ndarray = None
    comp_spike_begin: np.ndarray = None
    spike_target: np.ndarray = None
    comp_poisson_begin: np.ndarray = None
    poisson_rate: np.ndarray = None
    poisson_syn_type: np.ndarray = None
    poisson_weight: np.ndarray = None
    poisson_instance: np.ndarray = None
    poisson_rng_offset: np.ndarray = None
    record_comp: np.ndarray = None
    event_comp: np.ndarray = None
    lems: object = None                # eden_amd.lems.LemsSynapses: generic synapse types
    syn_peer: np.ndarray = None        # per synapse instance: the gap junction's peer compartment, or -1
    record_gate: np.ndarray = None     # per recorded column: -1 V, j >= 0 the compartment's j-th gate variable
                                       # (a point neuron's j-th state variable)
    lems_cells: object = None          # eden_amd.lems_cells.LemsCells, or None
    cell_lems_type: np.ndarray = None  # per cell: its point-neuron type, or -1
    # per-instance parameter vectors (the reference's EdenCustomSetup statements): per generic
    # synapse instance / per point-neuron cell an index into the list of vectors, -1 for its
    # type's; each vector has its type's layout
    syn_param_index: np.ndarray = None
    syn_param_vecs: list = None
    cell_param_index: np.ndarray = None
    cell_param_vecs: list = None
    # firing-synapse inputs (timedSynapticInput): per synapse instance the offset of its spike
    # list in spike_times, or -1; the lists' times (us), each ended by +inf
    syn_spike_list: np.ndarray = None
    spike_times: np.ndarray = None
    # Poisson firing-synapse inputs on the generic path: per synapse instance an index into the
    # spiker table, or -1; the table: dict of equal-length arrays rate (1/us), transient (0/1),
    # delay, duration (us), instance, rng_offset
    syn_poisson_spiker: np.ndarray = None
    poisson_spikers: dict = None
    # per channel type, per gate slot (EDENHIP_MAX_GATES of them): its kind (0 alpha/beta
    # rates, 1 tau/inf), or None for rates everywhere
    gate_kind: np.ndarray = None
    lems_rates: object = None          # eden_amd.lems_rates.LemsRates: user gate rate ComponentTypes, or None
    chan_scaling: np.ndarray = None    # per channel type: conductance_scaling (q10ConductanceScaling), or None
    gate_extra_rate: list = None       # per channel type, per gate slot: the third rate (N.Rate) of a
                                       # rates+tau / rates+inf gate, or None
    name: str = "model"

    def __post_init__(self):
        nc = self.n_comps
        z = lambda: np.zeros(nc + 1, np.int64)
        for f in ("comp_pulse_begin", "comp_syn_begin", "comp_spike_begin", "comp_poisson_begin"):
            if getattr(self, f) is None:
                setattr(self, f, z())
        f32 = lambda: np.zeros(0, np.float32)
        i32 = lambda: np.zeros(0, np.int32)
        for f in ("pulse_start", "pulse_duration", "pulse_amplitude", "pulse_weight", "syn_weight",
                  "syn_delay", "poisson_rate", "poisson_weight"):
            if getattr(self, f) is None:
                setattr(self, f, f32())
        for f in ("syn_type", "poisson_syn_type", "poisson_instance", "poisson_rng_offset"):
            if getattr(self, f) is None:
                setattr(self, f, i32())
        if self.spike_target is None:
            self.spike_target = np.zeros(0, np.int64)
        if self.record_comp is None:
            self.record_comp = np.zeros(0, np.int64)
        if self.event_comp is None:
            self.event_comp = np.zeros(0, np.int64)
        if self.cell_rng_seed is None:
            self.cell_rng_seed = cell_rng_seeds(0, self.n_cells)
        self._normalise()

    def _normalise(self):
        conv = {
            np.int64: ("cell_comp_begin", "comp_chan_begin", "comp_pulse_begin", "comp_syn_begin",
                       "comp_spike_begin", "comp_poisson_begin", "spike_target", "record_comp", "event_comp"),
            np.int32: ("parent", "chan_type", "syn_type", "poisson_syn_type", "poisson_instance",
                       "poisson_rng_offset"),
            np.float32: ("v0", "capacitance", "area", "inv_rc_axial", "r_axial", "v_threshold", "chan_gbase",
                         "chan_erev", "pulse_start", "pulse_duration", "pulse_amplitude", "pulse_weight",
                         "syn_weight", "syn_delay", "poisson_rate", "poisson_weight"),
            np.uint32: ("cell_rng_seed",),
        }
        for dt, names in conv.items():
            for n in names:
                setattr(self, n, np.ascontiguousarray(getattr(self, n), dtype=dt))
        if self.gate_kind is not None:
            self.gate_kind = np.ascontiguousarray(self.gate_kind, dtype=np.int32).reshape(-1)
            if self.gate_kind.shape != (len(self.channel_types) * N.MAX_GATES,):
                raise ValueError("gate_kind: MAX_GATES entries per channel type")
        if self.record_gate is not None:
            self.record_gate = np.ascontiguousarray(self.record_gate, dtype=np.int32)
            if self.record_gate.shape != self.record_comp.shape:
                raise ValueError("record_gate: one entry per recorded column")

    @property
    def n_cells(self):
        return len(self.cell_comp_begin) - 1

    @property
    def n_comps(self):
        return int(self.cell_comp_begin[-1])

    @property
    def n_synapses(self):
        return int(self.comp_syn_begin[-1])

    def comps_per_cell(self):
        return np.diff(self.cell_comp_begin)

    def subset(self, cells):
        """The model restricted to the given cells (in the given order): their compartments,
        channels, pulses, recorded and event columns, RNG seeds. Models without synapses and
        Poisson inputs only (cells are then independent, so each cell's trace does not depend
        on the others). Returns (model, record_cols, event_cols): the columns of this model's
        trace and event log that the subset's columns come from."""
        if self.n_synapses or len(self.poisson_rate) or len(self.spike_target):
            raise ValueError("subset: cells with synapses or inputs from other cells are not independent")
        cells = np.asarray(cells, np.int64)
        cb = self.cell_comp_begin
        comps = np.concatenate([np.arange(cb[c], cb[c + 1]) for c in cells]).astype(np.int64)
        new_of = np.full(self.n_comps, -1, np.int64)
        new_of[comps] = np.arange(len(comps))

        def ranges(begin, items):
            idx = np.concatenate([np.arange(begin[g], begin[g + 1]) for g in items] + [np.zeros(0, np.int64)])
            counts = np.array([begin[g + 1] - begin[g] for g in items], np.int64)
            return idx.astype(np.int64), _csr(counts)

        ch, ch_b = ranges(self.comp_chan_begin, comps)
        pu, pu_b = ranges(self.comp_pulse_begin, comps)
        rec_cols = np.flatnonzero(new_of[self.record_comp] >= 0)
        ev_cols = np.flatnonzero(new_of[self.event_comp] >= 0)
        # keep the full model's column order among the chosen cells' columns
        m = Model(
            dt=self.dt, t_final=self.t_final, cell_comp_begin=_csr(cb[cells + 1] - cb[cells]),
            v0=self.v0[comps], capacitance=self.capacitance[comps], area=self.area[comps],
            inv_rc_axial=self.inv_rc_axial[comps], r_axial=self.r_axial[comps], v_threshold=self.v_threshold[comps],
            parent=self.parent[comps], channel_types=self.channel_types, comp_chan_begin=ch_b,
            chan_type=self.chan_type[ch], chan_gbase=self.chan_gbase[ch], chan_erev=self.chan_erev[ch],
            cell_rng_seed=self.cell_rng_seed[cells], comp_pulse_begin=pu_b, pulse_start=self.pulse_start[pu],
            pulse_duration=self.pulse_duration[pu], pulse_amplitude=self.pulse_amplitude[pu],
            pulse_weight=self.pulse_weight[pu], syn_types=self.syn_types,
            record_comp=new_of[self.record_comp[rec_cols]], event_comp=new_of[self.event_comp[ev_cols]],
            record_gate=None if self.record_gate is None else self.record_gate[rec_cols], gate_kind=self.gate_kind,
            name=f"{self.name}_subset{len(cells)}")
        return m, rec_cols, ev_cols

    def total_rows(self):
        """floor(t_final/dt) + 1 with the reference's double accumulation of float dt."""
        t, n, dt = 0.0, 0, float(np.float32(self.dt))
        # the accumulation is exact for the dt values used here; count directly then verify the tail
        n = int(np.floor(self.t_final / dt)) + 1
        while (n - 1) * dt > self.t_final:
            n -= 1
        while n * dt <= self.t_final:
            n += 1
        return n

    def desc(self):
        d = N.ModelDesc()
        d.abi_version = N.ABI_VERSION
        d.dt = self.dt
        d.t_final = self.t_final
        d.n_cells = self.n_cells
        d.n_comps = self.n_comps
        self._ct = (N.ChannelType * max(1, len(self.channel_types)))(*self.channel_types)
        self._st = (N.ExpTwoType * max(1, len(self.syn_types)))(*self.syn_types)
        d.n_channel_types = len(self.channel_types)
        d.channel_types = C.cast(self._ct, C.c_void_p)
        d.n_syn_types = len(self.syn_types)
        d.syn_types = C.cast(self._st, C.c_void_p)
        for name, _ in N.ModelDesc._fields_:
            if name in ("abi_version", "dt", "t_final", "n_cells", "n_comps", "n_channel_types", "channel_types",
                        "n_syn_types", "syn_types", "n_record", "n_event_cols", "lems", "lems_cells",
                        "lems_rates_source", "gate_extra_rate", "chan_scaling"):
                continue
            setattr(d, name, N.ptr(getattr(self, name)))
        if self.lems is not None:
            L = self.lems
            pb = np.zeros(len(L.params) + 1, np.int32)
            np.cumsum([len(p) for p in L.params], out=pb[1:])
            self._lems_arrays = (pb, np.concatenate(L.params + [np.zeros(1, np.float32)]).astype(np.float32),
                                 np.array(L.n_state, np.int32), L.source_hip.encode())
            peer = None
            if self.syn_peer is not None:
                peer = np.ascontiguousarray(self.syn_peer, np.int64)
                self._lems_peer = peer
            self._lems_flags = np.array(L.type_flags or [0] * len(L.params), np.int32)
            prm, off = _with_instance_vectors(L.params, self.syn_param_vecs, self.syn_param_index, self.n_synapses)
            self._lems_arrays = self._lems_arrays[:1] + (prm,) + self._lems_arrays[2:] + (off,)
            sl, stimes = None, None
            if self.syn_spike_list is not None:
                sl = np.ascontiguousarray(self.syn_spike_list, np.int64)
                stimes = np.ascontiguousarray(self.spike_times, np.float32)
                if sl.shape != (self.n_synapses,):
                    raise ValueError("syn_spike_list: one entry per synapse instance")
            ps, pk = None, {}
            if self.syn_poisson_spiker is not None:
                ps = np.ascontiguousarray(self.syn_poisson_spiker, np.int64)
                if ps.shape != (self.n_synapses,):
                    raise ValueError("syn_poisson_spiker: one entry per synapse instance")
                P = self.poisson_spikers
                pk = {k: np.ascontiguousarray(P[k], dt) for k, dt in
                      (("rate", np.float32), ("transient", np.int32), ("delay", np.float32), ("duration", np.float32),
                       ("instance", np.int32), ("rng_offset", np.int32))}
                if len({len(v) for v in pk.values()}) != 1:
                    raise ValueError("poisson_spikers: arrays of equal length")
            self._lems_spikes = (sl, stimes, ps, pk)
            self._lems_desc = N.LemsSynapsesDesc(len(L.params), self._lems_arrays[3], N.ptr(self._lems_arrays[0]),
                                                 N.ptr(self._lems_arrays[1]), N.ptr(self._lems_arrays[2]),
                                                 N.ptr(peer), N.ptr(self._lems_flags), N.ptr(off), len(prm),
                                                 N.ptr(sl), N.ptr(stimes), 0 if stimes is None else len(stimes),
                                                 N.ptr(ps), len(pk.get("rate", ())), N.ptr(pk.get("rate")),
                                                 N.ptr(pk.get("transient")), N.ptr(pk.get("delay")),
                                                 N.ptr(pk.get("duration")), N.ptr(pk.get("instance")),
                                                 N.ptr(pk.get("rng_offset")))
            d.lems = C.pointer(self._lems_desc)
        if self.lems_cells is not None:
            LC = self.lems_cells
            pb = np.zeros(len(LC.params) + 1, np.int32)
            np.cumsum([len(p) for p in LC.params], out=pb[1:])
            ctype = np.ascontiguousarray(self.cell_lems_type, np.int32)
            self._lc_arrays = (pb, np.concatenate(LC.params + [np.zeros(1, np.float32)]).astype(np.float32),
                               np.array(LC.n_state, np.int32), np.array(LC.v_state, np.int32), ctype,
                               LC.source_hip.encode())
            prm, off = _with_instance_vectors(LC.params, self.cell_param_vecs, self.cell_param_index, self.n_cells)
            self._lc_arrays = self._lc_arrays[:1] + (prm,) + self._lc_arrays[2:] + (off,)
            a = self._lc_arrays
            self._lc_desc = N.LemsCellsDesc(len(LC.params), a[5], N.ptr(a[0]), N.ptr(a[1]), N.ptr(a[2]), N.ptr(a[3]),
                                            N.ptr(a[4]), N.ptr(off), len(prm))
            d.lems_cells = C.pointer(self._lc_desc)
        if self.chan_scaling is not None:
            self._cs = np.ascontiguousarray(self.chan_scaling, np.float32)
            if self._cs.shape != (len(self.channel_types),):
                raise ValueError("chan_scaling: one entry per channel type")
            d.chan_scaling = N.ptr(self._cs)
        if self.gate_extra_rate is not None:
            if len(self.gate_extra_rate) != len(self.channel_types) * N.MAX_GATES:
                raise ValueError("gate_extra_rate: MAX_GATES entries per channel type")
            self._gx = (N.Rate * len(self.gate_extra_rate))(
                *[r if r is not None else N.Rate(N.RATE_FIXED, 0.0, 0.0, 0.0) for r in self.gate_extra_rate])
            d.gate_extra_rate = C.cast(self._gx, C.c_void_p)
        if self.lems_rates is not None:
            self._rates_src = self.lems_rates.source_hip.encode()
            d.lems_rates_source = self._rates_src
        d.n_record = len(self.record_comp)
        d.n_event_cols = len(self.event_comp)
        return d


def _with_instance_vectors(type_params, vecs, index, n):
    """The parameter table of the types' vectors followed by the per-instance vectors, and
    each instance's offset into it (-1: its type's), or None when no instance has its own."""
    base = [np.asarray(p, np.float32) for p in type_params]
    total = sum(len(p) for p in base)
    if not vecs or index is None:
        return np.concatenate(base + [np.zeros(1, np.float32)]).astype(np.float32), None
    index = np.asarray(index, np.int64)
    if index.shape != (n,):
        raise ValueError("per-instance parameter index: one entry per instance")
    starts = total + np.concatenate([[0], np.cumsum([len(v) for v in vecs])[:-1]]).astype(np.int64)
    off = np.where(index >= 0, starts[np.maximum(index, 0)], -1).astype(np.int64)
    prm = np.concatenate(base + [np.asarray(v, np.float32) for v in vecs]).astype(np.float32)
    return prm, off


class EdenHipError(RuntimeError):
    pass


class Simulation:
    """One model on one GPU. Rows are recorded steps (state at t = row * dt)."""

    def __init__(self, model, threads_per_cell=0, persistent=1, chunk_steps=0, profile=False, device=0,
                 log_synapse_events=0, network_window=0, size_limits=None, use_graph=False, channel_wave=None):
        self.lib = N.load_library()
        self.model = model
        self._desc = model.desc()
        limits = (C.c_int32 * 4)(*(size_limits or (0, 0, 0, 0)))
        waves = (C.c_int32 * 4)(*(channel_wave or (0, 0, 0, 0)))
        opt = N.Options(threads_per_cell, persistent, chunk_steps, 1 if use_graph else 0, 1 if profile else 0, device,
                        log_synapse_events, network_window, limits, waves)
        h = C.c_void_p()
        rc = self.lib.edenhip_create(C.byref(self._desc), C.byref(opt), C.byref(h))
        self.h = h
        if rc != 0:
            msg = self.lib.edenhip_last_error(h).decode() if h.value else "create failed"
            self.close()
            raise EdenHipError(f"edenhip_create: {msg} ({rc})")
        self.total_rows = int(self.lib.edenhip_total_rows(h))
        self.n_record = len(model.record_comp)

    def _check(self, rc, what):
        if rc != 0:
            raise EdenHipError(f"{what}: {self.lib.edenhip_last_error(self.h).decode()} ({rc})")

    def run(self, n_rows=-1):
        done = C.c_int64()
        self._check(self.lib.edenhip_run(self.h, n_rows, C.byref(done)), "edenhip_run")
        return done.value

    def run_static(self, n_rows=-1):
        """run() through edenhip_run_static, the entry point that allocates nothing (its
        scratch is a static device array shared by every call on the GPU)."""
        done = C.c_int64()
        self._check(self.lib.edenhip_run_static(self.h, n_rows, C.byref(done)), "edenhip_run_static")
        return done.value

    def get_state(self):
        """The simulation state as bytes (edenhip_get_state): restore it with set_state() in
        a Simulation of the same model and options."""
        n = C.c_int64()
        self._check(self.lib.edenhip_state_size(self.h, C.byref(n)), "edenhip_state_size")
        buf = np.empty(n.value, np.uint8)
        self._check(self.lib.edenhip_get_state(self.h, N.ptr(buf), n.value), "edenhip_get_state")
        return buf.tobytes()

    def set_state(self, blob):
        buf = np.frombuffer(blob, np.uint8).copy()
        self._check(self.lib.edenhip_set_state(self.h, N.ptr(buf), len(buf)), "edenhip_set_state")

    def reset(self):
        self._check(self.lib.edenhip_reset(self.h), "edenhip_reset")

    def read_trace(self):
        rows, row0 = C.c_int64(), C.c_int64()
        self._check(self.lib.edenhip_read_trace(self.h, None, 0, C.byref(rows), C.byref(row0)), "read_trace")
        out = np.empty((max(1, self.total_rows), max(1, self.n_record)), np.float32)
        self._check(self.lib.edenhip_read_trace(self.h, N.ptr(out), out.shape[0], C.byref(rows), C.byref(row0)),
                    "read_trace")
        return out[: rows.value, : self.n_record], row0.value

    def run_all(self):
        """Run every row; return the full trace [rows x n_record]."""
        parts = []
        while True:
            n = self.run(-1)
            if n == 0:
                break
            tr, _ = self.read_trace()
            parts.append(tr.copy())
            if sum(p.shape[0] for p in parts) >= self.total_rows:
                break
        if not parts:
            return np.zeros((0, self.n_record), np.float32)
        return np.concatenate(parts, axis=0)

    def read_events(self):
        n = C.c_int64()
        self._check(self.lib.edenhip_read_events(self.h, None, None, 0, C.byref(n)), "read_events")
        col = np.empty(max(1, n.value), np.int64)
        row = np.empty(max(1, n.value), np.int64)
        self._check(self.lib.edenhip_read_events(self.h, N.ptr(col), N.ptr(row), n.value, C.byref(n)), "read_events")
        return col[: n.value], row[: n.value]

    def read_synapse_events(self):
        """(steps, synapses) of the logged synaptic events, sorted by (step, synapse)."""
        n = C.c_int64()
        self._check(self.lib.edenhip_read_synapse_events(self.h, None, None, 0, C.byref(n)), "read_synapse_events")
        syn = np.empty(max(1, n.value), np.int64)
        step = np.empty(max(1, n.value), np.int64)
        self._check(self.lib.edenhip_read_synapse_events(self.h, N.ptr(syn), N.ptr(step), n.value, C.byref(n)),
                    "read_synapse_events")
        return step[: n.value], syn[: n.value]

    def voltages(self):
        v = np.empty(self.model.n_comps, np.float32)
        self._check(self.lib.edenhip_get_voltages(self.h, N.ptr(v)), "get_voltages")
        return v

    def kernel_stats(self):
        ms, nl = C.c_double(), C.c_int64()
        name = C.create_string_buffer(64)
        self._check(self.lib.edenhip_kernel_stats(self.h, C.byref(ms), C.byref(nl), name, 64), "kernel_stats")
        return ms.value, nl.value, name.value.decode()

    def kernel_infos(self):
        """Per-kernel launch statistics: list of dicts (name, threads_per_block, grid, cells,
        comps, lds_bytes, launches, total_ms, cell_index array)."""
        out = []
        for i in range(self.lib.edenhip_kernel_count(self.h)):
            ki = N.KernelInfo()
            self._check(self.lib.edenhip_get_kernel_info(self.h, i, C.byref(ki)), "get_kernel_info")
            cells = np.empty(max(1, ki.n_cells), np.int32)
            self._check(self.lib.edenhip_get_kernel_cells(self.h, i, N.ptr(cells), len(cells)), "get_kernel_cells")
            out.append(dict(name=ki.name.decode(), threads_per_block=ki.threads_per_block, grid=ki.grid,
                            cells=ki.n_cells, comps=ki.n_comps, lds_bytes=ki.lds_bytes, launches=ki.launches,
                            total_ms=ki.total_ms,
                            cell_index=cells[: ki.n_cells]))
        return out

    def network_stats(self):
        """(steps per window, windows, windows run again step by step, synaptic events received)."""
        w, n, r, e = C.c_int32(), C.c_int64(), C.c_int64(), C.c_int64()
        self._check(self.lib.edenhip_network_stats(self.h, C.byref(w), C.byref(n), C.byref(r), C.byref(e)),
                    "network_stats")
        return w.value, n.value, r.value, e.value

    def launch_shape(self):
        t, k = C.c_int32(), C.c_int32()
        self._check(self.lib.edenhip_launch_shape(self.h, C.byref(t), C.byref(k)), "launch_shape")
        return t.value, k.value

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.edenhip_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
