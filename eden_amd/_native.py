"""ctypes mirror of include/edenhip.h and include/edenhip_setup.h, and the library loader.

The HIP library is required: loading fails loudly when ``eden_amd/libedenhip.so`` is
missing (build it with ``python -m eden_amd.build``). There is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(ROOT, "libedenhip.so")

ABI_VERSION = 4
MAX_GATES = 4

RATE_EXP, RATE_EXPLINEAR, RATE_SIGMOID, RATE_FIXED, RATE_LEMS = 0, 1, 2, 3, 4
GATE_RATES, GATE_TAU_INF, GATE_INSTANTANEOUS, GATE_RATES_TAU, GATE_RATES_INF = 0, 1, 2, 3, 4

E_OK, E_INVALID, E_DEVICE, E_LIMIT, E_STATE = 0, -1, -2, -3, -4


class Rate(C.Structure):
    _fields_ = [("form", C.c_int32), ("rate", C.c_float), ("midpoint", C.c_float), ("scale", C.c_float)]


class Gate(C.Structure):
    _fields_ = [("instances", C.c_int32), ("q10", C.c_float), ("alpha", Rate), ("beta", Rate)]


class ChannelType(C.Structure):
    _fields_ = [("n_gates", C.c_int32), ("gates", Gate * MAX_GATES)]


class ExpTwoType(C.Structure):
    _fields_ = [("gbase", C.c_float), ("erev", C.c_float), ("tau_rise", C.c_float),
                ("tau_decay", C.c_float), ("weight_prop", C.c_float)]


_P = C.c_void_p


class LemsSynapsesDesc(C.Structure):
    _fields_ = [("n_types", C.c_int32), ("source", C.c_char_p), ("param_begin", _P), ("params", _P),
                ("n_state", _P), ("peer", _P), ("type_flags", _P), ("instance_param", _P),
                ("n_params", C.c_int64), ("spike_list", _P), ("spike_times", _P), ("n_spike_times", C.c_int64),
                ("poisson_spiker", _P), ("n_poisson_spikers", C.c_int64), ("spiker_rate", _P),
                ("spiker_transient", _P), ("spiker_delay", _P), ("spiker_duration", _P), ("spiker_instance", _P),
                ("spiker_rng_offset", _P)]


class LemsCellsDesc(C.Structure):
    _fields_ = [("n_types", C.c_int32), ("source", C.c_char_p), ("param_begin", _P), ("params", _P),
                ("n_state", _P), ("v_state", _P), ("cell_type", _P), ("cell_param", _P),
                ("n_params", C.c_int64)]


class ModelDesc(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int32), ("dt", C.c_float), ("t_final", C.c_double),
        ("n_cells", C.c_int32), ("n_comps", C.c_int64), ("cell_comp_begin", _P), ("cell_rng_seed", _P),
        ("v0", _P), ("capacitance", _P), ("area", _P), ("inv_rc_axial", _P), ("r_axial", _P),
        ("v_threshold", _P), ("parent", _P),
        ("n_channel_types", C.c_int32), ("channel_types", _P), ("comp_chan_begin", _P), ("chan_type", _P),
        ("chan_gbase", _P), ("chan_erev", _P),
        ("comp_pulse_begin", _P), ("pulse_start", _P), ("pulse_duration", _P), ("pulse_amplitude", _P),
        ("pulse_weight", _P),
        ("n_syn_types", C.c_int32), ("syn_types", _P),
        ("comp_syn_begin", _P), ("syn_type", _P), ("syn_weight", _P), ("syn_delay", _P),
        ("comp_spike_begin", _P), ("spike_target", _P),
        ("comp_poisson_begin", _P), ("poisson_rate", _P), ("poisson_syn_type", _P), ("poisson_weight", _P),
        ("poisson_instance", _P), ("poisson_rng_offset", _P),
        ("n_record", C.c_int64), ("record_comp", _P), ("n_event_cols", C.c_int64), ("event_comp", _P),
        ("lems", C.POINTER(LemsSynapsesDesc)),
        ("record_gate", _P),
        ("lems_cells", C.POINTER(LemsCellsDesc)),
        ("gate_kind", _P),
        ("lems_rates_source", C.c_char_p),
        ("gate_extra_rate", _P),
        ("chan_scaling", _P),
    ]


class Options(C.Structure):
    _fields_ = [("threads_per_cell", C.c_int32), ("persistent", C.c_int32), ("chunk_steps", C.c_int32),
                ("use_graph", C.c_int32), ("profile", C.c_int32), ("device", C.c_int32),
                ("log_synapse_events", C.c_int32), ("network_window", C.c_int32),
                ("size_limits", C.c_int32 * 4), ("channel_wave", C.c_int32 * 4)]


class KernelInfo(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("threads_per_block", C.c_int32), ("grid", C.c_int32),
                ("n_cells", C.c_int32), ("lds_bytes", C.c_int32), ("n_comps", C.c_int64),
                ("launches", C.c_int64), ("total_ms", C.c_double)]


class AdjOptions(C.Structure):
    _fields_ = [("checkpoint_steps", C.c_int32), ("threads_per_cell", C.c_int32), ("device", C.c_int32),
                ("store_states", C.c_int32)]


EXPORTS = {
    # name: (restype, argtypes)
    "edenhip_create": (C.c_int, [C.POINTER(ModelDesc), C.POINTER(Options), C.POINTER(C.c_void_p)]),
    "edenhip_total_rows": (C.c_int64, [C.c_void_p]),
    "edenhip_run": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]),
    "edenhip_run_static": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]),
    "edenhip_reset": (C.c_int, [C.c_void_p]),
    "edenhip_state_size": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "edenhip_get_state": (C.c_int, [C.c_void_p, _P, C.c_int64]),
    "edenhip_set_state": (C.c_int, [C.c_void_p, _P, C.c_int64]),
    "edenhip_read_trace": (C.c_int, [C.c_void_p, _P, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "edenhip_trace_device_ptr": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "edenhip_read_events": (C.c_int, [C.c_void_p, _P, _P, C.c_int64, C.POINTER(C.c_int64)]),
    "edenhip_get_voltages": (C.c_int, [C.c_void_p, _P]),
    "edenhip_read_synapse_events": (C.c_int, [C.c_void_p, _P, _P, C.c_int64, C.POINTER(C.c_int64)]),
    "edenhip_kernel_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.c_char_p, C.c_int32]),
    "edenhip_kernel_count": (C.c_int32, [C.c_void_p]),
    "edenhip_get_kernel_info": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(KernelInfo)]),
    "edenhip_get_kernel_cells": (C.c_int, [C.c_void_p, C.c_int32, _P, C.c_int32]),
    "edenhip_adj_create": (C.c_int, [C.POINTER(ModelDesc), C.POINTER(AdjOptions), C.POINTER(C.c_void_p)]),
    "edenhip_adj_sizes": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                    C.POINTER(C.c_int64)]),
    "edenhip_adj_forward": (C.c_int, [C.c_void_p, _P, _P, _P]),
    "edenhip_adj_forward_w": (C.c_int, [C.c_void_p, _P, _P, _P, _P]),
    "edenhip_adj_backward_w": (C.c_int, [C.c_void_p, _P, _P, _P, _P]),
    "edenhip_adj_set_events": (C.c_int, [C.c_void_p, _P, _P, C.c_int64]),
    "edenhip_adj_n_synapses": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64)]),
    "edenhip_adj_backward": (C.c_int, [C.c_void_p, _P, _P, _P]),
    "edenhip_adj_timing": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "edenhip_adj_last_error": (C.c_char_p, [C.c_void_p]),
    "edenhip_adj_destroy": (None, [C.c_void_p]),
    "edenhip_write_trajectory": (C.c_int, [C.c_char_p, C.c_int, _P, _P, C.c_int64, C.c_int64, _P]),
    "edenhip_write_events": (C.c_int, [C.c_char_p, C.c_int, C.c_int32, _P, _P, _P, C.c_int64, _P, C.c_int64]),
    "edenhip_format_number": (C.c_int, [C.c_double, C.c_char_p, C.c_int32]),
    "edenhip_diag_divide": (C.c_int, [C.c_uint64, C.c_int64, _P, C.c_int32, C.POINTER(C.c_int64)]),
    "edenhip_network_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                        C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "edenhip_launch_shape": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "edenhip_last_error": (C.c_char_p, [C.c_void_p]),
    "edenhip_destroy": (None, [C.c_void_p]),
    "edenhip_cable_setup": (C.c_int, [C.c_int32, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "edenhip_cable_setup_slices": (C.c_int, [C.c_int32, _P, _P, _P, _P, C.c_int32, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
}

_lib = None


def load_library(path=LIB_PATH):
    """Load libedenhip.so; raise (never fall back) when it is absent."""
    global _lib
    if _lib is not None and path == LIB_PATH:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"eden_amd: HIP library not built ({path} missing); run `python -m eden_amd.build`")
    lib = C.CDLL(path)
    for name, (res, args) in EXPORTS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path == LIB_PATH:
        _lib = lib
    return lib


def ptr(a):
    """Data pointer of a numpy array (or None)."""
    if a is None:
        return None
    return C.c_void_p(a.ctypes.data)


def cable_setup(parent, proximal, distal, cm, ra):
    """Per-compartment (area um^2, C pF, R-to-parent MOhm, 1/RC 1/us) of one cell with one
    compartment per segment; see include/edenhip_setup.h."""
    lib = load_library()
    n = len(parent)
    parent = np.ascontiguousarray(parent, dtype=np.int32)
    proximal = np.ascontiguousarray(proximal, dtype=np.float32).reshape(n, 4)
    distal = np.ascontiguousarray(distal, dtype=np.float32).reshape(n, 4)
    cm = np.ascontiguousarray(np.broadcast_to(np.asarray(cm, np.float32), (n,)))
    ra = np.ascontiguousarray(np.broadcast_to(np.asarray(ra, np.float32), (n,)))
    area = np.empty(n, np.float32)
    cap = np.empty(n, np.float32)
    rax = np.empty(n, np.float32)
    inv = np.empty(n, np.float32)
    rc = lib.edenhip_cable_setup(n, ptr(parent), ptr(proximal), ptr(distal), ptr(cm), ptr(ra),
                                 ptr(area), ptr(cap), ptr(rax), ptr(inv))
    if rc != 0:
        raise ValueError(f"edenhip_cable_setup failed ({rc}): malformed tree or zero axial resistance")
    return area, cap, rax, inv


def cable_setup_slices(proximal, distal, cm, ra, parent, slice_begin, slice_seg, slice_from, slice_to):
    """Per-compartment (area, C, R to parent, 1/RC) of one cell whose compartments are made of
    segment slices; see include/edenhip_setup.h."""
    lib = load_library()
    ns = len(cm)
    n = len(parent)
    proximal = np.ascontiguousarray(proximal, dtype=np.float32).reshape(ns, 4)
    distal = np.ascontiguousarray(distal, dtype=np.float32).reshape(ns, 4)
    cm = np.ascontiguousarray(cm, dtype=np.float32)
    ra = np.ascontiguousarray(ra, dtype=np.float32)
    parent = np.ascontiguousarray(parent, dtype=np.int32)
    slice_begin = np.ascontiguousarray(slice_begin, dtype=np.int32)
    slice_seg = np.ascontiguousarray(slice_seg, dtype=np.int32)
    slice_from = np.ascontiguousarray(slice_from, dtype=np.float32)
    slice_to = np.ascontiguousarray(slice_to, dtype=np.float32)
    out = [np.empty(n, np.float32) for _ in range(4)]
    rc = lib.edenhip_cable_setup_slices(ns, ptr(proximal), ptr(distal), ptr(cm), ptr(ra), n, ptr(parent),
                                        ptr(slice_begin), ptr(slice_seg), ptr(slice_from), ptr(slice_to),
                                        *[ptr(o) for o in out])
    if rc != 0:
        raise ValueError(f"edenhip_cable_setup_slices failed ({rc}): malformed compartments or zero axial resistance")
    return tuple(out)
