"""EDEN's recorded-output files for the rows and events of a run (include/edenhip_output.h).

OutputFile: one line per recorded row, the time in seconds then each column's value times
its scale (1e-3 turns the engine's mV into the volts a classic OutputFile holds), every
number printed "%+16.8g" in a 16-character field and the columns joined by one space
(FixedWidthNumberPrinter(16, '\t', 0), eden/Eden.cpp:144-181, 10570-10571, 10843-10884).
EventOutputFile: TIME_ID / ID_TIME lines per event, or one ascii_v0 line per row with
events (eden/Eden.cpp:10886-10940).
"""
import numpy as np

from . import _native as N

FORMATS = {"TIME_ID": 0, "ID_TIME": 1, "ascii_v0": 2}
MV_TO_V = 1e-3


def format_number(v):
    buf = (N.C.c_char * 32)()
    rc = N.load_library().edenhip_format_number(float(v), buf, 32)
    if rc != 0:
        raise ValueError("format_number failed")
    return buf.value.decode()


def write_trajectory(path, time_us, trace, scale=MV_TO_V, append=False):
    """Write rows of `trace` (rows x cols float32, engine units) stamped with time_us[row]."""
    trace = np.ascontiguousarray(trace, np.float32)
    if trace.ndim != 2:
        raise ValueError("trace must be 2-D")
    time_us = np.ascontiguousarray(time_us, np.float64)
    if len(time_us) < trace.shape[0]:
        raise ValueError("fewer timestamps than rows")
    sc = np.ascontiguousarray(np.broadcast_to(np.asarray(scale, np.float64), (trace.shape[1],)))
    rc = N.load_library().edenhip_write_trajectory(str(path).encode(), int(append), N.ptr(time_us), N.ptr(trace),
                                                   trace.shape[0], trace.shape[1], N.ptr(sc))
    if rc != 0:
        raise OSError(f"write_trajectory({path}) failed ({rc})")


def column_scales(model):
    """Per recorded column, the factor from engine units to the file's SI units: 1e-3 for a
    membrane voltage (mV to V), 1 for a gate variable (dimensionless, LEMS_Pure in
    eden/NeuroML.cpp:1016), and for a point neuron's state variable the native unit of its
    dimension (nA: 1e-9, us: 1e-6, ...)."""
    sc = np.full(len(model.record_comp), MV_TO_V)
    if model.record_gate is not None:
        sc[np.asarray(model.record_gate) >= 0] = 1.0
    LC = getattr(model, "lems_cells", None)
    if LC is not None:
        cell_of = np.searchsorted(model.cell_comp_begin, model.record_comp, side="right") - 1
        for r, c in enumerate(cell_of):
            ty = int(model.cell_lems_type[c])
            if ty < 0:
                continue
            j = -1 if model.record_gate is None else int(model.record_gate[r])
            k = LC.v_state[ty] if j < 0 else j
            sc[r] = 10.0 ** LC.state_pow10[ty][k]
    return sc


def write_events(path, time_us, ev_row, ev_col, ids, fmt="TIME_ID", append=False):
    """Write events (row, column), sorted by row then column; column c is printed as ids[c]."""
    ev_row = np.ascontiguousarray(ev_row, np.int64)
    ev_col = np.ascontiguousarray(ev_col, np.int64)
    ids = np.ascontiguousarray(ids, np.int64)
    time_us = np.ascontiguousarray(time_us, np.float64)
    rc = N.load_library().edenhip_write_events(str(path).encode(), int(append), FORMATS[fmt], N.ptr(time_us),
                                               N.ptr(ev_row), N.ptr(ev_col), len(ev_row), N.ptr(ids), len(ids))
    if rc != 0:
        raise OSError(f"write_events({path}) failed ({rc})")


def row_times_us(dt_us, rows):
    """Row timestamps as the engine accumulates them: time += dt in double, dt in float."""
    t = np.empty(rows, np.float64)
    acc = 0.0
    d = float(np.float32(dt_us))
    for r in range(rows):
        t[r] = acc
        acc += d
    return t


def run_to_files(model, trajectory_path=None, event_path=None, event_format="TIME_ID", event_ids=None,
                 chunk_rows=0, **sim_options):
    """Run `model` on the GPU and stream its recorded rows to an OutputFile (all recorded
    columns, mV -> V) chunk by chunk, then its spike events to an EventOutputFile (event
    column c printed as event_ids[c], default c). Returns (rows, events)."""
    from .model import Simulation
    rows_total = model.total_rows()
    t_us = row_times_us(model.dt, rows_total + 1)
    n_ev = 0
    with Simulation(model, chunk_steps=chunk_rows, **sim_options) as sim:
        done = 0
        first = True
        while done < rows_total:
            n = sim.run(-1)
            if n <= 0:
                break
            if trajectory_path is not None:
                tr, row0 = sim.read_trace()
                write_trajectory(trajectory_path, t_us[row0:row0 + len(tr)], tr, append=not first)
            first = False
            done += n
        if event_path is not None:
            col, row = sim.read_events()
            ids = np.arange(len(model.event_comp)) if event_ids is None else np.asarray(event_ids)
            write_events(event_path, t_us, row, col, ids, event_format)
            n_ev = len(row)
    return done, n_ev
