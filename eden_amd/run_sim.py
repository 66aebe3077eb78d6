"""Run a LEMS / NeuroML file on the GPU and write its output files, with EDEN's Python API.

`run_lems(path)` is the GPU counterpart of `eden nml <file>` (eden/Eden.cpp:10379-11968): it
reads the model (eden_amd.neuroml), runs it on the GPU in chunks of rows and writes every
OutputFile and EventOutputFile in EDEN's text format, then reports EDEN's
`Config: Setup: Run:` timings. Recorded rows are formatted on a writer thread while the GPU
runs the next chunk: the GPU and the text formatting overlap, where EDEN formats rows on
its master thread inside the time loop (eden/Eden.cpp:11549-11638).

`runEden(...)` mirrors the reference's Python wrapper
(testing/python_package/eden_simulator/run_sim.py:22-106): same arguments, plus `backend`;
it returns the same OrderedDicts, reloaded from the files the run wrote as the reference's
`reload_saved_data` does (run_sim.py:110-252).
"""
import os
import queue
import threading
import time
from collections import OrderedDict
from datetime import datetime

import numpy as np

from . import neuroml, output


def run_lems(path, rng_seed=None, chunk_rows=None, verbose=False, out_dir=None, **sim_options):
    """Run the simulation of a LEMS / NeuroML file on the GPU and write its output files.
    Returns a dict of timings (config_s, setup_s, run_s) and counts (rows, events)."""
    from .model import Simulation
    t0 = time.time()
    model, outs = neuroml.read(path, rng_seed=rng_seed)
    t1 = time.time()
    rows_total = model.total_rows()
    if chunk_rows is None:
        # chunks of about 64 MiB of recorded values, so that formatting overlaps the GPU run
        chunk_rows = max(1, min(rows_total, (64 << 20) // max(4, 4 * len(model.record_comp))))
    t_us = output.row_times_us(model.dt, rows_total + 1)
    where = (lambda f: os.path.join(out_dir, f)) if out_dir else (lambda f: f)
    traj = [(where(f.file_name), [c[2] for c in f.columns]) for f in outs.trajectories]
    scales = output.column_scales(model)
    with Simulation(model, chunk_steps=chunk_rows, **sim_options) as sim:
        t2 = time.time()
        # writer thread: formats and writes chunk k while the GPU runs chunk k + 1 (the
        # library's ctypes calls release the interpreter lock)
        q = queue.Queue(maxsize=2)
        err = []

        def writer():
            while True:
                item = q.get()
                if item is None:
                    return
                if err:
                    continue
                tr, row0, first = item
                try:
                    for fname, cols in traj:
                        output.write_trajectory(fname, t_us[row0:row0 + len(tr)], tr[:, cols], scale=scales[cols],
                                                append=not first)
                except Exception as e:   # reported after the run
                    err.append(e)

        th = threading.Thread(target=writer, daemon=True)
        th.start()
        done, first = 0, True
        try:
            while done < rows_total:
                n = sim.run(-1)
                if n <= 0:
                    break
                if traj:
                    tr, row0 = sim.read_trace()
                    q.put((tr.copy(), row0, first))
                first = False
                done += n
        finally:
            q.put(None)
            th.join()
        if err:
            raise err[0]
        n_ev = 0
        if outs.events:
            col, row = sim.read_events()
            n_ev = len(row)
            for f in outs.events:
                cols = [s[2] for s in f.selections]
                ids = np.array([s[0] for s in f.selections], np.int64)
                keep = np.isin(col, cols)
                remap = {c: i for i, c in enumerate(cols)}
                fc = np.array([remap[int(c)] for c in col[keep]], np.int64)
                fr = row[keep]
                output.write_events(where(f.file_name), t_us, fr, fc, ids, f.format)
        t3 = time.time()
    res = dict(config_s=t1 - t0, setup_s=t2 - t1, run_s=t3 - t2, rows=done, events=n_ev, model=model, outputs=outs)
    if verbose:
        print(f"{model.n_cells} cells, {model.n_comps} compartments, {model.n_synapses} synapses, {done} rows")
    return res


def runEden(example_lems_file, reload_events=False, threads=None, extra_cmdline_args=None, executable_path=None,
            verbose=False, full_cmdline=None, backend="hip"):
    """Run a LEMS / NeuroML2 file and reload its recorded outputs, as the reference's
    runEden does. backend "hip" runs the GPU path in this process; the arguments that
    configure EDEN's CPU binary (threads, executable_path, full_cmdline) are accepted and do
    not apply to it. extra_cmdline_args understands EDEN's `rng_seed N`."""
    if backend != "hip":
        raise ValueError(f"backend {backend!r}: this package runs the GPU path only ('hip')")
    rng_seed = None
    args = list(extra_cmdline_args or [])
    if "rng_seed" in args:
        i = args.index("rng_seed")
        rng_seed = int(args[i + 1])
    t_run = datetime.now()
    tic = time.time()
    try:
        run_lems(example_lems_file, rng_seed=rng_seed, verbose=verbose)
    except neuroml.NeuroMLError as e:
        raise RuntimeError(f"Could not run simulation: {e}") from e
    if verbose:
        print("Ran EDEN (GPU) in %.2f seconds" % (time.time() - tic))
    results = reload_saved_data(example_lems_file, reload_events=reload_events, t_run=t_run)
    if reload_events:
        traje, event = results
        return OrderedDict(sorted(traje.items())), OrderedDict(sorted(event.items()))
    return OrderedDict(sorted(results.items()))


def _find(name, base_dir, lems_dir):
    for d in (base_dir, lems_dir, os.getcwd()):
        p = os.path.join(d, name)
        if os.path.isfile(p):
            return p
    raise OSError(f"Could not find simulation output file {name}")


def reload_saved_data(lems_file_name, base_dir=".", t_run=datetime(1900, 1, 1), reload_events=False):
    """The recorded outputs of a run, read back from its files as the reference's
    reload_saved_data reads them: traces keyed by OutputColumn quantity (and 't'), events
    keyed by EventSelection select, times in seconds."""
    import xml.etree.ElementTree as ET
    lems_dir = os.path.dirname(os.path.realpath(lems_file_name))
    root = ET.parse(lems_file_name).getroot()
    strip = lambda t: t.rsplit("}", 1)[-1]
    sim = next((e for e in root.iter() if strip(e.tag) == "Simulation"), None)
    if sim is None:
        sim = next((e for e in root.iter() if strip(e.tag) == "Component" and e.get("type") == "Simulation"), None)
    traces, events = OrderedDict(), OrderedDict()
    if reload_events:
        for of in (e for e in sim if strip(e.tag) == "EventOutputFile"):
            fname = _find(of.get("fileName"), base_dir, lems_dir)
            fmt = of.get("format")
            sel = {}
            for col in (e for e in of if strip(e.tag) == "EventSelection"):
                events[col.get("select")] = []
                sel[int(col.get("id"))] = col.get("select")
            with open(fname) as f:
                for line in f:
                    v = line.split()
                    if fmt == "TIME_ID":
                        t, i = float(v[0]), int(v[1])
                    elif fmt == "ID_TIME":
                        i, t = int(v[0]), float(v[1])
                    else:
                        raise ValueError(fmt)
                    events[sel[i]].append(t)
    for of in (e for e in sim if strip(e.tag) == "OutputFile"):
        traces["t"] = []
        fname = _find(of.get("fileName"), base_dir, lems_dir)
        if datetime.fromtimestamp(os.path.getmtime(fname)) < t_run:
            raise Exception(f"Expected output file {fname} has not been modified since the run started")
        cols = ["t"] + [c.get("quantity") for c in of if strip(c.tag) == "OutputColumn"]
        with open(fname) as f:
            data = np.array(f.read().split(), float).reshape((-1, len(cols)), order="C")
        for i, c in enumerate(cols):
            traces[c] = data[:, i]
    if reload_events:
        return traces, events
    return traces
