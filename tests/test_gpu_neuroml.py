"""NeuroML / LEMS files through the GPU path: read (eden_amd.neuroml), run on the GPU, output
files written (eden_amd.run_sim) and compared byte for byte with the text the reference's
printer gives for the oracle's run of the same instantiated model (tests/eden_text.py).

* the reference's single-compartment HH example (tests/golden/neuroml), by run_lems, by
  runEden (the reference's return shape) and by the command line;
* the reference's cable discretisation model (tests/golden/neuroml), whose cable groups the
  reader cuts into compartments, and its inhomogeneous-parameters model (channel densities
  that vary with path length);
* synthetic workloads written as LEMS (config 1's construction, a tree-cell / point-cell
  network with synapses on dendrites): the GPU run of the file equals the GPU run of the
  configs model it was written from, and the oracle's run of the model read back.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from eden_amd import configs, neuroml, nml_export, output
from eden_amd.model import Simulation
from eden_amd.run_sim import run_lems, runEden
from eden_text import assert_same_text, events_text, trajectory_text
from oracle import oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
HH = os.path.join(HERE, "golden", "neuroml", "LEMS_HHCellExample.xml")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _expected_files(model, outs):
    tr, ec, er, _ = oracle.run(model)
    t_us = output.row_times_us(model.dt, model.total_rows() + 1)
    files = {}
    scales = output.column_scales(model)
    for f in outs.trajectories:
        cols = [c[2] for c in f.columns]
        files[f.file_name] = trajectory_text(t_us, tr[:, cols], scales[cols])
    for f in outs.events:
        cols = [s[2] for s in f.selections]
        keep = np.isin(ec, cols)
        remap = {c: i for i, c in enumerate(cols)}
        fc = np.array([remap[int(c)] for c in ec[keep]], np.int64)
        files[f.file_name] = events_text(t_us, er[keep], fc, [s[0] for s in f.selections], f.format)
    return files, len(er)


def test_reference_hh_example_files(tmp_path):
    res = run_lems(HH, out_dir=str(tmp_path), chunk_rows=7000)
    m, outs = res["model"], res["outputs"]
    _say(f"[neuroml] HH example: {res['rows']} rows, {res['events']} spikes, run {res['run_s']:.3f} s")
    want, n_spikes = _expected_files(m, outs)
    assert n_spikes > 3
    for name, text in want.items():
        assert_same_text((tmp_path / name).read_text(), text)


def test_reference_cable_discretisation_files(tmp_path):
    """The reference's EdenTest_CableDiscretisation model (cable groups cut into compartments
    by the reader): output file byte-equal to the oracle's run of the same model; the single
    segment cut in five, the five segments in a five-compartment cable and the five free
    segments give identical voltages (their compartments are identical), which respond to the
    current steps; the merged cable's three recording points are one compartment."""
    path = os.path.join(HERE, "golden", "neuroml", "LEMS_EdenTest_CableDiscretisation.xml")
    res = run_lems(path, out_dir=str(tmp_path))
    m, outs = res["model"], res["outputs"]
    _say(f"[neuroml] cable discretisation: {m.n_comps} compartments, {res['rows']} rows, run {res['run_s']:.3f} s")
    want, _ = _expected_files(m, outs)
    for name, text in want.items():
        assert_same_text((tmp_path / name).read_text(), text)
    with Simulation(m) as sim:
        tr = sim.run_all()
    assert np.array_equal(tr[:, 0], tr[:, 1]) and np.array_equal(tr[:, 0], tr[:, 2])
    sub, expl, eff = tr[:, 3], tr[:, 4:9], tr[:, 9:14]
    assert np.array_equal(expl, eff)
    assert np.array_equal(sub, expl[:, 2])
    assert float(np.ptp(expl[:, 0])) > 1.0


def test_reference_inhomogeneous_parameters_files(tmp_path):
    """The reference's EdenTest_InhomogeneousParameters model (channel densities from
    path-length expressions): output file byte-equal to the oracle's run; the cell whose
    non-uniform density is a constant gives the uniform cell's voltages exactly, and the
    cells with decaying densities differ from it."""
    path = os.path.join(HERE, "golden", "neuroml", "LEMS_EdenTest_InhomogeneousParameters.xml")
    res = run_lems(path, out_dir=str(tmp_path))
    m, outs = res["model"], res["outputs"]
    _say(f"[neuroml] inhomogeneous parameters: {m.n_comps} compartments, {res['rows']} rows")
    want, _ = _expected_files(m, outs)
    for name, text in want.items():
        assert_same_text((tmp_path / name).read_text(), text)
    with Simulation(m) as sim:
        tr = sim.run_all()
    assert np.array_equal(tr[:, 0:3], tr[:, 3:6])
    assert float(np.abs(tr[:, 6:9] - tr[:, 0:3]).max()) > 1.0


@pytest.mark.parametrize("name", ["LEMS_EdenTest2.xml", "LEMS_EdenTest1.xml"])
def test_reference_gate_recording_files(tmp_path, name):
    """The reference's EdenTest1 (a network of multi-compartment cells recording the m gate
    of a dendrite compartment) and EdenTest2 (one HH cell recording m, h and n next to V):
    output files byte-equal to the oracle's run, the GPU trace equal to the oracle's."""
    path = os.path.join(HERE, "golden", "neuroml", name)
    res = run_lems(path, out_dir=str(tmp_path))
    m, outs = res["model"], res["outputs"]
    _say(f"[neuroml] {name}: {m.n_comps} compartments, gate columns {list(m.record_gate)}, {res['rows']} rows")
    want, n_ev = _expected_files(m, outs)
    for fname, text in want.items():
        assert_same_text((tmp_path / fname).read_text(), text)
    with Simulation(m) as sim:
        tr = sim.run_all()
    tr_r, _, _, _ = oracle.run(m)
    assert np.array_equal(tr, tr_r)
    assert (m.record_gate >= 0).any()


def test_runeden_return_shape(tmp_path, monkeypatch):
    import shutil
    for f in ("LEMS_HHCellExample.xml", "NML2_SingleCompHHCell.nml"):
        shutil.copy(os.path.join(HERE, "golden", "neuroml", f), tmp_path / f)
    monkeypatch.chdir(tmp_path)
    traces, events = runEden("LEMS_HHCellExample.xml", reload_events=True)
    assert list(traces) == ["hhpop/0/hhcell/0/v", "t"]
    assert list(events) == ["hhpop/0/hhcell"]
    m, _ = neuroml.read("LEMS_HHCellExample.xml")
    tr, ec, er, _ = oracle.run(m)
    np.testing.assert_allclose(traces["hhpop/0/hhcell/0/v"], tr[:, 0].astype(np.float64) * 1e-3, rtol=1e-7)
    np.testing.assert_allclose(traces["t"], np.arange(len(tr)) * 1e-5, rtol=1e-7, atol=1e-12)
    assert len(events["hhpop/0/hhcell"]) == len(er)
    only_traces = runEden("LEMS_HHCellExample.xml")
    assert list(only_traces) == ["hhpop/0/hhcell/0/v", "t"]


def test_command_line(tmp_path):
    import shutil
    for f in ("LEMS_HHCellExample.xml", "NML2_SingleCompHHCell.nml"):
        shutil.copy(os.path.join(HERE, "golden", "neuroml", f), tmp_path / f)
    root = os.path.dirname(HERE)
    env = dict(os.environ, PYTHONPATH=root + os.pathsep + os.environ.get("PYTHONPATH", ""))
    p = subprocess.run([sys.executable, "-m", "eden_amd", "nml", "LEMS_HHCellExample.xml", "gpu"], cwd=tmp_path,
                       env=env, capture_output=True, text=True, timeout=300)
    _say(f"[neuroml] cli: {p.stdout.strip().splitlines()[-2:]}")
    assert p.returncode == 0, p.stdout + p.stderr
    assert "Config: " in p.stdout and " Run: " in p.stdout
    m, outs = neuroml.read(str(tmp_path / "LEMS_HHCellExample.xml"))
    want, _ = _expected_files(m, outs)
    for name, text in want.items():
        assert_same_text((tmp_path / name).read_text(), text)


@pytest.mark.parametrize("which", ["config1", "mixed"])
def test_exported_workload_runs_as_its_model(tmp_path, which):
    m = configs.hh_population(n_cells=100, t_final_ms=60.0) if which == "config1" else \
        configs.mixed_network(tree_sizes=(40, 90, 150, 300), n_point=16, n_conn=250, t_final_ms=60.0, seed=21)
    path = nml_export.write_lems(m, str(tmp_path / f"{which}.xml"), m.morphologies)
    res = run_lems(path, out_dir=str(tmp_path))
    r, outs = res["model"], res["outputs"]
    _say(f"[neuroml] {which}: {r.n_cells} cells, {r.n_comps} comps, {res['rows']} rows, {res['events']} spikes")
    want, n_spikes = _expected_files(r, outs)
    assert n_spikes > 0
    for name, text in want.items():
        assert_same_text((tmp_path / name).read_text(), text)
    with Simulation(m) as sim:
        direct = sim.run_all()
    got = np.array(open(tmp_path / outs.trajectories[0].file_name).read().split(), float)
    got = got.reshape(len(direct), -1)[:, 1:]
    np.testing.assert_array_equal(got, np.array([float("%+16.8g" % (float(v) * 1e-3)) for v in direct.ravel()]).reshape(got.shape))
