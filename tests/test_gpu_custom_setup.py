"""Per-instance parameter settings (EdenCustomSetup) on the GPU: generic synapse instances and
point neurons read their own parameter vectors (edenhip_lems_synapses.instance_param,
edenhip_lems_cells.cell_param); native inputs carry theirs in the input tables.

* The reference's EdenTest_Extension_CustomSetup model through the reader, the library and the
  output files: bit for bit the oracle's, and the reference's own criteria
  (testing/validation_tests/validation_tests.py:29-41, each customised trace equal to its
  distinct counterpart) exact on the GPU's traces.
* Native pulse inputs and an expTwoSynapse connection set per instance, in the plain and
  graph launch modes, bit for bit the oracle's.
"""
import os
import shutil
import sys

import numpy as np
import pytest

from eden_amd import output
from eden_amd.model import Simulation
from eden_amd.run_sim import run_lems
from eden_text import assert_same_text, trajectory_text
from oracle import oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "neuroml")
PAIRS = [(5, 0), (6, 1), (7, 2), (8, 3), (9, 4), (10, 12), (11, 13)]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_reference_custom_setup_files(tmp_path):
    for f in ("EdenTest_Extension_CustomSetup.nml", "EdenTest_Extension_CustomSetup.txt"):
        shutil.copy(os.path.join(GOLD, f), tmp_path / f)
    res = run_lems(str(tmp_path / "EdenTest_Extension_CustomSetup.nml"), out_dir=str(tmp_path))
    m, outs = res["model"], res["outputs"]
    _say(f"[setup] CustomSetup: {m.n_cells} cells, {len(m.cell_param_vecs)} cell and "
         f"{len(m.syn_param_vecs)} synapse vectors, {res['rows']} rows")
    tr_r, ec, er, _ = oracle.run(m)
    t_us = output.row_times_us(m.dt, m.total_rows() + 1)
    scales = output.column_scales(m)
    for f in outs.trajectories:
        cols = [c[2] for c in f.columns]
        assert_same_text((tmp_path / f.file_name).read_text(), trajectory_text(t_us, tr_r[:, cols], scales[cols]))
    with Simulation(m) as sim:
        tr = sim.run_all()
    assert np.array_equal(tr, tr_r)
    for a, b in PAIRS:
        assert np.array_equal(tr[:, a], tr[:, b]), (a, b)


@pytest.mark.parametrize("opts", [{}, dict(use_graph=True)])
def test_native_inputs_and_synapse_settings_bit_exact(tmp_path, opts):
    from test_custom_setup_cpu import _native_model
    m, _ = _native_model(tmp_path, "set input L 7,3 delay multi ms\nvalues 3 8\n"
                                    "set input L all duration multi ms\nvalues 20 4\n"
                                    "set input L 7 amplitude 20 pA\nset input L 3 amplitude -1e-11 A\n"
                                    "set synapse A all post gbase 2 nS\nset synapse A 0 post erev -20 mV\n"
                                    "set synapse A 0 post tauRise 1 ms\nset synapse A all post tauDecay 6 ms\n")
    _say(f"[setup] native model, opts {opts}")
    tr_r, ec, er, _ = oracle.run(m)
    with Simulation(m, **opts) as sim:
        tr = sim.run_all()
    assert np.array_equal(tr, tr_r), float(np.nanmax(np.abs(tr - tr_r)))
    assert np.array_equal(tr[:, 2], tr[:, 0]) and np.array_equal(tr[:, 3], tr[:, 1])
    assert np.array_equal(tr[:, 4], tr[:, 5])
