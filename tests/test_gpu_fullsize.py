"""Parity at the bench workloads' full sizes.

config 4 (4,096 cells, 1.57M compartments, 20,001 rows): cells are independent, so a cell's
trace must not depend on the batch it runs in. A dozen cells of every workgroup size are
taken from the full run and compared, bit for bit, with the oracle run on those cells alone
and with a GPU run of those cells alone.

config 2 in full (512 cells, 10,001 rows): every cell's trace and spike raster against the
oracle on the same model, exactly.

config 3 at full scale (100,000 cells, 10.1M synapses), the first 15 ms: recorded voltages,
spike raster and every synaptic delivery (step, synapse instance) against the oracle on the
same model, exactly.
"""
import sys

import numpy as np
import pytest

from eden_amd import configs
from eden_amd.model import Simulation
from oracle import oracle

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_config4_full_run_cells_equal_oracle_and_standalone():
    m = configs.config4()
    _say(f"[fullsize] {m.name}: {m.n_cells} cells, {m.n_comps} compartments, {m.total_rows()} rows")
    with Simulation(m) as sim:
        tr = sim.run_all()
        ev_col, ev_row = sim.read_events()
    sizes = m.comps_per_cell()
    order = np.argsort(sizes, kind="stable")
    # three cells from each workgroup size class and the extremes
    pick = []
    for lo, hi in ((2, 256), (257, 512), (513, 1024), (1025, 2047)):
        cand = order[(sizes[order] >= lo) & (sizes[order] <= hi)]
        if len(cand):
            pick += [int(cand[0]), int(cand[len(cand) // 2]), int(cand[-1])]
    sub, rc, ecols = m.subset(pick)
    _say(f"[fullsize] subset of {len(pick)} cells, sizes {sorted(sizes[pick].tolist())}")
    ref, oc, orow, _ = oracle.run(sub)
    assert ref.shape == (tr.shape[0], len(rc))
    assert np.array_equal(tr[:, rc], ref), float(np.abs(tr[:, rc] - ref).max())
    # events of the chosen cells, in the subset's column numbering
    col_new = {int(c): i for i, c in enumerate(ecols)}
    keep = np.isin(ev_col, ecols)
    g_ev = sorted(zip(ev_row[keep].tolist(), [col_new[int(c)] for c in ev_col[keep]]))
    o_ev = sorted(zip(orow.tolist(), oc.tolist()))
    assert g_ev == o_ev and len(o_ev) > 0
    with Simulation(sub) as s2:
        alone = s2.run_all()
    assert np.array_equal(alone, ref)
    _say(f"[fullsize] {len(pick)} cells x {tr.shape[0]} rows identical; {len(o_ev)} spikes")


def test_config2_full_equals_oracle():
    m = configs.config2()
    _say(f"[fullsize] {m.name}: {m.n_cells} cells, {m.n_comps} compartments, {m.total_rows()} rows")
    with Simulation(m) as sim:
        tr = sim.run_all()
        ev = sim.read_events()
    ref, ec, er, _ = oracle.run(m)
    _say(f"[fullsize] config 2: spikes gpu {len(ev[1])} oracle {len(er)}")
    assert len(er) > 0
    assert np.array_equal(tr, ref), float(np.abs(tr - ref).max())
    assert np.array_equal(ev[0], ec) and np.array_equal(ev[1], er)


def test_config3_full_scale_first_15ms_bit_exact():
    m = configs.spiking_network(t_final_ms=15.0)
    _say(f"[fullsize] {m.name}: {m.n_cells} cells, {m.n_synapses} synapses, {m.total_rows()} rows")
    with Simulation(m, log_synapse_events=20_000_000) as sim:
        tr = sim.run_all()
        ev = sim.read_events()
        g_step, g_syn = sim.read_synapse_events()
        w, nwin, rerun, deliv = sim.network_stats()
    (ref, ec, er, _), (o_step, o_syn) = oracle.run_with_synapse_log(m)
    _say(f"[fullsize] window {w} steps x {nwin} ({rerun} rerun); deliveries gpu {len(g_step)} oracle {len(o_step)}")
    assert len(o_step) > 1000
    assert np.array_equal(tr, ref)
    assert np.array_equal(ev[0], ec) and np.array_equal(ev[1], er)
    assert np.array_equal(g_step, o_step) and np.array_equal(g_syn, o_syn)
