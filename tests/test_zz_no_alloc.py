"""edenhip_run_static, the entry point that allocates nothing, against edenhip_run through
the C-ABI: identical traces and spike events (bit for bit) on two calls in a row, the second
on a smaller model, and on a model with point cells and every tree workgroup size; models it
does not take are rejected with an error."""
import sys

import numpy as np
import pytest

from eden_amd import configs
from eden_amd.model import EdenHipError, Simulation

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _runs(model, static):
    with Simulation(model) as sim:
        parts = []
        while True:
            n = sim.run_static(-1) if static else sim.run(-1)
            if n == 0:
                break
            parts.append(sim.read_trace()[0].copy())
            if sum(len(p) for p in parts) >= sim.total_rows:
                break
        ev = sim.read_events()
    return np.concatenate(parts), ev


def test_static_entry_point_equals_run_on_two_calls_second_smaller():
    big = configs.random_trees(6, sizes=[30, 200, 300, 600, 1100, 2000], t_final_ms=15.0, seed=21)
    small = configs.random_trees(3, sizes=[25, 140, 400], t_final_ms=10.0, seed=22)
    for m in (big, small):
        _say(f"[no_alloc] {m.name}: {m.n_cells} cells, {m.n_comps} compartments")
        tr_s, ev_s = _runs(m, True)
        tr_r, ev_r = _runs(m, False)
        assert np.array_equal(tr_s, tr_r)
        assert np.array_equal(ev_s[0], ev_r[0]) and np.array_equal(ev_s[1], ev_r[1])


def test_static_entry_point_chunks_point_cells_and_networks():
    m = configs.mixed_network(tree_sizes=(40, 300), n_point=10, n_conn=80, t_final_ms=20.0, seed=4)
    _say(f"[no_alloc] {m.name}: per-step network with tree and point cells")
    tr_s, ev_s = _runs(m, True)
    tr_r, ev_r = _runs(m, False)
    assert np.array_equal(tr_s, tr_r)
    assert np.array_equal(ev_s[1], ev_r[1])
    p = configs.hh_population(n_cells=50, t_final_ms=10.0)
    with Simulation(p, chunk_steps=100) as a, Simulation(p, chunk_steps=100) as b:
        while True:
            na, nb = a.run_static(-1), b.run(-1)
            assert na == nb
            if na == 0:
                break
            assert np.array_equal(a.read_trace()[0], b.read_trace()[0])


def test_static_entry_point_rejects_what_it_cannot_run():
    net = configs.spiking_network(n_cells=300, in_degree=10, t_final_ms=5.0)
    with Simulation(net) as sim:
        with pytest.raises(EdenHipError, match="windowed network path"):
            sim.run_static(-1)
    m = configs.random_trees(2, sizes=[30, 40], t_final_ms=5.0)
    with Simulation(m, profile=True) as sim:
        with pytest.raises(EdenHipError, match="profile"):
            sim.run_static(-1)
