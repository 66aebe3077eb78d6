"""bench.py's contract on the host: one GPU only, and the algorithmic bytes behind the
roofline (per unit: 56 B per HH tree compartment-update, 24 B passive, 40 B per
one-compartment HH cell, 4 B per recorded tree column-row, 32 B per synapse-step)."""
import os
import subprocess
import sys

import numpy as np

from eden_amd import configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_algorithmic_bytes_of_point_cells():
    m = configs.hh_population(n_cells=10, t_final_ms=1.0)
    rows = m.total_rows()
    assert bench.algorithmic_bytes(m, rows) == rows * 10 * 40


def test_algorithmic_bytes_of_trees():
    sizes = [12, 30]
    m = configs.random_trees(len(sizes), sizes=sizes, t_final_ms=1.0, seed=3)
    rows = m.total_rows()
    # HH on each soma, passive elsewhere, somatic V recorded
    expect = rows * (2 * 56 + (sum(sizes) - 2) * 24 + 2 * 4)
    assert bench.algorithmic_bytes(m, rows) == expect
    # restricted to the second cell
    assert bench.algorithmic_bytes(m, rows, cells=[1]) == rows * (56 + 29 * 24 + 4)


def test_algorithmic_bytes_count_synapse_steps():
    m = configs.spiking_network(n_cells=50, in_degree=5, t_final_ms=1.0)
    rows = m.total_rows()
    n_syn = int(np.diff(m.comp_syn_begin).sum() + np.diff(m.comp_poisson_begin).sum())
    assert n_syn > 0
    assert bench.algorithmic_bytes(m, rows) == rows * (m.n_comps * 40 + n_syn * 32)
