"""The launch-shape autotuner (eden_amd/autotune.py) on a small batched model: it returns
limits from its candidate set that are no slower than the size rule in its own timings, and
a run with them gives the same trace as the size rule (the schedules differ, the operations
do not)."""
import sys

import numpy as np
import pytest

from eden_amd import autotune, configs
from eden_amd.model import Simulation

pytestmark = pytest.mark.gpu


def test_autotune_small_batch():
    sizes = [40, 130, 200, 260, 420, 600, 800, 1100]
    m = configs.random_trees(len(sizes), sizes=sizes, t_final_ms=5.0, seed=31)
    limits, report = autotune.autotune(m, steps=60, repeats=1)
    print(f"[autotune] {limits} from {len(report)} timings", flush=True)
    sys.stdout.flush()
    assert len(report) >= 2
    times = dict(report)
    assert times[tuple(limits)] <= times[autotune.DEFAULT]
    for k, v in enumerate(limits[:3]):
        assert v in autotune.CANDIDATES[k]
    with Simulation(m) as a, Simulation(m, size_limits=limits) as b:
        ta, tb = a.run_all(), b.run_all()
    np.testing.assert_array_equal(ta, tb)
