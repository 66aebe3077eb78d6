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
    # channel waves: the rule's choice per workgroup size, then each used size flipped
    waves, wreport = autotune.autotune_channel_wave(m, limits, steps=60, repeats=1)
    print(f"[autotune] channel waves {waves} from {len(wreport)} timings", flush=True)
    rule = autotune.rule_channel_wave(m, limits)
    assert len(wreport) == 1 + sum(r is not None for r in rule)
    assert all(w == 0 or rule[k] is not None for k, w in enumerate(waves))
    with Simulation(m, size_limits=limits, channel_wave=waves) as c:
        np.testing.assert_array_equal(c.run_all(), ta)
