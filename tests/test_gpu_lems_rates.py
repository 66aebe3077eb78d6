"""User gate rate / time course / steady state ComponentTypes on the GPU: the compartments
with such gates run their channels in the channel pre-pass compiled with hiprtc together with
the generated lems_rate. The HH cell with its rates given by user types equals the oracle bit
for bit and equals the built-in forms' run; a temperature-dependent tau/inf channel and a user
reverse rate on top also match the oracle, in the plain and graph launch modes."""
import sys

import numpy as np
import pytest

from eden_amd.model import Simulation
from oracle import oracle

pytestmark = pytest.mark.gpu

EXTRA = ('<channelDensity id="kva" ionChannel="kva" condDensity="5 mS_per_cm2" erev="-77mV"/>'
         '<channelDensity id="kdr" ionChannel="kdr" condDensity="3 mS_per_cm2" erev="-77mV"/>')


def _run(m, **opts):
    with Simulation(m, **opts) as sim:
        tr = sim.run_all()
        ev = sim.read_events()
    return tr, ev


@pytest.mark.parametrize("opts", [{}, dict(use_graph=True)])
def test_user_rate_types_bit_exact(tmp_path, opts):
    from test_lems_rates_cpu import hh_model
    mu, mb = hh_model(tmp_path, True), hh_model(tmp_path, False)
    print(f"[rates] user rate instances {len(mu.lems_rates.entries)}, opts {opts}", flush=True)
    sys.stdout.flush()
    tu, evu = _run(mu, **opts)
    tb, evb = _run(mb, **opts)
    tr_r, ec, er, _ = oracle.run(mu)
    assert np.array_equal(tu, tr_r) and np.array_equal(evu[1], er)
    assert np.array_equal(tu, tb)
    m2 = hh_model(tmp_path, True, extra_channel=EXTRA)
    t2, ev2 = _run(m2, **opts)
    tr2, ec2, er2, _ = oracle.run(m2)
    assert np.array_equal(t2, tr2), float(np.nanmax(np.abs(t2 - tr2)))
    assert np.array_equal(ev2[0], ec2) and np.array_equal(ev2[1], er2)
    assert not np.array_equal(t2, tu)          # the extra channels act
