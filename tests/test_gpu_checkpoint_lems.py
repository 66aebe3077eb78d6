"""Checkpoint and restore of the state tables of the generic paths: the channel pre-pass's
gate variables (compartments with more than 4 channel distributions), the firing-synapse
inputs' spike-list positions and Poisson spikers, and random spike sources' state.
edenhip_get_state after some rows, edenhip_set_state into a fresh context of the same model:
the rows after the restore equal, bit for bit, those the first context runs on."""
import sys

import numpy as np
import pytest

from eden_amd.model import Simulation

pytestmark = pytest.mark.gpu


def _model(tmp_path, which):
    if which == "many_channels":
        from test_gpu_many_channels import _model as mc
        return mc(tmp_path)
    from test_firing_inputs_cpu import _list, _model as fi
    return fi(tmp_path,
              '<poissonFiringSynapse id="pg" averageRate="150 Hz" synapse="e1" spikeTarget="./e1"/>'
              '<transientPoissonFiringSynapse id="tp" averageRate="300 Hz" delay="20ms" duration="100ms" '
              'synapse="e2" spikeTarget="./e2"/>'
              '<timedSynapticInput id="ti" synapse="e1" spikeTarget="./e1"><spike id="0" time="5 ms"/>'
              '<spike id="1" time="30 ms"/><spike id="2" time="60 ms"/></timedSynapticInput>',
              _list("G", "pg", [0, 1]) + _list("T", "tp", [0]) + _list("S", "ti", [1]), length_ms=150.0)


@pytest.mark.parametrize("which", ["many_channels", "firing_inputs"])
def test_checkpoint_restore(tmp_path, which):
    m = _model(tmp_path, which)
    first = 1201
    with Simulation(m, chunk_steps=first) as a:
        a.run(-1)
        blob = a.get_state()
        rest = []
        while a.run(-1):
            rest.append(a.read_trace()[0].copy())
    with Simulation(m, chunk_steps=first) as b:
        b.set_state(blob)
        rest_b = []
        while b.run(-1):
            rest_b.append(b.read_trace()[0].copy())
    ra, rb = np.concatenate(rest), np.concatenate(rest_b)
    print(f"[state] {which}: {len(blob)} bytes, {len(ra)} rows after the restore", flush=True)
    sys.stdout.flush()
    assert len(ra) > 1000 and np.array_equal(ra, rb)
    assert float(np.ptp(ra[:, 0])) > 0.01
