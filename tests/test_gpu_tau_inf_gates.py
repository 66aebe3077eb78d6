"""Tau/inf gates (gateHHtauInf) on the GPU: in the point kernel (one-compartment cells) and
the tree kernel (a five-compartment cell, gated everywhere), beside alpha/beta HH gates,
recorded gate variables and voltages bit for bit against the oracle, in the plain, graph and
chunked launch modes."""
import sys

import numpy as np
import pytest

from eden_amd import neuroml
from eden_amd.model import Simulation
from oracle import oracle

pytestmark = pytest.mark.gpu


def _model(tmp_path):
    parent = lambda k: "" if k == 0 else '<parent segment="%d"/>' % (k - 1)
    segs = "".join(
        f'<segment id="{k}">{parent(k)}<proximal x="{20 * k}" y="0" z="0" diameter="4"/>'
        f'<distal x="{20 * k + 20}" y="0" z="0" diameter="4"/></segment>' for k in range(5))
    bio = ('<biophysicalProperties id="b{n}"><membraneProperties>'
           '<channelDensity id="l" ionChannel="leak" condDensity="0.3 mS_per_cm2" erev="-65mV"/>'
           '<channelDensity id="k" ionChannel="kti" condDensity="3 mS_per_cm2" erev="-80mV"/>'
           '<channelDensity id="na" ionChannel="naChan" condDensity="40 mS_per_cm2" erev="50mV"/>'
           '<specificCapacitance value="1.0 uF_per_cm2"/><initMembPotential value="-65mV"/></membraneProperties>'
           '<intracellularProperties><resistivity value="0.1 kohm_cm"/></intracellularProperties>'
           '</biophysicalProperties>')
    p = tmp_path / "tig.xml"
    p.write_text(
        '<Lems><Target component="s"/><neuroml>'
        '<ionChannelHH id="leak" conductance="10pS"/>'
        '<ionChannelHH id="kti" conductance="10pS" species="k">'
        '<gateHHtauInf id="n" instances="3"><q10Settings type="q10ExpTemp" q10Factor="3" experimentalTemp="6.3 degC"/>'
        '<timeCourse type="fixedTimeCourse" tau="2ms"/>'
        '<steadyState type="HHSigmoidVariable" rate="1" scale="6mV" midpoint="-50mV"/></gateHHtauInf>'
        '<gate id="h" type="gateHHtauInf" instances="1"><timeCourse type="fixedTimeCourse" tau="15ms"/>'
        '<steadyState type="HHExpLinearVariable" rate="0.2" scale="-8mV" midpoint="-60mV"/></gate></ionChannelHH>'
        '<ionChannelHH id="naChan" conductance="10pS" species="na">'
        '<gateHHrates id="m" instances="3"><forwardRate type="HHExpLinearRate" rate="1per_ms" midpoint="-40mV" '
        'scale="10mV"/><reverseRate type="HHExpRate" rate="4per_ms" midpoint="-65mV" scale="-18mV"/></gateHHrates>'
        '<gateHHtauInf id="h" instances="1"><timeCourse type="fixedTimeCourse" tau="1ms"/>'
        '<steadyState type="HHSigmoidVariable" rate="1" scale="-7mV" midpoint="-62mV"/></gateHHtauInf></ionChannelHH>'
        '<cell id="P"><morphology id="m1"><segment id="0"><proximal x="0" y="0" z="0" diameter="10"/>'
        '<distal x="0" y="0" z="0" diameter="10"/></segment></morphology>' + bio.format(n=1) + '</cell>'
        f'<cell id="T"><morphology id="m2">{segs}</morphology>' + bio.format(n=2) + '</cell>'
        '<pulseGenerator id="pg" delay="5ms" duration="60ms" amplitude="0.08nA"/>'
        '<network id="n" temperature="16.3 degC"><population id="pp" component="P" size="3"/>'
        '<population id="tt" component="T" size="2"/>'
        '<explicitInput target="pp[0]" input="pg"/><explicitInput target="pp[2]" input="pg"/>'
        '<explicitInput target="tt[1]" input="pg"/></network></neuroml>'
        '<Simulation id="s" length="100ms" step="0.025ms" target="n" seed="1">'
        '<OutputFile id="o" fileName="o.txt">'
        '<OutputColumn id="v0" quantity="pp/0/P/0/v"/><OutputColumn id="v2" quantity="pp/2/P/0/v"/>'
        '<OutputColumn id="t0" quantity="tt/1/T/0/v"/><OutputColumn id="t4" quantity="tt/1/T/4/v"/>'
        '<OutputColumn id="q1" quantity="pp/0/P/0/b1/membraneProperties/k/kti/n/q"/>'
        '<OutputColumn id="q2" quantity="tt/1/T/3/b2/membraneProperties/na/naChan/h/q"/>'
        '</OutputFile><EventOutputFile id="e" fileName="e.txt" format="TIME_ID">'
        '<EventSelection id="0" select="pp[0]" eventPort="spike"/><EventSelection id="1" select="tt[1]" eventPort="spike"/>'
        '</EventOutputFile></Simulation></Lems>')
    return neuroml.read(str(p))[0]


@pytest.mark.parametrize("opts", [{}, dict(use_graph=True), dict(chunk_steps=1111)])
def test_tau_inf_gates_bit_exact(tmp_path, opts):
    m = _model(tmp_path)
    print(f"[tau-inf] {m.n_cells} cells, {m.n_comps} compartments, opts {opts}", flush=True)
    sys.stdout.flush()
    tr_r, ec, er, _ = oracle.run(m)
    with Simulation(m, **opts) as sim:
        tr = sim.run_all()
        ev = sim.read_events()
    assert np.array_equal(tr, tr_r), float(np.nanmax(np.abs(tr - tr_r)))
    assert np.array_equal(ev[0], ec) and np.array_equal(ev[1], er)
    assert float(np.ptp(tr[:, 4])) > 0.05 and float(np.ptp(tr[:, 5])) > 0.05
    assert float(np.ptp(tr[:, 2])) > 5.0


@pytest.mark.parametrize("opts", [{}, dict(use_graph=True)])
def test_instantaneous_gates_bit_exact(tmp_path, opts):
    """gateHHInstantaneous: its channel runs in the channel pre-pass with the step's own
    voltage; traces bit for bit against the oracle."""
    from test_tau_inf_gates_cpu import _model as ti_model
    gate = ('<gateHHInstantaneous id="n" instances="2"><steadyState type="HHSigmoidVariable" rate="1" '
            'scale="4mV" midpoint="-60mV"/></gateHHInstantaneous>')
    m = ti_model(tmp_path, gate, record_q=False, n_cells=3)
    print(f"[tau-inf] instantaneous gate, opts {opts}", flush=True)
    tr_r, _, _, _ = oracle.run(m)
    with Simulation(m, **opts) as sim:
        tr = sim.run_all()
    assert np.array_equal(tr, tr_r), float(np.nanmax(np.abs(tr - tr_r)))
    assert float(np.ptp(tr[:, 0])) > 1.0


@pytest.mark.parametrize("kind", ["gateHHratesTau", "gateHHratesInf"])
def test_rates_with_a_third_rate_bit_exact(tmp_path, kind):
    from test_tau_inf_gates_cpu import _model as ti_model
    rates = ('<forwardRate type="HHSigmoidRate" rate="0.5per_ms" midpoint="-50mV" scale="5mV"/>'
             '<reverseRate type="HHExpLinearRate" rate="0.1per_ms" midpoint="-55mV" scale="10mV"/>')
    third = ('<timeCourse type="fixedTimeCourse" tau="3ms"/>' if kind == "gateHHratesTau" else
             '<steadyState type="HHSigmoidVariable" rate="1" midpoint="-50mV" scale="5mV"/>')
    m = ti_model(tmp_path, f'<{kind} id="n" instances="2">{rates}{third}</{kind}>', record_q=False, n_cells=2)
    print(f"[tau-inf] {kind}", flush=True)
    tr_r, _, _, _ = oracle.run(m)
    with Simulation(m) as sim:
        tr = sim.run_all()
    assert np.array_equal(tr, tr_r), float(np.nanmax(np.abs(tr - tr_r)))


def test_conductance_scaling_bit_exact(tmp_path):
    from test_tau_inf_gates_cpu import TAU_INF, _model as ti_model
    sc = '<q10ConductanceScaling q10Factor="3" experimentalTemp="-3.7 degC"/>'
    m = ti_model(tmp_path, sc + TAU_INF.format(q10=""), record_q=False, n_cells=2)
    print("[tau-inf] q10ConductanceScaling", flush=True)
    tr_r, _, _, _ = oracle.run(m)
    with Simulation(m) as sim:
        tr = sim.run_all()
    assert np.array_equal(tr, tr_r), float(np.nanmax(np.abs(tr - tr_r)))
