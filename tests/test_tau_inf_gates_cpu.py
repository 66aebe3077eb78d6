"""Gates given by their time course and steady state (gateHHtauInf, gateHHratesTauInf) on the
CPU: the reader's tables and the oracle's dynamics.

* A passive cell with a tau/inf potassium-like channel, depolarised by a pulse: the recorded
  gate follows the recurrence q' = clip(b q + (1 - b) inf(V)) with b = exp(-dt q10 / tau), the
  fixed time course tau taken as it is (not 1 / (alpha + beta)) and inf the sigmoid
  steady state rate / (1 + exp((mid - V) / scale)) of the step's start voltage, evaluated here
  in float32 from the recorded voltage trace (eden/Eden.cpp:5282-5383, 5236-5259).
* gateHHratesTauInf reads like gateHHtauInf (its rates play no part); gateHHratesTau and
  gateHHratesInf take their third rate in place of the one their rates would give.
* gateHHInstantaneous (eden/Eden.cpp:5269-5281): no state; in each step the channel's open
  fraction is the steady state at that step's own voltage. The first step from rest, by hand:
  V1 = (V0 + dt (I + V0 G) / C) / (1 + dt G / C) with the channel's I = g x(V0)^k (E - V0) 1e-5 A.
"""
import math
import os
import shutil

import numpy as np
import pytest

from eden_amd import _native as N
from eden_amd import neuroml
from oracle import oracle

HERE = os.path.dirname(os.path.abspath(__file__))


def _model(tmp_path, gate_xml, q10="", length_ms=60.0, record_q=True, n_cells=1):
    p = tmp_path / "ti.xml"
    p.write_text(
        '<Lems><Target component="s"/><neuroml>'
        '<ionChannelHH id="leak" conductance="10pS"/>'
        f'<ionChannelHH id="kti" conductance="10pS" species="k">{gate_xml}</ionChannelHH>'
        '<cell id="P"><morphology id="m"><segment id="0"><proximal x="0" y="0" z="0" diameter="10"/>'
        '<distal x="0" y="0" z="0" diameter="10"/></segment></morphology><biophysicalProperties id="b">'
        '<membraneProperties><channelDensity id="l" ionChannel="leak" condDensity="0.3 mS_per_cm2" erev="-65mV"/>'
        '<channelDensity id="k" ionChannel="kti" condDensity="2 mS_per_cm2" erev="-80mV"/>'
        '<specificCapacitance value="1.0 uF_per_cm2"/><initMembPotential value="-65mV"/></membraneProperties>'
        '<intracellularProperties><resistivity value="0.1 kohm_cm"/></intracellularProperties>'
        '</biophysicalProperties></cell>'
        '<pulseGenerator id="pg" delay="5ms" duration="30ms" amplitude="0.05nA"/>'
        f'<network id="n" temperature="6.3 degC"><population id="pp" component="P" size="{n_cells}"/>'
        '<explicitInput target="pp[0]" input="pg"/></network></neuroml>'
        f'<Simulation id="s" length="{length_ms}ms" step="0.025ms" target="n" seed="1">'
        '<OutputFile id="o" fileName="o.txt"><OutputColumn id="v" quantity="pp/0/P/0/v"/>'
        + ('<OutputColumn id="q" quantity="pp/0/P/0/b/membraneProperties/k/kti/n/q"/>' if record_q else '') +
        '</OutputFile>'
        '</Simulation></Lems>')
    return neuroml.read(str(p))[0]


TAU_INF = ('<gateHHtauInf id="n" instances="2">{q10}<timeCourse type="fixedTimeCourse" tau="4ms"/>'
           '<steadyState type="HHSigmoidVariable" rate="1" scale="5mV" midpoint="-55mV"/></gateHHtauInf>')


def test_tau_inf_gate_recurrence(tmp_path):
    q10 = '<q10Settings type="q10Fixed" fixedQ10="1.5"/>'
    m = _model(tmp_path, TAU_INF.format(q10=q10))
    assert m.gate_kind.tolist() == [0] * 4 + [N.GATE_TAU_INF, 0, 0, 0]
    g = m.channel_types[1].gates[0]
    assert (g.alpha.form, g.alpha.rate) == (N.RATE_FIXED, 4000.0)
    assert (g.beta.form, g.beta.rate, g.beta.midpoint, g.beta.scale) == (N.RATE_SIGMOID, 1.0, -55.0, 5.0)
    tr, _, _, _ = oracle.run(m)
    V, q = tr[:, 0], tr[:, 1]
    assert float(np.ptp(V)) > 3.0 and float(np.ptp(q)) > 0.05
    f32 = np.float32
    blend = f32(math.exp(-float(f32(f32(25.0) * f32(1.5)) / f32(4000.0))))
    want = [q[0]]
    for r in range(len(q) - 1):
        inf = f32(1.0) / (f32(1.0) + f32(math.exp(float(f32(f32(-55.0) - V[r]) / f32(5.0)))))
        x = blend * want[-1] + (f32(1.0) - blend) * inf
        want.append(min(max(x, f32(1e-6)), f32(1 - 1e-6)))
    np.testing.assert_allclose(q, np.array(want, np.float32), rtol=0, atol=2e-7)
    # the first row: the initial steps leave the gate at inf of the initial voltage
    assert q[0] == pytest.approx(1.0 / (1.0 + math.exp((-55.0 + 65.0) / 5.0)), abs=1e-7)


def test_rates_tau_inf_reads_as_tau_inf(tmp_path):
    m1 = _model(tmp_path, TAU_INF.format(q10=""))
    tr1, _, _, _ = oracle.run(m1)
    rti = ('<gateHHratesTauInf id="n" instances="2"><forwardRate type="HHExpRate" rate="1per_ms" midpoint="0mV" '
           'scale="10mV"/><reverseRate type="HHExpRate" rate="1per_ms" midpoint="0mV" scale="-10mV"/>'
           '<timeCourse type="fixedTimeCourse" tau="4ms"/>'
           '<steadyState type="HHSigmoidVariable" rate="1" scale="5mV" midpoint="-55mV"/></gateHHratesTauInf>')
    m2 = _model(tmp_path, rti)
    tr2, _, _, _ = oracle.run(m2)
    assert np.array_equal(tr1, tr2)


def test_rates_with_a_third_rate(tmp_path):
    """gateHHratesTau takes tau from its time course and inf from its rates, gateHHratesInf inf
    from its steady state and tau from its rates (eden/Eden.cpp:5282-5383). Two sigmoid rates
    summing to 0.5 per ms at every voltage give tau = 2 ms and inf = the sigmoid of rate 1: a
    rates+tau gate with a 2 ms time course and a rates+inf gate with that sigmoid as its steady
    state reproduce the plain rates gate to float rounding; a 20 ms time course changes the
    trace."""
    rates = ('<forwardRate type="HHSigmoidRate" rate="0.5per_ms" midpoint="-50mV" scale="5mV"/>'
             '<reverseRate type="HHSigmoidRate" rate="0.5per_ms" midpoint="-50mV" scale="-5mV"/>')
    # alpha + beta = 0.5 per ms at every V (the two sigmoids sum to 1), so tau = 2 ms, and
    # inf = alpha / 0.5 = the sigmoid of rate 1
    plain = _model(tmp_path, f'<gateHHrates id="n" instances="2">{rates}</gateHHrates>')
    with_tau = _model(tmp_path, f'<gateHHratesTau id="n" instances="2">{rates}'
                                '<timeCourse type="fixedTimeCourse" tau="2ms"/></gateHHratesTau>')
    with_inf = _model(tmp_path, f'<gateHHratesInf id="n" instances="2">{rates}'
                                '<steadyState type="HHSigmoidVariable" rate="1" midpoint="-50mV" scale="5mV"/>'
                                '</gateHHratesInf>')
    slow = _model(tmp_path, f'<gateHHratesTau id="n" instances="2">{rates}'
                            '<timeCourse type="fixedTimeCourse" tau="20ms"/></gateHHratesTau>')
    assert with_tau.gate_kind.tolist()[4] == N.GATE_RATES_TAU and with_inf.gate_kind.tolist()[4] == N.GATE_RATES_INF
    assert with_tau.gate_extra_rate[4].rate == 2000.0
    t0 = oracle.run(plain)[0]
    for m in (with_tau, with_inf):
        np.testing.assert_allclose(oracle.run(m)[0], t0, rtol=0, atol=2e-4)
    assert float(np.abs(oracle.run(slow)[0][:, 1] - t0[:, 1]).max()) > 1e-2


def test_instantaneous_gate_first_step_by_hand(tmp_path):
    gate = ('<gateHHInstantaneous id="n" instances="2"><steadyState type="HHSigmoidVariable" rate="1" '
            'scale="4mV" midpoint="-60mV"/></gateHHInstantaneous>')
    m = _model(tmp_path, gate, length_ms=2.0)
    assert m.gate_kind.tolist()[4] == N.GATE_INSTANTANEOUS
    tr, _, _, _ = oracle.run(m)
    V0 = -65.0
    x = 1.0 / (1.0 + math.exp((-60.0 - V0) / 4.0))
    A = float(m.area[0])
    C = float(m.capacitance[0])
    g_leak, g_k = 0.3 * 1e-5 * A, 2.0 * x * x * 1e-5 * A     # condDensity in mS/cm2 units of the engine
    I = g_leak * (-65.0 - V0) + g_k * (-80.0 - V0)
    G = g_leak + g_k
    # the reference's semi-implicit step for a one-compartment cell (eden/Eden.cpp:6057-6107,
    # 6114-6225 with N = 1): V' = V + dt (I + V G) / C, then V' / (1 + dt G / C)
    V1 = (V0 + 25.0 * (I + V0 * G) / C) / (1.0 + 25.0 * G / C)
    assert tr[1, 0] == pytest.approx(V1, abs=2e-5)
    assert float(np.ptp(tr[:, 0])) > 1.0


def test_instantaneous_gate_takes_the_step_s_own_voltage(tmp_path):
    """Step 2 from the recorded V1: the open fraction uses x(V1), not the previous step's x(V0)
    (which a gate updated after its use would give)."""
    gate = ('<gateHHInstantaneous id="n" instances="2"><steadyState type="HHSigmoidVariable" rate="1" '
            'scale="4mV" midpoint="-60mV"/></gateHHInstantaneous>')
    m = _model(tmp_path, gate, length_ms=2.0)
    tr, _, _, _ = oracle.run(m)
    A, C = float(m.area[0]), float(m.capacitance[0])

    def step(V, xv):
        g_leak, g_k = 0.3 * 1e-5 * A, 2.0 * xv * xv * 1e-5 * A
        I = g_leak * (-65.0 - V) + g_k * (-80.0 - V)
        G = g_leak + g_k
        return (V + 25.0 * (I + V * G) / C) / (1.0 + 25.0 * G / C)

    x = lambda V: 1.0 / (1.0 + math.exp((-60.0 - V) / 4.0))
    V1 = float(tr[1, 0])
    want, lagged = step(V1, x(V1)), step(V1, x(-65.0))
    assert tr[2, 0] == pytest.approx(want, abs=2e-5)
    assert abs(lagged - want) > 1e-4


def test_q10_conductance_scaling(tmp_path):
    """q10ConductanceScaling starts the channel's open fraction at q10Factor^((T - T0) / 10)
    (eden/Eden.cpp:5700-5721): with T - T0 = 10 K and a factor 3 the channel acts as one of
    three times its density (to float rounding of the two products)."""
    sc = '<q10ConductanceScaling q10Factor="3" experimentalTemp="-3.7 degC"/>'
    m = _model(tmp_path, sc + TAU_INF.format(q10=""), record_q=False)
    assert m.chan_scaling.tolist() == [1.0, pytest.approx(3.0, rel=1e-6)]
    tr = oracle.run(m)[0]
    src = (tmp_path / "ti.xml").read_text().replace(sc, "").replace('condDensity="2 mS_per_cm2"', 'condDensity="6 mS_per_cm2"')
    (tmp_path / "ti3.xml").write_text(src)
    tr3 = oracle.run(neuroml.read(str(tmp_path / "ti3.xml"))[0])[0]
    np.testing.assert_allclose(tr, tr3, rtol=0, atol=1e-3)
    tr1 = oracle.run(_model(tmp_path, TAU_INF.format(q10=""), record_q=False))[0]
    assert float(np.abs(tr1 - tr).max()) > 0.1
