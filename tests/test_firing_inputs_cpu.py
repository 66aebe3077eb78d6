"""Firing-synapse inputs on the generic path (poissonFiringSynapse with a synapse other than
expTwoSynapse, transientPoissonFiringSynapse) on the CPU.

* The generic spiker draws the same intervals as the native Poisson input: one cell with a
  native poissonFiringSynapse (expTwoSynapse) and a generic one (expOneSynapse), each the
  first instance of its input on the cell, so both draw randof(work item = the cell's seed,
  instance 0, step, invocation 0 or 1): the generic instance's synaptic events fall on exactly
  the native input's firing steps. The native input itself is pinned by the Poisson
  statistics of tests/test_oracle.py.
* transientPoissonFiringSynapse (eden/neuroml/LEMS_CoreComponents.inc.xml:674-735): no spike
  before its delay, none after delay + duration (the interval drawn at the last spike in the
  window is pushed 1e9 hours away), and a count inside a 4-sigma Poisson band of
  rate x duration.
* Per-instance settings reach the spiker (averageRate, delay, duration) and the synapse
  ("synapse/<parameter>", eden/NeuroML.cpp:1712-1716).
"""
import numpy as np
import pytest

from eden_amd import neuroml
from oracle import oracle

PAS = ('<ionChannelHH id="passiveChan" conductance="10pS"/>'
       '<cell id="P"><morphology id="m"><segment id="0"><proximal x="0" y="0" z="0" diameter="10"/>'
       '<distal x="0" y="0" z="0" diameter="10"/></segment></morphology><biophysicalProperties id="b">'
       '<membraneProperties><channelDensity id="leak" ionChannel="passiveChan" condDensity="0.3 mS_per_cm2" '
       'erev="-65mV"/><specificCapacitance value="1.0 uF_per_cm2"/><initMembPotential value="-65mV"/>'
       '</membraneProperties><intracellularProperties><resistivity value="0.1 kohm_cm"/>'
       '</intracellularProperties></biophysicalProperties></cell>')


def _model(tmp_path, inputs, lists, length_ms=500.0, setup=None, n_cells=2):
    sim_extra = '<EdenCustomSetup filename="s.txt"/>' if setup else ""
    if setup:
        (tmp_path / "s.txt").write_text(setup)
    p = tmp_path / "fi.xml"
    p.write_text(
        '<Lems><Target component="s"/><neuroml>' + PAS +
        '<expOneSynapse id="e1" gbase="0.5nS" erev="0mV" tauDecay="2ms"/>'
        '<expTwoSynapse id="e2" gbase="0.5nS" erev="0mV" tauRise="0.5ms" tauDecay="3ms"/>' + inputs +
        f'<network id="n"><population id="pp" component="P" size="{n_cells}"/>' + lists + '</network></neuroml>'
        f'<Simulation id="s" length="{length_ms}ms" step="0.025ms" target="n" seed="11">'
        '<OutputFile id="o" fileName="o.txt"><OutputColumn id="a" quantity="pp/0/P/0/v"/></OutputFile>'
        + sim_extra + '</Simulation></Lems>')
    return neuroml.read(str(p))[0]


def _list(lid, comp, cells):
    return (f'<inputList id="{lid}" component="{comp}" population="pp">' +
            "".join(f'<input id="{k}" target="../pp/{c}/P" destination="synapses"/>' for k, c in enumerate(cells)) +
            '</inputList>')


def test_generic_poisson_draws_the_native_intervals(tmp_path):
    m = _model(tmp_path,
               '<poissonFiringSynapse id="pn" averageRate="200 Hz" synapse="e2" spikeTarget="./e2"/>'
               '<poissonFiringSynapse id="pg" averageRate="200 Hz" synapse="e1" spikeTarget="./e1"/>',
               _list("N", "pn", [0]) + _list("G", "pg", [0]))
    assert m.lems is not None and len(m.poisson_rate) == 1                 # pn native, pg generic
    assert list(m.syn_poisson_spiker) == [0] and m.poisson_spikers["instance"].tolist() == [0]
    _, (pst, ppi) = oracle.run_with_poisson_log(m)
    _, (sst, ssy) = oracle.run_with_synapse_log(m)
    assert len(pst) > 60
    assert np.array_equal(np.sort(pst), np.sort(sst))


def test_transient_poisson_fires_inside_its_window(tmp_path):
    rate, delay, duration = 400.0, 100.0, 200.0
    m = _model(tmp_path,
               f'<transientPoissonFiringSynapse id="tp" averageRate="{rate} Hz" delay="{delay}ms" '
               f'duration="{duration}ms" synapse="e2" spikeTarget="./e2"/>',
               _list("T", "tp", [0, 1]))
    assert m.poisson_spikers["transient"].tolist() == [1, 1]
    _, (st, sy) = oracle.run_with_synapse_log(m)
    dt = 25.0
    for inst in (0, 1):
        steps = st[sy == inst]
        t_us = (steps - 1) * dt                       # the step's start time
        assert t_us.min() >= delay * 1000 and t_us.max() <= (delay + duration) * 1000 + 2 * dt
        mean = rate * duration / 1000
        assert abs(len(steps) - mean) < 4 * np.sqrt(mean), len(steps)
    # the two instances (on different cells) draw different intervals
    assert not np.array_equal(st[sy == 0], st[sy == 1])


def test_settings_reach_spikers_and_synapses(tmp_path):
    m = _model(tmp_path,
               '<transientPoissonFiringSynapse id="tp" averageRate="50 Hz" delay="10ms" duration="20ms" '
               'synapse="e1" spikeTarget="./e1"/>'
               '<timedSynapticInput id="ti" synapse="e1" spikeTarget="./e1"><spike id="0" time="5 ms"/>'
               '</timedSynapticInput>',
               _list("T", "tp", [0, 1]) + _list("S", "ti", [1]),
               setup="set input T 1 averageRate 0.3 kHz\nset input T all delay multi s\nvalues 0.02 0.04\n"
                     "set input T 0 synapse/tauDecay 4 ms\nset input S 0 synapse/gbase 2 nS\n")
    P = m.poisson_spikers
    assert P["rate"].tolist() == [np.float32(50e-6), np.float32(300e-6)]
    assert P["delay"].tolist() == [2e4, 4e4] and P["duration"].tolist() == [2e4, 2e4]
    vecs = {int(i): v for i, v in zip(np.flatnonzero(m.syn_param_index >= 0), m.syn_param_vecs)}
    assert len(vecs) == 2
    # expOneSynapse: gbase (uS), erev, tauDecay (us), weight
    got = sorted(tuple(v.tolist()) for v in vecs.values())
    f = lambda *x: tuple(float(np.float32(y)) for y in x)
    assert got == [f(5e-4, 0.0, 4000.0, 1.0), f(2e-3, 0.0, 2000.0, 1.0)]


def test_setting_refusals(tmp_path):
    common = ('<poissonFiringSynapse id="pn" averageRate="20 Hz" synapse="e2" spikeTarget="./e2"/>'
              '<timedSynapticInput id="ti" synapse="e1" spikeTarget="./e1"><spike id="0" time="5 ms"/>'
              '</timedSynapticInput>')
    lists = _list("N", "pn", [0]) + _list("S", "ti", [1])
    for text, msg in (("set input N 0 synapse/gbase 1 nS\n", "natively run"),
                      ("set input S 0 gbase 1 nS\n", "timedSynapticInput paths"),
                      ("set input N 0 isi 1 ms\n", "state variable")):
        with pytest.raises(neuroml.NeuroMLError, match=msg):
            _model(tmp_path, common, lists, setup=text)
