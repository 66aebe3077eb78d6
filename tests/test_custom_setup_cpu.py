"""Per-instance parameter settings (EdenCustomSetup files) on the CPU.

* The reference's EdenTest_Extension_CustomSetup model (tests/golden/neuroml, with its setup
  file): five generic Izhikevich cells and five generic pulseGeneratorDL inputs, set per
  instance to the parameters of five distinct cells and inputs, and two expOneSynapse
  connections set to the parameters of two other synapse types. The reference's own criteria
  (testing/validation_tests/validation_tests.py:29-41) hold bit for bit in the oracle: each
  customised trace equals its distinct counterpart exactly. Without the setup file the pairs
  differ, so the settings are what makes them equal.
* Native inputs: a pulseGenerator input list set per input (delay, duration, amplitude)
  reproduces separately declared pulses exactly; a poissonFiringSynapse list's averageRate
  lands in the Poisson table.
* A model of expTwoSynapse projections only takes the generic path when a statement sets a
  synapse parameter, and a connection set to another type's parameters reproduces it.
* Gap junctions: 'pre' and 'post' address the two ends of an electrical connection.
* The grammar and its errors: comments and blank lines, unit conversion of single and multi
  values, later statements overriding earlier ones, and the refusals (duplicate ids, unknown
  ids, a values line of the wrong length, a dimensional value without units, cable modes,
  state variables, 'pre' on a spiking projection, 'set cell' on physical cells).
"""
import os
import re
import shutil

import numpy as np
import pytest

from eden_amd import custom_setup, neuroml
from oracle import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "neuroml")
PAIRS = [(5, 0), (6, 1), (7, 2), (8, 3), (9, 4), (10, 12), (11, 13)]   # (distinct, customised) columns


def _reference_model(tmp_path, with_setup=True):
    src = open(os.path.join(GOLD, "EdenTest_Extension_CustomSetup.nml")).read()
    if not with_setup:
        src = re.sub(r"<EdenCustomSetup[^>]*>\s*</EdenCustomSetup>", "", src)
        assert "EdenCustomSetup" not in src
    (tmp_path / "m.nml").write_text(src)
    shutil.copy(os.path.join(GOLD, "EdenTest_Extension_CustomSetup.txt"), tmp_path)
    return neuroml.read(str(tmp_path / "m.nml"))


def test_reference_custom_setup_pairs_exact(tmp_path):
    m, outs = _reference_model(tmp_path)
    cols = [c[0] for c in outs.trajectories[0].columns]
    assert cols[:5] == ["iiA", "iiB", "iiC", "iiD", "iiE"] and cols[5:10] == ["izA", "izB", "izC", "izD", "izE"]
    # the five generic cells and inputs, and the two ExpWild connections, have their own vectors
    assert list(m.cell_param_index) == [-1] * 5 + [0, 1, 2, 3, 4] + [-1] * 5
    assert np.count_nonzero(m.syn_param_index >= 0) == 7
    tr, ec, er, _ = oracle.run(m)
    for a, b in PAIRS:
        assert np.array_equal(tr[:, a], tr[:, b]), (a, b)
        assert float(np.ptp(tr[:, a])) > 50.0     # every pair moves (spikes or a large PSP)
    m0, _ = _reference_model(tmp_path, with_setup=False)
    assert m0.cell_param_index is None and m0.syn_param_index is None
    tr0, _, _, _ = oracle.run(m0)
    for a, b in PAIRS:
        assert not np.array_equal(tr0[:, a], tr0[:, b]), (a, b)
    # the distinct cells do not depend on the settings
    assert np.array_equal(tr0[:, 5:12], tr[:, 5:12])


def test_setting_values_in_engine_units(tmp_path):
    m, _ = _reference_model(tmp_path)
    # pulseGeneratorDL (delay, duration in us; amplitude dimensionless; weight): 'sec' and 'ms'
    v = m.syn_param_vecs
    assert [list(x) for x in v[:5]] == [[1e4, 9e4, 14.0, 1.0], [2e4, 1.8e5, 0.5, 1.0], [2.2e4, 1.98e5, 15.0, 1.0],
                                         [2e4, 1.8e5, 0.6, 1.0], [1.6e4, 1.44e5, 10.0, 1.0]]
    # expOneSynapse (gbase uS, erev mV, tauDecay us; weight): 0.05 V and 5 / 0.5 msec
    assert [list(x) for x in v[5:]] == [[0.01, 50.0, 5000.0, 1.0], [0.01, 50.0, 500.0, 1.0]]
    # izhikevichCell (v0, a, b, c, d, thresh): v0 -64 mV, b 0.25 on cells 1 and 3; c -55 on
    # 2..4 then -50 on 2; d per cell
    want = [[-70, .02, .2, -65, 6, 30], [-64, .02, .25, -65, 6, 30], [-70, .02, .2, -50, 2, 30],
            [-64, .02, .25, -55, .05, 30], [-70, .02, .2, -55, 4, 30]]
    for got, w in zip(m.cell_param_vecs, want):
        assert list(got) == list(np.array(w, np.float32))


def _native_model(tmp_path, setup, extra_sim=""):
    shutil.copy(os.path.join(GOLD, "NML2_SingleCompHHCell.nml"), tmp_path / "NML2_SingleCompHHCell.nml")
    (tmp_path / "setup.txt").write_text(setup)
    pas = ('<cell id="P"><morphology id="m"><segment id="0"><proximal x="0" y="0" z="0" diameter="10"/>'
           '<distal x="0" y="0" z="0" diameter="10"/></segment></morphology><biophysicalProperties id="b">'
           '<membraneProperties><channelDensity id="leak" ionChannel="passiveChan" condDensity="0.3 mS_per_cm2" '
           'erev="-65mV"/><specificCapacitance value="1.0 uF_per_cm2"/><initMembPotential value="-65mV"/>'
           '</membraneProperties><intracellularProperties><resistivity value="0.1 kohm_cm"/>'
           '</intracellularProperties></biophysicalProperties></cell>')
    p = tmp_path / "native.xml"
    p.write_text(
        '<Lems><Target component="s"/><Include file="NML2_SingleCompHHCell.nml"/><neuroml>' + pas +
        '<pulseGenerator id="pg" delay="5ms" duration="10ms" amplitude="0.01nA"/>'
        '<pulseGenerator id="pa" delay="3ms" duration="20ms" amplitude="0.02nA"/>'
        '<pulseGenerator id="pb" delay="8ms" duration="4ms" amplitude="-0.01nA"/>'
        '<pulseGenerator id="hs" delay="2ms" duration="100ms" amplitude="0.1nA"/>'
        '<poissonFiringSynapse id="pf" averageRate="50 Hz" synapse="e2a" spikeTarget="./e2a"/>'
        '<expTwoSynapse id="e2a" gbase="1nS" erev="0mV" tauRise="0.5ms" tauDecay="3ms"/>'
        '<expTwoSynapse id="e2b" gbase="2nS" erev="-20mV" tauRise="1ms" tauDecay="6ms"/>'
        '<network id="n"><population id="hh" component="hhcell" size="1"/><population id="pp" component="P" size="7"/>'
        '<explicitInput target="hh[0]" input="hs"/>'
        '<explicitInput target="pp[0]" input="pa"/><explicitInput target="pp[1]" input="pb"/>'
        '<inputList id="L" component="pg" population="pp">'
        '<input id="7" target="../pp/2/P" destination="synapses"/>'
        '<input id="3" target="../pp/3/P" destination="synapses"/></inputList>'
        '<inputList id="Q" component="pf" population="pp"><input id="0" target="../pp/6/P" destination="synapses"/>'
        '</inputList>'
        '<projection id="A" presynapticPopulation="hh" postsynapticPopulation="pp" synapse="e2a">'
        '<connection id="0" preCellId="../hh/0/hhcell" postCellId="../pp/4/P"/></projection>'
        '<projection id="B" presynapticPopulation="hh" postsynapticPopulation="pp" synapse="e2b">'
        '<connection id="0" preCellId="../hh/0/hhcell" postCellId="../pp/5/P"/></projection>'
        '</network></neuroml>'
        '<Simulation id="s" length="60ms" step="0.025ms" target="n" seed="5">'
        '<OutputFile id="o" fileName="o.txt">' +
        "".join(f'<OutputColumn id="c{j}" quantity="pp/{j}/P/0/v"/>' for j in range(6)) +
        '</OutputFile>' + extra_sim + '<EdenCustomSetup filename="setup.txt"/></Simulation></Lems>')
    return neuroml.read(str(p))


def test_native_pulse_inputs_set_per_input(tmp_path):
    m, _ = _native_model(tmp_path, "# two pulses made equal to pa and pb\n\n"
                                    "set input L 7,3 delay multi ms\nvalues 3 8\n"
                                    "set input L all duration multi ms\nvalues 20 4\n"
                                    "set input L 7 amplitude 20 pA\nset input L 3 amplitude -1e-11 A\n")
    assert m.lems is None        # expTwoSynapse projections only: the native path
    tr, _, _, _ = oracle.run(m)
    assert np.array_equal(tr[:, 2], tr[:, 0]) and np.array_equal(tr[:, 3], tr[:, 1])
    assert float(np.ptp(tr[:, 2])) > 1.0


def test_poisson_rate_set_per_input(tmp_path):
    m, _ = _native_model(tmp_path, "set input Q 0 averageRate 0.2 kHz\n")
    assert list(m.poisson_rate) == [np.float32(200.0 * 1e-6)]


def test_synapse_parameters_set_per_connection(tmp_path):
    # projection A's connection gets e2b's parameters: pp[4] then follows pp[5], whose e2b
    # synapse is driven by the same presynaptic spikes
    m, _ = _native_model(tmp_path, "set synapse A all post gbase 2 nS\nset synapse A 0 post erev -0.02 V\n"
                                    "set synapse A 0 post tauRise multi ms\nvalues 1\n"
                                    "set synapse A all post tauDecay 6000 us\n")
    assert m.lems is not None and np.count_nonzero(m.syn_param_index >= 0) == 1
    tr, ec, er, _ = oracle.run(m)
    assert np.array_equal(tr[:, 4], tr[:, 5])
    assert float(np.ptp(tr[:, 4])) > 0.1


def test_gap_junction_ends(tmp_path):
    (tmp_path / "g.txt").write_text("set synapse G 1 pre conductance 4 nS\n")
    shutil.copy(os.path.join(GOLD, "NML2_SingleCompHHCell.nml"), tmp_path / "NML2_SingleCompHHCell.nml")
    p = tmp_path / "gj.xml"
    p.write_text(
        '<Lems><Target component="s"/><Include file="NML2_SingleCompHHCell.nml"/><neuroml>'
        '<gapJunction id="gj" conductance="2nS"/><pulseGenerator id="hs" delay="2ms" duration="100ms" amplitude="0.1nA"/>'
        '<network id="n"><population id="hh" component="hhcell" size="3"/><explicitInput target="hh[0]" input="hs"/>'
        '<electricalProjection id="G" presynapticPopulation="hh" postsynapticPopulation="hh">'
        '<electricalConnection id="0" preCell="0" preSegment="0" postCell="1" postSegment="0" synapse="gj"/>'
        '<electricalConnection id="1" preCell="0" preSegment="0" postCell="2" postSegment="0" synapse="gj"/>'
        '</electricalProjection></network></neuroml>'
        '<Simulation id="s" length="20ms" step="0.025ms" target="n" seed="5">'
        '<OutputFile id="o" fileName="o.txt"><OutputColumn id="a" quantity="hh/0/hhcell/0/v"/></OutputFile>'
        '<EdenCustomSetup filename="g.txt"/></Simulation></Lems>')
    m, _ = neuroml.read(str(p))
    # connection 1's pre end: on hh[0] (which holds the pre ends of both connections), its
    # peer hh[2]
    on = np.flatnonzero(m.syn_param_index >= 0)
    assert len(on) == 1
    i = int(on[0])
    assert np.searchsorted(m.comp_syn_begin, i, side="right") - 1 == 0            # on hh[0]
    assert int(m.syn_peer[i]) == 2                                                 # its peer: hh[2]
    assert m.syn_param_vecs[0][0] == np.float32(4e-3)                              # uS


@pytest.mark.parametrize("text, msg", [
    ("set cell pop0 1,1 all b 0.2\n", "more than once"),
    ("set cell pop0 9 all b 0.2\n", "could not resolve id 9"),
    ("set cell pop0 all all d multi\nvalues 1 2\n", "5 values were expected"),
    ("set cell pop0 all all d multi\n", "no 'values' line"),
    ("set cell pop0 all all v0 -60\n", "needs units"),
    ("set cell pop0 all all v0 -60 furlong\n", "unknown voltage unit"),
    ("set cell pop0 all all v0 cable\n", "not supported"),
    ("set cell pop0 all all U 0.5\n", "state variable"),
    ("set cell pop0 all 0 b 0.5\n", "must be 'all'"),
    ("set cell popp all all b 0.5\n", "physical cells"),
    ("set cell nopop all all b 0.5\n", "not found"),
    ("set synapse ToPopp_ExpWild all pre erev 0 mV\n", "'pre' specifier"),
    ("set synapse ToPopp_ExpWild all post nonesuch 0 mV\n", "no parameter nonesuch"),
    ("set input Input0 all amplitude pop0[0]/v\n", "references to state variables"),
    ("set widget x 0 y 1\n", "unknown type"),
    ("unset cell pop0 all all b 1\n", "unknown statement"),
])
def test_grammar_errors(tmp_path, text, msg):
    src = open(os.path.join(GOLD, "EdenTest_Extension_CustomSetup.nml")).read()
    (tmp_path / "m.nml").write_text(src)
    (tmp_path / "EdenTest_Extension_CustomSetup.txt").write_text("# a comment\n\n" + text)
    with pytest.raises(neuroml.NeuroMLError, match=re.escape(msg)):
        neuroml.read(str(tmp_path / "m.nml"))


def test_tokenizer_skips_comments_and_blank_lines():
    st = custom_setup.parse("  # c\n\n\tset  input L 1,2 delay multi ms  \nvalues 1 2\n#x\nset input L all delay 3 ms\n")
    assert [(s.kind, s.group, s.targets, s.param, s.multi, s.values, s.line) for s in st] == [
        ("input", "L", [1, 2], "delay", True, ["1", "2"], 3), ("input", "L", None, "delay", False, None, 6)]
