"""Compartments with more channel distributions (5-7) or gate variables (5-7) than the
compartment kernels hold: their channels run in ext_channels_kernel each step, in
distribution order, and the kernels take the totals. Point cells and tree cells (gated
compartments and passive-only dendrites), a synaptic connection between them, in the plain
and graph launch modes: traces, rasters and synaptic events bit for bit against the oracle,
which sums every compartment's channels in distribution order."""
import sys

import numpy as np
import pytest

from eden_amd import neuroml
from eden_amd.model import Simulation
from oracle import oracle

pytestmark = pytest.mark.gpu


def _model(tmp_path, record_gate=False):
    parent = lambda k: "" if k == 0 else '<parent segment="%d"/>' % (k - 1)
    segs = "".join(f'<segment id="{k}">{parent(k)}<proximal x="{20 * k}" y="0" z="0" diameter="4"/>'
                   f'<distal x="{20 * k + 20}" y="0" z="0" diameter="4"/></segment>' for k in range(4))
    segs += ('<segmentGroup id="soma"><member segment="0"/></segmentGroup>'
             '<segmentGroup id="dend"><member segment="1"/><member segment="2"/><member segment="3"/></segmentGroup>')
    many = ('<channelDensity id="l1" ionChannel="leak" condDensity="0.1 mS_per_cm2" erev="-65mV"{sg}/>'
            '<channelDensity id="na" ionChannel="naChan" condDensity="120 mS_per_cm2" erev="50mV"{sg}/>'
            '<channelDensity id="k" ionChannel="kChan" condDensity="36 mS_per_cm2" erev="-77mV"{sg}/>'
            '<channelDensity id="kti" ionChannel="kti" condDensity="2 mS_per_cm2" erev="-80mV"{sg}/>'
            '<channelDensity id="l2" ionChannel="leak" condDensity="0.1 mS_per_cm2" erev="-60mV"{sg}/>'
            '<channelDensity id="l3" ionChannel="leak" condDensity="0.1 mS_per_cm2" erev="-55mV"{sg}/>')
    passive5 = "".join(f'<channelDensity id="p{k}" ionChannel="leak" condDensity="0.05 mS_per_cm2" '
                       f'erev="{-70 + 3 * k}mV" segmentGroup="dend"/>' for k in range(5))
    bio = ('<biophysicalProperties id="b{n}"><membraneProperties>{chans}'
           '<specificCapacitance value="1.0 uF_per_cm2"/><initMembPotential value="-65mV"/></membraneProperties>'
           '<intracellularProperties><resistivity value="0.1 kohm_cm"/></intracellularProperties>'
           '</biophysicalProperties>')
    chans = ('<ionChannelHH id="leak" conductance="10pS"/>'
             '<ionChannelHH id="naChan" conductance="10pS" species="na">'
             '<gateHHrates id="m" instances="3"><forwardRate type="HHExpLinearRate" rate="1per_ms" midpoint="-40mV" '
             'scale="10mV"/><reverseRate type="HHExpRate" rate="4per_ms" midpoint="-65mV" scale="-18mV"/></gateHHrates>'
             '<gateHHrates id="h" instances="1"><forwardRate type="HHExpRate" rate="0.07per_ms" midpoint="-65mV" '
             'scale="-20mV"/><reverseRate type="HHSigmoidRate" rate="1per_ms" midpoint="-35mV" scale="10mV"/>'
             '</gateHHrates></ionChannelHH>'
             '<ionChannelHH id="kChan" conductance="10pS" species="k">'
             '<gateHHrates id="n" instances="4"><forwardRate type="HHExpLinearRate" rate="0.1per_ms" midpoint="-55mV" '
             'scale="10mV"/><reverseRate type="HHExpRate" rate="0.125per_ms" midpoint="-65mV" scale="-80mV"/>'
             '</gateHHrates></ionChannelHH>'
             '<ionChannelHH id="kti" conductance="10pS" species="k">'
             '<gateHHtauInf id="a" instances="2"><timeCourse type="fixedTimeCourse" tau="3ms"/>'
             '<steadyState type="HHSigmoidVariable" rate="1" scale="6mV" midpoint="-50mV"/></gateHHtauInf>'
             '<gateHHtauInf id="b" instances="1"><timeCourse type="fixedTimeCourse" tau="20ms"/>'
             '<steadyState type="HHSigmoidVariable" rate="1" scale="-7mV" midpoint="-70mV"/></gateHHtauInf>'
             '</ionChannelHH>')
    rec = ('<OutputColumn id="g" quantity="pp/0/P/0/b1/membraneProperties/na/naChan/m/q"/>' if record_gate else '')
    p = tmp_path / "mc.xml"
    p.write_text(
        '<Lems><Target component="s"/><neuroml>' + chans +
        '<cell id="P"><morphology id="m1"><segment id="0"><proximal x="0" y="0" z="0" diameter="15"/>'
        '<distal x="0" y="0" z="0" diameter="15"/></segment></morphology>'
        + bio.format(n=1, chans=many.format(sg="")) + '</cell>'
        f'<cell id="T"><morphology id="m2">{segs}</morphology>'
        + bio.format(n=2, chans=many.format(sg=' segmentGroup="soma"') + passive5) + '</cell>'
        '<expTwoSynapse id="syn" gbase="3nS" erev="0mV" tauRise="0.5ms" tauDecay="4ms"/>'
        '<pulseGenerator id="pg" delay="5ms" duration="80ms" amplitude="0.3nA"/>'
        '<network id="n" temperature="6.3 degC"><population id="pp" component="P" size="2"/>'
        '<population id="tt" component="T" size="2"/>'
        '<explicitInput target="pp[0]" input="pg"/><explicitInput target="tt[0]" input="pg"/>'
        '<projection id="pr" presynapticPopulation="pp" postsynapticPopulation="tt" synapse="syn">'
        '<connectionWD id="0" preCellId="../pp/0/P" postCellId="../tt/1/T" postSegmentId="2" weight="4" delay="2ms"/>'
        '</projection></network></neuroml>'
        '<Simulation id="s" length="100ms" step="0.025ms" target="n" seed="1">'
        '<OutputFile id="o" fileName="o.txt"><OutputColumn id="v0" quantity="pp/0/P/0/v"/>'
        '<OutputColumn id="v1" quantity="pp/1/P/0/v"/><OutputColumn id="t0" quantity="tt/0/T/0/v"/>'
        '<OutputColumn id="t3" quantity="tt/0/T/3/v"/><OutputColumn id="u2" quantity="tt/1/T/2/v"/>' + rec +
        '</OutputFile><EventOutputFile id="e" fileName="e.txt" format="TIME_ID">'
        '<EventSelection id="0" select="pp[0]" eventPort="spike"/><EventSelection id="1" select="tt[0]" eventPort="spike"/>'
        '</EventOutputFile></Simulation></Lems>')
    return neuroml.read(str(p))[0]


@pytest.mark.parametrize("opts", [{}, dict(use_graph=True)])
def test_many_channels_bit_exact(tmp_path, opts):
    m = _model(tmp_path)
    print(f"[many-channels] {m.n_cells} cells, {m.n_comps} compartments, "
          f"channels per compartment {np.diff(m.comp_chan_begin).tolist()}, opts {opts}", flush=True)
    sys.stdout.flush()
    (tr_r, ec, er, _), (ost, osy) = oracle.run_with_synapse_log(m)
    with Simulation(m, log_synapse_events=1 << 16, **opts) as sim:
        tr = sim.run_all()
        ev = sim.read_events()
        st, sy = sim.read_synapse_events()
        names = [k["name"] for k in sim.kernel_infos()]
    assert np.array_equal(tr, tr_r), float(np.nanmax(np.abs(tr - tr_r)))
    assert np.array_equal(ev[0], ec) and np.array_equal(ev[1], er)
    assert np.array_equal(st, ost) and np.array_equal(sy, osy)
    counts = np.bincount(ec, minlength=2)
    assert counts[0] > 2 and counts[1] > 0 and len(ost) > 2      # spikes, and deliveries to tt[1]
    assert float(np.ptp(tr[:, 4])) > 0.5


def test_gate_columns_of_such_compartments(tmp_path):
    """A gate variable of a compartment whose channels run in the pre-pass, recorded from the
    pre-pass's gate table: bit for bit the oracle's."""
    m = _model(tmp_path, record_gate=True)
    tr_r, _, _, _ = oracle.run(m)
    with Simulation(m) as sim:
        tr = sim.run_all()
    assert np.array_equal(tr, tr_r), float(np.nanmax(np.abs(tr - tr_r)))
    assert float(np.ptp(tr[:, 5])) > 0.1
