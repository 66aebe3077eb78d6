"""LEMS point neurons on the GPU: the library compiles the emitted cell functions with hiprtc
into lems_cell_kernel at context creation. Against the oracle running the C dialect of the
same text: traces, spike rasters and synaptic deliveries bit for bit.

* the reference's ArtificialCells model in full (tests/artcells.py), through the NeuroML
  reader and the output files;
* a network of point neurons and physical cells: an HH cell driving integrate-and-fire,
  Izhikevich and adaptive exponential cells through synapses, point neurons driving passive
  cells, spike sources (spikeArray, spikeGenerator) driving both kinds, a gap junction
  between a point neuron and a physical cell, state variables recorded, in the plain, graph
  and chunked launch modes.
"""
import os
import shutil
import sys

import numpy as np
import pytest

from eden_amd import neuroml, output
from eden_amd.model import Simulation
from eden_amd.run_sim import run_lems
from eden_text import assert_same_text, events_text, trajectory_text
from oracle import oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "neuroml")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _gpu(m, **kw):
    with Simulation(m, log_synapse_events=1 << 20, **kw) as sim:
        tr = sim.run_all()
        ev = sim.read_events()
        syn = sim.read_synapse_events()
    return tr, ev, syn


def test_reference_artificial_cells_files(tmp_path):
    from artcells import write_model
    path = write_model(str(tmp_path))
    res = run_lems(path, out_dir=str(tmp_path))
    m, outs = res["model"], res["outputs"]
    _say(f"[cells] ArtificialCells: {m.n_cells} cells, {len(m.lems_cells.type_names)} point-neuron types, "
         f"{res['rows']} rows")
    tr_r, ec, er, _ = oracle.run(m)
    t_us = output.row_times_us(m.dt, m.total_rows() + 1)
    scales = output.column_scales(m)
    for f in outs.trajectories:
        cols = [c[2] for c in f.columns]
        assert_same_text((tmp_path / f.file_name).read_text(), trajectory_text(t_us, tr_r[:, cols], scales[cols]))
    for f in outs.events:
        cols = [s[2] for s in f.selections]
        keep = np.isin(ec, cols)
        remap = {c: i for i, c in enumerate(cols)}
        fc = np.array([remap[int(c)] for c in ec[keep]], np.int64)
        assert_same_text((tmp_path / f.file_name).read_text(),
                         events_text(t_us, er[keep], fc, [s[0] for s in f.selections], f.format))
    tr, ev, _ = _gpu(m)
    assert np.array_equal(tr, tr_r)
    assert np.array_equal(ev[0], ec) and np.array_equal(ev[1], er)
    assert len(er) > 20


def _network_file(tmp_path):
    shutil.copy(os.path.join(GOLD, "NML2_SingleCompHHCell.nml"), tmp_path / "NML2_SingleCompHHCell.nml")
    pas = ('<cell id="{n}"><morphology id="m{n}"><segment id="0"><proximal x="0" y="0" z="0" diameter="10"/>'
           '<distal x="0" y="0" z="0" diameter="10"/></segment></morphology><biophysicalProperties id="b{n}">'
           '<membraneProperties><channelDensity id="leak" ionChannel="passiveChan" condDensity="0.3 mS_per_cm2" '
           'erev="-65mV"/><specificCapacitance value="1.0 uF_per_cm2"/><initMembPotential value="-65mV"/>'
           '</membraneProperties><intracellularProperties><resistivity value="0.1 kohm_cm"/>'
           '</intracellularProperties></biophysicalProperties></cell>')
    cells = (pas.format(n="P1") + pas.format(n="P2") +
             '<iafCell id="iaf" leakReversal="-65mV" thresh="-52mV" reset="-70mV" C="0.2nF" leakConductance="0.01uS"/>'
             '<izhikevich2007Cell id="iz" v0="-60mV" C="100 pF" k="0.7 nS_per_mV" vr="-60 mV" vt="-40 mV" '
             'vpeak="35 mV" a="0.03 per_ms" b="-2 nS" c="-50 mV" d="100 pA"/>'
             '<adExIaFCell id="adex" C="281pF" gL="30nS" EL="-70.6mV" reset="-48.5mV" VT="-50.4mV" thresh="-40.4mV" '
             'refract="0.2ms" delT="2mV" tauw="40ms" a="4nS" b="0.08nA"/>'
             '<spikeArray id="sa"><spike id="0" time="12 ms"/><spike id="1" time="30.01 ms"/><spike id="2" time="31 ms"/>'
             '</spikeArray><spikeGenerator id="sg" period="7ms"/>')
    syn = ('<expOneSynapse id="e1" gbase="3nS" erev="0mV" tauDecay="3ms"/>'
           '<expTwoSynapse id="e2" gbase="2nS" erev="0mV" tauRise="0.5ms" tauDecay="5ms"/>'
           '<alphaSynapse id="a1" gbase="1nS" erev="0mV" tau="2ms"/>'
           '<gapJunction id="gj" conductance="2nS"/>'
           '<pulseGenerator id="hhstim" delay="5ms" duration="200ms" amplitude="0.08nA"/>'
           '<pulseGenerator id="pstim" delay="20ms" duration="100ms" amplitude="0.15nA"/>')
    pops = "".join(f'<population id="p{n}" component="{n}" size="{k}"/>'
                   for n, k in (("P1", 2), ("P2", 1), ("iaf", 3), ("iz", 2), ("adex", 1), ("sa", 1), ("sg", 1)))
    conn = ('<projection id="{id}" presynapticPopulation="{pre}" postsynapticPopulation="{post}" synapse="{syn}">'
            '{c}</projection>')
    cw = '<connectionWD id="{k}" preCellId="../{pre}/{i}/{pc}" postCellId="../{post}/{j}/{qc}" weight="{w}" delay="{d}ms"/>'
    net = (
        conn.format(id="c1", pre="hhpop", post="piaf", syn="e1",
                    c="".join(cw.format(k=j, pre="hhpop", i=0, pc="hhcell", post="piaf", j=j, qc="iaf", w=4, d=1 + j)
                              for j in range(3)))
        + conn.format(id="c2", pre="hhpop", post="piz", syn="a1",
                      c="".join(cw.format(k=j, pre="hhpop", i=0, pc="hhcell", post="piz", j=j, qc="iz", w=20, d=2)
                                for j in range(2)))
        + conn.format(id="c3", pre="piaf", post="pP1", syn="e2",
                      c="".join(cw.format(k=j, pre="piaf", i=j, pc="iaf", post="pP1", j=j % 2, qc="P1", w=3, d=1)
                                for j in range(3)))
        + conn.format(id="c4", pre="piz", post="padex", syn="e2",
                      c=cw.format(k=0, pre="piz", i=0, pc="iz", post="padex", j=0, qc="adex", w=30, d=0.5))
        + conn.format(id="c5", pre="psa", post="pP2", syn="e2",
                      c=cw.format(k=0, pre="psa", i=0, pc="sa", post="pP2", j=0, qc="P2", w=10, d=1))
        + conn.format(id="c6", pre="psg", post="piaf", syn="e1",
                      c=cw.format(k=0, pre="psg", i=0, pc="sg", post="piaf", j=2, qc="iaf", w=5, d=0))
        + '<electricalProjection id="g1" presynapticPopulation="piz" postsynapticPopulation="pP2">'
          '<electricalConnection id="0" preCell="1" preSegment="0" postCell="0" postSegment="0" synapse="gj"/>'
          '</electricalProjection>'
        + '<explicitInput target="hhpop[0]" input="hhstim"/><explicitInput target="piaf[1]" input="pstim"/>'
          '<explicitInput target="padex[0]" input="pstim"/>')
    cols = ('<OutputColumn id="hh" quantity="hhpop/0/hhcell/0/v"/>'
            + "".join(f'<OutputColumn id="iaf{j}" quantity="piaf[{j}]/v"/>' for j in range(3))
            + '<OutputColumn id="iz0" quantity="piz/0/iz/v"/><OutputColumn id="u0" quantity="piz/0/iz/u"/>'
              '<OutputColumn id="iz1" quantity="piz[1]/v"/><OutputColumn id="ad" quantity="padex[0]/v"/>'
              '<OutputColumn id="w" quantity="padex[0]/w"/><OutputColumn id="ls" quantity="padex[0]/lastSpikeTime"/>'
              '<OutputColumn id="p1" quantity="pP1/0/P1/0/v"/><OutputColumn id="p2" quantity="pP2/0/P2/0/v"/>')
    evs = "".join(f'<EventSelection id="{k}" select="{s}" eventPort="spike"/>' for k, s in enumerate(
        ["hhpop[0]", "piaf[0]", "piaf[1]", "piaf[2]", "piz[0]", "piz[1]", "padex[0]", "psa[0]", "psg[0]"]))
    p = tmp_path / "cells_net.xml"
    p.write_text(
        '<Lems><Target component="s"/><Include file="NML2_SingleCompHHCell.nml"/><neuroml>' + syn + cells +
        '<network id="n2"><population id="hhpop" component="hhcell" size="1"/>' + pops + net + '</network>'
        '</neuroml><Simulation id="s" length="150ms" step="0.025ms" target="n2" seed="3">'
        f'<OutputFile id="o" fileName="o.txt">{cols}</OutputFile>'
        f'<EventOutputFile id="e" fileName="e.txt" format="TIME_ID">{evs}</EventOutputFile>'
        '</Simulation></Lems>')
    return str(p)


@pytest.mark.parametrize("opts", [{}, dict(use_graph=True), dict(chunk_steps=777)])
def test_point_neuron_network_bit_exact(tmp_path, opts):
    m, outs = neuroml.read(_network_file(tmp_path))
    _say(f"[cells] network: {m.n_cells} cells, types {m.lems_cells.type_names}, opts {opts}")
    (tr_r, ec, er, _), (ost, osy) = oracle.run_with_synapse_log(m)
    tr, ev, (st, sy) = _gpu(m, **opts)
    assert np.array_equal(tr, tr_r), float(np.nanmax(np.abs(tr - tr_r)))
    assert np.array_equal(ev[0], ec) and np.array_equal(ev[1], er)
    assert np.array_equal(st, ost) and np.array_equal(sy, osy)
    counts = np.bincount(ec, minlength=9)
    assert counts[0] > 0 and counts[1:7].sum() > 5           # the HH cell and the point neurons fire
    assert counts[7] == 3 and counts[8] == 21                 # the spike sources' listed and periodic spikes
    assert float(np.ptp(tr[:, 5])) > 0.01                      # iz's u moves
    assert float(np.abs(tr[:, 11] + 65.0).max()) > 0.05        # the gap junction moves P2


@pytest.mark.parametrize("opts", [{}, dict(use_graph=True)])
def test_random_spike_sources_bit_exact(tmp_path, opts):
    """spikeGeneratorRandom, spikeGeneratorPoisson, spikeGeneratorRefPoisson and
    SpikeSourcePoisson (random() stamped with each cell's seed and the step), driving an
    integrate-and-fire cell through a synapse: spike rasters, recorded state and the target's
    trace bit for bit."""
    p = tmp_path / "rs.xml"
    p.write_text(
        '<Lems><Target component="s"/><neuroml>'
        '<spikeGeneratorRandom id="r" minISI="10ms" maxISI="30ms"/>'
        '<spikeGeneratorPoisson id="p" averageRate="40 Hz"/>'
        '<spikeGeneratorRefPoisson id="q" averageRate="40 Hz" minimumISI="15ms"/>'
        '<SpikeSourcePoisson id="s" start="100ms" duration="150ms" rate="60 Hz"/>'
        '<iafCell id="iaf" leakReversal="-65mV" thresh="-52mV" reset="-70mV" C="0.2nF" leakConductance="0.01uS"/>'
        '<expOneSynapse id="e1" gbase="2nS" erev="0mV" tauDecay="3ms"/>'
        '<network id="n"><population id="a" component="r" size="2"/><population id="b" component="p" size="2"/>'
        '<population id="c" component="q" size="1"/><population id="d" component="s" size="1"/>'
        '<population id="x" component="iaf" size="1"/>'
        + "".join(f'<projection id="j{k}" presynapticPopulation="{pp}" postsynapticPopulation="x" synapse="e1">'
                  f'<connectionWD id="0" preCellId="../{pp}/0/{c}" postCellId="../x/0/iaf" weight="3" delay="1ms"/>'
                  '</projection>' for k, (pp, c) in enumerate((("a", "r"), ("b", "p"), ("c", "q"), ("d", "s"))))
        + '</network></neuroml><Simulation id="s" length="400ms" step="0.025ms" target="n" seed="9">'
        '<OutputFile id="o" fileName="o.txt"><OutputColumn id="v" quantity="x[0]/v"/>'
        '<OutputColumn id="i0" quantity="a[1]/isi"/><OutputColumn id="i1" quantity="b[1]/tnextUsed"/></OutputFile>'
        '<EventOutputFile id="e" fileName="e.txt" format="TIME_ID">'
        + "".join(f'<EventSelection id="{k}" select="{s}" eventPort="spike"/>'
                  for k, s in enumerate(("a[0]", "a[1]", "b[0]", "b[1]", "c[0]", "d[0]", "x[0]")))
        + '</EventOutputFile></Simulation></Lems>')
    m, _ = neuroml.read(str(p))
    _say(f"[cells] random sources, opts {opts}: types {m.lems_cells.type_names}")
    (tr_r, ec, er, _), (ost, osy) = oracle.run_with_synapse_log(m)
    tr, ev, (st, sy) = _gpu(m, **opts)
    assert np.array_equal(tr, tr_r), float(np.nanmax(np.abs(tr - tr_r)))
    assert np.array_equal(ev[0], ec) and np.array_equal(ev[1], er)
    assert np.array_equal(st, ost) and np.array_equal(sy, osy)
    counts = np.bincount(ec, minlength=7)
    # every source fires (the SpikeSourcePoisson, about 9 spikes in its 150 ms, at least once)
    assert counts[:6].min() > 0 and not np.array_equal(er[ec == 0], er[ec == 1])


def test_pinsky_rinzel_bit_exact(tmp_path):
    """pinskyRinzelCA3Cell (conditional derived variables, powers, dimension factors between
    current densities and specific capacitance): Vs, Vd and hs bit for bit over 500 ms of
    bursting."""
    from test_cells_cpu import pinsky_rinzel_model
    m, _ = neuroml.read(pinsky_rinzel_model(tmp_path, n=3))
    _say("[cells] pinskyRinzelCA3Cell x3")
    tr_r, _, _, _ = oracle.run(m)
    tr, _, _ = _gpu(m)
    assert np.array_equal(tr, tr_r), float(np.nanmax(np.abs(tr - tr_r)))
    assert float(tr[:, 0].max()) > 20.0
