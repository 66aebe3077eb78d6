"""Generic LEMS synapse types on the GPU: the library compiles the emitted functions with
hiprtc into its synapse kernel at context creation. Against the oracle running the C dialect
of the same text: traces, spike raster and every synaptic delivery bit for bit.

* the reference's CoreSynapses model (tests/coresyn.py): one synapse of each core type
  (the blocking/plastic one included), graded and electrical connections, through the
  NeuroML reader and the output files;
* expTwoSynapse as a generic type equals the library's native expTwo path;
* a network of tree and point cells with generic synapses on dendrites and somata;
* gap junctions (a pair of membranes; the reference's DomainDecomposition model, whose
  electrical projections join compartments of different tree cells);
* user-defined synapse ComponentTypes read from LEMS;
* LEMS current inputs (the reference's CoreInputs and EdenTest4 models; generators and a
  voltage clamp beside native inputs on point and tree compartments).
"""
import dataclasses
import os
import sys

import numpy as np
import pytest

from eden_amd import configs, lems, neuroml, output
from eden_amd.model import Simulation
from eden_amd.run_sim import run_lems
from eden_text import assert_same_text, trajectory_text
from oracle import oracle

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _gpu(m):
    with Simulation(m, log_synapse_events=1_000_000, profile=True) as sim:
        tr = sim.run_all()
        ev = sim.read_events()
        st, sy = sim.read_synapse_events()
        names = [k["name"] for k in sim.kernel_infos()]
    return tr, ev, (st, sy), names


def _check(m):
    tr, ev, (st, sy), names = _gpu(m)
    (rt, ec, er, _), (ost, osy) = oracle.run_with_synapse_log(m)
    _say(f"[lems] {m.name}: kernels {names}, spikes {len(er)}, synaptic events {len(ost)}")
    assert np.array_equal(tr, rt), float(np.abs(tr - rt).max())
    assert np.array_equal(ev[0], ec) and np.array_equal(ev[1], er)
    assert np.array_equal(st, ost) and np.array_equal(sy, osy)
    return tr, er, ost


def test_reference_core_synapses(tmp_path):
    from coresyn import write_model
    path = write_model(str(tmp_path))
    m, outs = neuroml.read(path)
    tr, er, ost = _check(m)
    assert sorted(set(oracle.run_with_synapse_log(m)[1][1].tolist())) == list(range(11))
    res = run_lems(path, out_dir=str(tmp_path))
    t_us = output.row_times_us(m.dt, m.total_rows() + 1)
    f = outs.trajectories[0]
    assert_same_text((tmp_path / f.file_name).read_text(), trajectory_text(t_us, tr[:, [c[2] for c in f.columns]]))
    _say(f"[lems] CoreSynapses: {res['rows']} rows, run {res['run_s']:.3f} s")


def test_generic_exptwo_equals_native_on_gpu():
    m = configs.mixed_network(tree_sizes=(40, 90, 300), n_point=16, n_conn=200, t_final_ms=60.0, seed=12)
    types = [("expTwoSynapse", dict(gbase=st.gbase, erev=st.erev, tauRise=st.tau_rise, tauDecay=st.tau_decay),
              st.weight_prop) for st in m.syn_types]
    g = dataclasses.replace(m, lems=lems.build(types), name="mixed_generic_exptwo")
    a = _gpu(m)
    b = _gpu(g)
    assert any("lems" in n or True for n in b[3])
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1][1], b[1][1]) and np.array_equal(a[2][0], b[2][0])


@pytest.mark.parametrize("hh", [False, True])
def test_tree_and_point_network_with_generic_synapses(hh):
    m = configs.mixed_network(tree_sizes=(40, 150, 700), n_point=20, n_conn=300, t_final_ms=60.0, seed=30 + hh,
                              hh_everywhere=hh)
    types = [("alphaSynapse", dict(gbase=4e-3, erev=0.0, tau=1500.0), 1.0),
             ("expThreeSynapse", dict(gbase1=3e-3, gbase2=2e-3, erev=-80.0, tauRise=400.0, tauDecay1=2000.0,
                                      tauDecay2=6000.0), 1.0)]
    g = dataclasses.replace(m, lems=lems.build(types), name=f"mixed_generic_{hh}")
    tr, er, ost = _check(g)
    assert len(ost) > 50


def test_gap_junction_pair(tmp_path):
    """Two membranes joined by a gap junction (tests/test_lems_cpu.py pins the oracle's
    closed form): the GPU equals the oracle bit for bit."""
    from test_lems_cpu import _gap_pair_file
    m, _ = neuroml.read(_gap_pair_file(tmp_path, -65.0, -55.0, 200.0))
    tr, _, _ = _check(m)
    assert abs(float(tr[-1, 0] - tr[-1, 1])) < 5.0


def test_reference_domain_decomposition_model(tmp_path):
    """The reference's EdenTest_DomainDecomposition model: tree cells joined by gap junctions
    between compartments of different cells, beside chemical synapses; output file byte-equal
    to the oracle's run of the same model."""
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "neuroml",
                        "LEMS_EdenTest_DomainDecomposition.xml")
    m, outs = neuroml.read(path)
    assert m.syn_peer is not None and int((m.syn_peer >= 0).sum()) >= 2
    tr, er, ost = _check(m)
    res = run_lems(path, out_dir=str(tmp_path))
    t_us = output.row_times_us(m.dt, m.total_rows() + 1)
    for f in outs.trajectories:
        assert_same_text((tmp_path / f.file_name).read_text(), trajectory_text(t_us, tr[:, [c[2] for c in f.columns]]))
    _say(f"[lems] domain decomposition: {m.n_cells} cells, {m.n_comps} compartments, {res['rows']} rows")


def test_user_defined_component_types(tmp_path):
    """User ComponentTypes (an exponential synapse and a gap junction written as LEMS
    dynamics) compiled by hiprtc: the GPU equals the oracle, and the user types' cells equal
    the core types' cells."""
    from test_lems_cpu import _user_types_file
    m, _ = neuroml.read(_user_types_file(tmp_path))
    tr, er, ost = _check(m)
    assert len(ost) > 2
    assert np.array_equal(tr[:, 1], tr[:, 2]) and np.array_equal(tr[:, 3:5], tr[:, 5:7])


def test_reference_core_inputs_files(tmp_path):
    """The reference's CoreInputs model (tests/coreinputs.py): LEMS current inputs (sine, ramp,
    voltage clamps, dimensionless generators on Izhikevich cells), its timedSynapticInput and
    spike sources; output files byte-equal to the oracle's, trace and raster equal."""
    from coreinputs import write_model
    path = write_model(str(tmp_path))
    res = run_lems(path, out_dir=str(tmp_path))
    m, outs = res["model"], res["outputs"]
    _say(f"[lems] CoreInputs: {m.n_cells} cells, types {m.lems.type_names}, {res['rows']} rows")
    tr_r, ec, er, _ = oracle.run(m)
    t_us = output.row_times_us(m.dt, m.total_rows() + 1)
    f = outs.trajectories[0]
    assert_same_text((tmp_path / f.file_name).read_text(), trajectory_text(t_us, tr_r[:, [c[2] for c in f.columns]]))
    tr, ev, _, _ = _gpu(m)
    assert np.array_equal(tr, tr_r)
    assert np.array_equal(ev[0], ec) and np.array_equal(ev[1], er)


def test_reference_custom_input_model_on_tree_cells():
    """The reference's EdenTest4 (two-compartment cells with a user-defined pulse input, user
    gap junction and synapse types beside the core ones): the GPU equals the oracle."""
    m, _ = neuroml.read(os.path.join(HERE, "golden", "neuroml", "LEMS_EdenTest4.xml"))
    tr, er, ost = _check(m)
    assert float(np.ptp(tr[:, 4])) > 0.5


@pytest.mark.parametrize("where", ["point", "tree"])
def test_lems_inputs_beside_native_inputs(tmp_path, where):
    """A sine and a ramp generator and a voltage clamp beside native pulse and Poisson inputs,
    on one-compartment cells and on the compartments of tree cells: bit for bit."""
    m = configs.mixed_network(tree_sizes=(40, 90), n_point=6, n_conn=40, t_final_ms=40.0, seed=21)
    types = [("expTwoSynapse", dict(gbase=st.gbase, erev=st.erev, tauRise=st.tau_rise, tauDecay=st.tau_decay),
              st.weight_prop) for st in m.syn_types]
    types += [("sineGenerator", dict(phase=0.1, delay=5000.0, duration=20000.0, amplitude=0.05, period=3000.0), 1.0),
              ("rampGenerator", dict(delay=2000.0, duration=30000.0, startAmplitude=0.0, finishAmplitude=0.08,
                                     baselineAmplitude=0.01), 1.0),
              ("voltageClamp", dict(delay=10000.0, duration=10000.0, simpleSeriesResistance=300.0,
                                    targetVoltage=-30.0), 1.0)]
    n_syn_types = len(m.syn_types)
    ls = lems.build(types)
    # three input instances on chosen compartments, appended after each compartment's synapses
    cb = m.cell_comp_begin
    comps = [int(cb[2]), int(cb[3]), int(cb[4])] if where == "point" else [int(cb[0]) + 5, int(cb[1]) + 17,
                                                                               int(cb[0])]
    per = [list(range(m.comp_syn_begin[g], m.comp_syn_begin[g + 1])) for g in range(m.n_comps)]
    old_type, old_w, old_d = m.syn_type.copy(), m.syn_weight.copy(), m.syn_delay.copy()
    new_type, new_w, new_d, remap = [], [], [], {}
    extra = {g: [(n_syn_types + k, w) for k, (gg, w) in enumerate(zip(comps, (1.0, 0.7, 1.0))) if gg == g]
             for g in comps}
    for g in range(m.n_comps):
        for s in per[g]:
            remap[s] = len(new_type)
            new_type.append(old_type[s]); new_w.append(old_w[s]); new_d.append(old_d[s])
        for ty, w in extra.get(g, []):
            new_type.append(ty); new_w.append(w); new_d.append(0.0)
    counts = [len(per[g]) + len(extra.get(g, [])) for g in range(m.n_comps)]
    begin = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    g2 = dataclasses.replace(m, comp_syn_begin=begin, syn_type=np.array(new_type, np.int32),
                             syn_weight=np.array(new_w, np.float32), syn_delay=np.array(new_d, np.float32),
                             spike_target=np.array([remap[int(s)] for s in m.spike_target], np.int64),
                             lems=ls, name=f"lems_inputs_{where}")
    tr, er, ost = _check(g2)
    base = _check(dataclasses.replace(m, lems=lems.build(types[:n_syn_types]), name="no_inputs"))[0]
    assert not np.array_equal(tr, base)


@pytest.mark.parametrize("opts", [{}, dict(use_graph=True)])
def test_timed_synaptic_input(tmp_path, opts):
    """timedSynapticInput: spike lists firing their synapse instances in the same step, the
    position state advancing per step (two cells, one list with two times in one step):
    traces and the synaptic event log bit for bit, in the plain and graph launch modes, and
    after a reset the same run again."""
    from test_lems_cpu import _timed_model
    m = _timed_model(tmp_path, [0.003, 0.0102, 1.0, 2.001, 2.004, 3.5], both=True)
    _say(f"[lems] timed input, opts {opts}")
    (rt, ec, er, _), (ost, osy) = oracle.run_with_synapse_log(m)
    with Simulation(m, log_synapse_events=1 << 16, **opts) as sim:
        tr = sim.run_all()
        st, sy = sim.read_synapse_events()
        sim.reset()
        tr2 = sim.run_all()
    assert np.array_equal(tr, rt) and np.array_equal(tr2, rt)
    assert np.array_equal(st, ost) and np.array_equal(sy, osy)
    assert sorted(st[sy == 0].tolist()) == [1, 2, 101, 201, 351]


@pytest.mark.parametrize("opts", [{}, dict(use_graph=True)])
def test_poisson_firing_inputs_on_the_generic_path(tmp_path, opts):
    """A native poissonFiringSynapse beside a generic one (expOneSynapse), a
    transientPoissonFiringSynapse on two cells, and a timedSynapticInput: the spikers' RNG
    draws, the synaptic events and the traces bit for bit, in the plain and graph launch
    modes, and again after a reset."""
    from test_firing_inputs_cpu import _list, _model
    m = _model(tmp_path,
               '<poissonFiringSynapse id="pn" averageRate="200 Hz" synapse="e2" spikeTarget="./e2"/>'
               '<poissonFiringSynapse id="pg" averageRate="150 Hz" synapse="e1" spikeTarget="./e1"/>'
               '<transientPoissonFiringSynapse id="tp" averageRate="300 Hz" delay="50ms" duration="100ms" '
               'synapse="e2" spikeTarget="./e2"/>'
               '<timedSynapticInput id="ti" synapse="e1" spikeTarget="./e1"><spike id="0" time="5 ms"/>'
               '<spike id="1" time="7.5 ms"/></timedSynapticInput>',
               _list("N", "pn", [0]) + _list("G", "pg", [0, 1]) + _list("T", "tp", [0, 1]) + _list("S", "ti", [1]),
               length_ms=200.0)
    _say(f"[lems] firing inputs, opts {opts}: spikers {len(m.poisson_spikers['rate'])}")
    (rt, ec, er, _), (ost, osy) = oracle.run_with_synapse_log(m)
    with Simulation(m, log_synapse_events=1 << 16, **opts) as sim:
        tr = sim.run_all()
        st, sy = sim.read_synapse_events()
        sim.reset()
        tr2 = sim.run_all()
    assert np.array_equal(tr, rt) and np.array_equal(tr2, rt)
    assert np.array_equal(st, ost) and np.array_equal(sy, osy)
    assert len(ost) > 50
