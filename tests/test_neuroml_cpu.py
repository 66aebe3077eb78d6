"""The NeuroML / LEMS reader (eden_amd.neuroml), the exporter (eden_amd.nml_export) and the
command-line front end, on the CPU.

* The reference's own example model (examples/NML2_SingleCompHHCell.nml, kept under
  tests/golden/neuroml/ as input data, with a LEMS simulation file of this repository's)
  instantiates to the numbers its elements state, computed here independently: sphere area
  pi d^2, C = Cm A, densities in mS/cm^2, rates in 1/us, the pulse in us and nA
  (eden/NeuroML.cpp:305-425 fixes the engine units).
* Quantity conversion follows ParseQuantity's arithmetic (eden/NeuroML.cpp:3046-3077).
* Every synthetic workload written as LEMS by the exporter reads back into the same
  instantiated model, array for array (spike recipients per source as a set: their order
  does not change results, flags being idempotent).
* The reference's full EdenTest_CoreSynapses model (blocking/plastic synapse, continuous and
  electrical projections) reads in the reference's instance order, each continuous
  connection's postsynaptic component first, each end with the other as its peer.
* Cable discretisation: the reference's EdenTest_CableDiscretisation model gives identical
  compartments for a segment cut in five and five segments; tapering cables of uneven
  segments cut into 1, 3 and 7 compartments match the cone-section closed forms; locations
  by fractionAlong.
"""
import dataclasses
import os

import numpy as np
import pytest

from eden_amd import configs, neuroml, nml_export

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "neuroml")


def test_quantity_conversion():
    c = neuroml.convert
    assert c("-65mV", "voltage") == -65.0
    assert c("0.03 kohm_cm", "resistivity") == pytest.approx(0.03)
    assert np.float32(c("3.0 S_per_m2", "conductivity")) == np.float32(0.3)
    assert c("360 S_per_m2", "conductivity") == pytest.approx(36.0)
    assert c("1per_ms", "frequency") == 1e-3
    assert np.float32(c("5Hz", "frequency")) == np.float32(5e-6)   # 5 * 10^-6 in double, then float
    assert c("100ms", "time") == 100000.0
    assert c("0.08nA", "current") == 0.08
    assert c("20nS", "conductance") == pytest.approx(0.02)
    assert c("1.0 uF_per_cm2", "specific_capacitance") == 1.0
    assert c("0V", "voltage") == 0.0
    with pytest.raises(neuroml.NeuroMLError):
        c("5 furlongs", "time")
    with pytest.raises(neuroml.NeuroMLError):
        c("abc", "time")


def test_reference_hh_example_instantiates_to_its_stated_numbers():
    m, outs = neuroml.read(os.path.join(GOLD, "LEMS_HHCellExample.xml"))
    assert m.n_cells == 1 and m.n_comps == 1
    d = 17.841242
    area = np.float32(np.pi * np.float32(d) ** 2)
    assert m.area[0] == pytest.approx(area, rel=1e-6)
    assert m.capacitance[0] == pytest.approx(1.0 * area * 1e-2, rel=1e-6)     # uF/cm^2 * um^2 -> pF
    assert list(m.chan_type) == [0, 1, 2]                                     # passive, Na, K in file order
    np.testing.assert_allclose(m.chan_gbase, [0.3, 120.0, 36.0], rtol=1e-7)
    np.testing.assert_allclose(m.chan_erev, [-54.3, 50.0, -77.0], rtol=1e-7)
    na = m.channel_types[1]
    assert na.n_gates == 2 and na.gates[0].instances == 3 and na.gates[1].instances == 1
    assert na.gates[0].alpha.rate == np.float32(1e-3) and na.gates[0].alpha.midpoint == -40.0
    assert m.channel_types[2].gates[0].beta.rate == np.float32(0.125e-3)
    assert m.channel_types[0].n_gates == 0
    assert m.v0[0] == -65.0
    assert m.pulse_start[0] == 100000.0 and m.pulse_duration[0] == 100000.0
    assert m.pulse_amplitude[0] == np.float32(0.08)
    assert m.dt == 10.0 and m.t_final == 300000.0
    assert [c[1] for c in outs.trajectories[0].columns] == ["hhpop/0/hhcell/0/v"]
    assert outs.events[0].format == "TIME_ID"
    assert m.v_threshold[0] == -20.0


@pytest.mark.parametrize("which", ["config1", "config2", "config4", "config3", "mixed"])
def test_exported_workloads_read_back_to_the_same_model(tmp_path, which):
    m = {
        "config1": lambda: configs.hh_population(n_cells=30, t_final_ms=20.0),
        "config2": lambda: configs.random_trees(6, sizes=[20, 45, 80, 130, 260, 400], t_final_ms=20.0),
        "config4": lambda: configs.config4(n_cells=8, t_final_ms=10.0),
        "config3": lambda: configs.spiking_network(n_cells=400, in_degree=12, t_final_ms=10.0),
        "mixed": lambda: configs.mixed_network(),
    }[which]()
    path = nml_export.write_lems(m, str(tmp_path / f"{which}.xml"), m.morphologies)
    r, outs = neuroml.read(path)
    for f in dataclasses.fields(m):
        a, b = getattr(m, f.name), getattr(r, f.name)
        if f.name in ("name", "spike_target"):
            continue
        if f.name in ("channel_types", "syn_types"):
            assert len(a) == len(b)
            assert all(bytes(x) == bytes(y) for x, y in zip(a, b)), f.name
            continue
        if isinstance(a, np.ndarray):
            assert a.shape == b.shape, f.name
            assert np.array_equal(a, b, equal_nan=a.dtype.kind == "f"), f.name
        else:
            assert a == b, f.name
    for g in range(m.n_comps):
        s0, s1 = m.comp_spike_begin[g], m.comp_spike_begin[g + 1]
        assert sorted(m.spike_target[s0:s1]) == sorted(r.spike_target[s0:s1])


def test_reference_core_synapses_full_model_reads_in_instance_order():
    """The full model: the blocking/plastic synapse (composed with its Mg block and
    depression-facilitation mechanisms, their parameters from the child elements), the
    continuous projections (linearGraded / silent / graded ends) and the gap junction. Each
    continuous or electrical connection puts its postsynaptic component on the post cell
    first, then its presynaptic one on the pre cell, each with the other end as its peer
    (eden/Eden.cpp:9233-9270)."""
    m, _ = neuroml.read(os.path.join(GOLD, "LEMS_EdenTest_CoreSynapses.xml"))
    names = m.lems.type_names
    assert names[5] == "blockingPlasticSynapse/voltageConcDepBlockMechanism/tsodyksMarkramDepFacMechanism"
    assert names[6:10] == ["silentSynapse", "gapJunction", "linearGradedSynapse", "gradedSynapse"]
    p = m.lems.params[5]
    # gbase, erev, tauRise, tauDecay, blockConcentration, scalingConc, scalingVolt (uM, mV),
    # initReleaseProb, tauRec, tauFac (us), weight
    np.testing.assert_allclose(p, [0.01, 0.0, 500.0, 1000.0, 1.0, 2000.0, 150.0, 0.25, 15000.0, 5000.0, 1.0],
                               rtol=1e-6)
    comp_of = np.repeat(np.arange(m.n_comps), np.diff(m.comp_syn_begin))
    graded = [(int(comp_of[k]), names[m.syn_type[k]], int(m.syn_peer[k])) for k in range(m.n_synapses)
              if m.syn_peer[k] >= 0]
    # ContPre / ContPost cells: the projections connect cell k of each (k = 1, 2, 3, 4)
    pre = {k: int(m.cell_comp_begin[31 + k]) for k in range(1, 5)}
    post = {k: int(m.cell_comp_begin[51 + k]) for k in range(1, 5)}
    want = [(post[1], "silentSynapse", pre[1]), (post[2], "gapJunction", pre[2]),
            (post[3], "linearGradedSynapse", pre[3]), (post[4], "gradedSynapse", pre[4]),
            (pre[1], "linearGradedSynapse", post[1]), (pre[2], "gapJunction", post[2]),
            (pre[3], "silentSynapse", post[3]), (pre[4], "silentSynapse", post[4])]
    assert sorted(graded) == sorted(want)


def test_reader_errors_name_the_problem(tmp_path):
    p = tmp_path / "bad.xml"
    p.write_text('<Lems><Target component="s"/><neuroml><network id="n"><population id="a" component="nope" '
                 'size="2"/></network></neuroml><Simulation id="s" length="1ms" step="0.025ms" target="n" '
                 'seed="1"/></Lems>')
    with pytest.raises(neuroml.NeuroMLError, match="nope"):
        neuroml.read(str(p))


def test_cli_reports_missing_model_and_unknown_tokens(capsys):
    from eden_amd.__main__ import main
    assert main(["frobnicate"]) == 2
    out = capsys.readouterr().out
    assert 'skipping unknown token "frobnicate"' in out
    assert "NeuroML model not selected" in out
    assert main(["nml", "/nonexistent/model.xml"]) == 1


# --- cable discretisation (segment groups with neuroLexId sao864921383) -------------------

def test_reference_cable_discretisation_model():
    """The reference's EdenTest_CableDiscretisation model (input data under tests/golden/):
    one 1000 um segment cut into 5 compartments by its cable group, five 200 um segments
    in a 5-compartment cable, and the same five segments with no cable give identical
    compartments, inputs on the same compartments (the inputs' fractionAlong 0.1 and 0.7 on
    the single segment fall in its first and fourth fifths), and recordings at fractionAlong
    0.5; the three segments of the default (nseg 1) cable merge into one compartment whose
    area is their summed frustum area."""
    m, outs = neuroml.read(os.path.join(GOLD, "LEMS_EdenTest_CableDiscretisation.xml"))
    cb = m.cell_comp_begin
    assert list(m.comps_per_cell()) == [1, 5, 5, 5]
    span = [slice(cb[c], cb[c + 1]) for c in range(4)]
    for arr in (m.area, m.capacitance, m.inv_rc_axial, m.parent, np.nan_to_num(m.r_axial)):
        a = np.asarray(arr)
        assert np.array_equal(a[span[1]], a[span[2]]) and np.array_equal(a[span[1]], a[span[3]])
    assert np.allclose(np.asarray(m.area)[span[1]], np.pi * 1.0 * 200.0, rtol=1e-6)     # d = 1 um cylinders
    pulses = np.repeat(np.arange(m.n_comps), np.diff(m.comp_pulse_begin))
    assert list(pulses) == [0, 0, 1, 4, 6, 9, 11, 14]
    assert list(m.record_comp) == [0, 0, 0, 3, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15]
    # the merged cable's area: the three segments' frustum surfaces
    import xml.etree.ElementTree as ET
    root = ET.parse(os.path.join(GOLD, "EdenTest_CableDiscretisation.nml")).getroot()
    cell = next(c for c in root if c.tag.endswith("cell") and c.get("id") == "Cable_Unified")
    segs = [s for s in cell.iter() if s.tag.endswith("}segment")]
    pts, total = {}, 0.0
    for s in segs:
        get = lambda tag: next((c for c in s if c.tag.endswith(tag)), None)
        d = get("distal")
        p = get("proximal")
        par = get("parent")
        q = [float(d.get(k)) for k in ("x", "y", "z", "diameter")]
        a = [float(p.get(k)) for k in ("x", "y", "z", "diameter")] if p is not None else pts[par.get("segment")]
        pts[s.get("id")] = q
        L = np.linalg.norm(np.subtract(q[:3], a[:3]))
        total += np.pi / 2 * (a[3] + q[3]) * np.sqrt(((a[3] - q[3]) / 2) ** 2 + L * L)
    assert np.isclose(float(m.area[0]), total, rtol=1e-6)


def _cable_file(tmp_path, lengths, d0, d1, nseg, ra_kohm_cm=0.5):
    """A LEMS file of one cell: an unbranched cable along x of the given segment lengths,
    diameter tapering linearly from d0 to d1, in one cable group of nseg compartments."""
    x = np.concatenate([[0.0], np.cumsum(lengths)])
    d = d0 + (d1 - d0) * x / x[-1]
    v = lambda a: repr(float(a))
    segs = [f'<segment id="0"><proximal x="0" y="0" z="0" diameter="{v(d[0])}"/>'
            f'<distal x="{v(x[1])}" y="0" z="0" diameter="{v(d[1])}"/></segment>']
    for k in range(1, len(lengths)):
        segs.append(f'<segment id="{k}"><parent segment="{k - 1}"/><distal x="{v(x[k + 1])}" y="0" z="0" '
                    f'diameter="{v(d[k + 1])}"/></segment>')
    members = "".join(f'<member segment="{k}"/>' for k in range(len(lengths)))
    p = tmp_path / f"cable_{nseg}.xml"
    p.write_text(
        '<Lems><Target component="s"/><neuroml>'
        '<ionChannelHH id="pas" conductance="10pS"/>'
        f'<cell id="c"><morphology id="m">{"".join(segs)}'
        f'<segmentGroup id="axon" neuroLexId="sao864921383"><property tag="numberInternalDivisions" '
        f'value="{nseg}"/>{members}</segmentGroup></morphology>'
        '<biophysicalProperties id="b"><membraneProperties>'
        '<channelDensity id="leak" ionChannel="pas" condDensity="0.3 mS_per_cm2" erev="-65mV"/>'
        '<specificCapacitance value="1.0 uF_per_cm2"/><initMembPotential value="-65mV"/>'
        f'</membraneProperties><intracellularProperties><resistivity value="{ra_kohm_cm} kohm_cm"/>'
        '</intracellularProperties></biophysicalProperties></cell>'
        '<network id="n"><population id="p" component="c" size="1"/></network></neuroml>'
        '<Simulation id="s" length="1ms" step="0.025ms" target="n" seed="1">'
        '<OutputFile id="o" fileName="o.txt"><OutputColumn id="v" quantity="p/0/c/0/v"/></OutputFile>'
        '</Simulation></Lems>')
    return str(p), x, d


@pytest.mark.parametrize("nseg", [1, 3, 7])
def test_cable_compartments_match_frustum_closed_forms(tmp_path, nseg):
    """A tapering cable of uneven segments cut into nseg compartments of equal path length:
    a compartment's area, capacitance and the resistance between neighbouring midpoints are
    those of the cone sections they span (lateral area pi/2 (d_a + d_b) sqrt(((d_a - d_b)/2)^2
    + l^2), C = Cm A, R = Ra 4 l / (pi d_a d_b), all additive over sections), whatever the
    segment boundaries inside a compartment; the compartments form a chain."""
    lengths = [100.0, 37.0, 250.0, 13.0, 60.0]
    path, x, d = _cable_file(tmp_path, lengths, 4.0, 1.0, nseg)
    m, _ = neuroml.read(path)
    assert m.n_comps == nseg
    L = x[-1]
    dia = lambda xx: 4.0 + (1.0 - 4.0) * xx / L
    edges = np.linspace(0.0, L, nseg + 1)
    area = [np.pi / 2 * (dia(a) + dia(b)) * np.sqrt(((dia(a) - dia(b)) / 2) ** 2 + (b - a) ** 2)
            for a, b in zip(edges[:-1], edges[1:])]
    np.testing.assert_allclose(m.area, area, rtol=2e-5)
    np.testing.assert_allclose(m.capacitance, np.asarray(area) * 1e-2, rtol=2e-5)   # 1 uF/cm^2 x um^2 in pF
    assert list(m.parent) == [-1] + list(range(nseg - 1))
    mid = (edges[:-1] + edges[1:]) / 2
    r = [0.5 * 10.0 * 4 * (b - a) / (np.pi * dia(a) * dia(b)) for a, b in zip(mid[:-1], mid[1:])]   # MOhm
    np.testing.assert_allclose(np.asarray(m.r_axial)[1:], r, rtol=2e-5)


def test_cable_locations_by_fraction_along(tmp_path):
    """Locations on a segment map to the compartment whose span holds them; a compartment
    boundary belongs to the compartment it ends."""
    path, x, d = _cable_file(tmp_path, [1000.0], 2.0, 2.0, 5)
    doc_cells = neuroml._Doc(path)
    c = neuroml._cell(doc_cells, next(el for role, el in doc_cells.items.values() if role == "cell"), None, {"pas": 0})
    assert [c.comp_of_segment(0, "t", f) for f in (0.0, 0.1, 0.2, 0.2001, 0.5, 0.7, 0.99, 1.0)] == [0, 0, 0, 1, 2, 3, 4, 4]


def test_reference_inhomogeneous_parameters_model():
    """The reference's EdenTest_InhomogeneousParameters model: a channelDensityNonUniform
    whose value is a constant 1200 S/m^2 equals the uniform 120 mS/cm^2 cell; '500 *
    exp(-p/20)' over the dendrite's compartments, p their midpoints' path length from the
    root (15 and 25 um for 10 um compartments after a 10 um soma), and the normalised form
    (p - min) / (max - min) give the closed-form densities."""
    m, _ = neuroml.read(os.path.join(GOLD, "LEMS_EdenTest_InhomogeneousParameters.xml"))
    assert list(m.comps_per_cell()) == [3, 3, 3, 3]
    cb, chb = m.cell_comp_begin, m.comp_chan_begin

    def dens(cell):
        return [list(map(float, m.chan_gbase[chb[g]:chb[g + 1]])) for g in range(cb[cell], cb[cell + 1])]

    flat, inhf, inho, inhn = (dens(c) for c in range(4))
    assert flat == inhf == [[], [120.0, 36.0], [120.0, 36.0]]
    np.testing.assert_allclose([inho[1][0], inho[2][0]], [50 * np.exp(-15 / 20), 50 * np.exp(-25 / 20)], rtol=1e-6)
    np.testing.assert_allclose([inhn[1][0], inhn[2][0]], [50.0, 50 * np.exp(-1 / 20)], rtol=1e-6)
    assert inho[1][1] == inho[2][1] == inhn[1][1] == 36.0


def _hh_rates(v):
    """Hodgkin-Huxley rates (1/ms) of the m, h and n gates at v (mV), as the reference's
    EdenTestNetwork2 channels state them (HHExpLinearRate, HHExpRate, HHSigmoidRate)."""
    am = 1.0 * ((v + 40) / 10) / (1 - np.exp(-(v + 40) / 10))
    bm = 4.0 * np.exp((v + 65) / -18)
    ah = 0.07 * np.exp((v + 65) / -20)
    bh = 1.0 / (1 + np.exp(-(v + 35) / 10))
    an = 0.1 * ((v + 55) / 10) / (1 - np.exp(-(v + 55) / 10))
    bn = 0.125 * np.exp((v + 65) / -80)
    return np.array([[am, bm], [ah, bh], [an, bn]])


def test_reference_gate_recording_model_reads_its_gate_columns():
    """The reference's EdenTest2 model records m, h and n of the naChans / kChans
    distributions next to V: the paths resolve to the compartment's gate variables 0, 1 and
    2 (the leak distribution before them has no gates), read at scale 1 (V at 1e-3)."""
    from eden_amd import output
    m, outs = neuroml.read(os.path.join(GOLD, "LEMS_EdenTest2.xml"))
    assert list(m.record_comp) == [0, 0, 0, 0]
    assert list(m.record_gate) == [-1, 0, 1, 2]
    assert list(output.column_scales(m)) == [1e-3, 1.0, 1.0, 1.0]
    m1, _ = neuroml.read(os.path.join(GOLD, "LEMS_EdenTest1.xml"))
    # pop0/2/MultiCompCell/1 is compartment 1 of the third cell; m is its first gate variable
    assert list(m1.record_comp) == [0, 2, 5, 4, 5] and list(m1.record_gate) == [-1, -1, -1, -1, 0]
    assert neuroml._quantity_path("p/0/c/3/bio/membraneProperties/d/ch/g/q") == ("p", 0, 3, ("bio", "d", "ch", "g"))


def test_gate_paths_that_name_nothing_are_refused(tmp_path):
    src = open(os.path.join(GOLD, "LEMS_EdenTest2.xml")).read()
    import shutil
    shutil.copy(os.path.join(GOLD, "EdenTestNetwork2.nml"), tmp_path / "EdenTestNetwork2.nml")
    for bad, why in (("naChans/naChan/x/q", "no gate x"), ("naChans/kChan/n/q", "holds ion channel naChan"),
                     ("caChans/caChan/m/q", "no channel distribution caChans"),
                     ("naChans/naChan/m/tau", "can be recorded")):
        p = tmp_path / "LEMS_bad.xml"
        p.write_text(src.replace("naChans/naChan/m/q", bad))
        with pytest.raises(neuroml.NeuroMLError, match=why):
            neuroml.read(str(p))


def test_oracle_gate_columns_are_hh_steady_states_then_cnexp_steps():
    """Oracle gate columns of EdenTest2: row 0 holds the textbook resting values
    alpha / (alpha + beta) at -65 mV (m 0.0529, h 0.5961, n 0.3177); every later row follows
    from the one before by the exponential-Euler gate step at that row's recorded V
    (q' = q_inf + (q - q_inf) exp(-dt (alpha + beta))), so V and the gates of a row belong
    to the same instant. A gate column taken from the wrong channel instance or gate, or a
    row late or early, fails one of these."""
    from oracle import oracle
    m, _ = neuroml.read(os.path.join(GOLD, "LEMS_EdenTest2.xml"))
    tr, _, _, _ = oracle.run(m)
    tr = tr.astype(np.float64)
    r = _hh_rates(-65.0)
    np.testing.assert_allclose(tr[0, 1:], r[:, 0] / r.sum(1), rtol=2e-6)
    np.testing.assert_allclose(tr[0, 1:], [0.0529325, 0.5961208, 0.3176769], rtol=1e-5)
    dt = float(np.float32(m.dt)) * 1e-3   # ms
    for row in range(len(tr) - 1):
        r = _hh_rates(tr[row, 0])
        tot = r.sum(1)
        inf = r[:, 0] / tot
        want = inf + (tr[row, 1:] - inf) * np.exp(-dt * tot)
        np.testing.assert_allclose(tr[row + 1, 1:], want, atol=2e-6, err_msg=f"row {row + 1}")
    assert tr[:, 0].max() > 0.0    # the cell fires: the gates move far from rest
