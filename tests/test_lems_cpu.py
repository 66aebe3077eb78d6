"""Generic LEMS synapse types (eden_amd.lems) on the CPU.

* Expression parsing: precedence and associativity of the printed, fully parenthesised text.
* Each type's emitted step function (the C dialect, compiled with gcc) against closed forms
  of its discrete dynamics: after one event from rest, the sum over steps of the conductance
  (or current) times dt equals a value fixed by the type's parameters. For a variable that
  decays by Euler steps, x_{n+1} = x_n (1 - dt / tau), the sum of x_n dt is x_0 tau exactly
  (a geometric series); for the alpha forms, g_{n+1} = g_n + dt (e A_n - g_n) / tau, summing
  both sides over all steps gives sum g_n dt = e sum A_n dt = e A_0 tau; for expTwo and
  expThree the charge is the weighted difference of two such series. A wrong coefficient,
  a missing factor e or a swapped time constant fails one of them.
* expTwoSynapse emitted from its LEMS definition equals the library's native expTwo in the
  oracle, bit for bit, over a network run.
* The reference's CoreSynapses model reads into one generic type per core synapse (the
  blocking/plastic one composed with its mechanisms), in the reference's instance order,
  and runs in the oracle.
* Graded synapses and the blocking / plastic synapse's mechanisms against closed forms.
* Gap junctions: a coupled pair of membranes follows the discrete closed form.
* User-defined ComponentTypes with the dynamics of core types run as the core types do.
"""
import ctypes as C
import dataclasses
import math
import os

import numpy as np
import pytest

from eden_amd import configs, lems, neuroml
from oracle import oracle

E = math.e


def test_parser_precedence_and_associativity():
    names = {k: k for k in "abcdx"}
    assert lems.infix(lems.parse("a - b - c"), names) == "( ( ( ( a - b ) ) - c ) )"
    assert lems.infix(lems.parse("a / b / c"), names) == "( ( ( ( a / b ) ) / c ) )"
    assert lems.infix(lems.parse("a + b * c"), names) == "( ( a + ( ( b * c ) ) ) )"
    assert lems.infix(lems.parse("-a / b"), names) == "( ( ( -a ) / b ) )"
    assert lems.infix(lems.parse("exp(-x)"), names) == "EXP( ( -x ) )"
    assert lems.infix(lems.parse("2^3"), names) == "POW( ( (float)2.0) , ( (float)3.0) )"
    with pytest.raises(lems.LemsError):
        lems.parse("a + (b")
    with pytest.raises(lems.LemsError):
        lems.infix(lems.parse("q * 2"), names)


def _stepper(tname, vals, weight=1.0):
    ls = lems.build([(tname, vals, weight)])
    lib = oracle.lems_library(ls.source_c)
    fn = lib.lems_syn
    fn.restype = None
    fn.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_long, C.c_float, C.c_float, C.c_float, C.c_int,
                   C.c_int, C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    P = ls.params[0]
    S = np.zeros(ls.n_state[0], np.float32)

    def step(v, init, event, dt):
        i, g = C.c_float(), C.c_float()
        fn(0, P.ctypes.data, S.ctypes.data, S.ctypes.data, 1, v, v, 0.0, init, event, dt, C.byref(i), C.byref(g))
        return i.value, g.value
    return step


def _charge(step, v, dt, n=400_000, use="g"):
    step(v, 1, 0, dt)                    # an initialisation step
    tot, first = 0.0, None
    for k in range(n):
        i, g = step(v, 0, 1 if k == 0 else 0, dt)
        x = g if use == "g" else i
        tot += x * dt
    return tot


DT = 25.0


@pytest.mark.parametrize("tname,vals,use,expected", [
    ("expOneSynapse", dict(gbase=2e-3, erev=0.0, tauDecay=3000.0), "g", lambda p: p["gbase"] * p["tauDecay"]),
    ("alphaSynapse", dict(gbase=5e-3, erev=0.0, tau=2000.0), "g", lambda p: E * p["gbase"] * p["tau"]),
    ("alphaCurrentSynapse", dict(tau=4000.0, ibase=0.1), "i", lambda p: E * p["ibase"] * p["tau"]),
    ("expTwoSynapse", dict(gbase=2e-2, erev=0.0, tauRise=500.0, tauDecay=1000.0), "g", None),
    ("expThreeSynapse", dict(gbase1=3e-2, gbase2=3e-2, erev=0.0, tauRise=500.0, tauDecay1=1500.0, tauDecay2=5000.0),
     "g", None),
    ("expCurrSynapse", dict(tau_syn=2.0), "i", lambda p: 1.0 * p["tau_syn"] * 1000.0),
    ("alphaCurrSynapse", dict(tau_syn=2.0), "i", lambda p: 2.7182818 * p["tau_syn"] * 1000.0),
])
def test_charge_after_one_event(tname, vals, use, expected):
    step = _stepper(tname, vals, weight=1.0)
    got = _charge(step, -65.0, DT, use=use)
    if tname == "expTwoSynapse":
        tr, td = vals["tauRise"], vals["tauDecay"]
        tp = math.log(td / tr) * tr * td / (td - tr)
        wf = 1.0 / (-math.exp(-tp / tr) + math.exp(-tp / td))
        want = vals["gbase"] * wf * (td - tr)
    elif tname == "expThreeSynapse":
        tr, d1, d2, g1, g2 = vals["tauRise"], vals["tauDecay1"], vals["tauDecay2"], vals["gbase1"], vals["gbase2"]
        wf = [1.0 / (-math.exp(-(math.log(d / tr) * tr * d / (d - tr)) / tr) +
                     math.exp(-(math.log(d / tr) * tr * d / (d - tr)) / d)) for d in (d1, d2)]
        A0 = (g1 * wf[0] + g2 * wf[1]) / (g1 + g2)
        want = g1 * (wf[0] * d1 - A0 * tr) + g2 * (wf[1] * d2 - A0 * tr)
    else:
        want = expected(vals)
    assert got == pytest.approx(want, rel=2e-3), (tname, got, want)


def test_conductance_synapse_current_sign_and_size():
    """i = g (erev - v): at v = erev the current vanishes, below it is inward (positive)."""
    for tname, vals in (("alphaSynapse", dict(gbase=5e-3, erev=0.0, tau=2000.0)),
                        ("expCondSynapse", dict(tau_syn=2.0, e_rev=0.0))):
        step = _stepper(tname, vals)
        step(-65.0, 1, 0, DT)
        step(-65.0, 0, 1, DT)
        for _ in range(20):
            i_low, g = step(-65.0, 0, 0, DT)
        assert i_low > 0
    step = _stepper("expOneSynapse", dict(gbase=2e-3, erev=-10.0, tauDecay=3000.0))
    step(-10.0, 1, 0, DT)
    step(-10.0, 0, 1, DT)
    i, g = step(-10.0, 0, 0, DT)
    assert i == 0.0 and g > 0


def test_generic_exptwo_equals_native_exptwo():
    m = configs.mixed_network(tree_sizes=(40, 90), n_point=10, n_conn=100, t_final_ms=40.0, seed=5)
    types = [("expTwoSynapse", dict(gbase=st.gbase, erev=st.erev, tauRise=st.tau_rise, tauDecay=st.tau_decay),
              st.weight_prop) for st in m.syn_types]
    g = dataclasses.replace(m, lems=lems.build(types), name="generic")
    (a, ea, ra, _), (sa, ya) = oracle.run_with_synapse_log(m)
    (b, eb, rb, _), (sb, yb) = oracle.run_with_synapse_log(g)
    assert len(sa) > 50
    assert np.array_equal(a, b) and np.array_equal(ra, rb) and np.array_equal(sa, sb) and np.array_equal(ya, yb)


def test_reference_core_synapses_model(tmp_path):
    from coresyn import write_model
    m, outs = neuroml.read(write_model(str(tmp_path)))
    # every synapse element in reading order, the blocking/plastic one composed with its
    # mechanisms
    assert m.lems.type_names == ["alphaCurrentSynapse", "alphaSynapse", "expOneSynapse", "expTwoSynapse",
                                 "expThreeSynapse",
                                 "blockingPlasticSynapse/voltageConcDepBlockMechanism/tsodyksMarkramDepFacMechanism",
                                 "silentSynapse", "gapJunction", "linearGradedSynapse", "gradedSynapse",
                                 "expCondSynapse", "alphaCondSynapse", "expCurrSynapse", "alphaCurrSynapse",
                                 "sineGenerator"]       # an input type, unused by the network
    # ChemPost cells 1..10 and 21 receive one synapse each, of the projections' types in
    # order; then the graded and electrical ends on ContPre 1-4 and ContPost 1-4
    comp_of = np.repeat(np.arange(m.n_comps), np.diff(m.comp_syn_begin))
    assert list(comp_of) == [2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 22, 32, 33, 34, 35, 52, 53, 54, 55]
    assert list(m.syn_type) == [0, 1, 2, 3, 4, 5, 10, 11, 12, 13, 0, 8, 7, 6, 6, 6, 7, 8, 9]
    assert m.syn_weight[10] == np.float32(1.7) and m.syn_delay[10] == 5000.0
    (tr, ec, er, _), (st, sy) = oracle.run_with_synapse_log(m)
    assert sorted(set(sy.tolist())) == list(range(11))       # every chemical synapse received events
    # every postsynaptic cell left rest
    cols = [c[2] for c in outs.trajectories[0].columns][1:12]
    assert np.all(np.abs(tr[:, cols] - (-65.0)).max(axis=0) > 0.05)


def _gap_pair_file(tmp_path, v0a, v0b, g_ps, cm_uf=1.0, diam=10.0):
    """Two one-compartment cells with no membrane conductance, coupled by one gap junction."""
    p = tmp_path / "gap_pair.xml"
    cell = (f'<cell id="{{}}"><morphology id="m{{}}"><segment id="0"><proximal x="0" y="0" z="0" diameter="{diam}"/>'
            f'<distal x="0" y="0" z="0" diameter="{diam}"/></segment></morphology>'
            '<biophysicalProperties id="b{}"><membraneProperties>'
            '<channelDensity id="leak" ionChannel="pas" condDensity="0 mS_per_cm2" erev="-65mV"/>'
            f'<specificCapacitance value="{cm_uf} uF_per_cm2"/><initMembPotential value="{{}}mV"/>'
            '</membraneProperties><intracellularProperties><resistivity value="0.1 kohm_cm"/>'
            '</intracellularProperties></biophysicalProperties></cell>')
    p.write_text(
        '<Lems><Target component="s"/><neuroml><ionChannelHH id="pas" conductance="10pS"/>'
        f'<gapJunction id="gj" conductance="{g_ps}pS"/>'
        + cell.format("A", "A", "A", v0a) + cell.format("B", "B", "B", v0b) +
        '<network id="n"><population id="pa" component="A" size="1"/><population id="pb" component="B" size="1"/>'
        '<electricalProjection id="e" presynapticPopulation="pa" postsynapticPopulation="pb">'
        '<electricalConnection id="0" preCell="0" preSegment="0" postCell="0" postSegment="0" synapse="gj"/>'
        '</electricalProjection></network></neuroml>'
        '<Simulation id="s" length="20ms" step="0.025ms" target="n" seed="1">'
        '<OutputFile id="o" fileName="o.txt"><OutputColumn id="a" quantity="pa/0/A/0/v"/>'
        '<OutputColumn id="b" quantity="pb/0/B/0/v"/></OutputFile></Simulation></Lems>')
    return str(p)


def test_gap_junction_closed_form(tmp_path):
    """Two membranes with no conductance of their own joined by a gap junction g: with the
    current g (Vpeer - V) taken from the start of each step, the difference x = Va - Vb obeys
    x' = x (1 - 2 g dt / C) exactly per step and Va + Vb stays fixed; each end holds one
    instance whose peer is the other end."""
    m, _ = neuroml.read(_gap_pair_file(tmp_path, -65.0, -55.0, 200.0))
    assert m.lems.type_names == ["gapJunction"]
    assert list(m.syn_peer) == [1, 0] and list(np.diff(m.comp_syn_begin)) == [1, 1]
    tr, _, _, _ = oracle.run(m, precision="f32")
    C = float(m.capacitance[0])
    g = 200e-6                                  # uS
    x = tr[:, 0].astype(np.float64) - tr[:, 1].astype(np.float64)
    k = 1 - 2 * g * m.dt / C
    pred = -10.0 * k ** np.arange(len(x))
    np.testing.assert_allclose(x, pred, rtol=1e-4, atol=1e-4)
    assert x[-1] / x[0] < 0.5                   # the voltages approach each other
    np.testing.assert_allclose(tr[:, 0].astype(np.float64) + tr[:, 1], -120.0, atol=1e-3)


_USER_TYPES = """
<ComponentType name="myGapJunction" extends="baseGradedSynapse">
    <Parameter name="conductance" dimension="conductance"/>
    <Dynamics>
        <DerivedVariable name="i" exposure="i" dimension="current" value="(vpeer - v) * conductance"/>
    </Dynamics>
</ComponentType>
<ComponentType name="myExpOneSynapse" extends="baseConductanceBasedSynapse">
    <Property name="weight" dimension="none" defaultValue="1"/>
    <Parameter name="tauDecay" dimension="time"/>
    <Dynamics>
        <StateVariable name="g" dimension="conductance"/>
        <DerivedVariable name="i" exposure="i" dimension="current" value="g * (erev - v)"/>
        <DerivedVariable name="g0" exposure="g" dimension="conductance" value="g"/>
        <TimeDerivative variable="g" value="-g / tauDecay"/>
        <OnStart><StateAssignment variable="g" value="0"/></OnStart>
        <OnEvent port="in"><StateAssignment variable="g" value="g + (weight * gbase)"/></OnEvent>
    </Dynamics>
</ComponentType>
"""


def _user_types_file(tmp_path):
    """An HH cell (the reference's example cell, tests/golden/neuroml) spiking under a current
    step, projecting through expOneSynapse onto one passive cell and through a user-defined
    ComponentType of the same dynamics onto another; two passive pairs joined by gapJunction
    and by a user-defined gap junction type."""
    import shutil
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "neuroml")
    shutil.copy(os.path.join(gold, "NML2_SingleCompHHCell.nml"), tmp_path / "NML2_SingleCompHHCell.nml")
    pas = ('<cell id="{n}"><morphology id="m{n}"><segment id="0"><proximal x="0" y="0" z="0" diameter="10"/>'
           '<distal x="0" y="0" z="0" diameter="10"/></segment></morphology><biophysicalProperties id="b{n}">'
           '<membraneProperties><channelDensity id="leak" ionChannel="passiveChan" condDensity="0.3 mS_per_cm2" '
           'erev="-65mV"/><specificCapacitance value="1.0 uF_per_cm2"/><initMembPotential value="{v}mV"/>'
           '</membraneProperties><intracellularProperties><resistivity value="0.1 kohm_cm"/>'
           '</intracellularProperties></biophysicalProperties></cell>')
    cells = "".join(pas.format(n=n, v=v) for n, v in (("P1", -65), ("P2", -65), ("G1", -65), ("G2", -50),
                                                        ("G3", -65), ("G4", -50)))
    pops = "".join(f'<population id="p{n}" component="{n}" size="1"/>' for n in ("P1", "P2", "G1", "G2", "G3", "G4"))
    conn = ('<projection id="{id}" presynapticPopulation="hhpop" postsynapticPopulation="{post}" synapse="{syn}">'
            '<connectionWD id="0" preCellId="../hhpop/0/hhcell" postCellId="../{post}/0/{cell}" weight="2" '
            'delay="1ms"/></projection>')
    gap = ('<electricalProjection id="{id}" presynapticPopulation="{a}" postsynapticPopulation="{b}">'
           '<electricalConnection id="0" preCell="0" preSegment="0" postCell="0" postSegment="0" synapse="{syn}"/>'
           '</electricalProjection>')
    net = (conn.format(id="c1", post="pP1", cell="P1", syn="e1") + conn.format(id="c2", post="pP2", cell="P2", syn="e2")
           + gap.format(id="g1", a="pG1", b="pG2", syn="gj1") + gap.format(id="g2", a="pG3", b="pG4", syn="gj2"))
    cols = "".join(f'<OutputColumn id="{n}" quantity="p{n}/0/{n}/0/v"/>' for n in ("P1", "P2", "G1", "G2", "G3", "G4"))
    p = tmp_path / "user_types.xml"
    p.write_text(
        '<Lems><Target component="s"/><Include file="NML2_SingleCompHHCell.nml"/><neuroml>' + _USER_TYPES +
        '<expOneSynapse id="e1" gbase="0.5nS" erev="0mV" tauDecay="3ms"/>'
        '<myExpOneSynapse id="e2" gbase="0.5nS" erev="0mV" tauDecay="3ms"/>'
        '<gapJunction id="gj1" conductance="300pS"/><myGapJunction id="gj2" conductance="300pS"/>'
        + cells + '<network id="n2"><population id="hhpop" component="hhcell" size="1"/>' + pops +
        '<explicitInput target="hhpop[0]" input="pulseGen1"/>' + net + '</network></neuroml>'
        '<Simulation id="s" length="300ms" step="0.025ms" target="n2" seed="1">'
        f'<OutputFile id="o" fileName="o.txt"><OutputColumn id="hh" quantity="hhpop/0/hhcell/0/v"/>{cols}</OutputFile>'
        '</Simulation></Lems>')
    return str(p)


def test_user_defined_component_types_equal_the_core_types(tmp_path):
    """User ComponentTypes with the dynamics of expOneSynapse and gapJunction (as the
    reference's EdenTestNetwork4 defines them) read into generic types whose runs equal the
    core types' bit for bit; their parameters come from the base types they extend."""
    m, outs = neuroml.read(_user_types_file(tmp_path))
    assert m.lems.type_names == ["expOneSynapse", "myExpOneSynapse", "gapJunction", "myGapJunction"]
    assert [list(p) for p in m.lems.params[:2]] == [list(m.lems.params[0])] * 2
    (tr, ec, er, _), (st, sy) = oracle.run_with_synapse_log(m)
    assert len(st) > 2                                           # the HH cell's spikes arrived
    assert np.array_equal(tr[:, 1], tr[:, 2])                    # expOne vs the user type
    assert np.array_equal(tr[:, 3:5], tr[:, 5:7])                # gap junction pair vs the user type
    assert float(np.abs(tr[:, 1] + 65.0).max()) > 0.1 and float(np.abs(tr[:, 3] + 65.0).max()) > 0.1


def test_state_exposure_reads_the_value_an_event_assigned():
    """A state variable exposed directly is read after the step's assignments (the
    reference's exposures follow the update and read the assigned present value), a derived
    one before them: on an event step the first shows the jump, the second does not."""
    src = ('<Lems><ComponentType name="s1" extends="baseConductanceBasedSynapse">'
           '<Property name="weight" dimension="none" defaultValue="1"/><Parameter name="tauDecay" dimension="time"/>'
           '<Dynamics><StateVariable name="g" dimension="conductance" exposure="g"/>'
           '<DerivedVariable name="i" exposure="i" dimension="current" value="g * (erev - v)"/>'
           '<TimeDerivative variable="g" value="-g / tauDecay"/>'
           '<OnEvent port="in"><StateAssignment variable="g" value="g + (weight * gbase)"/></OnEvent>'
           '</Dynamics></ComponentType></Lems>')
    import xml.etree.ElementTree as ET
    cts = {el.get("name"): el for el in ET.fromstring(src)}
    st, w = lems.synapse_type_from_lems("s1", cts, neuroml.convert)
    vals = dict(gbase=2e-3, erev=0.0, tauDecay=3000.0)
    for tname, custom, jump_seen in (("s1", {"s1": st}, True), ("expOneSynapse", None, False)):
        step = _stepper_peer(tname, vals, custom)
        step(-65.0, -65.0, 1, 0)
        i, g, S = step(-65.0, -65.0, 0, 1)
        assert S[0] == np.float32(2e-3) and i == 0.0
        assert g == (np.float32(2e-3) if jump_seen else 0.0), tname


def _stepper_peer(tname, vals, custom=None, weight=1.0):
    """Step function of one instance taking (v, vpeer, init, event, dt); returns (i, g, state)."""
    ls = lems.build([(tname, vals, weight)], custom)
    lib = oracle.lems_library(ls.source_c)
    fn = lib.lems_syn
    fn.restype = None
    fn.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_long, C.c_float, C.c_float, C.c_float, C.c_int,
                   C.c_int, C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    P = ls.params[0]
    S = np.zeros(max(1, ls.n_state[0]), np.float32)

    def step(v, vpeer, init, event, dt=DT, t=0.0):
        i, g = C.c_float(), C.c_float()
        fn(0, P.ctypes.data, S.ctypes.data, S.ctypes.data, 1, v, vpeer, t, init, event, dt, C.byref(i), C.byref(g))
        return i.value, g.value, S.copy()
    return step


def test_graded_synapses_follow_their_closed_forms():
    """silentSynapse carries no current; linearGradedSynapse is conductance (vpeer - v) in
    float arithmetic; gradedSynapse's s relaxes by Euler steps towards
    inf = 1 / (1 + exp((Vth - vpeer) / delta)) with tau = (1 - inf) / k, so from s = 0,
    s_n = inf (1 - (1 - dt / tau)^n) and i = conductance s (erev - v); where 1 - inf < 1e-4 the
    OnCondition sets s = inf in one step (the reference's guard against tau ~ 0). Swapping
    v and vpeer, or the sign of the exponent, fails these."""
    step = _stepper_peer("silentSynapse", {})
    assert step(-40.0, 10.0, 0, 0)[:2] == (0.0, 0.0)
    step = _stepper_peer("linearGradedSynapse", dict(conductance=1e-3))
    i, g, _ = step(-65.0, -20.0, 0, 0)
    assert i == np.float32(1e-3) * np.float32(-20.0 - -65.0) and g == 0.0
    vals = dict(conductance=1e-3, delta=5.0, k=2.5e-5, Vth=-55.0, erev=0.0)
    step = _stepper_peer("gradedSynapse", vals)
    v, vpeer = -65.0, -50.0
    inf = 1.0 / (1.0 + math.exp((vals["Vth"] - vpeer) / vals["delta"]))
    tau = (1.0 - inf) / vals["k"]
    step(v, vpeer, 1, 0)
    for n in range(1, 2001):
        i, _, S = step(v, vpeer, 0, 0)
        want_s = inf * (1.0 - (1.0 - DT / tau) ** n)
        assert abs(S[0] - want_s) <= 2e-5 * inf, (n, S[0], want_s)
    # the current of a step uses the state at its start
    s_prev = S[0]
    i, _, S = step(v, vpeer, 0, 0)
    assert i == pytest.approx(vals["conductance"] * s_prev * (vals["erev"] - v), rel=1e-6)
    # saturated activation: 1 - inf < 1e-4 -> s = inf after one step, no Euler blow-up
    step = _stepper_peer("gradedSynapse", vals)
    step(v, 0.0, 1, 0)
    _, _, S = step(v, 0.0, 0, 0)
    assert S[0] == np.float32(1.0) / (np.float32(1.0) + np.float32(math.exp((vals["Vth"] - 0.0) / vals["delta"])))


def _bps(block, plast, extra):
    st = lems.blocking_plastic_type(block, plast)
    key = f"bps/{block}/{plast}"
    vals = dict(gbase=2e-2, erev=0.0, tauRise=500.0, tauDecay=1000.0, **extra)
    return _stepper_peer(key, vals, {key: st}), vals


def _event_charges(step, v, events, gap):
    """Charge sum(g dt) after each of `events` events `gap` steps apart."""
    step(v, v, 1, 0)
    out = []
    for e in range(events):
        q = 0.0
        for k in range(gap):
            _, g, _ = step(v, v, 0, 1 if k == 0 else 0)
            q += g * DT
        out.append(q)
    return out


def test_blocking_factor_scales_the_exptwo_charge():
    """With the Mg block alone, one event's charge is the expTwo charge gbase wf (tauDecay -
    tauRise) times 1 / (1 + (c / c_s) exp(-v / v_s)) at the (clamped) membrane voltage."""
    extra = dict(blockConcentration=1.0, scalingConc=2000.0, scalingVolt=150.0)
    for v in (-65.0, -20.0, 10.0):
        step, vals = _bps("voltageConcDepBlockMechanism", None, extra)
        q = _event_charges(step, v, 1, 20000)[0]
        tr, td = vals["tauRise"], vals["tauDecay"]
        tp = math.log(td / tr) * tr * td / (td - tr)
        wf = 1.0 / (-math.exp(-tp / tr) + math.exp(-tp / td))
        block = 1.0 / (1.0 + (1.0 / 2000.0) * math.exp(-v / 150.0))
        assert q == pytest.approx(vals["gbase"] * wf * (td - tr) * block, rel=2e-3), v


def test_tsodyks_markram_release_follows_the_event_maps():
    """Each event releases in proportion to R U of the present state. Without recovery
    (time constants far beyond the run), depression-facilitation has U_k = 1 - (1 - p)^(k+1)
    and R_k = prod_{j<k} (1 - U_j) at event k, so event k's charge over event 0's is
    R_k U_k / p.
    Depression alone with recovery between periodic events reaches the steady state
    R* = (1 - a^m) / (1 - (1 - p) a^m), a = 1 - dt / tauRec, m Euler steps between events
    (the event step's own update is the reset), which the late events' charges show."""
    p = 0.25
    step, _ = _bps(None, "tsodyksMarkramDepFacMechanism", dict(initReleaseProb=p, tauRec=1e30, tauFac=1e30))
    q = _event_charges(step, -65.0, 5, 4000)
    U = [1.0 - (1.0 - p) ** (k + 1) for k in range(5)]             # present U at event k
    R = [float(np.prod([1.0 - U[j] for j in range(k)])) for k in range(5)]
    for k in range(5):
        assert q[k] / q[0] == pytest.approx(R[k] * U[k] / p, rel=1e-5), k
    step, _ = _bps(None, "tsodyksMarkramDepMechanism", dict(initReleaseProb=p, tauRec=20000.0))
    gap = 4000
    q = _event_charges(step, -65.0, 12, gap)
    a = 1.0 - DT / 20000.0
    m = gap - 1
    r_ss = (1.0 - a ** m) / (1.0 - (1.0 - p) * a ** m)
    assert q[-1] / q[0] == pytest.approx(r_ss, rel=1e-4)
    assert q[1] / q[0] == pytest.approx(1.0 - p * a ** m, rel=1e-4)     # 1 - R recovers as a^m


def _input_stepper(tname, vals, custom=None):
    return _stepper_peer(tname, vals, custom)


def test_input_generators_follow_their_definitions():
    """The LEMS inputs' current at time t (their state is assigned by the OnCondition of the
    window t falls in): sine amplitude sin(phase + 2 pi (t - delay) / period) on
    [delay, delay + duration), 0 outside; ramp from start to finish amplitude over the window,
    baseline outside (and at initialisation); pulseGeneratorDL the amplitude on the window;
    voltage clamps (target - v) / R while active, with the conditioning, testing and return
    voltages for the triple clamp."""
    sine = _input_stepper("sineGenerator", dict(phase=0.3, delay=20000.0, duration=50000.0, amplitude=1.0,
                                                period=5000.0))
    pi = 3.14159265                # the definition's literal
    for t, want in ((10000.0, 0.0), (20000.0, math.sin(0.3)), (21250.0, math.sin(0.3 + pi / 2)),
                    (69999.0, math.sin(0.3 + 2 * pi * 49999.0 / 5000.0)), (70000.0, 0.0)):
        i, g, _ = sine(-65.0, -65.0, 0, 0, t=t)
        # the float argument near 63 carries an ulp of 4e-6
        assert i == pytest.approx(want, abs=1e-5), t
    ramp = _input_stepper("rampGenerator", dict(delay=30000.0, duration=170000.0, startAmplitude=-0.032,
                                                finishAmplitude=0.05, baselineAmplitude=-0.032))
    assert ramp(-65.0, -65.0, 1, 0, t=0.0)[2][0] == np.float32(-0.032)          # OnStart
    for t, want in ((0.0, -0.032), (30000.0, -0.032), (115000.0, -0.032 + 0.082 * 0.5), (200000.0, -0.032)):
        assert ramp(-65.0, -65.0, 0, 0, t=t)[0] == pytest.approx(want, abs=1e-7), t
    dl = _input_stepper("pulseGeneratorDL", dict(delay=10000.0, duration=20000.0, amplitude=14.0))
    assert [dl(-65.0, -65.0, 0, 0, t=t)[0] for t in (9999.0, 10000.0, 29999.0, 30000.0)] == [0.0, 14.0, 14.0, 0.0]
    vc = _input_stepper("voltageClamp", dict(delay=10000.0, duration=30000.0, simpleSeriesResistance=100.0,
                                             targetVoltage=-20.0))
    assert vc(-65.0, -65.0, 0, 0, t=5000.0)[0] == 0.0
    assert vc(-65.0, -65.0, 0, 0, t=10000.0)[0] == np.float32(np.float32(45.0) / np.float32(100.0))
    assert vc(-65.0, -65.0, 0, 0, t=40000.0)[0] != 0.0                          # t = delay + duration: still on
    assert vc(-65.0, -65.0, 0, 0, t=40025.0)[0] == 0.0
    tc = _input_stepper("voltageClampTriple", dict(delay=10000.0, duration=30000.0, simpleSeriesResistance=100.0,
                                                   active=1.0, conditioningVoltage=-63.77, testingVoltage=0.0,
                                                   returnVoltage=-40.77))
    for t, target in ((0.0, -63.77), (10000.0, 0.0), (40025.0, -40.77)):
        assert tc(-65.0, -65.0, 0, 0, t=t)[0] == pytest.approx((target + 65.0) / 100.0, rel=1e-6), t


def test_voltage_clamp_holds_a_passive_cell_near_its_target(tmp_path):
    """A passive compartment under a voltage clamp with series resistance R settles where the
    leak and clamp currents cancel: V = (g E + V_t / R) / (g + 1 / R)."""
    p = tmp_path / "clamp.xml"
    p.write_text(
        '<Lems><Target component="s"/><neuroml><ionChannelHH id="pas" conductance="10pS"/>'
        '<voltageClamp id="vc" delay="5ms" duration="200ms" targetVoltage="-20mV" simpleSeriesResistance="1e8ohm"/>'
        '<cell id="c"><morphology id="m"><segment id="0"><proximal x="0" y="0" z="0" diameter="17.841242"/>'
        '<distal x="0" y="0" z="0" diameter="17.841242"/></segment></morphology><biophysicalProperties id="b">'
        '<membraneProperties><channelDensity id="leak" ionChannel="pas" condDensity="3.0 S_per_m2" erev="-54.3mV"/>'
        '<specificCapacitance value="1.0 uF_per_cm2"/><initMembPotential value="-65mV"/></membraneProperties>'
        '<intracellularProperties><resistivity value="0.03 kohm_cm"/></intracellularProperties>'
        '</biophysicalProperties></cell><network id="n"><population id="p" component="c" size="1"/>'
        '<inputList id="il" population="p" component="vc"><input id="0" target="../p/0/c" destination="synapses"/>'
        '</inputList></network></neuroml>'
        '<Simulation id="s" length="150ms" step="0.025ms" target="n" seed="1"><OutputFile id="o" fileName="o.txt">'
        '<OutputColumn id="v" quantity="p/0/c/0/v"/></OutputFile></Simulation></Lems>')
    m, _ = neuroml.read(str(p))
    assert m.lems.type_flags == [1]
    tr, _, _, _ = oracle.run(m)
    g = 0.3 * 1e-5 * float(m.area[0])        # 0.3 mS/cm^2 (3 S/m^2) over the area (um^2), in uS
    want = (g * -54.3 + -20.0 / 100.0) / (g + 1.0 / 100.0)
    assert float(tr[-1, 0]) == pytest.approx(want, abs=2e-3)
    # before the clamp's delay the membrane relaxes to the leak reversal by the engine's
    # implicit step: V_n = E + (V_0 - E) / (1 + dt g / C)^n
    C = float(m.capacitance[0])
    n = np.arange(150)
    want0 = -54.3 + (-65.0 + 54.3) / (1.0 + 25.0 * g / C) ** n
    np.testing.assert_allclose(tr[:150, 0], want0, atol=1e-3)   # float32 parameters and steps


def test_reference_core_inputs_and_custom_input_models_read(tmp_path):
    """The reference's CoreInputs model: one LEMS input type per generator element, flagged as
    inputs, each instance after the synapse types in its compartment's list, and the
    timedSynapticInput as an input instance of its expOneSynapse with the spike list (the
    model's unused Poisson inputs give their types too);
    EdenTest4's user-defined myPulseGenerator reads as an input type and its tree cells carry
    the instances; both run in the oracle."""
    from coreinputs import write_model
    m, _ = neuroml.read(write_model(str(tmp_path)))
    assert m.lems.type_names == ["expOneSynapse", "expTwoSynapse", "pulseGeneratorDL", "sineGenerator",
                                 "sineGeneratorDL", "rampGenerator", "rampGeneratorDL", "voltageClamp",
                                 "voltageClampTriple", "timedSynapticInput/expOneSynapse",
                                 "poissonFiringSynapse/expOneSynapse", "transientPoissonFiringSynapse/expTwoSynapse"]
    assert m.lems.type_flags == [0, 0, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1]
    comp_of = np.repeat(np.arange(m.n_comps), np.diff(m.comp_syn_begin))
    # InputDemos/02-06 (compartments 2-6): sine, ramp, clamp, triple clamp, the timed synaptic
    # input; 11, 12: the spike sources' synapses; InputDemosDL/01-03: the dimensionless generators
    assert list(comp_of) == [2, 3, 4, 5, 6, 11, 12, 18, 19, 20]
    assert list(m.syn_type) == [3, 5, 7, 8, 9, 0, 0, 2, 4, 6]
    assert list(m.syn_spike_list) == [-1, -1, -1, -1, 0, -1, -1, -1, -1, -1]
    assert m.spike_times[-1] == np.inf and len(m.spike_times) == 17
    tr, ec, er, _ = oracle.run(m)
    assert np.bincount(ec).tolist() == [5, 16]
    m4, _ = neuroml.read(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "neuroml",
                                      "LEMS_EdenTest4.xml"))
    assert m4.lems.type_names[-2:] == ["myPulseGenerator", "myPulseGenerator"] and m4.lems.type_flags[-2:] == [1, 1]
    tr4, _, _, _ = oracle.run(m4)
    assert np.array_equal(tr4[:, 0], tr4[:, 1])                 # a native pulse and ... the same pulse
    assert float(np.ptp(tr4[:, 4])) > 0.5


def _timed_model(tmp_path, times_ms, dt_ms=0.01, length_ms=5.0, both=False):
    """One passive cell driven by a timedSynapticInput of an expOneSynapse (and, with both,
    a second cell whose pulse current equals nothing, to check the input total's order)."""
    spikes = "".join(f'<spike id="{k}" time="{t} ms"/>' for k, t in enumerate(times_ms))
    pas = ('<cell id="P"><morphology id="m"><segment id="0"><proximal x="0" y="0" z="0" diameter="10"/>'
           '<distal x="0" y="0" z="0" diameter="10"/></segment></morphology><biophysicalProperties id="b">'
           '<membraneProperties><channelDensity id="leak" ionChannel="passiveChan" condDensity="0.3 mS_per_cm2" '
           'erev="-65mV"/><specificCapacitance value="1.0 uF_per_cm2"/><initMembPotential value="-65mV"/>'
           '</membraneProperties><intracellularProperties><resistivity value="0.1 kohm_cm"/>'
           '</intracellularProperties></biophysicalProperties></cell><ionChannelHH id="passiveChan" conductance="10pS"/>')
    p = tmp_path / "timed.xml"
    p.write_text(
        '<Lems><Target component="s"/><neuroml>' + pas +
        '<expOneSynapse id="e1" gbase="2nS" erev="0mV" tauDecay="1ms"/>'
        f'<timedSynapticInput id="ti" synapse="e1" spikeTarget="./e1">{spikes}</timedSynapticInput>'
        '<network id="n"><population id="pp" component="P" size="2"/>'
        '<inputList id="L" component="ti" population="pp"><input id="0" target="../pp/0/P" destination="synapses"/>'
        + ('<input id="1" target="../pp/1/P" destination="synapses"/>' if both else '') +
        '</inputList></network></neuroml>'
        f'<Simulation id="s" length="{length_ms}ms" step="{dt_ms}ms" target="n" seed="1">'
        '<OutputFile id="o" fileName="o.txt"><OutputColumn id="a" quantity="pp/0/P/0/v"/>'
        '<OutputColumn id="b" quantity="pp/1/P/0/v"/></OutputFile></Simulation></Lems>')
    return neuroml.read(str(p))[0]


def test_timed_synaptic_input_fires_on_the_listed_steps(tmp_path):
    """The spike list fires on every step whose end passes its next listed times, all of them
    consumed at once (eden/Eden.cpp:3780-3788), and the event reaches the synapse in that step:
    time T on step floor(T / dt) + 1 (t = (step - 1) dt, fired when t + dt > T); two times in
    one step give one event; a time equal to a step's end fires on the next step."""
    times = [0.003, 0.0102, 1.0, 2.001, 2.004, 3.5]
    m = _timed_model(tmp_path, times, both=True)
    assert list(m.syn_spike_list) == [0, 0] and list(m.spike_times[:6]) == [np.float32(t * 1000) for t in times]
    (tr, _, _, _), (st, sy) = oracle.run_with_synapse_log(m)
    for inst in (0, 1):
        steps = sorted(int(x) for x in st[sy == inst])
        assert steps == [1, 2, 101, 201, 351], (inst, steps)
    # expOneSynapse exposes its pre-event conductance: the event step (row 1) adds no current,
    # the next step depolarises the cell from rest
    assert tr[1, 0] == np.float32(-65.0) and tr[2, 0] > np.float32(-65.0)
    assert np.array_equal(tr[:, 0], tr[:, 1])


def test_timed_synaptic_input_refuses_custom_settings(tmp_path):
    """The reference has no per-instance paths into a timedSynapticInput
    (eden/NeuroML.cpp:1746-1749); neither does the reader."""
    m = _timed_model(tmp_path, [1.0])
    src = (tmp_path / "timed.xml").read_text().replace("</Simulation>", '<EdenCustomSetup filename="c.txt"/></Simulation>')
    (tmp_path / "timed.xml").write_text(src)
    (tmp_path / "c.txt").write_text("set input L all gbase 1 nS\n")
    with pytest.raises(neuroml.NeuroMLError, match="timedSynapticInput paths"):
        neuroml.read(str(tmp_path / "timed.xml"))
    assert m.lems.type_flags == [0, 1]
