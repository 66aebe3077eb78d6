"""Pins of the oracle (the CPU restatement) against what the mathematics fixes.

The reference ships no golden vectors for this path, and its build needs generated sources
the image cannot produce, so each oracle function is checked against an independent route:
an inverse or multiplicative formulation (RNG hash), textbook HH integrated by scipy, closed
forms of the discrete recurrences (passive membrane, synapse charge), a dense linear solve
(Hines), a direct recursive DFS (elimination order) and hand-enumerated event tables (delay
rule). A plausible slip (a wrong shift, a dropped unit factor, a transposed index, an
off-by-one step) fails at least one of them.
"""
import numpy as np
import pytest
from scipy.integrate import solve_ivp

from eden_amd import _native as N
from eden_amd import configs
from eden_amd.model import Model, channel_type, hh_squid_types, rate
from oracle import oracle

M64 = (1 << 64) - 1


# ---------------------------------------------------------------- RNG (eden/Eden.cpp:3200-3233)

def _xorshift_right_inverse(y, k):
    x = y
    for _ in range(64 // k + 1):
        x = y ^ (x >> k)
    return x & M64


def _hash_inverse(h):
    """Inverse of Thomas Wang's hash64shift, step by step: the add-shift steps are
    multiplications mod 2^64 (by 2^31+1, 21, 265, 2^21-1 after the complement) and the
    xor-shifts are inverted by iteration."""
    k = h
    k = (k * pow((1 << 31) + 1, -1, 1 << 64)) & M64
    k = _xorshift_right_inverse(k, 28)
    k = (k * pow(21, -1, 1 << 64)) & M64
    k = _xorshift_right_inverse(k, 14)
    k = (k * pow(265, -1, 1 << 64)) & M64
    k = _xorshift_right_inverse(k, 24)
    # (~key) + (key << 21) = key * (2^21 - 1) - 1
    k = ((k + 1) * pow((1 << 21) - 1, -1, 1 << 64)) & M64
    return k


def test_hash64shift_is_inverted_by_published_steps():
    rng = np.random.default_rng(1)
    keys = [0, 1, 2, 0xFFFFFFFFFFFFFFFF, 0x8000000000000000] + [int(x) for x in rng.integers(0, 2**63, 200)]
    for key in keys:
        h = oracle.hash64shift(key)
        assert 0 <= h <= M64
        assert _hash_inverse(h) == key


def test_randof_uniform_and_periodic():
    vals = np.array([oracle.randof(1.0, 12345, inst, step, inv)
                     for inst in range(4) for step in range(-3, 500) for inv in range(2)])
    assert vals.min() >= 0.0 and vals.max() < 1.0
    assert abs(vals.mean() - 0.5) < 0.02
    # resolution 2^-23
    assert np.all(np.round(vals * 2**23) == vals * 2**23)
    # stamp fields wrap: step modulo 2^40, instance modulo 2^24
    for step in (-3, -1, 0, 7, 123456):
        assert oracle.randof(1.0, 7, 3, step, 1) == oracle.randof(1.0, 7, 3, step + (1 << 40), 1)
    assert oracle.randof(1.0, 7, 3, 9, 0) == oracle.randof(1.0, 7, 3 + (1 << 24), 9, 0)
    # the scale argument multiplies
    assert oracle.randof(4.0, 7, 3, 9, 0) == pytest.approx(4.0 * oracle.randof(1.0, 7, 3, 9, 0), rel=1e-7)
    # distinct stamps give distinct streams
    a = [oracle.randof(1.0, 7, 3, s, 0) for s in range(64)]
    b = [oracle.randof(1.0, 7, 3, s, 1) for s in range(64)]
    assert a != b


def test_cell_rng_seed_reverses_bits():
    from eden_amd.model import cell_rng_seeds, reverse_bits32
    assert reverse_bits32(1) == 0x80000000
    assert reverse_bits32(0x80000000) == 1
    assert reverse_bits32(0x0000FFFF) == 0xFFFF0000
    s = cell_rng_seeds(3, 5)
    assert list(s) == [reverse_bits32(3) ^ g for g in range(5)]


# ---------------------------------------------------------------- rate forms (eden/Eden.cpp:5126-5200)

@pytest.mark.parametrize("precision,tol", [("f64", 1e-12), ("f32", 3e-6)])
def test_rate_forms_match_direct_evaluation(precision, tol):
    V = np.linspace(-100.0, 60.0, 321)
    f = lambda x: float(np.float32(x))      # rates are stored as fp32 (1/us)
    cases = [
        (N.RATE_EXP, 4.0, -65.0, -18.0, lambda v: f(4e-3) * np.exp((v + 65.0) / -18.0)),
        (N.RATE_EXPLINEAR, 1.0, -40.0, 10.0,
         lambda v: np.where(v == -40.0, f(1e-3), f(1e-3) * ((v + 40.0) / 10.0) / (1.0 - np.exp(-(v + 40.0) / 10.0)))),
        (N.RATE_SIGMOID, 1.0, -35.0, 10.0, lambda v: f(1e-3) / (1.0 + np.exp((-35.0 - v) / 10.0))),
        (N.RATE_FIXED, 0.5, 0.0, 1.0, lambda v: np.full_like(v, f(5e-4))),
    ]
    for form, r, mid, sc, ref in cases:
        rr = rate(form, r, mid, sc)
        got = np.array([oracle.rate(rr, v, precision) for v in V])
        with np.errstate(invalid="ignore", divide="ignore"):
            want = ref(V)
        np.testing.assert_allclose(got, want, rtol=tol, atol=0)
    # EXPLINEAR is continuous through its removable singularity
    rr = rate(N.RATE_EXPLINEAR, 1.0, -40.0, 10.0)
    assert oracle.rate(rr, -40.0, "f64") == pytest.approx(oracle.rate(rr, -40.0 + 1e-7, "f64"), rel=1e-6)


# ---------------------------------------------------------------- HH vs textbook ODE

def _hh_textbook_rhs(I_density_uA_cm2):
    def f(t, y):
        v, m, h, n = y
        am = 0.1 * (v + 40.0) / (1.0 - np.exp(-(v + 40.0) / 10.0))
        bm = 4.0 * np.exp(-(v + 65.0) / 18.0)
        ah = 0.07 * np.exp(-(v + 65.0) / 20.0)
        bh = 1.0 / (1.0 + np.exp(-(v + 35.0) / 10.0))
        an = 0.01 * (v + 55.0) / (1.0 - np.exp(-(v + 55.0) / 10.0))
        bn = 0.125 * np.exp(-(v + 65.0) / 80.0)
        Iapp = I_density_uA_cm2(t)
        dv = Iapp - 120.0 * m**3 * h * (v - 50.0) - 36.0 * n**4 * (v + 77.0) - 0.3 * (v + 54.387)
        return [dv / 1.0, am * (1 - m) - bm * m, ah * (1 - h) - bh * h, an * (1 - n) - bn * n]
    return f


def _spike_times(t, v, thr=-20.0):
    i = np.nonzero((v[:-1] < thr) & (v[1:] >= thr))[0]
    return t[i] + (thr - v[i]) / (v[i + 1] - v[i]) * (t[i + 1] - t[i])


def test_hh_oracle_converges_to_textbook_hh():
    """Single-compartment HH squid axon (units, gate powers, rate constants, the 1e-5
    density factor, area and capacitance) against scipy on the textbook equations."""
    amp = 0.15
    m = configs.hh_population(n_cells=1, t_final_ms=40.0, amp_max_nA=1.0)
    m.pulse_amplitude[:] = amp
    area = float(m.area[0])
    # oracle FP64 at dt/25 = 1 us
    tr, ec, er, _ = oracle.run(m, precision="f64", substeps=25)
    t_or = np.arange(tr.shape[0]) * 0.025
    dens = amp / area * 1e5   # nA/um^2 -> uA/cm^2

    def I(t):
        return dens if 10.0 <= t <= 90.0 else 0.0
    v0 = -65.0
    am = 0.1 * (v0 + 40) / (1 - np.exp(-(v0 + 40) / 10)); bm = 4 * np.exp(-(v0 + 65) / 18)
    ah = 0.07 * np.exp(-(v0 + 65) / 20); bh = 1 / (1 + np.exp(-(v0 + 35) / 10))
    an = 0.01 * (v0 + 55) / (1 - np.exp(-(v0 + 55) / 10)); bn = 0.125 * np.exp(-(v0 + 65) / 80)
    y0 = [v0, am / (am + bm), ah / (ah + bh), an / (an + bn)]
    sol = solve_ivp(_hh_textbook_rhs(I), (0.0, 40.0), y0, method="LSODA", rtol=1e-10, atol=1e-10,
                    max_step=0.01, dense_output=True)
    v_ref = sol.sol(t_or)[0]
    s_or = _spike_times(t_or, tr[:, 0])
    s_ref = _spike_times(t_or, v_ref)
    assert len(s_ref) >= 2 and len(s_or) == len(s_ref)
    np.testing.assert_allclose(s_or, s_ref, atol=0.05)
    # before the pulse both sit at rest
    assert np.abs(tr[: int(9.5 / 0.025), 0] - v_ref[: int(9.5 / 0.025)]).max() < 0.05


# ---------------------------------------------------------------- passive membrane closed form

def _one_passive_cell(g=0.3, E=-54.387, v0=-54.387, amp=0.05, start_ms=5.0, dur_ms=1e6, t_final_ms=30.0,
                      diameter=20.0):
    prox = np.array([[0, 0, 0, diameter]], np.float32)
    area, cap, rax, inv = N.cable_setup(np.array([-1]), prox, prox, 1.0, 0.1)
    return Model(
        dt=25.0, t_final=t_final_ms * 1000.0, cell_comp_begin=np.array([0, 1]),
        v0=[v0], capacitance=cap, area=area, inv_rc_axial=[0.0], r_axial=[np.nan], v_threshold=[np.nan],
        parent=[-1], channel_types=[channel_type([])], comp_chan_begin=[0, 1], chan_type=[0],
        chan_gbase=[g], chan_erev=[E], comp_pulse_begin=[0, 1], pulse_start=[start_ms * 1000.0],
        pulse_duration=[dur_ms * 1000.0], pulse_amplitude=[amp], pulse_weight=[1.0], record_comp=[0])


def test_passive_membrane_matches_backward_euler_closed_form():
    g, E, amp = 0.3, -54.387, 0.05
    m = _one_passive_cell(g=g, E=E, amp=amp)
    G = g * float(m.area[0]) * 1e-5            # uS
    C = float(m.capacitance[0])                 # pF
    a = 25.0 * G / C
    Vinf = E + amp / G
    k0 = int(5.0 / 0.025)                       # first row whose time is inside the pulse
    r = np.arange(m.total_rows())
    want = np.where(r <= k0, E, Vinf + (E - Vinf) * (1.0 + a) ** (-(r - k0).astype(float)))
    tr64, _, _, _ = oracle.run(m, precision="f64")
    np.testing.assert_allclose(tr64[:, 0], want, atol=1e-9)
    tr32, _, _, _ = oracle.run(m, precision="f32")
    # fp32: per-step rounding accumulates over the 1,200 steps (about 1e-3 mV at most)
    np.testing.assert_allclose(tr32[:, 0], want, atol=2e-3)
    # the pulse window is inclusive at both ends: a pulse ending exactly on a grid time still
    # acts at that step
    m2 = _one_passive_cell(g=g, E=E, amp=amp, start_ms=5.0, dur_ms=1.0)
    tr, _, _, _ = oracle.run(m2, precision="f64")
    k1 = int(6.0 / 0.025)
    want2 = np.array(want)
    want2[k1 + 1:] = E + (want[k1 + 1] - E) * (1.0 + a) ** (-(r[k1 + 1:] - (k1 + 1)).astype(float))
    np.testing.assert_allclose(tr[:, 0], want2, atol=1e-9)


# ---------------------------------------------------------------- Hines vs dense solve

def _random_passive_tree(n, seed):
    rng = np.random.default_rng(seed)
    par, prox, dist = configs.random_tree(rng, n)
    area, cap, rax, inv = N.cable_setup(par, prox, dist, 1.0, 0.1)
    v0 = rng.uniform(-80.0, -40.0, n).astype(np.float32)
    g = rng.uniform(0.1, 1.0, n).astype(np.float32)
    E = rng.uniform(-70.0, -50.0, n).astype(np.float32)
    m = Model(dt=25.0, t_final=100.0, cell_comp_begin=np.array([0, n]), v0=v0, capacitance=cap, area=area,
              inv_rc_axial=inv, r_axial=rax, v_threshold=np.full(n, np.nan), parent=par,
              channel_types=[channel_type([])], comp_chan_begin=np.arange(n + 1), chan_type=np.zeros(n),
              chan_gbase=g, chan_erev=E, record_comp=np.arange(n))
    return m


@pytest.mark.parametrize("n,seed", [(2, 1), (7, 2), (33, 3), (64, 4)])
def test_hines_step_matches_dense_solve(n, seed):
    m = _random_passive_tree(n, seed)
    tr, _, _, _ = oracle.run(m, precision="f64", n_rows=2)
    dt = 25.0
    Cm = m.capacitance.astype(float)
    G = m.chan_gbase.astype(float) * m.area.astype(float) * 1e-5
    A = np.diag(Cm / dt + G)
    for i in range(1, n):
        j = m.parent[i]
        y = 1.0 / float(m.r_axial[i])
        A[i, i] += y; A[j, j] += y; A[i, j] -= y; A[j, i] -= y
    b = Cm / dt * m.v0.astype(float) + G * m.chan_erev.astype(float)
    v1 = np.linalg.solve(A, b)
    np.testing.assert_allclose(tr[0], m.v0, atol=0)
    np.testing.assert_allclose(tr[1], v1, rtol=0, atol=2e-5)   # the trace is stored in fp32
    # the per-compartment sum of 1/RC matches the adjacency of the tree
    want_inv = np.zeros(n)
    for i in range(1, n):
        j = m.parent[i]
        want_inv[i] += 1.0 / (m.r_axial[i] * Cm[i])
        want_inv[j] += 1.0 / (m.r_axial[i] * Cm[j])
    np.testing.assert_allclose(m.inv_rc_axial, want_inv, rtol=1e-5)


def test_dfs_order_is_post_order_with_ascending_children():
    rng = np.random.default_rng(5)
    for n in (1, 2, 5, 40, 300):
        par = np.full(n, -1, np.int32)
        for i in range(1, n):
            par[i] = rng.integers(0, i)
        children = [[] for _ in range(n)]
        for i in range(1, n):
            children[par[i]].append(i)
        want = []

        def visit(v):
            for c in sorted(children[v]):
                visit(c)
            want.append(v)
        import sys
        sys.setrecursionlimit(10000)
        visit(0)
        assert list(oracle.dfs_order(par)) == want


# ---------------------------------------------------------------- cable geometry (host setup)

def test_cable_setup_closed_forms():
    from scipy.integrate import quad
    d, L = 2.0, 30.0
    # two identical cylinders in a row
    prox = np.array([[0, 0, 0, d], [L, 0, 0, d]], np.float32)
    dist = np.array([[L, 0, 0, d], [2 * L, 0, 0, d]], np.float32)
    area, cap, rax, inv = N.cable_setup(np.array([-1, 0]), prox, dist, 1.0, 0.1)
    np.testing.assert_allclose(area, np.pi * d * L, rtol=1e-6)
    np.testing.assert_allclose(cap, np.pi * d * L * 1e-2, rtol=1e-6)          # uF/cm^2 * um^2 -> pF
    R_half = 0.1 * 1e4 * (L / 2) / (np.pi * (d / 2) ** 2) * 1e-6 * 1e-2 * 1e6   # Ohm: rho[Ohm cm]*len/area
    # rho = 0.1 kOhm cm = 100 Ohm cm; length and area in cm; result in MOhm
    rho = 100.0
    R_half = rho * (L / 2 * 1e-4) / (np.pi * (d / 2 * 1e-4) ** 2) / 1e6
    np.testing.assert_allclose(rax[1], 2 * R_half, rtol=1e-6)
    np.testing.assert_allclose(inv, [1 / (2 * R_half * cap[0]), 1 / (2 * R_half * cap[1])], rtol=1e-6)
    # sphere (zero-length segment): area pi d^2
    p = np.array([[0, 0, 0, 17.8]], np.float32)
    a, c, r, i = N.cable_setup(np.array([-1]), p, p, 1.0, 0.1)
    np.testing.assert_allclose(a, np.pi * 17.8 ** 2, rtol=1e-6)
    # tapered frustum: lateral area and resistance by numerical integration
    d1, d2, L = 2.0, 0.5, 40.0
    prox = np.array([[0, 0, 0, 3.0], [0, 0, 0, d1]], np.float32)
    dist = np.array([[0, 0, 0, 3.0], [0, L, 0, d2]], np.float32)
    a, c, r, i = N.cable_setup(np.array([-1, 0]), prox, dist, 1.0, 0.1)
    slant = np.sqrt(((d1 - d2) / 2) ** 2 + L ** 2)
    np.testing.assert_allclose(a[1], np.pi * (d1 + d2) / 2 * slant, rtol=1e-6)

    def radius(x):
        return (d1 + (d2 - d1) * x / L) / 2
    # proximal half of the frustum plus the distal half of the zero-length parent (a cylinder
    # of length d = 3 um)
    r_prox = quad(lambda x: rho * 1e-4 / (np.pi * (radius(x) * 1e-4) ** 2), 0, L / 2)[0] / 1e6
    r_par = rho * (1.5e-4) / (np.pi * (1.5e-4) ** 2) / 1e6
    np.testing.assert_allclose(r[1], r_prox + r_par, rtol=1e-5)


# ---------------------------------------------------------------- expTwo synapse charge

def test_exptwo_charge_matches_closed_form():
    """One event on an expTwo synapse of a cell with a huge capacitance (V ~ constant): the
    voltage step equals the synaptic charge / C, and with forward Euler the discrete sums of A
    and B are tau_rise and tau_decay times their jumps, so the charge is
    gbase * weight * wf * (tau_decay - tau_rise) * (E - V)."""
    tr_, td_ = 500.0, 5000.0
    w_conn = 2.0
    gbase = 1e-3
    m = configs.spiking_network(n_cells=2, in_degree=1, t_final_ms=400.0, poisson_hz=0.0)
    # cell 0: presynaptic, fires once on a strong short pulse; cell 1: huge capacitance, no channels
    nc = 2
    m.comp_pulse_begin = np.array([0, 1, 1])
    m.pulse_start = np.array([1000.0], np.float32)
    m.pulse_duration = np.array([1000.0], np.float32)
    m.pulse_amplitude = np.array([2.0], np.float32)
    m.pulse_weight = np.array([1.0], np.float32)
    m.comp_poisson_begin = np.zeros(nc + 1, np.int64)
    m.poisson_rate = np.zeros(0, np.float32)
    m.comp_chan_begin = np.array([0, 3, 3])
    m.chan_type = m.chan_type[:3]
    m.chan_gbase = m.chan_gbase[:3]
    m.chan_erev = m.chan_erev[:3]
    m.capacitance = np.array([m.capacitance[0], 1e7], np.float32)
    m.v0 = np.array([-65.0, 0.0], np.float32)       # target near 0 mV: fp32 resolves the tiny step
    m.syn_types = [N.ExpTwoType(gbase, 65.0, tr_, td_, 1.0)]
    m.comp_syn_begin = np.array([0, 0, 1])
    m.syn_type = np.array([0])
    m.syn_weight = np.array([w_conn], np.float32)
    m.syn_delay = np.array([1000.0], np.float32)
    m.comp_spike_begin = np.array([0, 1, 1])
    m.spike_target = np.array([0])
    m.record_comp = np.array([1])
    m.event_comp = np.array([0])
    m.poisson_syn_type = np.zeros(0, np.int32)
    m.poisson_weight = np.zeros(0, np.float32)
    m.poisson_instance = np.zeros(0, np.int32)
    m.poisson_rng_offset = np.zeros(0, np.int32)
    m._normalise()
    (tr, ec, er, _), (sstep, ssyn) = oracle.run_with_synapse_log(m, precision="f64")
    assert len(er) >= 1 and len(sstep) == len(er)
    peak = np.log(td_ / tr_) * (tr_ * td_) / (td_ - tr_)
    wf = 1.0 / (-np.exp(-peak / tr_) + np.exp(-peak / td_))
    charge = gbase * w_conn * wf * (td_ - tr_) * (65.0 - 0.0) * len(sstep)   # uS * mV * us = nA * us
    dv = charge / 1e7                                                       # nA * us / pF = mV
    assert tr[-1, 0] - tr[0, 0] == pytest.approx(dv, rel=2e-3)


# ---------------------------------------------------------------- delay rule, hand-enumerated

def _two_cell_delay_model(delay_ms, amp=1.0):
    m = configs.spiking_network(n_cells=2, in_degree=1, t_final_ms=150.0, poisson_hz=0.0)
    m.comp_pulse_begin = np.array([0, 1, 1])
    m.pulse_start = np.array([5000.0], np.float32)
    m.pulse_duration = np.array([1e6], np.float32)
    m.pulse_amplitude = np.array([amp], np.float32)
    m.pulse_weight = np.array([1.0], np.float32)
    m.comp_poisson_begin = np.zeros(3, np.int64)
    m.poisson_rate = np.zeros(0, np.float32)
    m.poisson_syn_type = np.zeros(0, np.int32)
    m.poisson_weight = np.zeros(0, np.float32)
    m.poisson_instance = np.zeros(0, np.int32)
    m.poisson_rng_offset = np.zeros(0, np.int32)
    m.comp_syn_begin = np.array([0, 0, 1])
    m.syn_type = np.array([0])
    m.syn_weight = np.array([1e-6], np.float32)      # negligible effect on the target
    m.syn_delay = np.array([delay_ms * 1000.0], np.float32)
    m.comp_spike_begin = np.array([0, 1, 1])
    m.spike_target = np.array([0])
    m.event_comp = np.array([0])
    m.record_comp = np.array([0, 1])
    m._normalise()
    return m


def expected_deliveries(spike_steps, D):
    """Single pending slot per synapse (eden/Eden.cpp:3658-3685), in integer steps: the spike
    detected while computing step s sets the trigger read at step s + 1, whose arrival is
    s + 1 + D (D = delay / dt); a trigger is dropped while an arrival is still pending beyond
    the current step."""
    out = []
    pending = None
    for s in spike_steps:
        k = s + 1
        if pending is not None and pending > k:   # next_spike >= t + dt: keep the first
            continue
        pending = k + D
        out.append(pending)
    return out


@pytest.mark.parametrize("delay_ms", [1.0, 2.5, 15.0])
def test_delay_rule_hand_enumerated(delay_ms):
    m = _two_cell_delay_model(delay_ms)
    (tr, ec, er, _), (sstep, ssyn) = oracle.run_with_synapse_log(m)
    assert len(er) >= 4
    # event rows: a spike detected while computing step s is logged at row s
    spike_steps = list(er)
    D = int(round(delay_ms / 0.025))
    want = [k for k in expected_deliveries(spike_steps, D) if k <= m.total_rows()]
    assert list(sstep) == want
    assert set(ssyn.tolist()) <= {0}
    if delay_ms == 15.0:
        # inter-spike intervals are shorter than the delay: some spikes must be dropped
        assert len(want) < len(spike_steps)


def test_model_subset_runs_as_in_the_full_batch():
    """Model.subset keeps each cell's data: the oracle's trace and events of a subset equal the
    full model's columns for those cells (cells without synapses are independent)."""
    from eden_amd import configs
    m = configs.random_trees(6, sizes=[20, 30, 40, 50, 60, 70], t_final_ms=12.0, seed=7, record_all=True)
    sub, rc, ec = m.subset([4, 1, 5])
    assert sub.n_cells == 3 and sub.n_comps == 60 + 30 + 70
    tr, c_full, r_full, _ = oracle.run(m)
    ts, c_sub, r_sub, _ = oracle.run(sub)
    np.testing.assert_array_equal(tr[:, rc], ts)
    keep = np.isin(c_full, ec)
    new = {int(c): i for i, c in enumerate(ec)}
    assert sorted(zip(r_full[keep].tolist(), [new[int(c)] for c in c_full[keep]])) == \
        sorted(zip(r_sub.tolist(), c_sub.tolist()))


# ---------------------------------------------------------------- Poisson input process
# poissonFiringSynapse (LEMS_CoreComponents.inc.xml:623-672; emitted at eden/Eden.cpp:3859-3912):
# OnStart tsince = 0, isi = -averageIsi log(1 - random(1)); every step tsince += dt; when
# tsince > isi: tsince = 0, a new isi, a spike. After a firing at step s0 the step s0 + 1 sees
# tsince = 0 and step s0 + 1 + j sees j dt, so the next firing is floor(isi / dt) + 2 steps
# later: the firing interval in steps is k = floor(isi / dt) + 2 with isi ~ Exp(rate). Then
# (k - 2 + U) dt, U ~ U(0, 1), is Exp(rate) up to the shape of the density inside one dt
# (a relative error of order rate * dt), and the mean interval is
# dt (e^{-r dt} / (1 - e^{-r dt}) + 2). OnStart draws the first isi at step 0, so step 0
# starts an interval too. Intervals still open at the end of the run are never seen, which
# would bias the sample to short ones: only intervals that start at least 15 mean intervals
# before the end are used (the choice depends on the start alone, so it selects intervals
# without regard to their length; one would still be open with probability e^-15).

def _poisson_model(rate_hz, n_cells=200, t_final_ms=None):
    if t_final_ms is None:
        t_final_ms = 5000.0 if rate_hz < 50 else 2500.0
    nc = n_cells
    passive = [channel_type([])]
    return Model(
        dt=25.0, t_final=t_final_ms * 1000.0, cell_comp_begin=np.arange(nc + 1, dtype=np.int64),
        v0=np.full(nc, -65.0), capacitance=np.full(nc, 10.0), area=np.full(nc, 1000.0),
        inv_rc_axial=np.zeros(nc), r_axial=np.full(nc, np.nan), v_threshold=np.full(nc, np.nan),
        parent=np.full(nc, -1), channel_types=passive, comp_chan_begin=np.arange(nc + 1, dtype=np.int64),
        chan_type=np.zeros(nc), chan_gbase=np.full(nc, 0.3), chan_erev=np.full(nc, -65.0),
        cell_rng_seed=configs.cell_rng_seeds(7, nc),
        syn_types=[N.ExpTwoType(1e-6, 0.0, 500.0, 5000.0, 1.0)],
        comp_poisson_begin=np.arange(nc + 1, dtype=np.int64), poisson_rate=np.full(nc, rate_hz * 1e-6),
        poisson_syn_type=np.zeros(nc), poisson_weight=np.ones(nc), poisson_instance=np.zeros(nc),
        poisson_rng_offset=np.zeros(nc), name=f"poisson_{rate_hz}Hz")


def _poisson_stats(model, rate_hz, library=None):
    """(count, expected count, KS p-value of the dt-corrected intervals against Exp(rate))."""
    from scipy import stats
    _, (step, inp) = oracle.run_with_poisson_log(model, library=library)
    dt = 25.0
    r = rate_hz * 1e-6
    T = (model.total_rows() - 1) * dt
    q = np.exp(-r * dt)
    mean_isi = dt * (q / (1.0 - q) + 2.0)
    expected = model.n_cells * T / mean_isi
    t_cut = T - 15.0 / r
    ks = []
    for c in range(model.n_cells):
        f = np.concatenate([[0], step[inp == c]])
        ks.append(np.diff(f)[(f[:-1] * dt < t_cut)[: len(f) - 1]])
    k = np.concatenate(ks)
    u = np.random.default_rng(3).random(len(k))
    x = (k - 2 + u) * dt
    p = stats.kstest(x, "expon", args=(0.0, 1.0 / r)).pvalue if len(x) > 10 else 0.0
    return len(step), expected, p


@pytest.fixture(scope="module")
def planted_oracles(tmp_path_factory):
    """The oracle's source built with each planted mistake (test-only macros)."""
    import os
    import subprocess
    src = os.path.join(os.path.dirname(oracle.__file__), "eden_oracle.c")
    out = {}
    d = tmp_path_factory.mktemp("planted")
    for name, macro in (("rate_per_ms", "ORACLE_PLANT_POISSON_RATE_PER_MS"), ("no_reset", "ORACLE_PLANT_POISSON_NO_RESET")):
        so = str(d / f"oracle_{name}.so")
        subprocess.run(["gcc", "-O2", "-std=c11", "-fPIC", "-fopenmp", "-shared", "-DREAL=float", "-DSFX=f32",
                        f"-D{macro}", src, "-o", so, "-lm"], check=True)
        out[name] = oracle.load_variant(so)
    return out


@pytest.mark.parametrize("rate_hz", [5.0, 200.0])
def test_poisson_input_count_and_interval_distribution(rate_hz):
    m = _poisson_model(rate_hz)
    n, expected, p = _poisson_stats(m, rate_hz)
    assert m.total_rows() >= 100_000
    assert abs(n - expected) <= 4.0 * np.sqrt(expected), (n, expected)
    assert p > 1e-3, p


@pytest.mark.parametrize("mistake", ["rate_per_ms", "no_reset"])
def test_poisson_checks_fail_on_planted_mistakes(planted_oracles, mistake):
    for rate_hz in (5.0, 200.0):
        m = _poisson_model(rate_hz, t_final_ms=500.0)
        n, expected, p = _poisson_stats(m, rate_hz, library=planted_oracles[mistake])
        count_ok = abs(n - expected) <= 4.0 * np.sqrt(expected)
        assert not (count_ok and p > 1e-3), (mistake, rate_hz, n, expected, p)


def test_randof_field_periods():
    """Each stamp field wraps at its width (eden/Eden.cpp:3223-3226): the invocation field at
    2^24 (it is shifted by 40 in a 64-bit word), the instance at 2^24, the step at 2^40, and a
    negative work item (sign-extended) differs from its 32-bit unsigned reading."""
    for inv in (0, 1, 5):
        assert oracle.randof(1.0, 9, 4, 77, inv) == oracle.randof(1.0, 9, 4, 77, inv + (1 << 24))
        assert oracle.randof(1.0, 9, 4, 77, inv) != oracle.randof(1.0, 9, 4, 77, inv + (1 << 23))
    assert oracle.randof(1.0, 9, 5, 77, 0) != oracle.randof(1.0, 9, 5 + (1 << 23), 77, 0)
    assert oracle.randof(1.0, 9, 5, 77, 0) != oracle.randof(1.0, 9, 5, 77 + (1 << 39), 0)
    assert oracle.randof(1.0, -1, 5, 77, 0) != oracle.randof(1.0, 0x7FFFFFFF, 5, 77, 0)
