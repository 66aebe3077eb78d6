"""Synthetic workloads (instantiated models) for the five benchmark configurations.

Every generator takes its sizes as arguments so tests can build scaled-down versions of the
same construction. Random streams are numpy ``default_rng(seed)``. Geometry goes through
``edenhip_cable_setup`` (the library's restatement of EDEN's cable analysis), one
compartment per segment.

  hh_population      1,000 unconnected one-compartment HH cells + pulse (config 1)
  random_trees       random binary-tree cells, HH soma + passive dendrites (configs 2, 4, 5)
  spiking_network    one-compartment HH network, expTwo synapses, Poisson drive (config 3)
"""
import numpy as np

from . import _native as N
from .model import Model, cell_rng_seeds, hh_squid_types

MS = 1000.0            # us per ms
CM = 1.0               # uF/cm^2
RA = 0.1               # kOhm*cm (100 Ohm*cm)
HH_G = dict(leak=0.3, na=120.0, k=36.0)            # mS/cm^2
HH_E = dict(leak=-54.387, na=50.0, k=-77.0)        # mV
V0 = -65.0
VT = -20.0
DT = 25.0              # us


def _hh_channels(n_comp, hh_mask, g_scale=None):
    """Channel CSR: leak everywhere, then Na and K where hh_mask (distribution order
    leak, Na, K as in the HH example's biophysics). g_scale: optional per-comp (3,) factors."""
    counts = np.where(hh_mask, 3, 1)
    begin = np.zeros(n_comp + 1, np.int64)
    np.cumsum(counts, out=begin[1:])
    total = int(begin[-1])
    ctype = np.empty(total, np.int32)
    g = np.empty(total, np.float32)
    e = np.empty(total, np.float32)
    s = np.ones((n_comp, 3), np.float32) if g_scale is None else np.asarray(g_scale, np.float32)
    pos = begin[:-1]
    ctype[pos] = 0
    g[pos] = np.float32(HH_G["leak"]) * s[:, 0]
    e[pos] = HH_E["leak"]
    hp = pos[hh_mask]
    ctype[hp + 1] = 1
    g[hp + 1] = np.float32(HH_G["na"]) * s[hh_mask, 1]
    e[hp + 1] = HH_E["na"]
    ctype[hp + 2] = 2
    g[hp + 2] = np.float32(HH_G["k"]) * s[hh_mask, 2]
    e[hp + 2] = HH_E["k"]
    return begin, ctype, g, e


def hh_population(n_cells=1000, t_final_ms=100.0, seed=1001, amp_max_nA=0.2, diameter=17.8, sim_seed=1,
                  record_events=True):
    """Config 1: unconnected one-compartment HH cells, soma diameter 17.8 um (a zero-length
    segment: EDEN takes it as a sphere, area pi d^2), pulse delay 10 ms, duration 80 ms,
    amplitude U(0, amp_max) nA."""
    rng = np.random.default_rng(seed)
    amp = rng.uniform(0.0, amp_max_nA, n_cells).astype(np.float32)
    prox = np.array([[0, 0, 0, diameter]], np.float32)
    area, cap, rax, inv = N.cable_setup(np.array([-1]), prox, prox, CM, RA)
    nc = n_cells
    ch_b, ch_t, ch_g, ch_e = _hh_channels(nc, np.ones(nc, bool))
    m = Model(
        dt=DT, t_final=t_final_ms * MS,
        cell_comp_begin=np.arange(nc + 1, dtype=np.int64),
        v0=np.full(nc, V0), capacitance=np.full(nc, cap[0]), area=np.full(nc, area[0]),
        inv_rc_axial=np.zeros(nc), r_axial=np.full(nc, np.nan), v_threshold=np.full(nc, VT),
        parent=np.full(nc, -1),
        channel_types=hh_squid_types(), comp_chan_begin=ch_b, chan_type=ch_t, chan_gbase=ch_g, chan_erev=ch_e,
        cell_rng_seed=cell_rng_seeds(sim_seed, nc),
        comp_pulse_begin=np.arange(nc + 1), pulse_start=np.full(nc, 10.0 * MS), pulse_duration=np.full(nc, 80.0 * MS),
        pulse_amplitude=amp, pulse_weight=np.ones(nc),
        record_comp=np.arange(nc), event_comp=np.arange(nc) if record_events else None,
        name=f"hh_population_{n_cells}")
    m.morphologies = [(np.array([-1], np.int32), prox, prox)] * nc
    return m


def random_binary_parents(rng, n):
    """Node k attaches to a uniform node with fewer than two children."""
    parent = np.full(n, -1, np.int32)
    nchild = np.zeros(n, np.int32)
    open_nodes = [0]
    for k in range(1, n):
        i = int(rng.integers(len(open_nodes)))
        p = open_nodes[i]
        parent[k] = p
        nchild[p] += 1
        if nchild[p] == 2:
            open_nodes[i] = open_nodes[-1]
            open_nodes.pop()
        open_nodes.append(k)
    return parent


def tree_geometry(rng, parent):
    """Segments for a parent array: length U(10, 50) um, random direction, diameter linear in
    path distance from 2 um at the root to 0.5 um at the farthest distal point. Returns
    proximal, distal (x, y, z, diameter)."""
    n = len(parent)
    length = rng.uniform(10.0, 50.0, n)
    direc = rng.normal(size=(n, 3))
    direc /= np.linalg.norm(direc, axis=1, keepdims=True)
    direc[0] = (1.0, 0.0, 0.0)
    pstart = np.zeros((n, 3))
    pend = np.zeros((n, 3))
    path0 = np.zeros(n)
    path1 = np.zeros(n)
    for k in range(n):
        if parent[k] >= 0:
            pstart[k] = pend[parent[k]]
            path0[k] = path1[parent[k]]
        pend[k] = pstart[k] + length[k] * direc[k]
        path1[k] = path0[k] + length[k]
    pmax = path1.max()
    d0 = 2.0 - 1.5 * path0 / pmax
    d1 = 2.0 - 1.5 * path1 / pmax
    d0[0] = 2.0
    d1[0] = 2.0
    prox = np.concatenate([pstart, d0[:, None]], axis=1).astype(np.float32)
    dist = np.concatenate([pend, d1[:, None]], axis=1).astype(np.float32)
    return prox, dist


def random_tree(rng, n):
    """One random binary tree morphology (random_binary_parents, tree_geometry). Returns
    parent, proximal, distal."""
    parent = random_binary_parents(rng, n)
    prox, dist = tree_geometry(rng, parent)
    return parent, prox, dist


def random_trees(n_cells=512, sizes=None, t_final_ms=250.0, seed=2001, hh_everywhere=False,
                 density_scale=False, amp_range=(0.1, 0.5), sim_seed=1, record_all=False, parents=None):
    """Configs 2/4/5: random-tree cells with HH on the soma (or everywhere) and leak on every
    compartment; somatic current step U(amp_range) nA from 10 ms to the end.
    sizes: compartment count per cell (defaults to config 2's lognormal(ln 100, 0.5) clipped
    to [20, 400]); parents: explicit parent arrays (cell-local, parents before children)
    instead of random binary trees."""
    rng = np.random.default_rng(seed)
    if parents is not None:
        sizes = [len(p) for p in parents]
    if sizes is None:
        sizes = np.clip(np.round(rng.lognormal(np.log(100.0), 0.5, n_cells)), 20, 400).astype(np.int64)
    sizes = np.asarray(sizes, np.int64)
    n_cells = len(sizes)
    begin = np.zeros(n_cells + 1, np.int64)
    np.cumsum(sizes, out=begin[1:])
    nc = int(begin[-1])
    area = np.empty(nc, np.float32)
    cap = np.empty(nc, np.float32)
    rax = np.empty(nc, np.float32)
    inv = np.empty(nc, np.float32)
    parent = np.empty(nc, np.int32)
    morphologies = []
    for c in range(n_cells):
        if parents is not None:
            par = np.asarray(parents[c], np.int32)
            prox, dist = tree_geometry(rng, par)
        else:
            par, prox, dist = random_tree(rng, int(sizes[c]))
        morphologies.append((par, prox, dist))
        a, cp, r, iv = N.cable_setup(par, prox, dist, CM, RA)
        sl = slice(begin[c], begin[c + 1])
        area[sl], cap[sl], rax[sl], inv[sl], parent[sl] = a, cp, r, iv, par
    soma = np.zeros(nc, bool)
    soma[begin[:-1]] = True
    hh_mask = np.ones(nc, bool) if hh_everywhere else soma
    g_scale = None
    if density_scale:
        f = rng.uniform(0.5, 1.5, (n_cells, 3)).astype(np.float32)
        g_scale = np.repeat(f, sizes, axis=0)
        g_scale[:, 0] = 1.0                 # scale the HH (Na, K) densities only
    ch_b, ch_t, ch_g, ch_e = _hh_channels(nc, hh_mask, g_scale)
    amp = rng.uniform(amp_range[0], amp_range[1], n_cells).astype(np.float32)
    t_final = t_final_ms * MS
    pulse_begin = np.zeros(nc + 1, np.int64)
    pulse_begin[begin[:-1] + 1] = 1
    pulse_begin = np.cumsum(pulse_begin)
    vthr = np.full(nc, np.nan, np.float32)
    vthr[soma] = VT
    m = Model(
        dt=DT, t_final=t_final, cell_comp_begin=begin,
        v0=np.full(nc, V0), capacitance=cap, area=area, inv_rc_axial=inv, r_axial=rax, v_threshold=vthr,
        parent=parent, channel_types=hh_squid_types(), comp_chan_begin=ch_b, chan_type=ch_t, chan_gbase=ch_g,
        chan_erev=ch_e, cell_rng_seed=cell_rng_seeds(sim_seed, n_cells),
        comp_pulse_begin=pulse_begin, pulse_start=np.full(n_cells, 10.0 * MS),
        pulse_duration=np.full(n_cells, t_final), pulse_amplitude=amp, pulse_weight=np.ones(n_cells),
        record_comp=np.arange(nc) if record_all else begin[:-1].copy(), event_comp=begin[:-1].copy(),
        name=f"random_trees_{n_cells}")
    m.morphologies = morphologies
    return m


def config2_sizes(n_cells=512, seed=2001):
    rng = np.random.default_rng(seed)
    return np.clip(np.round(rng.lognormal(np.log(100.0), 0.5, n_cells)), 20, 400).astype(np.int64)


def config4_sizes(n_cells=4096, seed=4001, lo=10, hi=2000):
    rng = np.random.default_rng(seed)
    return np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), n_cells))).astype(np.int64)


def config2(n_cells=512, t_final_ms=250.0):
    return random_trees(n_cells, sizes=config2_sizes(n_cells), t_final_ms=t_final_ms, seed=2001)


def config4(n_cells=4096, t_final_ms=500.0, hi=2000):
    return random_trees(n_cells, sizes=config4_sizes(n_cells, hi=hi), t_final_ms=t_final_ms, seed=4001,
                        density_scale=True)


def spiking_network(n_cells=100_000, in_degree=100, t_final_ms=1000.0, seed=3001, sim_seed=3,
                    frac_exc=0.8, poisson_hz=5.0, n_record=100, weight_mu=-1.0, weight_sigma=0.5,
                    delay_ms=(1.0, 5.0), diameter=17.8):
    """Config 3: one-compartment HH cells (80% excitatory), fixed in-degree with uniform
    presynaptic partners (no self-connections), expTwo synapses (tau_rise 0.5 ms, tau_decay
    5 ms, E 0 / -80 mV, gbase 1 nS), lognormal weights, delays U(1, 5) ms quantised to dt,
    a 5 Hz poissonFiringSynapse on the excitatory synapse type per cell."""
    rng = np.random.default_rng(seed)
    nc = n_cells
    n_exc = int(round(frac_exc * nc))
    prox = np.array([[0, 0, 0, diameter]], np.float32)
    area, cap, rax, inv = N.cable_setup(np.array([-1]), prox, prox, CM, RA)
    ch_b, ch_t, ch_g, ch_e = _hh_channels(nc, np.ones(nc, bool))
    k = min(in_degree, nc - 1)
    # presynaptic partners without replacement, excluding self: sample from nc-1 and shift
    pre = np.empty((nc, k), np.int64)
    for t in range(nc):
        p = rng.choice(nc - 1, size=k, replace=False)
        p[p >= t] += 1
        pre[t] = p
    w = rng.lognormal(weight_mu, weight_sigma, (nc, k)).astype(np.float32)
    steps = np.round(rng.uniform(delay_ms[0], delay_ms[1], (nc, k)) * MS / DT)
    delay = (steps * DT).astype(np.float32)
    is_inh = pre >= n_exc
    # instance order per target: excitatory type (0) first, then inhibitory (1), each in
    # connection order
    order = np.argsort(is_inh, axis=1, kind="stable")
    pre = np.take_along_axis(pre, order, 1)
    w = np.take_along_axis(w, order, 1)
    delay = np.take_along_axis(delay, order, 1)
    stype = np.take_along_axis(is_inh, order, 1).astype(np.int32)
    syn_begin = np.arange(nc + 1, dtype=np.int64) * k
    # spike recipients by presynaptic cell, in target order
    flat_pre = pre.ravel()
    src_order = np.argsort(flat_pre, kind="stable")
    spk_begin = np.zeros(nc + 1, np.int64)
    np.cumsum(np.bincount(flat_pre, minlength=nc), out=spk_begin[1:])
    syn_types = [N.ExpTwoType(1e-3, 0.0, 0.5 * MS, 5.0 * MS, 1.0), N.ExpTwoType(1e-3, -80.0, 0.5 * MS, 5.0 * MS, 1.0)]
    m = Model(
        dt=DT, t_final=t_final_ms * MS, cell_comp_begin=np.arange(nc + 1, dtype=np.int64),
        v0=np.full(nc, V0), capacitance=np.full(nc, cap[0]), area=np.full(nc, area[0]),
        inv_rc_axial=np.zeros(nc), r_axial=np.full(nc, np.nan), v_threshold=np.full(nc, VT),
        parent=np.full(nc, -1), channel_types=hh_squid_types(), comp_chan_begin=ch_b, chan_type=ch_t,
        chan_gbase=ch_g, chan_erev=ch_e, cell_rng_seed=cell_rng_seeds(sim_seed, nc),
        syn_types=syn_types, comp_syn_begin=syn_begin, syn_type=stype.ravel(), syn_weight=w.ravel(),
        syn_delay=delay.ravel(), comp_spike_begin=spk_begin, spike_target=src_order.astype(np.int64),
        comp_poisson_begin=np.arange(nc + 1, dtype=np.int64),
        poisson_rate=np.full(nc, poisson_hz * 1e-6), poisson_syn_type=np.zeros(nc),
        poisson_weight=np.ones(nc), poisson_instance=np.zeros(nc), poisson_rng_offset=np.zeros(nc),
        record_comp=np.arange(min(n_record, nc)), event_comp=np.arange(nc),
        name=f"spiking_network_{n_cells}")
    m.morphologies = [(np.array([-1], np.int32), prox, prox)] * nc
    return m


def config5(n_cells=256, n_comps=100, t_final_ms=500.0, seed=2001, init_seed=5001):
    """Config 5 (adjoint fitting): n_cells random trees of n_comps compartments built as in
    config 2, HH on every compartment, somatic current steps. Returns (model, gbase_init):
    the model carries the ground-truth densities (per compartment gNa, gK, gLeak in its
    channel distributions); gbase_init = ground truth x U(0.5, 1.5), independently per
    channel distribution (numpy default_rng(init_seed))."""
    m = random_trees(n_cells, sizes=[n_comps] * n_cells, t_final_ms=t_final_ms, seed=seed, hh_everywhere=True)
    rng = np.random.default_rng(init_seed)
    g0 = (np.asarray(m.chan_gbase, np.float64) * rng.uniform(0.5, 1.5, len(m.chan_gbase)))
    m.name = f"config5_{n_cells}x{n_comps}"
    return m, g0


def mixed_network(tree_sizes=(40, 90, 150), n_point=8, n_conn=60, t_final_ms=60.0, seed=11, sim_seed=5,
                  delay_ms=(1.0, 4.0), weight=(2.0, 6.0), hh_everywhere=False, amp_range=(0.3, 0.6)):
    """Tree cells (as random_trees, somatic current steps) and one-compartment HH cells in one
    network: n_conn connections from a random cell's soma to a random compartment (dendrites
    included) of another cell, expTwo synapses (excitatory and inhibitory types, each target
    compartment's instances sorted by type, then in connection order), delays U(delay_ms)
    quantised to dt, weights U(weight). Spikes of tree somata and of point cells both reach
    synapses on both kinds of cell."""
    rng = np.random.default_rng(seed)
    trees = random_trees(len(tree_sizes), sizes=list(tree_sizes), t_final_ms=t_final_ms, seed=seed,
                         hh_everywhere=hh_everywhere, amp_range=amp_range)
    pts = hh_population(n_cells=n_point, t_final_ms=t_final_ms, seed=seed + 1, amp_max_nA=0.15)
    # concatenate the two models' compartment tables
    nt, npt = trees.n_comps, pts.n_comps
    nc = nt + npt
    cat = lambda a, b: np.concatenate([np.asarray(a), np.asarray(b)])
    cell_begin = cat(trees.cell_comp_begin, pts.cell_comp_begin[1:] + nt)
    somata = cell_begin[:-1].copy()
    ch_b = cat(trees.comp_chan_begin, pts.comp_chan_begin[1:] + trees.comp_chan_begin[-1])
    pu_b = cat(trees.comp_pulse_begin, pts.comp_pulse_begin[1:] + trees.comp_pulse_begin[-1])
    # connections: (source soma, target compartment, type, weight, delay)
    src_cell = rng.integers(0, len(somata), n_conn)
    tgt_comp = rng.integers(0, nc, n_conn)
    tgt_cell = np.searchsorted(cell_begin, tgt_comp, side="right") - 1
    keep = src_cell != tgt_cell
    src_cell, tgt_comp = src_cell[keep], tgt_comp[keep]
    n = len(src_cell)
    stype = (rng.random(n) < 0.25).astype(np.int32)
    w = rng.uniform(weight[0], weight[1], n).astype(np.float32)
    delay = (np.round(rng.uniform(delay_ms[0], delay_ms[1], n) * MS / DT) * DT).astype(np.float32)
    # synapse instances per target compartment: by type, then connection order
    order = np.lexsort((np.arange(n), stype, tgt_comp))
    syn_begin = np.zeros(nc + 1, np.int64)
    np.cumsum(np.bincount(tgt_comp, minlength=nc), out=syn_begin[1:])
    inst_of_conn = np.empty(n, np.int64)
    inst_of_conn[order] = np.arange(n)
    # spike recipients per source compartment, in connection order
    src_comp = somata[src_cell]
    sorder = np.argsort(src_comp, kind="stable")
    spk_begin = np.zeros(nc + 1, np.int64)
    np.cumsum(np.bincount(src_comp, minlength=nc), out=spk_begin[1:])
    syn_types = [N.ExpTwoType(2e-3, 0.0, 0.5 * MS, 5.0 * MS, 1.0), N.ExpTwoType(4e-3, -80.0, 0.5 * MS, 5.0 * MS, 1.0)]
    vthr = cat(trees.v_threshold, pts.v_threshold)
    m = Model(
        dt=DT, t_final=t_final_ms * MS, cell_comp_begin=cell_begin,
        v0=cat(trees.v0, pts.v0), capacitance=cat(trees.capacitance, pts.capacitance),
        area=cat(trees.area, pts.area), inv_rc_axial=cat(trees.inv_rc_axial, pts.inv_rc_axial),
        r_axial=cat(trees.r_axial, pts.r_axial), v_threshold=vthr, parent=cat(trees.parent, pts.parent),
        channel_types=hh_squid_types(), comp_chan_begin=ch_b, chan_type=cat(trees.chan_type, pts.chan_type),
        chan_gbase=cat(trees.chan_gbase, pts.chan_gbase), chan_erev=cat(trees.chan_erev, pts.chan_erev),
        cell_rng_seed=cell_rng_seeds(sim_seed, len(somata)),
        comp_pulse_begin=pu_b, pulse_start=cat(trees.pulse_start, pts.pulse_start),
        pulse_duration=cat(trees.pulse_duration, pts.pulse_duration),
        pulse_amplitude=cat(trees.pulse_amplitude, pts.pulse_amplitude),
        pulse_weight=cat(trees.pulse_weight, pts.pulse_weight),
        syn_types=syn_types, comp_syn_begin=syn_begin, syn_type=stype[order], syn_weight=w[order],
        syn_delay=delay[order], comp_spike_begin=spk_begin, spike_target=inst_of_conn[sorder],
        record_comp=np.concatenate([somata, tgt_comp[:8]]), event_comp=somata,
        name=f"mixed_network_{len(tree_sizes)}t{n_point}p")
    m.morphologies = trees.morphologies + pts.morphologies
    return m
