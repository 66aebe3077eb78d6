"""GPU discrete adjoint (include/edenhip_adjoint.h) against finite differences.

* the adjoint's FP64 forward against the oracle's FP64 flavour (same model and operations,
  except the double exp/log implementations and the Hines solve's multiplications by
  reciprocal diagonals where the oracle divides): 1e-8 mV;
* dLoss/dGbase and dLoss/dAmplitude against central finite differences in FP64 of the
  GPU's own FP64 forward (h = 1e-6 relative): 1e-4 relative;
* the same gradient against central finite differences of the oracle's FP64 CPU path;
* the checkpoint interval does not change the gradient (recomputation is deterministic; only
  the order in which per-segment gradient sums are added differs, so equal to rounding).

Models: config 5's construction (random trees, HH on every compartment, somatic current
steps) at 3 cells x 24 compartments and 40 ms, so the oracle runs in well under a second, in
each workgroup size the adjoint kernels have (64, 128 and 256 threads per cell); and config 5
itself (256 cells x 100 compartments, 20,001 rows, checkpoints every 500 steps), whose cells
take the 256-thread kernels by the size rule.
"""
import sys

import numpy as np
import pytest
import torch

from eden_amd import configs
from eden_amd.adjoint import Adjoint, trace
from oracle import oracle

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _model(t_final_ms=40.0):
    m, g0 = configs.config5(n_cells=3, n_comps=24, t_final_ms=t_final_ms)
    # float-representable starting parameters, so the oracle (fp32 inputs) sees the same values
    return m, np.asarray(g0, np.float32).astype(np.float64)


def _loss(tr, target):
    return ((tr - target) ** 2).sum() / tr.numel()


def test_adjoint_forward_matches_oracle_f64():
    m, _ = _model()
    _say(f"[adjoint] forward vs oracle f64: cells={m.n_cells} comps={m.n_comps} rows={m.total_rows()}")
    with Adjoint(m, checkpoint_steps=100) as adj:
        tr = adj.forward().cpu().numpy()
    ref = oracle.run_f64_trace(m)
    assert tr.shape == ref.shape
    err = float(np.max(np.abs(tr - ref)))
    _say(f"[adjoint] max |V_gpu64 - V_oracle64| = {err:.3g} mV, V range {ref.min():.1f}..{ref.max():.1f}")
    assert ref.max() > 0.0          # the cells spike
    assert err < 1e-8


def _gradients(adj, g0, target):
    g = torch.tensor(g0, dtype=torch.float64, device=DEV, requires_grad=True)
    _, a0 = adj.default_parameters()
    a = a0.clone().requires_grad_(True)
    L = _loss(trace(adj, g, a), target)
    L.backward()
    return g.grad.detach().cpu().numpy(), a.grad.detach().cpu().numpy(), a0


@pytest.mark.parametrize("tpb", [0, 128, 256])
def test_adjoint_gradient_matches_central_fd(tpb):
    m, g0 = _model()
    _say(f"[adjoint] central fd, threads_per_cell {tpb}")
    with Adjoint(m, checkpoint_steps=37, threads_per_cell=tpb) as adj:
        gt, a0 = adj.default_parameters()
        target = adj.forward(gt, a0)
        gg, ga, a0 = _gradients(adj, g0, target)
        _say(f"[adjoint] forward {adj.timing()[0]:.2f} ms, backward {adj.timing()[1]:.2f} ms")

        def L(g, a):
            return float(_loss(adj.forward(torch.tensor(g, dtype=torch.float64, device=DEV), a), target))

        rng = np.random.default_rng(0)
        big = np.argsort(-np.abs(gg))[:6]
        pick = np.unique(np.concatenate([big, rng.choice(len(g0), 6, replace=False)]))
        scale = float(np.abs(gg).max())
        for i in pick:
            h = 1e-6 * abs(g0[i])
            gp, gm = g0.copy(), g0.copy()
            gp[i] += h
            gm[i] -= h
            fd = (L(gp, a0) - L(gm, a0)) / (gp[i] - gm[i])
            _say(f"[adjoint] gbase[{i}] = {g0[i]:.4g}: adjoint {gg[i]:.9g}  fd {fd:.9g}")
            assert abs(gg[i] - fd) <= 1e-4 * abs(fd) + 1e-9 * scale
        for p in range(len(ga)):
            ap, am = a0.clone(), a0.clone()
            h = 1e-6 * abs(float(a0[p]))
            ap[p] += h
            am[p] -= h
            fd = (L(g0, ap) - L(g0, am)) / float(ap[p] - am[p])
            _say(f"[adjoint] amplitude[{p}] = {float(a0[p]):.4g}: adjoint {ga[p]:.9g}  fd {fd:.9g}")
            assert abs(ga[p] - fd) <= 1e-4 * abs(fd)


def test_adjoint_gradient_matches_fd_of_the_oracle():
    """Central differences of the oracle's FP64 CPU path (fp32 parameter inputs, so the step
    is the exact float difference)."""
    m, g0 = _model()
    target = oracle.run_f64_trace(m)
    with Adjoint(m) as adj:
        gg, _, _ = _gradients(adj, g0, torch.tensor(target, device=DEV))
    base = m.chan_gbase.copy()

    def L(gf32):
        m.chan_gbase = gf32
        try:
            tr = oracle.run_f64_trace(m)
        finally:
            m.chan_gbase = base
        return float(((tr - target) ** 2).sum() / tr.size)

    for i in np.argsort(-np.abs(gg))[:3]:
        gp = g0.astype(np.float32).copy()
        gm = gp.copy()
        gp[i] = np.float32(g0[i] * (1 + 1e-4))
        gm[i] = np.float32(g0[i] * (1 - 1e-4))
        fd = (L(gp) - L(gm)) / (float(gp[i]) - float(gm[i]))
        _say(f"[adjoint] oracle fd gbase[{i}]: adjoint {gg[i]:.9g}  fd {fd:.9g}")
        assert abs(gg[i] - fd) <= 1e-3 * abs(fd)


def test_checkpoint_interval_does_not_change_the_gradient():
    """Checkpoint intervals 7 and 500, each with the state history kept in device memory and
    with checkpoints and recomputation: recomputation repeats the forward pass's operations,
    so for one interval both forms give the same bits; across intervals only the order in
    which per-segment sums are added differs (equal to rounding)."""
    m, g0 = _model(t_final_ms=20.0)
    res = {}
    for K in (7, 500):
        for store in (1, -1):
            with Adjoint(m, checkpoint_steps=K, store_states=store) as adj:
                gt, a0 = adj.default_parameters()
                target = adj.forward(gt, a0)
                res[(K, store)] = _gradients(adj, g0, target)[:2]
    for K in (7, 500):
        assert np.array_equal(res[(K, 1)][0], res[(K, -1)][0]) and np.array_equal(res[(K, 1)][1], res[(K, -1)][1])
    a, b = res[(7, 1)], res[(500, 1)]
    np.testing.assert_allclose(a[0], b[0], rtol=1e-11, atol=1e-14 * np.abs(b[0]).max())
    np.testing.assert_allclose(a[1], b[1], rtol=1e-11, atol=1e-14 * np.abs(b[1]).max())


def test_config5_full_size_forward_and_gradient():
    """Config 5 at full size: the FP64 forward equals the oracle's FP64 flavour on 16 of its
    cells (the cells are independent), and the gradient agrees with central differences
    (h = 1e-6 relative) of the GPU's FP64 forward on 8 channel densities and 2 amplitudes."""
    m, g0 = configs.config5()
    g0 = np.asarray(g0, np.float32).astype(np.float64)
    _say(f"[adjoint] config 5: {m.n_cells} cells x {m.n_comps // m.n_cells} comps, {m.total_rows()} rows")
    with Adjoint(m, checkpoint_steps=500) as adj:
        gt, a0 = adj.default_parameters()
        target = adj.forward(gt, a0)
        tr = target.cpu().numpy()
        cells = np.arange(0, m.n_cells, m.n_cells // 16)
        sub, rc, _ = m.subset(cells)
        ref = oracle.run_f64_trace(sub)
        err = float(np.abs(tr[:, rc] - ref).max())
        _say(f"[adjoint] config 5 forward, 16 cells: max |V_gpu64 - V_oracle64| = {err:.3g} mV")
        assert err < 1e-8
        gg, ga, a0 = _gradients(adj, g0, target)
        fwd_ms, bwd_ms = adj.timing()
        _say(f"[adjoint] config 5 forward {fwd_ms:.1f} ms, backward {bwd_ms:.1f} ms")

        def L(g, a):
            return float(_loss(adj.forward(torch.tensor(g, dtype=torch.float64, device=DEV), a), target))

        rng = np.random.default_rng(1)
        big = np.argsort(-np.abs(gg))[:4]
        pick = np.unique(np.concatenate([big, rng.choice(len(g0), 4, replace=False)]))
        scale = float(np.abs(gg).max())
        for i in pick:
            gp, gm = g0.copy(), g0.copy()
            gp[i] += 1e-6 * abs(g0[i])
            gm[i] -= 1e-6 * abs(g0[i])
            fd = (L(gp, a0) - L(gm, a0)) / (gp[i] - gm[i])
            _say(f"[adjoint] config 5 gbase[{i}]: adjoint {gg[i]:.9g}  fd {fd:.9g}")
            assert abs(gg[i] - fd) <= 1e-4 * abs(fd) + 1e-9 * scale
        for p in np.argsort(-np.abs(ga))[:2]:
            ap, am = a0.clone(), a0.clone()
            ap[p] += 1e-6 * abs(float(a0[p]))
            am[p] -= 1e-6 * abs(float(a0[p]))
            fd = (L(g0, ap) - L(g0, am)) / float(ap[p] - am[p])
            _say(f"[adjoint] config 5 amplitude[{p}]: adjoint {ga[p]:.9g}  fd {fd:.9g}")
            assert abs(ga[p] - fd) <= 1e-4 * abs(fd)


# --- synaptic weights (expTwo synapses, fixed event schedule) ---------------------------------

def _net_model():
    """A small network of tree and one-compartment HH cells joined by excitatory and
    inhibitory expTwo synapses (40 ms; the cells spike and the synapses receive events)."""
    m = configs.mixed_network(tree_sizes=(20, 30), n_point=4, n_conn=30, t_final_ms=40.0, seed=3)
    # the adjoint records a compartment once: keep the first of repeated recording columns
    _, first = np.unique(m.record_comp, return_index=True)
    m.record_comp = np.asarray(m.record_comp)[np.sort(first)]
    return m


def _schedule(m, weight=None):
    """The oracle's FP64 run: its synaptic event schedule (step, synapse) and its FP64 trace."""
    base = m.syn_weight
    if weight is not None:
        m.syn_weight = weight
    try:
        _, (st, sy) = oracle.run_with_synapse_log(m, precision="f64")
        tr = oracle.run_f64_trace(m)
    finally:
        m.syn_weight = base
    return np.asarray(st), np.asarray(sy), tr


def test_adjoint_synapses_forward_matches_oracle_f64():
    """With the oracle's own event schedule, the adjoint's FP64 forward of a network
    reproduces the oracle's FP64 network run, every cell's trace."""
    m = _net_model()
    st, sy, ref = _schedule(m)
    _say(f"[adjoint] network forward: cells={m.n_cells} comps={m.n_comps} synapses={len(m.syn_weight)} "
         f"events={len(st)} rows={m.total_rows()}")
    assert len(st) > 10 and len(np.unique(sy)) > 5
    with Adjoint(m, checkpoint_steps=100) as adj:
        assert adj.n_syn == len(m.syn_weight)
        adj.set_events(st, sy)
        tr = adj.forward().cpu().numpy()
        # without the schedule the synapses stay silent and the trace differs
        adj.set_events([], [])
        quiet = adj.forward().cpu().numpy()
    err = float(np.max(np.abs(tr - ref)))
    _say(f"[adjoint] network: max |V_gpu64 - V_oracle64| = {err:.3g} mV; without events {np.abs(quiet - ref).max():.3g} mV")
    assert err < 1e-8
    assert np.abs(quiet - ref).max() > 1.0


@pytest.mark.parametrize("tpb", [0, 128])
def test_adjoint_weight_gradient_matches_central_fd(tpb):
    """dLoss/dWeight (and dLoss/dGbase with synapses present) against central differences of
    the GPU's FP64 forward with the same schedule, h = 1e-6 relative: 1e-4 relative."""
    m = _net_model()
    st, sy, _ = _schedule(m)
    w0 = np.asarray(m.syn_weight, np.float64)
    with Adjoint(m, checkpoint_steps=37, threads_per_cell=tpb) as adj:
        adj.set_events(st, sy)
        g0, a0 = adj.default_parameters()
        wt = torch.tensor(w0 * np.linspace(0.6, 1.4, len(w0)), dtype=torch.float64, device=DEV)
        target = adj.forward(g0, a0, wt)
        w = torch.tensor(w0, dtype=torch.float64, device=DEV, requires_grad=True)
        g = g0.clone().requires_grad_(True)
        L = _loss(trace(adj, g, a0, w), target)
        L.backward()
        gw = w.grad.detach().cpu().numpy()
        gg = g.grad.detach().cpu().numpy()

        def Lw(wv, gv=g0):
            return float(_loss(adj.forward(gv, a0, torch.tensor(wv, dtype=torch.float64, device=DEV)), target))

        active = np.unique(sy)
        assert np.all(gw[np.setdiff1d(np.arange(len(w0)), active)] == 0.0)   # no event, no gradient
        rng = np.random.default_rng(2)
        big = np.argsort(-np.abs(gw))[:5]
        pick = np.unique(np.concatenate([big, rng.choice(active, 5, replace=False)]))
        scale = float(np.abs(gw).max())
        assert scale > 0.0
        for i in pick:
            wp, wm = w0.copy(), w0.copy()
            wp[i] += 1e-6 * abs(w0[i])
            wm[i] -= 1e-6 * abs(w0[i])
            fd = (Lw(wp) - Lw(wm)) / (wp[i] - wm[i])
            _say(f"[adjoint] weight[{i}] = {w0[i]:.4g}: adjoint {gw[i]:.9g}  fd {fd:.9g}")
            assert abs(gw[i] - fd) <= 1e-4 * abs(fd) + 1e-9 * scale
        gn = g0.cpu().numpy()
        for i in np.argsort(-np.abs(gg))[:3]:
            gp, gm = gn.copy(), gn.copy()
            gp[i] += 1e-6 * abs(gn[i])
            gm[i] -= 1e-6 * abs(gn[i])
            fd = (Lw(w0, torch.tensor(gp, device=DEV)) - Lw(w0, torch.tensor(gm, device=DEV))) / (gp[i] - gm[i])
            _say(f"[adjoint] gbase[{i}] with synapses: adjoint {gg[i]:.9g}  fd {fd:.9g}")
            assert abs(gg[i] - fd) <= 1e-4 * abs(fd)


def test_adjoint_weight_gradient_matches_fd_of_the_oracle_network():
    """Central differences of the oracle's FP64 network run (fp32 weights, steps of 1e-4
    relative) where the perturbed runs keep the unperturbed event schedule: there the fixed
    schedule is the network's own, and the adjoint gradient is the network's gradient."""
    m = _net_model()
    st, sy, ref = _schedule(m)
    w0 = np.asarray(m.syn_weight, np.float32)
    target = ref * 0.9
    with Adjoint(m) as adj:
        adj.set_events(st, sy)
        g0, a0 = adj.default_parameters()
        w = torch.tensor(w0.astype(np.float64), dtype=torch.float64, device=DEV, requires_grad=True)
        _loss(trace(adj, g0, a0, w), torch.tensor(target, device=DEV)).backward()
        gw = w.grad.detach().cpu().numpy()
    checked = 0
    for i in np.argsort(-np.abs(gw))[:4]:
        wp, wm = w0.copy(), w0.copy()
        wp[i] = np.float32(w0[i] * (1 + 1e-4))
        wm[i] = np.float32(w0[i] * (1 - 1e-4))
        sp, yp, tp = _schedule(m, wp)
        sm, ym, tm = _schedule(m, wm)
        same = all(np.array_equal(a, b) for a, b in ((sp, st), (yp, sy), (sm, st), (ym, sy)))
        fd = (float(((tp - target) ** 2).sum() / tp.size) - float(((tm - target) ** 2).sum() / tm.size)) / (
            float(wp[i]) - float(wm[i]))
        _say(f"[adjoint] oracle fd weight[{i}]: adjoint {gw[i]:.9g}  fd {fd:.9g}  schedule kept: {same}")
        if same:
            assert abs(gw[i] - fd) <= 1e-3 * abs(fd)
            checked += 1
    assert checked >= 2
