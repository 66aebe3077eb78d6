"""GPU discrete adjoint (include/edenhip_adjoint.h) against finite differences.

* the adjoint's FP64 forward against the oracle's FP64 flavour (same model, same order of
  operations; only the double exp/log implementations differ);
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
