"""GPU parity: the HIP path (through the C-ABI) against the oracle on the same seeded inputs.

Float tolerance: rasters equal up to one dt, voltage max-abs <= 1e-3 mV wherever the rasters
agree (tests/parity.py). Sizes are scaled down so the oracle finishes in seconds.
"""
import sys

import numpy as np
import pytest

from eden_amd import configs
from eden_amd.model import EdenHipError, Simulation
from oracle import oracle
from parity import TOL_MV, compare  # noqa: E402

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _gpu_run(model, **kw):
    with Simulation(model, **kw) as sim:
        tr = sim.run_all()
        ev = sim.read_events()
    return tr, ev


def _check(model, network=False, **kw):
    _say(f"[parity] {model.name} cells={model.n_cells} comps={model.n_comps} opts={kw}")
    tr_g, ev_g = _gpu_run(model, **kw)
    tr_r, ec, er, _ = oracle.run(model)
    assert tr_g.shape == tr_r.shape
    rec2ev = None
    if not network and len(model.event_comp):
        ev_of_comp = {int(c): i for i, c in enumerate(model.event_comp)}
        rec2ev = [ev_of_comp.get(int(c), 0) for c in model.record_comp]
    res = compare(tr_g, ev_g, tr_r, (ec, er), len(model.event_comp), rec2ev, network=network)
    _say(f"[parity] {model.name}: {res}")
    assert res["max_abs_mV"] <= TOL_MV, res
    # stronger than the tolerance: the HIP path performs the oracle's float operations in the
    # same order (exp/log rounded once from double on both sides), so traces are identical
    assert res["max_abs_mV"] == 0.0, res
    return res, (tr_g, ev_g), (tr_r, (ec, er))


def test_hh_population_small():
    res, g, r = _check(configs.hh_population(n_cells=200, t_final_ms=100.0))
    assert res["spikes_ref"] > 0
    assert res["raster_mismatch_cols"] == 0, res


def test_point_cells_with_four_gate_variables():
    """Uncoupled one-compartment cells with four gate variables (leak, Na, K and a second K
    distribution at half the density): every lane of a cell's quad advances a gate."""
    import dataclasses
    m = configs.hh_population(n_cells=300, t_final_ms=40.0, seed=5)
    nc = m.n_cells
    g3, e3 = m.chan_gbase[:3], m.chan_erev[:3]
    m4 = dataclasses.replace(
        m, comp_chan_begin=np.arange(nc + 1, dtype=np.int64) * 4,
        chan_type=np.tile(np.array([0, 1, 2, 2], np.int32), nc),
        chan_gbase=np.tile(np.array([g3[0], g3[1], g3[2], g3[2] * np.float32(0.5)], np.float32), nc),
        chan_erev=np.tile(np.array([e3[0], e3[1], e3[2], e3[2]], np.float32), nc), name="hh_population_4gates")
    with Simulation(m4, profile=True) as sim:
        sim.run_all()
        names = [k["name"] for k in sim.kernel_infos()]
    assert names == ["point_quad_kernel"], names
    res, g, r = _check(m4)
    assert res["spikes_ref"] > 0
    assert res["raster_mismatch_cols"] == 0, res


def test_random_trees_small():
    sizes = [20, 35, 64, 65, 100, 128, 129, 200, 256, 257, 400, 31]
    res, g, r = _check(configs.random_trees(len(sizes), sizes=sizes, t_final_ms=60.0, seed=7))
    assert res["raster_mismatch_cols"] == 0, res


def test_random_trees_all_compartments_recorded():
    sizes = [23, 77, 150]
    m = configs.random_trees(len(sizes), sizes=sizes, t_final_ms=30.0, seed=11, record_all=True)
    res, g, r = _check(m)
    assert res["raster_mismatch_cols"] == 0, res


def test_ragged_cells_one_to_max_compartments():
    """One-, two- and three-compartment cells beside a tree at the size limit in one model: the
    one-compartment cells take the point path, the rest the tree buckets, in the same run."""
    sizes = [1, 2, 3, 1, 2047, 5, 64, 2]
    m = configs.random_trees(len(sizes), sizes=sizes, t_final_ms=60.0, seed=17)
    with Simulation(m, profile=True) as sim:
        sim.run_all()
        names = {k["name"] for k in sim.kernel_infos()}
    assert any(n.startswith("point") for n in names), names
    assert any(n.startswith("tree_kernel<512") for n in names), names
    res, g, r = _check(m)
    # the one- to three-compartment cells spike (25 spikes in 60 ms)
    assert res["spikes_ref"] > 0
    assert res["raster_mismatch_cols"] == 0, res


def test_run_shorter_than_one_step():
    """t_final below dt: one row, the initial state, and no step taken."""
    m = configs.random_trees(2, sizes=[1, 40], t_final_ms=0.01, seed=3)
    assert m.total_rows() == 1
    res, g, r = _check(m)
    assert g[0].shape == (1, len(m.record_comp))


def test_large_trees_workgroup_per_cell():
    sizes = [513, 1000, 1500, 2000, 700]
    res, g, r = _check(configs.random_trees(len(sizes), sizes=sizes, t_final_ms=25.0, seed=13))
    assert res["raster_mismatch_cols"] == 0, res


def _shapes():
    rng = np.random.default_rng(99)
    n = 2047
    chain = np.arange(-1, n - 1, dtype=np.int32)                  # unbranched cable, depth 2046
    star = np.zeros(n, np.int32)
    star[0] = -1                                                   # root with 2046 children
    broom = np.concatenate([np.arange(-1, 999), np.full(n - 1000, 999)]).astype(np.int32)
    recursive = np.array([-1] + [int(rng.integers(i)) for i in range(1, n)], np.int32)   # any fan-out
    return [chain, star, broom, recursive, np.array([-1, 0], np.int32), np.array([-1, 0, 0, 1, 1, 2], np.int32)]


def test_extreme_tree_shapes_at_the_size_limit():
    """2,047-compartment chain, star, broom and random recursive tree, plus 2- and 6-node
    cells: deepest schedules, widest folds (the root fold over 2,046 children), every
    compartment recorded."""
    m = configs.random_trees(parents=_shapes(), t_final_ms=3.0, seed=77, record_all=True)
    res, g, r = _check(m)
    assert res["raster_mismatch_cols"] == 0, res


def test_hh_everywhere_trees():
    sizes = [100] * 6
    res, g, r = _check(configs.random_trees(len(sizes), sizes=sizes, t_final_ms=40.0, seed=5001,
                                            hh_everywhere=True))
    assert res["raster_mismatch_cols"] == 0, res


def test_hh_everywhere_large_trees():
    """Gated channels on every compartment (no constant subtrees) in 2-, 4- and 8-wave
    workgroups, the last with its separate channel wave."""
    sizes = [300, 700, 1100]
    res, g, r = _check(configs.random_trees(len(sizes), sizes=sizes, t_final_ms=15.0, seed=5003,
                                            hh_everywhere=True, record_all=True))
    assert res["raster_mismatch_cols"] == 0, res


def test_many_small_trees_without_channel_wave():
    """More one-wave cells than 4 per CU: the 64-thread bucket then updates gates on its cable
    wave (no separate channel wave), the form that runs the small cells of config 4."""
    rng = np.random.default_rng(29)
    sizes = [int(x) for x in rng.integers(10, 60, 1100)]
    m = configs.random_trees(len(sizes), sizes=sizes, t_final_ms=5.0, seed=29)
    with Simulation(m, profile=True) as sim:
        sim.run_all()
        infos = sim.kernel_infos()
    _say(f"[parity] kernels {[(k['name'], k['threads_per_block'], k['grid'], k['lds_bytes']) for k in infos]}")
    assert [k["name"] for k in infos] == ["tree_kernel<64, false, false>"]
    assert infos[0]["threads_per_block"] == 64 and infos[0]["lds_bytes"] > 0
    res, g, r = _check(m)
    assert res["raster_mismatch_cols"] == 0, res


def test_size_limits_agree():
    """Cells moved to larger workgroups by size limits (other schedules, same operations)."""
    sizes = [60, 150, 300, 500, 900, 1300]
    m = configs.random_trees(len(sizes), sizes=sizes, t_final_ms=15.0, seed=23)
    res, g, r = _check(m, size_limits=(128, 384, 768, 2047))
    assert res["raster_mismatch_cols"] == 0, res


@pytest.mark.parametrize("wave", [1, -1])
@pytest.mark.parametrize("hh", [False, True])
def test_channel_wave_settings_agree(wave, hh):
    """Every workgroup size with the separate channel wave forced on, then off (the kernels'
    other variants; same operations), HH on the soma only and on every compartment."""
    sizes = [40, 200, 400, 800, 1500]
    m = configs.random_trees(len(sizes), sizes=sizes, t_final_ms=15.0, seed=37, hh_everywhere=hh)
    with Simulation(m, profile=True, channel_wave=(wave,) * 4) as sim:
        names = sorted(k["name"] for k in sim.kernel_infos())
    _say(f"[parity] channel_wave {wave}: {names}")
    flag = "true" if wave > 0 else "false"
    assert names and all(n.split(", ")[1] == flag for n in names)
    res, g, r = _check(m, channel_wave=(wave,) * 4)
    assert res["raster_mismatch_cols"] == 0, res


def test_size_limits_rejected_when_invalid():
    m = configs.random_trees(2, sizes=[30, 40], t_final_ms=1.0, seed=3)
    with pytest.raises(Exception, match="size_limits"):
        Simulation(m, size_limits=(300, 0, 0, 0))
    with pytest.raises(Exception, match="channel_wave"):
        Simulation(m, channel_wave=(2, 0, 0, 0))


@pytest.mark.parametrize("tpb", [64, 128, 256, 512])
def test_launch_shapes_agree(tpb):
    sizes = [60, 120, 250]
    m = configs.random_trees(len(sizes), sizes=sizes, t_final_ms=20.0, seed=17)
    _check(m, threads_per_cell=tpb)


def test_graph_replays_match_eager():
    """One launch per step captured in hipGraphs (two graphs for 405 steps), replayed three
    times with eager runs between the replays: every replay equals the persistent run."""
    sizes = [40, 90, 300]
    m = configs.random_trees(len(sizes), sizes=sizes, t_final_ms=10.0, seed=41)
    with Simulation(m) as sim:
        ref, ev_ref = sim.run_all(), sim.read_events()
    with Simulation(m, persistent=-1, use_graph=True) as g, Simulation(m, persistent=-1) as e:
        for rep in range(3):
            _say(f"[graph] replay {rep}")
            g.reset()
            tr = g.run_all()
            np.testing.assert_array_equal(tr, ref)
            for a, b in zip(g.read_events(), ev_ref):
                np.testing.assert_array_equal(a, b)
            e.reset()
            np.testing.assert_array_equal(e.run_all(), ref)


def test_chunked_run_matches_single_run():
    m = configs.random_trees(4, sizes=[40, 80, 120, 160], t_final_ms=20.0, seed=19)
    with Simulation(m) as s:
        full = s.run_all()
    with Simulation(m, chunk_steps=97) as s:
        parts = s.run_all()
    assert np.array_equal(full, parts)


def test_spiking_network_small():
    m = configs.spiking_network(n_cells=300, in_degree=20, t_final_ms=60.0, poisson_hz=200.0,
                                weight_mu=2.0, n_record=50)
    res, g, r = _check(m, network=True)
    assert res["spikes_ref"] > 0


@pytest.mark.parametrize("window", [0, -1, -2])
def test_synaptic_events_bit_exact(window):
    """Integer event path: every (step, synapse instance) at which a synapse receives its
    event, GPU against the oracle, exactly; for the windowed ring-buffer path (0), the
    one-launch-per-step path (-1) and windows all run again step by step (-2)."""
    m = configs.spiking_network(n_cells=400, in_degree=30, t_final_ms=80.0, poisson_hz=150.0,
                                weight_mu=2.0, n_record=10)
    _say(f"[parity] synaptic events {m.name} synapses={m.n_synapses} window mode {window}")
    with Simulation(m, log_synapse_events=5_000_000, network_window=window) as sim:
        sim.run_all()
        g_step, g_syn = sim.read_synapse_events()
        ev_g = sim.read_events()
        w, nwin, rerun, deliv = sim.network_stats()
    _say(f"[parity] window {w} steps, {nwin} windows, {rerun} run again, {deliv} events counted")
    if window == 0:
        assert w > 1
    if window == -2:
        assert rerun == nwin
    if window != -1:
        assert deliv == len(g_step)
    (_, ec, er, _), (o_step, o_syn) = oracle.run_with_synapse_log(m)
    _say(f"[parity] synaptic events: gpu {len(g_step)} oracle {len(o_step)} spikes {len(er)}")
    assert len(o_step) > 100
    assert np.array_equal(ev_g[0], ec) and np.array_equal(ev_g[1], er)
    assert np.array_equal(g_step, o_step)
    assert np.array_equal(g_syn, o_syn)


@pytest.mark.parametrize("hh_everywhere", [False, True])
def test_network_with_tree_cells(hh_everywhere):
    """Tree cells and one-compartment cells in one network: synapses on dendrites and somata
    of tree cells, spikes sent by tree somata and by point cells. Traces, spike raster and
    every synaptic delivery equal the oracle's, bit for bit."""
    m = configs.mixed_network(tree_sizes=(40, 90, 150, 300, 700), n_point=24, n_conn=400, t_final_ms=80.0,
                              seed=13 + int(hh_everywhere), hh_everywhere=hh_everywhere)
    _say(f"[parity] {m.name} cells={m.n_cells} comps={m.n_comps} synapses={m.n_synapses}")
    with Simulation(m, log_synapse_events=1_000_000, profile=True) as sim:
        tr_g = sim.run_all()
        ev_g = sim.read_events()
        g_step, g_syn = sim.read_synapse_events()
        names = [k["name"] for k in sim.kernel_infos()]
    _say(f"[parity] kernels {names}")
    assert all(n == "point_kernel" or n.endswith(", false, true>") for n in names), names
    (tr_r, ec, er, _), (o_step, o_syn) = oracle.run_with_synapse_log(m)
    _say(f"[parity] spikes gpu {len(ev_g[1])} oracle {len(er)}, synaptic events gpu {len(g_step)} oracle {len(o_step)}")
    assert np.array_equal(ev_g[0], ec) and np.array_equal(ev_g[1], er)
    assert np.array_equal(tr_g, tr_r)
    assert np.array_equal(g_step, o_step) and np.array_equal(g_syn, o_syn)
    # the case is exercised: tree somata spiked, and synapses on tree dendrites received events
    n_tree = 5
    assert np.any(ec < n_tree), "no tree soma spiked"
    comp_of_syn = np.repeat(np.arange(m.n_comps), np.diff(m.comp_syn_begin))
    tgt = comp_of_syn[o_syn]
    dendrite = np.isin(tgt, np.setdiff1d(np.arange(m.cell_comp_begin[n_tree]), m.cell_comp_begin[:n_tree]))
    assert dendrite.sum() > 10, dendrite.sum()


def test_tree_synapses_without_recipients():
    """Synapses on tree compartments and no spike recipients anywhere: the NET tree kernels
    in persistent launches; identical to the oracle."""
    import dataclasses
    m = configs.mixed_network(tree_sizes=(60, 120), n_point=0, n_conn=30, t_final_ms=20.0, seed=3)
    m = dataclasses.replace(m, comp_spike_begin=np.zeros(m.n_comps + 1, np.int64), spike_target=np.zeros(0, np.int64),
                            name="tree_synapses_no_recipients")
    _check(m)


@pytest.mark.parametrize("delay_ms", [1.0, 15.0])
def test_delay_rule_on_gpu(delay_ms):
    from test_oracle import _two_cell_delay_model, expected_deliveries
    m = _two_cell_delay_model(delay_ms)
    with Simulation(m, log_synapse_events=100_000) as sim:
        sim.run_all()
        step, syn = sim.read_synapse_events()
        col, row = sim.read_events()
    D = int(round(delay_ms / 0.025))
    want = [k for k in expected_deliveries(list(row), D) if k <= m.total_rows()]
    assert list(step) == want


def test_deviation_from_fp64_dt10_is_the_oracles():
    """Both fp32 paths (GPU and oracle) against the FP64 run at dt/10: the deviations are
    reported and, the paths being bit-identical here, equal."""
    m = configs.hh_population(n_cells=50, t_final_ms=60.0)
    tr_g, ev_g = _gpu_run(m)
    tr_o, ec, er, _ = oracle.run(m)
    tr_64, ec64, er64, _ = oracle.run(m, precision="f64", substeps=10)
    dev_g = float(np.abs(tr_g.astype(np.float64) - tr_64).max())
    dev_o = float(np.abs(tr_o.astype(np.float64) - tr_64).max())
    spikes_64 = len(er64)
    _say(f"[parity] deviation from FP64 dt/10: gpu {dev_g:.4g} mV, oracle fp32 {dev_o:.4g} mV, "
         f"spikes fp32 {len(er)} fp64/10 {spikes_64}")
    assert dev_g == dev_o


@pytest.mark.parametrize("which", ["config1", "config4_sample"])
def test_deviation_from_the_eden_build(which):
    """Against the oracle built as EDEN builds its generated code (gcc -Ofast: reassociation
    and FMA contraction, glibc expf/logf): spike rasters identical, and the GPU's voltage
    deviation from that build is the exact-arithmetic oracle's own (the GPU equals the
    latter bit for bit). Its size is the rounding that -Ofast reorders, amplified on spike
    upstrokes: measured 0.37 mV on config 1's cells and 0.036 mV on config 4's, with glibc
    expf/logf alone 0.024 mV; bounded here at 0.5 and 0.1 mV."""
    if which == "config1":
        m, bound = configs.hh_population(n_cells=200, t_final_ms=100.0), 0.5
    else:
        full = configs.config4()
        sizes = full.comps_per_cell()
        order = np.argsort(sizes, kind="stable")
        m, _, _ = full.subset(order[np.linspace(0, len(order) - 1, 16).astype(np.int64)])
        bound = 0.1
    tr_g, ev_g = _gpu_run(m)
    tr_o, ec, er, _ = oracle.run(m)
    tr_f, ecf, erf, _ = oracle.run(m, precision="f32ofast")
    ev_of = {int(c): i for i, c in enumerate(m.event_comp)}
    rec2ev = [ev_of.get(int(c), 0) for c in m.record_comp]
    res = compare(tr_g, ev_g, tr_f, (ecf, erf), len(m.event_comp), rec2ev)
    ref = compare(tr_o, (ec, er), tr_f, (ecf, erf), len(m.event_comp), rec2ev)
    _say(f"[parity] {m.name} vs the -Ofast build: {res}")
    assert res["raster_mismatch_cols"] == 0 and res["spikes_ref"] > 0
    assert res["max_abs_mV"] == ref["max_abs_mV"]
    assert res["max_abs_mV"] <= bound


@pytest.mark.parametrize("which", ["trees", "network_window", "mixed_network"])
def test_state_checkpoint_and_restore(which):
    """edenhip_get_state after some rows, edenhip_set_state into a fresh context of the same
    model: the rows run after the restore equal, bit for bit, those the first context runs
    on (tree cells; the windowed network path with its delivery ring; a per-step network of
    tree and point cells)."""
    if which == "trees":
        m = configs.random_trees(5, sizes=[30, 80, 200, 600, 1500], t_final_ms=40.0, seed=8)
    elif which == "network_window":
        m = configs.spiking_network(n_cells=2000, in_degree=20, t_final_ms=60.0, poisson_hz=200.0)
    else:
        m = configs.mixed_network(tree_sizes=(40, 300), n_point=12, n_conn=120, t_final_ms=40.0, seed=6)
    first = 701
    with Simulation(m, chunk_steps=first) as a:
        a.run(-1)
        blob = a.get_state()
        rest = []
        while True:
            n = a.run(-1)
            if n == 0:
                break
            rest.append(a.read_trace()[0].copy())
        ev_a = a.read_events()
    with Simulation(m, chunk_steps=first) as b:
        b.set_state(blob)
        rest_b = []
        while True:
            n = b.run(-1)
            if n == 0:
                break
            rest_b.append(b.read_trace()[0].copy())
        ev_b = b.read_events()
    ra, rb = np.concatenate(rest), np.concatenate(rest_b)
    _say(f"[state] {which}: {len(blob)} bytes, {len(ra)} rows after the restore, spikes {len(ev_b[1])}")
    assert np.array_equal(ra, rb)
    # the restored context logs the spikes from the restore on
    keep = ev_a[1] >= first + 1
    assert np.array_equal(ev_a[1][keep], ev_b[1]) and np.array_equal(ev_a[0][keep], ev_b[0])
    with Simulation(configs.random_trees(2, sizes=[30, 40], t_final_ms=5.0)) as other:
        with pytest.raises(EdenHipError, match="does not belong"):
            other.set_state(blob)


def test_output_files_from_a_gpu_run(tmp_path):
    """OutputFile / EventOutputFile written from a chunked GPU run, byte-compared with the
    text the reference's printer gives for the oracle's trace and events (tests/eden_text.py,
    built from FixedWidthNumberPrinter's arithmetic, not from the library's writer)."""
    from eden_amd import output
    from eden_text import assert_same_text, events_text, trajectory_text
    m = configs.hh_population(n_cells=40, t_final_ms=30.0)
    tp, ep = tmp_path / "v.dat", tmp_path / "spikes.dat"
    rows, nev = output.run_to_files(m, tp, ep, "TIME_ID", chunk_rows=333)
    _say(f"[output] rows {rows} events {nev}")
    tr, ec, er, _ = oracle.run(m)
    t_us = output.row_times_us(m.dt, rows + 1)
    assert nev > 0
    assert_same_text(tp.read_text(), trajectory_text(t_us, tr))
    assert_same_text(ep.read_text(), events_text(t_us, er, ec, np.arange(len(m.event_comp)), "TIME_ID"))


def _with_gate_columns(m, comps, name):
    """The model also recording every gate variable of the given compartments, after its own
    V columns (and V of those compartments again, as a column of its own)."""
    import dataclasses
    rc, rg = list(map(int, m.record_comp)), [-1] * len(m.record_comp)
    for g in comps:
        ng = sum(m.channel_types[int(t)].n_gates for t in m.chan_type[m.comp_chan_begin[g]:m.comp_chan_begin[g + 1]])
        rc += [int(g)] * (ng + 1)
        rg += list(range(ng)) + [-1]
    return dataclasses.replace(m, record_comp=np.array(rc, np.int64), record_gate=np.array(rg, np.int32), name=name)


@pytest.mark.parametrize("kind", ["point", "trees", "trees_graph_chunks", "network"])
def test_gate_columns_bit_exact(kind):
    """Gate variables recorded next to V (point cells from T.q, tree cells from the gate
    table, a network of both): every column equals the oracle's exactly, the V columns equal
    those of the same model recording V only (run in its default persistent mode), and the
    gate columns hold values in [0, 1] that move."""
    if kind == "point":
        m = configs.hh_population(n_cells=40, t_final_ms=30.0, seed=5)
        comps, network, kw = [0, 7, 39], False, {}
    elif kind == "network":
        m = configs.mixed_network(t_final_ms=30.0, hh_everywhere=True)
        comps, network, kw = [0, 5, 41, m.n_comps - 1], True, {}
    else:
        sizes = [40, 90, 300]
        m = configs.random_trees(len(sizes), sizes=sizes, t_final_ms=15.0, seed=41, hh_everywhere=True)
        comps, network = [0, 3, 40, 129, 200], False
        kw = {} if kind == "trees" else dict(use_graph=True, chunk_steps=97)
    n_v = len(m.record_comp)
    with Simulation(m) as sim:
        v_only = sim.run_all()
    mg = _with_gate_columns(m, comps, f"{m.name}_gates")
    _say(f"[gates] {kind}: {len(mg.record_comp) - n_v} extra columns, opts={kw}")
    tr_g, ev_g = _gpu_run(mg, **kw)
    tr_r, ec, er, _ = oracle.run(mg)
    assert np.array_equal(tr_g, tr_r)
    assert np.array_equal(tr_g[:, :n_v], v_only)
    gate = mg.record_gate[n_v:] >= 0
    q = tr_g[:, n_v:][:, gate]
    assert q.min() >= 0.0 and q.max() <= 1.0 and float(np.ptp(q, axis=0).min()) > 1e-3
    # the extra V columns repeat the compartments' V
    again = np.flatnonzero(mg.record_gate[n_v:] < 0) + n_v
    assert np.array_equal(tr_g[:, again], tr_r[:, again])
    if network:
        assert len(er) > 0


def test_gate_columns_refused_when_out_of_range():
    m = configs.hh_population(n_cells=4, t_final_ms=1.0, seed=5)
    import dataclasses
    for j in (3, -2):   # the HH compartment has gate variables 0, 1, 2
        bad = dataclasses.replace(m, record_gate=np.array([j] + [-1] * (len(m.record_comp) - 1), np.int32))
        with pytest.raises(EdenHipError, match="record_gate"):
            Simulation(bad)
