"""Benchmark of the per-timestep simulation loop on one MI355X.

Workload (default): the batched uneven-instance case, 4,096 independent random-tree cells
with compartment counts log-uniform over 10-2,000 (1.54M compartments), HH soma with
per-cell density factors U(0.5, 1.5), passive dendrites, somatic current steps, dt 0.025 ms,
20,001 recorded rows (500 ms). A bench step = one full simulation of that workload from the
initial state, inputs already resident in HBM (the library's reset rewinds state on the
device). Other workloads (--workload config1/config2/config3) are the 1,000-cell HH
population, the 512-tree cable case and the 100,000-cell spiking network.

Prints one JSON line: compartment-updates/s (rows x compartments / s), the roofline of the
dominant kernel (its algorithmic bytes per launch over its HIP-event-timed duration), the
oracle's CPU throughput on a bounded sample (cpu_baseline) and a trace-tolerance check.

The "adjoint" object (skip with --no-adjoint) is measured on the adjoint-fitting case: 256
random trees of 100 compartments with HH on every compartment, 20,001 rows, trainable
Gbase per channel distribution initialised at U(0.5, 1.5) x ground truth, MSE on somatic V
against the ground-truth trace, checkpoints every 500 steps. It reports the FP64
differentiable forward, the backward (recompute + adjoint) and the bit-exact fp32 forward
of the same model, and gradient evaluations per second.

--calibrate-copy MiB times a plain device-to-device copy of that size instead (the
reference kernel for scaling the FETCH_SIZE / WRITE_SIZE profiler counters to bytes).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "compartment-updates/s vs EDEN CPU at trace tol; adjoint grads/s vs forward"
HBM_PEAK = 8.0e12
TOL_MV = 1e-3

# algorithmic bytes per compartment update (fp32 state), per unit of work:
B_PASSIVE = 24   # V r+w 8, C 4, R to parent 4, parent index 4, g_leak*A 4
B_HH = 56        # 24 + 3 gates r+w 24 + gNa*A, gK*A 8
B_RECORD = 4     # recorded V per column-row


B_HH_POINT = 40  # one-compartment HH cell: V, m, h, n r+w 32 + per-instance input 4 + recorded V 4
B_SYNAPSE = 32   # per synapse-step, scan formulation: A, B r+w 16, weight 4, delay 4, pending slot r+w 8


def algorithmic_bytes(model, rows, cells=None):
    """Bytes per full run over the given cells (all by default): per compartment-row 24 B
    (passive tree compartment), 56 B (HH tree compartment) or 40 B (one-compartment HH
    cell), plus 4 B per recorded column-row of tree cells, plus 32 B per synapse-row
    (chemical and Poisson-input synapses)."""
    n_chan = np.diff(model.comp_chan_begin)
    is_hh = n_chan > 1
    cb = model.cell_comp_begin
    comp_cell = np.repeat(np.arange(model.n_cells), np.diff(cb))
    point = np.diff(cb)[comp_cell] == 1
    sel = np.ones(model.n_comps, bool) if cells is None else np.isin(comp_cell, np.asarray(cells))
    hh_point = int((is_hh & point & sel).sum())
    hh = int((is_hh & ~point & sel).sum())
    passive = int((sel & ~is_hh).sum())
    rec_comp = np.asarray(model.record_comp, np.int64)
    rec = int((sel[rec_comp] & ~point[rec_comp]).sum())
    n_syn = np.diff(model.comp_syn_begin) + np.diff(model.comp_poisson_begin)
    syn = int(n_syn[sel].sum())
    return rows * (B_PASSIVE * passive + B_HH * hh + B_HH_POINT * hh_point + B_RECORD * rec + B_SYNAPSE * syn)


def calibrate_copy(mib, reps=10):
    """Plain device-to-device copy of `mib` MiB, `reps` times: the profiler's FETCH_SIZE and
    WRITE_SIZE per dispatch of this kernel should each read `mib` * 1024 KiB."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.arange(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / reps
    assert torch.equal(dst, src)
    print(json.dumps({"copy_mib": mib, "bytes_read": n * 4, "bytes_written": n * 4, "reps": reps,
                      "ms_per_copy": el * 1e3, "gbs_read_plus_write": 2 * n * 4 / el / 1e9}), flush=True)


def build_workload(name):
    from eden_amd import configs
    if name == "config2":
        m = configs.config2()
        desc = ("config2: 512 random-tree cells, compartments lognormal(median 100) clipped 20-400, "
                "HH soma + passive dendrites, dt 0.025 ms, 250 ms (10,001 rows)")
    elif name == "config4":
        m = configs.config4()
        desc = ("config4: 4,096 random-tree cells, compartments log-uniform 10-2000, HH soma x U(0.5,1.5), "
                "passive dendrites, dt 0.025 ms, 500 ms (20,001 rows)")
    elif name == "config3":
        m = configs.spiking_network()
        desc = ("config3: 100,000 one-compartment HH cells (80% E), in-degree 100, expTwo synapses, lognormal "
                "weights, delays U(1,5) ms, 5 Hz Poisson drive, dt 0.025 ms, 1 s (40,001 rows)")
    elif name == "config1":
        m = configs.hh_population()
        desc = "config1: 1,000 one-compartment HH cells + pulse, dt 0.025 ms, 100 ms (4,001 rows)"
    else:
        raise SystemExit(f"unknown workload {name}")
    return m, desc


def cpu_baseline(model, threads, budget_s=10.0):
    """The oracle (CPU restatement, OpenMP over cells) on a bounded sample of the same
    workload: all cells, the first R rows, R chosen for about budget_s seconds. Timed in the
    flavour built as EDEN builds its generated code (gcc -Ofast, glibc expf/logf; the
    baseline `value`), and in the exact-arithmetic flavour the GPU reproduces bit for bit."""
    from oracle import oracle
    total = model.total_rows()
    probe = min(total, 100)

    def timed(precision):
        t0 = time.perf_counter()
        oracle.run(model, n_rows=probe, threads=threads, want_trace=False, precision=precision)
        dtp = time.perf_counter() - t0
        rows = int(min(total, max(probe, probe * budget_s / max(dtp, 1e-6))))
        t0 = time.perf_counter()
        oracle.run(model, n_rows=rows, threads=threads, want_trace=False, precision=precision)
        return rows, time.perf_counter() - t0

    rows, el = timed("f32ofast")
    rows_x, el_x = timed("f32")
    # the 4 initialisation steps are included in the sample time and not counted as rows
    return dict(value=model.n_comps * rows / el, unit="compartment-updates/s", cores=threads, kind="port",
                build="gcc -Ofast -march=x86-64-v3, glibc expf/logf (EDEN's generated-code flags)",
                sample=f"all {model.n_cells} cells, first {rows} of {total} rows ({el:.1f} s, OpenMP static)",
                exact_flavour=dict(value=model.n_comps * rows_x / el_x,
                                   build="gcc -O2, no contraction, exp/log rounded once from double",
                                   sample=f"first {rows_x} rows ({el_x:.1f} s)"))


def parity_sample(model_name, model=None, n_cells=16):
    """GPU vs oracle over all rows on a sample of the workload: for models of independent
    cells, cells of the workload itself (spread over the sizes, `Model.subset`); otherwise a
    sub-population built the same way."""
    from eden_amd import configs
    from eden_amd.model import Simulation
    from oracle import oracle
    if model is not None and model.n_synapses == 0 and len(model.poisson_rate) == 0:
        sizes = model.comps_per_cell()
        order = np.argsort(sizes, kind="stable")
        pick = order[np.linspace(0, len(order) - 1, min(n_cells, len(order))).astype(np.int64)]
        m, _, _ = model.subset(pick)
        what = "cells of the workload"
    elif model_name == "config2":
        sizes = configs.config2_sizes()[:n_cells]
        m = configs.random_trees(n_cells, sizes=sizes, t_final_ms=250.0, seed=2001)
        what = "sub-population"
    else:
        m = configs.hh_population(n_cells=n_cells)
        what = "sub-population"
    with Simulation(m) as s:
        tg = s.run_all()
        eg = s.read_events()
    tr, ec, er, _ = oracle.run(m)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from parity import compare, compare_fine
    ev_of = {int(c): i for i, c in enumerate(m.event_comp)}
    rec2ev = [ev_of.get(int(c), 0) for c in m.record_comp]
    res = compare(tg, eg, tr, (ec, er), len(m.event_comp), rec2ev)
    # EDEN's own build of its generated code: -Ofast (reassociation, FMA contraction), glibc
    # expf/logf; the spread among CPU builds: glibc expf/logf alone at -O2
    to, eco, ero, _ = oracle.run(m, precision="f32ofast")
    r_of = compare(tg, eg, to, (eco, ero), len(m.event_comp), rec2ev)
    tl, ecl, erl, _ = oracle.run(m, precision="f32libm")
    r_lm = compare(to, (eco, ero), tl, (ecl, erl), len(m.event_comp), rec2ev)
    # the FP64 run at dt/10, compared wherever the rasters agree within one dt
    t64, ec64, er64, _ = oracle.run(m, precision="f64", substeps=10)
    d_g = compare_fine(tg, eg, t64, (ec64, er64), 10, len(m.event_comp), rec2ev)
    d_o = compare_fine(to, (eco, ero), t64, (ec64, er64), 10, len(m.event_comp), rec2ev)
    return dict(sample=what, cells=int(m.n_cells), compartments=int(m.n_comps), rows=int(res["rows"]),
                max_abs_mV=res["max_abs_mV"], raster_mismatch_cells=res["raster_mismatch_cols"],
                spikes=res["spikes_ref"], within_tol=bool(res["max_abs_mV"] <= TOL_MV),
                vs_eden_build=dict(build="oracle at gcc -Ofast -march=x86-64-v3, glibc expf/logf",
                                   max_abs_mV=r_of["max_abs_mV"], raster_mismatch_cells=r_of["raster_mismatch_cols"],
                                   identical_raster_cells=r_of["identical_raster_cols"],
                                   cpu_builds_spread_mV=r_lm["max_abs_mV"]),
                fp64_dt10_dev_mV=dict(gpu=d_g["max_abs_mV"], eden_build=d_o["max_abs_mV"],
                                      compared_samples=d_g["compared_samples"],
                                      raster_mismatch_cells=d_g["raster_mismatch_cols"]))


def adjoint_bench(reps=2):
    """Adjoint-fitting case: fp32 forward time, FP64 forward and backward times."""
    import torch
    from eden_amd import configs
    from eden_amd.adjoint import Adjoint, trace
    from eden_amd.model import Simulation
    m, g0 = configs.config5()
    with Simulation(m, profile=True) as sim:
        fwd32 = []
        for _ in range(reps):
            sim.reset()
            t0 = time.perf_counter()
            sim.run_all()
            fwd32.append(time.perf_counter() - t0)
    with Adjoint(m, checkpoint_steps=500) as adj:
        gt, a0 = adj.default_parameters()
        target = adj.forward(gt, a0)
        g = torch.tensor(g0, dtype=torch.float64, device=adj.device, requires_grad=True)
        f64, bwd, tot = [], [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            loss = ((trace(adj, g, a0) - target) ** 2).mean()
            loss.backward()
            torch.cuda.synchronize()
            tot.append(time.perf_counter() - t0)
            fms, bms = adj.timing()
            f64.append(fms / 1e3)
            bwd.append(bms / 1e3)
            g.grad = None
    f32 = float(np.min(fwd32))
    grad = float(np.min(tot))
    return dict(workload="256 random trees x 100 compartments, HH on every compartment, 20,001 rows, "
                         "Gbase per channel distribution trainable (76,800 parameters), MSE on somatic V, "
                         "every step's state and Jacobian factors kept in device memory (checkpoints every 500 steps "
                         "as fallback)",
                # headline: one gradient (FP64 forward + backward) against one fp32 forward pass of
                # the product path; target <= 3
                gradient_over_forward_fp32=grad / f32, target=3.0,
                forward_fp32_s=f32, forward_fp64_s=float(np.min(f64)), backward_s=float(np.min(bwd)),
                gradient_s=grad, grads_per_s=1.0 / grad,
                backward_over_forward_fp64=float(np.min(bwd)) / float(np.min(f64)),
                gradient_over_forward_fp64=grad / float(np.min(f64)),
                loss=float(loss.detach()))


def lds_entry(kernel, workload):
    """The dominant kernel's LDS figures from the committed SQ counter summary of the same
    workload (profiles/*_sq_counters.md), or None: LDS utilisation = LDS-active cycles over
    the CU cycles of the dispatch (the summary's LDS_UTIL), against a peak of 1 (the LDS busy
    every cycle); the share of those cycles lost to bank conflicts; LDS instructions per
    launch; and the shares of wave cycles spent waiting and issuing vector ALU work, which say
    how far the kernel is latency-bound instead."""
    import glob
    from eden_amd.profiling import read_counters
    for f in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_sq_counters.md"))):
        c = read_counters(f, kernel, workload)
        if not c or "LDS_UTIL" not in c or not c.get("SQ_LDS_IDX_ACTIVE"):
            continue
        active = c["SQ_LDS_IDX_ACTIVE"]
        util = c["LDS_UTIL"]
        out = dict(bound="lds", metric="LDS-active cycles / CU cycles", achieved=util, peak=1.0, frac=util,
                   bank_conflict_share=c.get("SQ_LDS_BANK_CONFLICT", 0.0) / active,
                   source=os.path.join("profiles", os.path.basename(f)))
        if "SQ_INSTS_LDS" in c:
            out["lds_insts_per_launch"] = c["SQ_INSTS_LDS"]
        if c.get("SQ_WAVE_CYCLES"):
            out["wave_wait_share"] = c.get("SQ_WAIT_ANY", 0.0) / c["SQ_WAVE_CYCLES"]
            out["valu_active_share"] = c.get("SQ_ACTIVE_INST_VALU", 0.0) / c["SQ_WAVE_CYCLES"]
        return out
    return None


def plain_launch_bench(model, rows):
    """The same workload with one kernel launch per step (state back in HBM between steps)
    instead of the persistent in-kernel time loop: launched one by one, and captured in
    hipGraphs of 256 steps (timed on the second run, which replays the graphs)."""
    from eden_amd.model import Simulation

    def timed(sim):
        sim.reset()
        t0 = time.perf_counter()
        done = 0
        while done < rows:
            n = sim.run(-1)
            if n == 0:
                break
            done += n
        return time.perf_counter() - t0

    with Simulation(model, persistent=-1, profile=True) as sim:
        el = timed(sim)
        launches = sum(k["launches"] for k in sim.kernel_infos())
    with Simulation(model, persistent=-1, use_graph=True) as sim:
        capture = timed(sim)    # captures and instantiates the graphs, then runs them
        replay = timed(sim)
    return dict(seconds=el, compartment_updates_per_s=model.n_comps * rows / el, launches=int(launches),
                graph_seconds=replay, graph_first_run_seconds=capture)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="config4")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MiB")
    ap.add_argument("--threads-per-cell", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip eden_amd.autotune (tree-cell size limits by timed runs) before the timed run")
    ap.add_argument("--channel-wave", default="",
                    help="channel wave per workgroup size a,b,c,d for 64/128/256/512 threads: 0 rule, 1 on, "
                         "-1 off (skips the channel-wave search)")
    ap.add_argument("--size-limits", default="",
                    help="tree-cell size limits a,b,c,d for 64/128/256/512 threads (skips the autotuner)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: the simulation runs on one GPU; multi-GPU runs are not supported (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if not args.no_adjoint:
        # torch's device runtime comes up before the library's own HIP use in this process
        import torch
        torch.cuda.init()

    from eden_amd.model import Simulation

    print(f"[bench] building workload {args.workload}", file=sys.stderr, flush=True)
    model, desc = build_workload(args.workload)
    rows = model.total_rows()
    limits, tune, waves = None, None, None
    if args.channel_wave:
        waves = tuple(int(x) for x in args.channel_wave.split(","))
    if args.size_limits:
        limits = tuple(int(x) for x in args.size_limits.split(","))
        tune = dict(size_limits=list(limits), given=True)
        if waves is not None:
            tune["channel_wave"] = list(waves)
    elif not args.no_autotune and args.threads_per_cell == 0 and model.n_comps > model.n_cells:
        from eden_amd import autotune
        t0 = time.perf_counter()
        limits, report = autotune.autotune(model, steps=None, repeats=2, verbose=True)
        best_t = min(t for l, t in report if tuple(l) == tuple(limits))
        if waves is None:
            waves, wreport = autotune.autotune_channel_wave(model, limits, repeats=2, verbose=True, baseline=best_t)
        else:
            wreport = []
        tune = dict(size_limits=list(limits), channel_wave=list(waves), seconds=time.perf_counter() - t0,
                    timings=[dict(size_limits=list(l), ms_per_run=t * 1e3) for l, t in report]
                    + [dict(size_limits=list(limits), channel_wave=list(w), ms_per_run=t * 1e3) for w, t in wreport])
        print(f"[bench] autotuned size limits {limits}, channel waves {waves}", file=sys.stderr, flush=True)
    # timed runs without per-launch HIP events; one more run with them gives the per-kernel
    # durations for the roofline
    sim = Simulation(model, threads_per_cell=args.threads_per_cell, size_limits=limits, channel_wave=waves)
    tpb, slots = sim.launch_shape()

    def one_step(s):
        s.reset()
        done = 0
        while done < rows:
            n = s.run(-1)
            if n == 0:
                break
            done += n
        return done

    for w in range(args.warmup):
        one_step(sim)
        print(f"[bench] warmup {w} done", file=sys.stderr, flush=True)
    times = []
    for s in range(args.steps):
        t0 = time.perf_counter()
        one_step(sim)
        el = time.perf_counter() - t0
        times.append(el)
        print(f"[bench] step {s}: {el * 1e3:.2f} ms wall", file=sys.stderr, flush=True)
    kern = {}
    with Simulation(model, threads_per_cell=args.threads_per_cell, profile=True, size_limits=limits,
                    channel_wave=waves) as psim:
        one_step(psim)    # first launches
        one_step(psim)
        infos = psim.kernel_infos()
        for ki in infos:
            acc = kern.setdefault(ki["name"], dict(ki, total_ms=0.0, launches=0))
            acc["total_ms"] += ki["total_ms"]
            acc["launches"] += ki["launches"]
        msg = ", ".join(f"{ki['name']} {ki['total_ms']:.2f} ms/{ki['launches']} (grid {ki['grid']}, "
                        f"LDS {ki['lds_bytes']} B)" for ki in infos)
        print(f"[bench] profiled run: {msg}", file=sys.stderr, flush=True)
    net = None
    if model.n_synapses > 0:
        w, nwin, rerun, deliv = sim.network_stats()
        net = dict(window_steps=w, windows=nwin, windows_rerun=rerun, synaptic_events=deliv,
                   synapse_instances=int(model.n_synapses + len(model.poisson_rate)))
    sim.close()
    t_step = float(np.mean(times))
    updates = model.n_comps * rows
    value = updates / t_step
    # dominant kernel: the one that advances the most compartments (the tree buckets run
    # concurrently, so each one's duration spans most of the run; the work share tells them
    # apart), then the longest
    dom = max(kern.values(), key=lambda k: (k["comps"], k["total_ms"]))
    per_launch_ms = dom["total_ms"] / max(1, dom["launches"])
    launches_per_step = max(1, dom["launches"])   # launches of one (the profiled) run
    bytes_per_launch = algorithmic_bytes(model, rows, dom["cell_index"]) / launches_per_step
    achieved = bytes_per_launch / (per_launch_ms * 1e-3)
    # HBM bytes per launch of that kernel on this workload, from the committed PMC summary;
    # with tree cells the bucket contents depend on the size limits, so the summary's
    # workload names them and only a profile taken with the limits used here applies
    import glob
    from eden_amd.autotune import DEFAULT as DEFAULT_LIMITS
    from eden_amd.profiling import read_traffic
    wkey = args.workload
    if model.n_comps > model.n_cells:
        wkey += " (size limits " + ",".join(str(x) for x in (limits or DEFAULT_LIMITS))
        if waves and any(waves):
            wkey += ", channel waves " + ",".join(str(x) for x in waves)
        wkey += ")"
    traffic = None
    for f in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_traffic.md"))):
        traffic = traffic or read_traffic(f, dom["name"], wkey)
    lds = lds_entry(dom["name"], wkey)
    roofline = dict(bound="hbm", achieved=achieved / 1e9, peak=HBM_PEAK / 1e9, unit="GB/s",
                    frac=achieved / HBM_PEAK, traffic=traffic, kernel=dom["name"],
                    kernel_cells=int(dom["cells"]), kernel_comps=int(dom["comps"]),
                    bytes_per_launch=bytes_per_launch, launch_ms=per_launch_ms,
                    all_kernels={k["name"]: dict(ms_per_launch=k["total_ms"] / max(1, k["launches"]),
                                                 cells=int(k["cells"]), comps=int(k["comps"]))
                                 for k in kern.values()},
                    # every kernel of the run together: algorithmic bytes of the whole workload
                    # over the mean wall time of a run (kernels overlap)
                    aggregate=dict(achieved=algorithmic_bytes(model, rows) / t_step / 1e9,
                                   frac=algorithmic_bytes(model, rows) / t_step / HBM_PEAK))
    if lds is not None:
        roofline["lds"] = lds
    out = {
        "metric": METRIC, "value": value, "unit": "compartment-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": t_step * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": desc, "cells": model.n_cells, "compartments": model.n_comps, "rows": rows,
                   "launch_shape": {"threads_per_cell": tpb, "compartments_per_thread": slots}},
        "roofline": roofline,
    }
    if tune is not None:
        out["autotune"] = tune
    if net is not None:
        if net["synaptic_events"] >= 0:
            net["synaptic_events_per_s"] = net["synaptic_events"] / t_step
        net["synapse_steps_per_s"] = net["synapse_instances"] * rows / t_step
        out["network"] = net
    if not args.no_parity:
        print("[bench] parity sample", file=sys.stderr, flush=True)
        out["parity"] = parity_sample(args.workload, model)
    if not args.no_plain and args.workload == "config4":
        print("[bench] plain launch (one launch per step)", file=sys.stderr, flush=True)
        pl = plain_launch_bench(model, rows)
        pl["persistent_speedup"] = pl["seconds"] / t_step
        pl["persistent_speedup_over_graph"] = pl["graph_seconds"] / t_step
        out["plain_launch"] = pl
    if not args.no_adjoint:
        print("[bench] adjoint (256 x 100-compartment trees)", file=sys.stderr, flush=True)
        out["adjoint"] = adjoint_bench()
    if not args.no_cpu_baseline:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
        print(f"[bench] cpu baseline on {threads} threads", file=sys.stderr, flush=True)
        cb = cpu_baseline(model, threads, args.cpu_budget)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = value / cb["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
