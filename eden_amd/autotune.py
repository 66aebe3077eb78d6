"""Launch-shape autotuner for tree cells: picks `size_limits` (the largest cell given 64, 128,
256 and 512 threads) and, per workgroup size, whether a separate channel wave runs the gates
(`channel_wave`), by timing runs of the model itself.

Each candidate is a full context (its own Hines schedules) advanced `steps` steps (default:
the whole run), timed from the host around `edenhip_run` after one warm-up run; the state is
rewound between runs. Short runs are cheaper but their timings need not carry over to the
whole run, where the buckets' tails overlap differently. The
search is coordinate-wise over the three movable limits, from the size rule (256, 512, 1024):
a limit moves down when the larger workgroup is faster for the cells between the two values
by more than `margin` (default 2 %: whole runs of one context vary by about 1 % from run to
run, so a smaller gain is not told apart from noise and the present limits are kept).

    from eden_amd.autotune import autotune, autotune_channel_wave
    limits, report = autotune(model)
    waves, report2 = autotune_channel_wave(model, limits)
    with Simulation(model, size_limits=limits, channel_wave=waves) as sim: ...

The channel-wave search starts from the library's rule (on for 512 threads, and for 64 when
the bucket has at most 4 cells per CU) and flips one workgroup size at a time, keeping a flip
that is faster by more than `margin`.
"""
import re
import copy
import sys
import time

from .model import Simulation

DEFAULT = (256, 512, 1024, 2047)
# candidate values per movable limit (each at most 4 compartments per thread)
CANDIDATES = ((128, 192, 256), (384, 448, 512), (768, 896, 1024))


def _shortened(model, steps):
    if steps is None:
        return model
    m = copy.copy(model)
    m.t_final = float(steps) * float(model.dt)
    return m


def time_limits(model, limits, steps=None, repeats=2, device=0, channel_wave=None):
    """Seconds per `steps` steps of `model` with the given size limits and channel-wave
    settings (best of `repeats`)."""
    m = _shortened(model, steps)
    with Simulation(m, size_limits=limits, device=device, channel_wave=channel_wave) as sim:
        rows = m.total_rows()

        def run_all():
            done = 0
            while done < rows:
                n = sim.run(-1)
                if n == 0:
                    break
                done += n

        run_all()            # warm-up (also the first-launch costs)
        best = float("inf")
        for _ in range(repeats):
            sim.reset()
            t0 = time.perf_counter()
            run_all()
            best = min(best, time.perf_counter() - t0)
    return best


def autotune(model, steps=None, repeats=2, device=0, verbose=False, margin=0.02):
    """Returns (limits, report): the fastest size limits found and every timing taken."""
    if model.n_comps == model.n_cells:
        return DEFAULT, []          # only one-compartment cells: nothing to tune
    report = []

    def measure(limits):
        t = time_limits(model, limits, steps, repeats, device)
        report.append((tuple(limits), t))
        if verbose:
            what = "the whole run" if steps is None else f"{steps} steps"
            print(f"[autotune] limits {tuple(limits)}: {t * 1e3:.2f} ms / {what}", file=sys.stderr, flush=True)
        return t

    best = list(DEFAULT)
    best_t = measure(best)
    for k, cands in enumerate(CANDIDATES):
        for v in cands:
            if v == best[k]:
                continue
            trial = list(best)
            trial[k] = v
            if any(trial[i] > trial[i + 1] for i in range(3)):
                continue
            t = measure(trial)
            if t < best_t * (1.0 - margin):
                best, best_t = trial, t
    return tuple(best), report


def rule_channel_wave(model, limits, device=0):
    """Per workgroup size (64, 128, 256, 512): None when no cell takes that size, else whether
    the library's rule gives it a channel wave (read from the kernel names of a context)."""
    out = [None] * 4
    with Simulation(model, size_limits=limits, device=device) as sim:
        for ki in sim.kernel_infos():
            m = re.match(r"tree_kernel<(\d+), (true|false), ", ki["name"])
            if m and ki["cells"] > 0:
                out[(64, 128, 256, 512).index(int(m.group(1)))] = m.group(2) == "true"
    return out


def autotune_channel_wave(model, limits, steps=None, repeats=2, device=0, verbose=False, margin=0.02,
                          baseline=None):
    """Returns (channel_wave, report): per workgroup size 0 (the rule) or the flipped setting
    (1 on, -1 off) where that was faster by more than `margin`; report lists (channel_wave,
    seconds). `baseline`: the time of the rule's settings when already measured."""
    if model.n_comps == model.n_cells:
        return (0, 0, 0, 0), []
    rule = rule_channel_wave(model, limits, device)
    report = []

    def measure(waves):
        t = time_limits(model, limits, steps, repeats, device, channel_wave=waves)
        report.append((tuple(waves), t))
        if verbose:
            print(f"[autotune] channel_wave {tuple(waves)}: {t * 1e3:.2f} ms", file=sys.stderr, flush=True)
        return t

    best = [0, 0, 0, 0]
    best_t = baseline if baseline is not None else measure(best)
    for k in range(4):
        if rule[k] is None:
            continue
        trial = list(best)
        trial[k] = -1 if rule[k] else 1
        t = measure(trial)
        if t < best_t * (1.0 - margin):
            best, best_t = trial, t
    return tuple(best), report
