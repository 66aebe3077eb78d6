"""Launch-shape autotuner for tree cells: picks `size_limits` (the largest cell given 64, 128,
256 and 512 threads) by timing short runs of the model itself.

Each candidate is a full context (its own Hines schedules) advanced `steps` steps (default:
the whole run), timed from the host around `edenhip_run` after one warm-up run; the state is
rewound between runs. Short runs are cheaper but their timings need not carry over to the
whole run, where the buckets' tails overlap differently. The
search is coordinate-wise over the three movable limits, from the size rule (256, 512, 1024):
a limit moves down when the larger workgroup is faster for the cells between the two values
by more than `margin` (default 2 %: whole runs of one context vary by about 1 % from run to
run, so a smaller gain is not told apart from noise and the present limits are kept).

    from eden_amd.autotune import autotune
    limits, report = autotune(model)
    with Simulation(model, size_limits=limits) as sim: ...
"""
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


def time_limits(model, limits, steps=None, repeats=2, device=0):
    """Seconds per `steps` steps of `model` with the given size limits (best of `repeats`)."""
    m = _shortened(model, steps)
    with Simulation(m, size_limits=limits, device=device) as sim:
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
