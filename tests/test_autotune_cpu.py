"""Host logic of the size-limit autotuner (eden_amd/autotune.py) with a fake timer in place of
GPU runs: the coordinate-wise search from the size rule, the ordering constraint between the
limits, and the margin a candidate must beat the present limits by."""
from eden_amd import autotune, configs


def _model():
    return configs.random_trees(4, sizes=[30, 200, 600, 1500], t_final_ms=1.0, seed=3)


def _tune(monkeypatch, timing, margin=0.02):
    calls = []

    def fake(model, limits, steps=None, repeats=2, device=0):
        calls.append(tuple(limits))
        return timing(tuple(limits))

    monkeypatch.setattr(autotune, "time_limits", fake)
    limits, report = autotune.autotune(_model(), margin=margin)
    assert [l for l, _ in report] == calls
    return limits, calls


def test_default_kept_when_no_candidate_is_faster(monkeypatch):
    limits, calls = _tune(monkeypatch, lambda l: 1.0)
    assert limits == autotune.DEFAULT
    assert calls[0] == autotune.DEFAULT
    # every candidate value of every limit was tried once, from the present limits
    assert len(calls) == 1 + sum(len([v for v in c if v != d]) for c, d in zip(autotune.CANDIDATES, autotune.DEFAULT))


def test_gain_within_the_margin_is_noise(monkeypatch):
    # 1 % faster with a smaller first limit: below the 2 % margin, the size rule stays
    limits, _ = _tune(monkeypatch, lambda l: 0.99 if l[0] == 192 else 1.0)
    assert limits == autotune.DEFAULT
    # with no margin the same timings move the limit
    limits, _ = _tune(monkeypatch, lambda l: 0.99 if l[0] == 192 else 1.0, margin=0.0)
    assert limits[0] == 192


def test_coordinate_search_keeps_each_improvement(monkeypatch):
    # faster with 128 for the first limit and, on top of it, 384 for the second
    def timing(l):
        t = 1.0
        if l[0] == 128:
            t -= 0.1
        if l[1] == 384:
            t -= 0.1
        return t

    limits, calls = _tune(monkeypatch, timing)
    assert limits == (128, 384, 1024, 2047)
    # the second limit's candidates were timed with the first limit already moved
    assert all(c[0] == 128 for c in calls if c[1] != 512)


def test_limits_stay_ordered(monkeypatch):
    # a candidate first limit above the second limit is never timed, however fast
    monkeypatch.setattr(autotune, "CANDIDATES", ((128, 600), (384, 448, 512), (768, 896, 1024)))
    limits, calls = _tune(monkeypatch, lambda l: 0.5 if l[0] > l[1] else 1.0)
    assert all(c[i] <= c[i + 1] for c in calls for i in range(3))
    assert (600, 512, 1024, 2047) not in calls
    assert limits == autotune.DEFAULT


def test_point_cells_only_skip_the_search(monkeypatch):
    def fail(*a, **k):
        raise AssertionError("no timing for one-compartment cells")

    monkeypatch.setattr(autotune, "time_limits", fail)
    limits, report = autotune.autotune(configs.hh_population(n_cells=10, t_final_ms=1.0))
    assert limits == autotune.DEFAULT and report == []


def _tune_waves(monkeypatch, timing, rule=(True, False, None, True), margin=0.02, baseline=None):
    calls = []

    def fake(model, limits, steps=None, repeats=2, device=0, channel_wave=None):
        calls.append(tuple(channel_wave))
        return timing(tuple(channel_wave))

    monkeypatch.setattr(autotune, "time_limits", fake)
    monkeypatch.setattr(autotune, "rule_channel_wave", lambda model, limits, device=0: list(rule))
    waves, report = autotune.autotune_channel_wave(_model(), autotune.DEFAULT, margin=margin, baseline=baseline)
    assert [w for w, _ in report] == calls
    return waves, calls


def test_channel_wave_search_flips_each_used_size_once(monkeypatch):
    waves, calls = _tune_waves(monkeypatch, lambda w: 1.0)
    assert waves == (0, 0, 0, 0)
    # the rule, then each size with cells flipped against the rule's choice; the 256 size has
    # no cells and is not timed
    assert calls == [(0, 0, 0, 0), (-1, 0, 0, 0), (0, 1, 0, 0), (0, 0, 0, -1)]


def test_channel_wave_flip_kept_when_faster_beyond_the_margin(monkeypatch):
    def timing(w):
        return 0.9 if w[1] == 1 else (0.995 if w[0] == -1 else 1.0)

    waves, calls = _tune_waves(monkeypatch, timing)
    assert waves == (0, 1, 0, 0)          # 10 % faster kept, 0.5 % faster is noise
    # later flips are timed on top of the kept one
    assert calls[-1] == (0, 1, 0, -1)


def test_channel_wave_baseline_not_timed_again(monkeypatch):
    waves, calls = _tune_waves(monkeypatch, lambda w: 1.0, baseline=1.0)
    assert (0, 0, 0, 0) not in calls


def test_channel_wave_point_cells_only_skip_the_search(monkeypatch):
    def fail(*a, **k):
        raise AssertionError("no timing for one-compartment cells")

    monkeypatch.setattr(autotune, "time_limits", fail)
    monkeypatch.setattr(autotune, "rule_channel_wave", fail)
    waves, report = autotune.autotune_channel_wave(configs.hh_population(n_cells=10, t_final_ms=1.0),
                                                   autotune.DEFAULT)
    assert waves == (0, 0, 0, 0) and report == []
