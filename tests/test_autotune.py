"""Autotuner host logic (no GPU): cache round trip, signatures, the search's choice."""
import numpy as np

from sph_amd import autotune as T
from sph_amd import workloads as W


def test_cache_round_trip(tmp_path):
    cache = T.TuneCache(str(tmp_path / "tune.json"))
    res = T.TuneResult(0.5, 128, 256, 0.123, ((1.0, 64, 64, 1, 0.2), (0.5, 128, 256, 4, 0.123)), 1.5, False, 4)
    cache.put("k", res)
    got = cache.get("k")
    assert got.cached and got.seconds == 0.0
    assert (got.cell_scale, got.density_block, got.force_block, got.ms_per_step, got.lanes) == (0.5, 128, 256, 0.123, 4)
    assert got.trials == res.trials
    assert cache.get("other") is None


def test_signature_tracks_workload_and_order():
    a = W.box_uniform(512, seed=1)
    b = W.box_uniform(512, seed=2, h=0.05)
    ka = T.signature("plain", [(a.params, a.constants, a.state)], "BOX", "NAIVE", 10, "cpu")
    kb = T.signature("plain", [(b.params, b.constants, b.state)], "BOX", "NAIVE", 10, "cpu")
    assert ka != kb
    assert ka == T.signature("plain", [(a.params, a.constants, a.state)], "BOX", "NAIVE", 10, "cpu")
    kab = T.signature("batch", [(a.params, a.constants, a.state), (b.params, b.constants, b.state)], "BOX", "NAIVE",
                      10, "cpu")
    kba = T.signature("batch", [(b.params, b.constants, b.state), (a.params, a.constants, a.state)], "BOX", "NAIVE",
                      10, "cpu")
    assert kab != kba


def test_signature_tracks_the_built_library(monkeypatch):
    """A rebuilt library (other kernels) must not reuse choices cached for the old one."""
    w = W.box_uniform(256, seed=1)
    inst = [(w.params, w.constants, w.state)]
    monkeypatch.setattr(T, "_LIBRARY_ID", "build-a")
    ka = T.signature("plain", inst, "BOX", "NAIVE", 20, "cpu")
    monkeypatch.setattr(T, "_LIBRARY_ID", "build-b")
    kb = T.signature("plain", inst, "BOX", "NAIVE", 20, "cpu")
    assert ka != kb


def test_search_picks_fastest_and_skips_rejected_grids(monkeypatch):
    # phase 1: (grid, lanes) at 128/128; phase 2: the other block pairs at the best of them
    cost = {(1.0, 128, 128, 1): 5.0, (1.0, 128, 128, 4): 4.5, (2.0, 128, 128, 1): 6.0, (2.0, 128, 128, 4): 7.0,
            (1.0, 64, 64, 4): 3.0, (1.0, 64, 128, 4): 2.0, (1.0, 128, 64, 4): 4.0}
    state = {}
    closed = []

    def make(scale):
        if scale == 0.25:
            raise RuntimeError("grid too fine")

        def configure(db, fb, ln):
            if ln == 16:
                raise RuntimeError("too many threads")
            state["cfg"] = (scale, db, fb, ln)
        return configure, lambda k: None, lambda: None, lambda: closed.append(scale)

    monkeypatch.setattr(T, "_time_steps", lambda step, reset, steps, warmup: cost[state["cfg"]])
    res = T._search(make, [1.0, 0.25, 2.0], (64, 128), 5, 1, lanes=(1, 4, 16))
    assert (res.cell_scale, res.density_block, res.force_block, res.lanes, res.ms_per_step) == (1.0, 64, 128, 4, 2.0)
    # two grid passes, the block pass, one re-timing pass for the three finalists (all on 1.0)
    assert len(res.trials) == 7 and closed == [1.0, 2.0, 1.0, 1.0]


def test_finalists_are_retimed(monkeypatch):
    # the first timing of the default is slow (warm-up); its second timing wins
    timings = {(1.0, 128, 128, 1): [9.0, 1.0, 1.0], (1.0, 128, 128, 2): [2.0, 2.0], (1.0, 64, 64, 1): [3.0, 3.0],
               (1.0, 64, 128, 1): [4.0], (1.0, 128, 64, 1): [5.0]}
    state = {}

    def make(scale):
        def configure(db, fb, ln):
            state["cfg"] = (scale, db, fb, ln)
        return configure, lambda k: None, lambda: None, lambda: None

    def fake_time(step, reset, steps, warmup):
        return timings[state["cfg"]].pop(0)

    monkeypatch.setattr(T, "_time_steps", fake_time)
    res = T._search(make, [1.0], (64, 128), 5, 1, lanes=(1, 2))
    assert (res.density_block, res.force_block, res.lanes, res.ms_per_step) == (128, 128, 1, 1.0)


def test_finalists_are_one_per_lane_count_and_retimed_over_the_run(monkeypatch):
    """The three fastest search trials are all one-lane block pairs; the finalists are the
    fastest trial of each lane count instead, and the runner-up block pair of the fastest
    lane count, re-timed over final_steps, where 4 lanes win (a workload whose cost per step
    changes as it evolves)."""
    search = {1: {(128, 128): 1.0, (64, 64): 1.1, (64, 128): 1.2, (128, 64): 1.3}, 4: {(128, 128): 1.5}, 2: {(128, 128): 1.4}}
    final = {1: 2.0, 2: 2.2, 4: 1.8}
    cfg, steps_seen = {}, []

    def make(scale):
        def configure(db, fb, ln):
            cfg["k"] = (db, fb, ln)
        return configure, lambda k: None, lambda: None, lambda: None

    def fake_time(step, reset, steps, warmup):
        db, fb, ln = cfg["k"]
        steps_seen.append((steps, cfg["k"]))
        return final[ln] if steps == 50 else search[ln][(db, fb)]

    monkeypatch.setattr(T, "_time_steps", fake_time)
    res = T._search(make, [1.0], (64, 128), 5, 1, lanes=(1, 2, 4), final_steps=50)
    assert (res.lanes, res.launch, res.ms_per_step) == (4, (128, 128), 1.8)
    finals = [k for n, k in steps_seen if n == 50]
    assert sorted(finals) == sorted([(128, 128, 1), (64, 64, 1), (128, 128, 2), (128, 128, 4)])


def test_grid_voxel_is_power_of_two():
    res = T.TuneResult(0.5, 64, 64, 1.0)
    g = res.grid_voxel(0.0457)
    assert g == 2.0 ** -6 and np.log2(g) == int(np.log2(g))
    assert T.TuneResult(None, 64, 64, 1.0).grid_voxel(0.0457) is None


def test_batch_result_round_trips_per_instance_values(tmp_path):
    cache = T.TuneCache(str(tmp_path / "tune.json"))
    res = T.TuneResult((1.0, 0.5, 1.0), 256, 64, 0.5, (("load>=512", 256, 256, "load/16", 0.6),), 2.0, False,
                       (1, 16, 2), ("load>=512", "load/16"))
    cache.put("b", res)
    got = cache.get("b")
    assert got.cell_scale == (1.0, 0.5, 1.0) and got.lanes == (1, 16, 2)
    assert got.rule == ("load>=512", "load/16") and got.trials == res.trials


def test_load_rules():
    load = np.array([0.5, 3.9, 4.0, 31.9, 64.0, 2627.0, np.nan])
    # largest power of two in 1..16 (or 1..cap) not above load / divisor; below 1 (or
    # unknown) one lane
    assert T.lanes_for_load(load, 4.0) == (1, 1, 1, 4, 16, 16, 1)
    assert T.lanes_for_load(load, 16.0) == (1, 1, 1, 1, 4, 16, 1)
    assert T.lanes_for_load(load, 128.0) == (1, 1, 1, 1, 1, 16, 1)
    assert T.lanes_for_load(load, 4.0, cap=64) == (1, 1, 1, 4, 16, 64, 1)
    assert T.scales_for_load(load, 64.0) == (1.0, 1.0, 1.0, 1.0, 0.5, 0.5, 1.0)
    assert T.scales_for_load(load, 0.0)[:6] == (0.5,) * 6
    assert T.scales_for_load(load, float("inf")) == (1.0,) * 7


def test_staged_search_times_grids_then_lanes(monkeypatch):
    """cross=False: every grid with the middle lanes key, then the other lanes keys on the
    fastest grid, then block pairs, then finalists."""
    cost = {("g0", "l0"): 3.0, ("g0", "l1"): 2.0, ("g0", "l2"): 2.5,
            ("g1", "l0"): 1.5, ("g1", "l1"): 1.2, ("g1", "l2"): 1.0}
    seen = []
    # drive _search through a make() whose timing is looked up from the configured key
    cfg = {}

    def make2(g):
        def configure(db, fb, ln):
            cfg["k"] = (g, ln, db, fb)
        return configure, lambda k: None, lambda: None, lambda: None

    def time_cfg(step, reset, steps, warmup):
        g, ln, db, fb = cfg["k"]
        seen.append((g, ln, db, fb))
        return cost[(g, ln)] + (0.0 if (db, fb) == (64, 64) else 0.1)

    monkeypatch.setattr(T, "_time_steps", time_cfg)
    res = T._search(make2, ["g0", "g1"], (64, 256), 5, 1, None, ["l0", "l1", "l2"], cross=False)
    first = seen[1:4]  # seen[0] is the discarded warm-up timing
    assert [(g, ln) for g, ln, _, _ in first] == [("g0", "l1"), ("g1", "l1"), ("g1", "l0")]
    assert (res.cell_scale, res.lanes, res.launch) == ("g1", "l2", (64, 64))
