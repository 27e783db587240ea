"""Sweep driver: interval arithmetic and per-point statistics (CPU), and the
surrogate-steered map against the plain method on a small grid (GPU)."""
import math

import numpy as np
import pytest

from ryr_hyperactivity_amd import sweep as sw
from ryr_hyperactivity_amd.sweep import (Cost, Point, PointResult, ReplicaStore, SteeredMap, SweepPaused, grid_points,
                                         holdout_points, holdout_summary, plan_batches, point_cost, point_seed,
                                         simulate_points, steered_map, summarize, wilson)


def test_wilson_interval():
    lo, hi = wilson(5, 10)
    assert lo == pytest.approx(0.2366, abs=1e-4) and hi == pytest.approx(0.7634, abs=1e-4)
    lo, hi = wilson(0, 20)
    assert lo == 0.0 and hi == pytest.approx(0.1611, abs=1e-4)
    # coverage: the interval holds p in about 95% of binomial draws
    rng = np.random.default_rng(0)
    hits = 0
    for _ in range(4000):
        k = rng.binomial(40, 0.3)
        lo, hi = wilson(int(k), 40)
        hits += lo <= 0.3 <= hi
    assert 0.93 < hits / 4000 < 0.97


def test_summarize_counts_beats_after_the_first():
    def beat(apd, ead, peak=1.0, jsr=700.0, sparks=10.0, lit=1):
        return [0, apd, 30, -80, ead, 60.0 if ead else -1, peak, jsr, sparks, 1, lit, 12.0 if lit else -1]
    reps = [np.array([beat(300, 1), beat(310, 1), beat(320, 0)]),        # first beat ignored
            np.array([beat(250, 0), beat(260, 0, lit=0), beat(-1, 1)])]  # a beat that never repolarised
    r = summarize(PointResult(Point(2.0, 0.0, 1000.0), reps), ncru=1000)
    assert r.ead == pytest.approx(0.5 * (0.5 + 0.5))
    assert r.ead_lit == pytest.approx(0.5 * (1.0 + 0.5))
    assert (r.ead_lo, r.ead_hi) == pytest.approx(wilson(2, 4))
    assert r.apd90 == pytest.approx(0.5 * (315.0 + 260.0))
    assert r.sparks == pytest.approx(10.0)


def test_seeds_and_costs_are_functions_of_the_point():
    a, b = Point(2.5, 0.25, 2000.0), Point(2.5, 0.25, 3000.0)
    assert point_seed(a, 3) == point_seed(Point(2.5, 0.25, 2000.0), 3)
    assert len({point_seed(a, r) for r in range(64)} | {point_seed(b, r) for r in range(64)}) == 128
    assert point_cost(a, 4, 5) == 4 * (round((100 + 5 * 2000) / 0.025) + 2)
    assert len(grid_points([1, 2], [0, 1], [1000, 2000, 3000])) == 12


def test_batch_plan_merges_cycle_lengths_only_when_latency_bound():
    pts = grid_points([1, 2], [0, 1], [1000.0, 2000.0, 4000.0])        # 4 points per cycle length
    small = plan_batches(pts, replicas=4, nbeats=5, ncru=640)          # 16 cells of 640 CRUs per length
    assert len(small) == 1 and len(small[0][0]) == 12                  # one batch at the longest length
    assert small[0][1] == round((100 + 5 * 4000) / 0.025) + 2
    big = plan_batches(pts, replicas=512, nbeats=5, ncru=20000)        # each length fills the chip
    assert [len(g) for g, _ in big] == [4, 4, 4] and [s for _, s in big] == sorted([s for _, s in big], reverse=True)
    assert [p for g, _ in big for p in g] and {p for g, _ in big for p in g} == set(pts)


def test_holdout_draw_and_summary():
    grid = grid_points([1, 2, 3, 4], [0, 0.5, 1], [1000.0, 2000.0])
    h = holdout_points(grid, 5, seed=3)
    assert len(set(h)) == 5 and set(h) <= set(grid) and h == holdout_points(grid, 5, seed=3)
    assert [grid.index(p) for p in h] == sorted(grid.index(p) for p in h)
    assert len(holdout_points(grid, 100)) == len(grid)
    # steered values: two simulated points (exact), two predictions
    a, b, c, d = grid[0], grid[1], grid[2], grid[3]
    sm = SteeredMap(values={a: 0.5, b: 0.0, c: 0.02, d: 0.9}, source={a: "train", b: "plain", c: "surrogate",
                                                                      d: "surrogate"},
                    cost=Cost(cell_steps=1000, gpu_s=3.0, train_s=1.0))
    hold = {p: PointResult(p, [], ead=v) for p, v in ((a, 0.5), (b, 0.0), (c, 0.1), (d, 0.7))}
    r = holdout_summary(grid, sm, hold, Cost(cell_steps=500, gpu_s=2.0), replicas=4, nbeats=3)
    plain = sum(point_cost(p, 4, 3) for p in grid)
    assert r["plain_cell_steps"] == plain
    assert r["plain_gpu_s_estimate"] == pytest.approx(2.0 / 500 * plain)
    assert r["gpu_time_ratio"] == pytest.approx(4.0 / (2.0 / 500 * plain))
    assert r["map_err_mean"] == pytest.approx((0.08 + 0.2) / 4) and r["map_err_max"] == pytest.approx(0.2)
    assert r["map_err_surrogate_mean"] == pytest.approx(0.14) and r["holdout_surrogate"] == 2
    assert r["resimulated_equal"]
    hold[b] = PointResult(b, [], ead=0.05)
    assert not holdout_summary(grid, sm, hold, Cost(cell_steps=500, gpu_s=2.0), 4, 3)["resimulated_equal"]


def _beats(rng, p, nb):
    """Synthetic completed-beat rows with EAD probability p per beat."""
    rows = np.zeros((nb, 12))
    rows[:, 1] = 300.0
    rows[:, 4] = (rng.random(nb) < p).astype(float)
    rows[:, 9] = 1.0
    rows[:, 10] = rows[:, 4]
    return rows


def test_replica_store_resumes_and_pauses(tmp_path):
    """Replicas in the store are taken with their recorded cost and the same
    rows (JSON keeps every bit); a missing replica whose batch would end past
    the deadline pauses the run before any engine is made."""
    path = str(tmp_path / "store.json")
    st = ReplicaStore(path)
    pts = [Point(1.0, 0.0, 1000.0), Point(2.0, 0.5, 1000.0)]
    rng = np.random.default_rng(0)
    rows = {}
    for pt in pts:
        for r in range(3):
            rows[pt, r] = _beats(rng, 0.5, 4) + rng.random((4, 12)) * 1e-9
            st.put(st.key("t", pt, r, 4, (5, 4, 3), 1, 0), rows[pt, r], 100, 0.25)
    st.save()
    again = ReplicaStore(path)
    res, cost = simulate_points(pts, 3, 4, (5, 4, 3), precision=0, store=again, tag="t")
    assert cost.cell_steps == 600 and cost.gpu_s == pytest.approx(1.5)
    for pt in pts:
        assert all(np.array_equal(res[pt].beats[r], rows[pt, r]) for r in range(3))
        assert res[pt].ead == summarize(PointResult(pt, [rows[pt, r] for r in range(3)]), 60).ead
    res, cost = simulate_points(pts, 3, 4, (5, 4, 3), precision=0, store=again, tag="t", r0=1)
    assert cost.cell_steps == 400 and len(res[pts[0]].beats) == 2
    with pytest.raises(SweepPaused):
        simulate_points(pts, 4, 4, (5, 4, 3), precision=0, store=again, tag="t", deadline=0.0)
    with pytest.raises(SweepPaused):        # another tag shares nothing
        simulate_points(pts, 3, 4, (5, 4, 3), precision=0, store=again, tag="u", deadline=0.0)


class _StubSurrogate:
    """Stands in for the GPU networks: incidence as a fixed function of the
    point, spread 0."""

    def __init__(self, *a, **k):
        pass

    def fit(self, *a, **k):
        return []

    def predict_point(self, ku2, ikr, cl, recs, ncru):
        return _truth(ku2, ikr), 0.0

    def close(self):
        pass


def _truth(ku2, ikr):
    return float(np.clip(0.9 - 2.0 * ikr + 0.05 * ku2, 0.0, 1.0))


def test_steered_map_tops_up_uncertain_training_points(monkeypatch):
    """Training points run on half the replicas; a training point whose own
    incidence or prediction is uncertain gets the other half, and its value
    is then the plain method's exactly; an uncertain new point runs every
    replica; the rest take the prediction.  Cost counts the replicas run."""
    import ryr_hyperactivity_amd.surrogate as surmod
    monkeypatch.setattr(surmod, "BeatSurrogate", _StubSurrogate)
    R, NB, lat = 8, 5, (5, 4, 3)
    grid = grid_points([1.0, 2.0, 3.0], [0.0, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6], [1000.0])
    train = grid_points([1.0, 3.0], [0.0, 0.2, 0.4, 0.6], [1000.0])
    rng = np.random.default_rng(3)
    plain = {p: [_beats(rng, _truth(p.ku2, p.ikr), NB) for _ in range(R)] for p in grid}
    calls = []

    def sim(pts, r0, r1):
        calls.append((sorted(pts, key=lambda q: (q.ku2, q.ikr)), r0, r1))
        return ({p: summarize(PointResult(p, plain[p][r0:r1]), 60) for p in pts},
                Cost(cell_steps=sum(point_cost(p, r1 - r0, NB) for p in pts)))

    sm = steered_map(grid, train, R, NB, lat, simulate=sim, train_replicas=R // 2)
    full = {p: summarize(PointResult(p, plain[p]), 60).ead for p in grid}
    half = {p: summarize(PointResult(p, plain[p][:R // 2]), 60).ead for p in grid}
    assert calls[0][1:] == (0, R // 2) and len(calls[0][0]) == len(train)
    for p in grid:
        inc = _truth(p.ku2, p.ikr)
        unsure = 0.1 <= inc <= 0.9
        if p in train:
            unsure = unsure or 0.1 <= half[p] <= 0.9
            assert sm.source[p] == ("train+plain" if unsure else "train")
            assert sm.values[p] == (full[p] if unsure else half[p])
        else:
            assert sm.source[p] == ("plain" if unsure else "surrogate")
            assert sm.values[p] == (full[p] if unsure else inc)
    top = [c for c in calls if c[1] == R // 2]
    assert len(top) == 1 and all(p in train for p in top[0][0])
    n_new = sum(1 for s in sm.source.values() if s == "plain")
    n_top = sum(1 for s in sm.source.values() if s == "train+plain")
    per = point_cost(grid[0], 1, NB)
    assert sm.cost.cell_steps == per * (len(train) * R // 2 + n_top * R // 2 + n_new * R)


@pytest.mark.gpu
def test_steered_map_spends_less_than_the_plain_map():
    """The learned part against the method without it, on a small grid: the
    plain method simulates all 18 points; the steered map simulates the 10
    training points and the points the surrogate is unsure of, in batches of
    its own, and takes the prediction elsewhere.  Both use the same seeds, so
    every replica the steered map simulates has the plain run's beat rows
    exactly (a cell's trajectory does not depend on its batch).  16 replicas x
    4 scored beats put the plain estimate's own noise near 0.06."""
    from ryr_hyperactivity_amd.sweep import Cost, simulate_points, steered_map
    print("case steered map 16x6x4", flush=True)
    lat, R, NB = (16, 6, 4), 16, 5
    ikr = [0.5, 0.75, 1.0, 1.25, 1.5, 1.75, 2.0, 2.25, 2.5]       # EADs fade out as I_Kr grows
    grid = grid_points([1.0, 4.0], ikr, [1000.0])
    plain, pcost = simulate_points(grid, R, NB, lat, progress=lambda s: None)
    train = grid_points([1.0, 4.0], ikr[::2], [1000.0])
    calls = []

    def resimulate(pts, r0, r1):
        calls.append(len(pts))
        res, cost = simulate_points(pts, r1, NB, lat, r0=r0, progress=lambda s: None)
        for p in pts:
            assert len(res[p].beats) == r1 - r0
            for a, b in zip(res[p].beats, plain[p].beats[r0:r1]):
                assert np.array_equal(a, b), p
        return res, cost

    sm = steered_map(grid, train, R, NB, lat, simulate=resimulate, train_steps=2000)
    srcs = [sm.source[p] for p in grid]
    n_sur = srcs.count("surrogate")
    ratio = sm.cost.cell_steps / pcost.cell_steps
    err = [abs(sm.values[p] - plain[p].ead) for p in grid if sm.source[p] == "surrogate"]
    print(f"surrogate points {n_sur}/{len(grid)}, cell-step ratio {ratio:.2f}, "
          f"surrogate-point error mean {np.mean(err) if err else 0:.3f} max {max(err) if err else 0:.3f}", flush=True)
    for p in grid:
        inc, spr = sm.predicted.get(p, (float("nan"), float("nan")))
        print(f"  ku2 {p.ku2:g} ikr {p.ikr:.3f}: plain {plain[p].ead:.3f} "
              f"[{plain[p].ead_lo:.2f}, {plain[p].ead_hi:.2f}] predicted {inc:.3f} spread {spr:.3f} "
              f"-> {sm.source[p]}", flush=True)
    assert calls[0] == len(train) and srcs.count("train") == len(train)
    assert n_sur >= 3 and ratio < 0.85
    assert np.mean(err) < 0.1 and max(err) < 0.25
    # every uncertain prediction at a new point went to the simulator (training points ran every replica)
    for p, (inc, spr) in sm.predicted.items():
        if p not in train and (0.1 <= inc <= 0.9 or spr > 0.1):
            assert sm.source[p] == "plain"
