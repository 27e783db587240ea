"""Ensemble sweeps of the CRU engine and the surrogate-steered EAD map.

Plain method: every point (Ku2, I_Kr scale, cycle length) of a grid is
simulated with `replicas` independent cells (distinct Philox seeds) for
`nbeats` beats; per point the driver reports the EAD incidence (fraction of
beats 2..nbeats with an EAD, averaged over replicas) with a 95% interval, and
the APD90, peak ci, JSR load and spark count of those beats.  Cells of one
cycle length are advanced together in one batched handle.

Steered method: the plain method runs on a coarse training subset, a small
ensemble of beat surrogates (surrogate.BeatSurrogate, trained on the GPU) is
fitted to those beats, and each remaining point takes the surrogate's
prediction unless it is uncertain (incidence in the band [0.1, 0.9] or the
networks disagree by more than `spread_max`); those points are simulated.
Cost is counted in simulated cell-steps (deterministic) and in GPU seconds
measured with HIP events around the batched runs.

The reference has no sweep driver; its per-run outputs are the traces the
engine also writes.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from dataclasses import dataclass, field

import numpy as np

from ._abi import BEAT_COLS, PREC_F32
from .engine import CellEngine
from ._abi import CellDesc, Params


@dataclass(frozen=True)
class Point:
    ku2: float
    ikr: float
    cl: float


@dataclass
class PointResult:
    point: Point
    beats: list                      # per replica: (nbeats x BEAT_COLS) completed-beat rows
    ead: float = float("nan")        # incidence over beats 2..n, mean over replicas
    ead_lit: float = float("nan")    # incidence of the literal flag (dV/dt > 0 above -40 mV in repolarisation)
    ead_lo: float = float("nan")     # 95% interval (Wilson, on the pooled beats)
    ead_hi: float = float("nan")
    apd90: float = float("nan")
    apd90_ci: float = float("nan")   # half width of the 95% interval over replicas
    peak_ci: float = float("nan")
    jsr_load: float = float("nan")
    sparks: float = float("nan")


@dataclass
class Cost:
    cell_steps: int = 0
    gpu_s: float = 0.0
    train_s: float = 0.0

    def add(self, other: "Cost"):
        self.cell_steps += other.cell_steps
        self.gpu_s += other.gpu_s
        self.train_s += other.train_s


def point_seed(p: Point, replica: int, seed0: int = 1) -> int:
    """Philox key of one replica: a fixed function of the point and replica,
    so a point gives the same cells whatever else runs beside it."""
    k = (int(round(p.ku2 * 1000)) * 1000003 + int(round(p.ikr * 1000)) * 10007 + int(round(p.cl))) & 0xFFFFFFFF
    return (seed0 << 40) ^ (k << 8) ^ replica


def wilson(k: int, n: int, z: float = 1.959964):
    if n == 0:
        return float("nan"), float("nan")
    ph = k / n
    den = 1 + z * z / n
    c = (ph + z * z / (2 * n)) / den
    h = z * math.sqrt(ph * (1 - ph) / n + z * z / (4 * n * n)) / den
    return max(0.0, c - h), min(1.0, c + h)


def summarize(res: PointResult, ncru: int) -> PointResult:
    per_rep, per_rep_lit, apds, pk, jl, sp = [], [], [], [], [], []
    k = n = 0
    for b in res.beats:
        tail = b[1:]
        if len(tail) == 0:
            continue
        e = (tail[:, 4] > 0).astype(float)
        per_rep.append(e.mean())
        per_rep_lit.append(float((tail[:, 10] > 0).mean()))
        k += int(e.sum())
        n += len(e)
        ok = tail[:, 1] > 0
        if ok.any():
            apds.append(tail[ok, 1].mean())
        pk.append(tail[:, 6].mean())
        jl.append(tail[:, 7].mean())
        sp.append(1000.0 * tail[:, 8].mean() / ncru)
    if per_rep:
        res.ead = float(np.mean(per_rep))
        res.ead_lit = float(np.mean(per_rep_lit))
        res.ead_lo, res.ead_hi = wilson(k, n)
    if apds:
        res.apd90 = float(np.mean(apds))
        res.apd90_ci = float(1.96 * np.std(apds, ddof=1) / math.sqrt(len(apds))) if len(apds) > 1 else float("nan")
    if pk:
        res.peak_ci, res.jsr_load, res.sparks = float(np.mean(pk)), float(np.mean(jl)), float(np.mean(sp))
    return res


# step time of a batched handle (FP32, MI355X): a latency floor per step, or
# the HBM/VALU-bound rate per CRU once the batch fills the chip
STEP_FLOOR_S = 30e-6
STEP_PER_CRU_S = 0.16e-9          # 256-cell batch: 0.47 ms per step for 3.07 M CRUs


def step_time_model(ncru_total: int) -> float:
    return max(STEP_FLOOR_S, ncru_total * STEP_PER_CRU_S)


def plan_batches(points, replicas: int, nbeats: int, ncru: int, dt: float = 0.025, stim_offset: float = 100.0):
    """Groups of points to advance together, and the steps of each group.
    Cells of one cycle length need the same number of steps; when those
    groups are too small to fill the chip the steps, not the cells, cost time,
    so merging groups (the shorter cycle lengths then run extra beats) can be
    cheaper.  Greedy over the cycle lengths in decreasing order, with
    step_time_model as the cost."""
    by_cl = {}
    for pt in points:
        by_cl.setdefault(pt.cl, []).append(pt)
    steps = {cl: int(round((stim_offset + nbeats * cl) / dt)) + 2 for cl in by_cl}
    groups = []
    for cl in sorted(by_cl, reverse=True):
        pts = by_cl[cl]
        if groups:
            g, gs = groups[-1]
            apart = step_time_model(len(g) * replicas * ncru) * gs + step_time_model(len(pts) * replicas * ncru) * steps[cl]
            merged = step_time_model((len(g) + len(pts)) * replicas * ncru) * gs
            if merged < apart:
                groups[-1] = (g + pts, gs)
                continue
        groups.append((list(pts), steps[cl]))
    return groups


class SweepPaused(Exception):
    """The time budget ran out before every requested replica was simulated;
    what was done is in the ReplicaStore."""


class ReplicaStore:
    """Per-replica results kept across processes: the completed-beat rows of
    each simulated (point, replica) and its share of the batch's cost, in a
    JSON file rewritten after every batch, so that a sweep longer than one
    process continues where it stopped.  Keys hold the run's settings, and a
    tag keeps apart runs that must not share results (a hold-out re-simulates
    points on purpose).  Rows round-trip through JSON exactly."""

    def __init__(self, path=None):
        self.path = path
        self.data = {"replicas": {}, "notes": {}}
        if path and os.path.exists(path):
            with open(path) as f:
                self.data = json.load(f)

    @staticmethod
    def key(tag, pt, r, nbeats, lattice, seed0, precision):
        return (f"{tag}|{lattice[0]}x{lattice[1]}x{lattice[2]}|{nbeats}|{seed0}|{precision}|"
                f"{pt.ku2!r}|{pt.ikr!r}|{pt.cl!r}|{r}")

    def get(self, k):
        e = self.data["replicas"].get(k)
        if e is None:
            return None
        return np.array(e["beats"], dtype=np.float64).reshape(-1, BEAT_COLS), e["cell_steps"], e["gpu_s"]

    def put(self, k, beats, cell_steps, gpu_s):
        self.data["replicas"][k] = {"beats": np.asarray(beats, dtype=np.float64).tolist(),
                                    "cell_steps": int(cell_steps), "gpu_s": float(gpu_s)}

    def note(self, name, value=None):
        if value is not None:
            self.data["notes"][name] = value
        return self.data["notes"].get(name)

    def save(self):
        if not self.path:
            return
        tmp = self.path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(self.data, f)
        os.replace(tmp, self.path)


def simulate_points(points, replicas: int, nbeats: int, lattice=(20, 8, 4), precision: int = PREC_F32,
                    seed0: int = 1, device: int = 0, max_cells: int = 4096, chunk_steps: int = 20000,
                    progress=None, r0: int = 0, store: ReplicaStore | None = None, tag: str = "",
                    deadline: float | None = None):
    """Plain method on `points` with replicas r0..replicas-1: returns
    ({point: PointResult}, Cost).  Replicas already in `store` (same tag and
    settings) are taken from it with their recorded cost; the others are
    simulated in batches of at most `max_cells` cells and recorded.  A batch
    is started only if its estimated duration (step_time_model) ends before
    `deadline` (time.time() seconds); otherwise SweepPaused is raised."""
    nx, ny, nz = lattice
    ncru = nx * ny * nz
    p = Params.default(precision=precision, lqt=0)           # lqt off: the I_Kr scale applies
    store = store or ReplicaStore()
    key = lambda pt, r: store.key(tag, pt, r, nbeats, lattice, seed0, precision)
    missing = [pt for pt in points if any(store.get(key(pt, r)) is None for r in range(r0, replicas))]
    for pts, nsteps in plan_batches(missing, replicas - r0, nbeats, ncru, p.dt, p.stim_offset):
        cl = max(q.cl for q in pts)
        todo = [(pt, r) for pt in pts for r in range(r0, replicas) if store.get(key(pt, r)) is None]
        for g0 in range(0, len(todo), max_cells):
            grp = todo[g0:g0 + max_cells]
            est = step_time_model(len(grp) * ncru) * nsteps
            if deadline is not None and time.time() + est > deadline:
                store.save()
                raise SweepPaused(f"{len(todo) - g0} replicas left at cl {cl:g} (next batch ~{est:.0f} s)")
            cells = [CellDesc.default(nx=nx, ny=ny, nz=nz, Ku2=pt.ku2, ikr_scale=pt.ikr, cl=pt.cl,
                                      seed=point_seed(pt, r, seed0)) for pt, r in grp]
            with CellEngine(p, cells, device=device) as eng:
                eng.kernel_time()
                done = 0
                while done < nsteps:
                    n = min(chunk_steps, nsteps - done)
                    eng.step(n)
                    done += n
                    if progress:
                        progress(f"cl {cl:g}: {done}/{nsteps} steps, {len(cells)} cells")
                eng.sync()
                ms, _ = eng.kernel_time()
                for i, (pt, r) in enumerate(grp):
                    b = eng.beats(i)
                    store.put(key(pt, r), b[b[:, 9] > 0][:nbeats], nsteps, ms * 1e-3 / len(grp))
            store.save()
    out, cost = {}, Cost()
    for pt in points:
        reps = []
        for r in range(r0, replicas):
            b, cs, gs = store.get(key(pt, r))
            reps.append(b)
            cost.cell_steps += cs
            cost.gpu_s += gs
        out[pt] = summarize(PointResult(pt, reps), ncru)
    return out, cost


def point_cost(pt: Point, replicas: int, nbeats: int, dt: float = 0.025, stim_offset: float = 100.0) -> int:
    """Cell-steps the plain method spends on one point."""
    return replicas * (int(round((stim_offset + nbeats * pt.cl) / dt)) + 2)


@dataclass
class SteeredMap:
    values: dict                     # point -> EAD incidence
    source: dict                     # point -> "train" | "train+plain" | "plain" | "surrogate"
    cost: Cost
    predicted: dict = field(default_factory=dict)   # point -> (incidence, spread) from the surrogate


def steered_map(grid, train_points, replicas: int, nbeats: int, lattice=(20, 8, 4), band=(0.1, 0.9),
                spread_max: float = 0.1, n_models: int = 3, train_steps: int = 3000, device: int = 0,
                seed0: int = 1, simulate=None, progress=None, n_near: int = 2, train_replicas: int | None = None,
                store: ReplicaStore | None = None, deadline: float | None = None) -> SteeredMap:
    """Surrogate-steered EAD map over `grid` (see the module docstring).

    The training points are simulated with the first `train_replicas`
    replicas (all of them by default).  After training, a point is simulated
    with all `replicas` replicas if the surrogate's incidence lies in `band`
    or its networks' spread exceeds `spread_max`; a training point also if its
    own incidence over its training replicas lies in `band`, and then only its
    remaining replicas are run.  A simulated point therefore has exactly the
    plain method's value (same seeds, same replicas); an unflagged training
    point keeps the incidence of its training replicas.

    `simulate(points, r0, r1)` may replace the plain runs of replicas
    r0..r1-1 (it must return what simulate_points returns).  With a `store`,
    the steering decisions are recorded beside the replicas, so a run resumed
    after SweepPaused takes the same decisions."""
    from .surrogate import BeatSurrogate, training_set

    tr = replicas if train_replicas is None else int(train_replicas)
    assert 1 <= tr <= replicas
    sim = simulate or (lambda pts, r0, r1: simulate_points(pts, r1, nbeats, lattice, seed0=seed0, device=device,
                                                           progress=progress, r0=r0, store=store, tag="steered",
                                                           deadline=deadline))
    ncru = lattice[0] * lattice[1] * lattice[2]
    train_res, cost = sim(list(train_points), 0, tr)
    tp = list(train_res)
    dkey = (f"decisions|{lattice[0]}x{lattice[1]}x{lattice[2]}|{replicas}|{tr}|{nbeats}|{seed0}|{band}|{spread_max}|"
            f"{n_models}|{train_steps}|{n_near}|" + ";".join(f"{q.ku2!r},{q.ikr!r},{q.cl!r}" for q in grid)
            + "|" + ";".join(f"{q.ku2!r},{q.ikr!r},{q.cl!r}" for q in tp))
    saved = store.note(dkey) if store is not None else None
    predicted, todo = {}, []
    if saved is not None:
        for row in saved["predicted"]:
            predicted[Point(*row[:3])] = (row[3], row[4])
        todo = [Point(*row) for row in saved["todo"]]
        cost.train_s += saved["train_s"]
    else:
        recs = [(pt.ku2, pt.ikr, pt.cl, ncru, b) for pt, r in train_res.items() for b in r.beats if len(b) > 1]
        x, y, t, mask = training_set(recs)
        sur = BeatSurrogate(n_models=n_models, device=device, seed=seed0)
        t0 = time.perf_counter()
        sur.fit(x, y, t, mask, steps=train_steps, seed=seed0)
        train_s = time.perf_counter() - t0
        cost.train_s += train_s
        # previous-beat records for a point: the beats (all but the last of
        # each replica) of its `n_near` nearest training points
        tkey = np.array([[q.ku2, q.ikr, q.cl / 1000.0] for q in tp])
        scale = np.maximum(tkey.std(0), 1e-9)
        prev = {q: [row for b in train_res[q].beats for row in b[:-1]] for q in tp}
        for pt in grid:
            d = np.sum(((tkey - [pt.ku2, pt.ikr, pt.cl / 1000.0]) / scale) ** 2, axis=1)
            recs = [row for j in np.argsort(d)[:n_near] for row in prev[tp[int(j)]]]
            inc, spr = sur.predict_point(pt.ku2, pt.ikr, pt.cl, recs, ncru)
            predicted[pt] = (inc, spr)
            unsure = band[0] <= inc <= band[1] or spr > spread_max
            if pt in train_res:
                unsure = tr < replicas and (unsure or band[0] <= train_res[pt].ead <= band[1])
            if unsure:
                todo.append(pt)
        sur.close()
        if store is not None:
            store.note(dkey, {"predicted": [[q.ku2, q.ikr, q.cl, v[0], v[1]] for q, v in predicted.items()],
                              "todo": [[q.ku2, q.ikr, q.cl] for q in todo], "train_s": train_s})
            store.save()

    values, source = {}, {}
    for pt in grid:
        if pt in train_res:
            values[pt], source[pt] = train_res[pt].ead, "train"
        elif pt not in todo:
            values[pt], source[pt] = predicted[pt][0], "surrogate"
    new = [pt for pt in todo if pt not in train_res]
    top = [pt for pt in todo if pt in train_res]
    if new:
        res, c = sim(new, 0, replicas)
        cost.add(c)
        for pt in new:
            values[pt], source[pt] = res[pt].ead, "plain"
    if top:
        res, c = sim(top, tr, replicas)
        cost.add(c)
        for pt in top:
            full = summarize(PointResult(pt, list(train_res[pt].beats) + list(res[pt].beats)), ncru)
            values[pt], source[pt] = full.ead, "train+plain"
    return SteeredMap(values, source, cost, predicted)


def holdout_points(grid, n: int, seed: int = 1):
    """`n` distinct grid points drawn at random (seeded), in grid order."""
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(len(grid), size=min(n, len(grid)), replace=False))
    return [grid[int(i)] for i in idx]


def holdout_summary(grid, sm: SteeredMap, hold: dict, hold_cost: Cost, replicas: int, nbeats: int) -> dict:
    """Map error of the steered map against a plain-method hold-out, and its
    GPU time against the plain map's.  The plain map is not run: its GPU time
    is the hold-out's GPU time per cell-step times the plain map's cell-steps
    (the hold-out runs in batches of the same kind).  Points the steered map
    simulated with all replicas carry the same seeds, so their values must
    equal the hold-out's exactly; `resimulated_equal` reports it."""
    plain_cells = sum(point_cost(p, replicas, nbeats) for p in grid)
    plain_gpu_s = hold_cost.gpu_s / hold_cost.cell_steps * plain_cells
    err = np.array([abs(sm.values[p] - hold[p].ead) for p in hold])
    sur = np.array([sm.source[p] == "surrogate" for p in hold])
    same = [sm.values[p] == hold[p].ead for p in hold if sm.source[p] in ("plain", "train+plain")]
    return dict(holdout=len(hold), holdout_surrogate=int(sur.sum()),
                map_err_mean=float(err.mean()), map_err_max=float(err.max()),
                map_err_surrogate_mean=float(err[sur].mean()) if sur.any() else 0.0,
                map_err_surrogate_max=float(err[sur].max()) if sur.any() else 0.0,
                resimulated_equal=bool(all(same)), plain_cell_steps=plain_cells,
                plain_gpu_s_estimate=plain_gpu_s,
                cell_step_ratio=sm.cost.cell_steps / plain_cells,
                gpu_time_ratio=(sm.cost.gpu_s + sm.cost.train_s) / plain_gpu_s)


def grid_points(ku2s, ikrs, cls):
    return [Point(float(k), float(i), float(c)) for k in ku2s for i in ikrs for c in cls]


def main(argv=None):
    ap = argparse.ArgumentParser(description="EAD-incidence sweeps: plain ensembles and the surrogate-steered map")
    ap.add_argument("--lattice", default="20x8x4")
    ap.add_argument("--ku2", default="1,1.5,2,2.5,3,3.5,4")
    ap.add_argument("--ikr", default="0,0.25,0.5,0.75,1")
    ap.add_argument("--cl", default="1000,2000,3000,4000")
    ap.add_argument("--replicas", type=int, default=16)
    ap.add_argument("--nbeats", type=int, default=6)
    ap.add_argument("--mode", choices=["plain", "steered", "compare", "holdout"], default="compare",
                    help="compare: the plain method on every grid point; holdout: on --holdout random points")
    ap.add_argument("--holdout", type=int, default=200, help="hold-out size of --mode holdout")
    ap.add_argument("--train-every", default="2",
                    help="training subset: every k-th value on each axis (one k, or k_ku2,k_ikr,k_cl)")
    ap.add_argument("--train-replicas", type=int, default=None,
                    help="replicas of a training point before steering (default: all)")
    ap.add_argument("--max-cells", type=int, default=4096, help="cells per batched handle")
    ap.add_argument("--store", default=None, help="JSON file of per-replica results; a paused run resumes from it")
    ap.add_argument("--budget-s", type=float, default=None,
                    help="start no batch that would end later than this many seconds after the start; the run "
                         "then pauses (exit status 3) and a later run with the same arguments continues")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    lat = tuple(int(v) for v in args.lattice.split("x"))
    ax = [[float(v) for v in s.split(",")] for s in (args.ku2, args.ikr, args.cl)]
    grid = grid_points(*ax)
    log = lambda s: print(s, flush=True)
    store = ReplicaStore(args.store)
    deadline = time.time() + args.budget_s if args.budget_s else None
    run = dict(seed0=args.seed, progress=log, store=store, deadline=deadline, max_cells=args.max_cells)
    report = {"lattice": lat, "replicas": args.replicas, "nbeats": args.nbeats, "grid": len(grid)}
    try:
        if args.mode == "plain":
            res, cost = simulate_points(grid, args.replicas, args.nbeats, lat, tag="plain", **run)
            report["points"] = [dict(ku2=p.ku2, ikr=p.ikr, cl=p.cl, ead=r.ead, ead_lit=r.ead_lit, ead_lo=r.ead_lo,
                                     ead_hi=r.ead_hi,
                                     apd90=r.apd90, apd90_ci=r.apd90_ci) for p, r in res.items()]
            report["cost"] = cost.__dict__
        else:
            ks = [int(v) for v in args.train_every.split(",")]
            ks = ks * 3 if len(ks) == 1 else ks
            sub = [a[::k] + ([a[-1]] if (len(a) - 1) % k else []) for a, k in zip(ax, ks)]
            train = grid_points(*sub)
            sim = lambda pts, r0, r1: simulate_points(pts, r1, args.nbeats, lat, r0=r0, tag="steered", **run)
            sm = steered_map(grid, train, args.replicas, args.nbeats, lat, seed0=args.seed, progress=log,
                             train_replicas=args.train_replicas, store=store, simulate=sim)
            n_sur = sum(1 for s in sm.source.values() if s == "surrogate")
            plain_cells = sum(point_cost(p, args.replicas, args.nbeats) for p in grid)
            report.update(steered_cost=sm.cost.__dict__, plain_cell_steps=plain_cells,
                          cell_step_ratio=sm.cost.cell_steps / plain_cells, n_surrogate=n_sur,
                          n_plain=sum(1 for s in sm.source.values() if s == "plain"),
                          n_train_topped_up=sum(1 for s in sm.source.values() if s == "train+plain"),
                          n_train=len(train), train_replicas=args.train_replicas or args.replicas)
            if args.mode == "holdout":
                hp = holdout_points(grid, args.holdout, args.seed)
                hold, hcost = simulate_points(hp, args.replicas, args.nbeats, lat, tag="holdout", **run)
                report.update(holdout_cost=hcost.__dict__,
                              **holdout_summary(grid, sm, hold, hcost, args.replicas, args.nbeats),
                              points=[dict(ku2=p.ku2, ikr=p.ikr, cl=p.cl, steered=sm.values[p], source=sm.source[p],
                                           predicted=sm.predicted.get(p, (None, None))[0],
                                           plain=hold[p].ead if p in hold else None,
                                           plain_lit=hold[p].ead_lit if p in hold else None) for p in grid])
            if args.mode == "compare":
                # the plain method on every grid point (same seeds), timed the same way
                res, cost = simulate_points(grid, args.replicas, args.nbeats, lat, tag="plain", **run)
                err = np.array([abs(sm.values[p] - res[p].ead) for p in grid])
                sur = np.array([sm.source[p] == "surrogate" for p in grid])
                report.update(plain_cost=cost.__dict__, map_err_max=float(err.max()), map_err_mean=float(err.mean()),
                              map_err_surrogate_max=float(err[sur].max()) if sur.any() else 0.0,
                              map_err_surrogate_mean=float(err[sur].mean()) if sur.any() else 0.0,
                              gpu_time_ratio=(sm.cost.gpu_s + sm.cost.train_s) / cost.gpu_s,
                              points=[dict(ku2=p.ku2, ikr=p.ikr, cl=p.cl, plain=res[p].ead, plain_lo=res[p].ead_lo,
                                           plain_hi=res[p].ead_hi, plain_lit=res[p].ead_lit, steered=sm.values[p],
                                           source=sm.source[p], apd90=res[p].apd90) for p in grid])
    except SweepPaused as e:
        store.save()
        print(json.dumps({"paused": str(e), "store": args.store}), flush=True)
        return 3
    print(json.dumps(report, indent=1), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            json.dump(report, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
