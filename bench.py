#!/usr/bin/env python3
"""Benchmark: simulated cell-ms per wall-second of the CRU update on one MI355X.

A "step" is one time step (dt = 0.025 ms) of the hot path over the whole
batch: the CRU kernel over every CRU, the in-kernel whole-cell reduction and
membrane ODE, and (every 100 steps) the output-row reduction.  Inputs live in
HBM before the timed region starts; nothing is copied to the host inside it.

Workloads (synthetic, model ICs; no datasets involved):
  cell  : one 100x20x10-CRU cell (20,000 CRUs), RyR multiplier Ku2 = 2.5,
          I_Kr scaled to 0, CL 2000 ms, FP32 production precision, seed 1
  batch : 256 independent cells, lattices drawn from {50x10x10, 75x15x10,
          100x20x10}, Ku2 ~ U[1,3], I_Kr scale ~ U[0,1] (lqt off, so the scale
          applies), CL in {1000,2000,3000}, seed = cell index, same precision
The JSON line reports the `cell` workload as `value` and the batch beside it.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "simulated cell-ms/wall-s (1 cell, 256-cell batch); EAD-map GPU-time vs plain"
HBM_PEAK_GBS = 8000.0


def alg_bytes_per_cru_step(precision):
    """Algorithmic HBM bytes of one interior CRU-step of the whole step: every
    live state word read once and written once (neighbour re-reads, ghost
    cells and the output-step extras excluded).  FP32 mode: 55 float words
    (8 ci, 5x8 fast buffers, cs, cp, Ancx, 4 sarcolemmal buffers) + 18 double
    words (8 cnsr, 8 slow troponin, cjsr, CSQN-bound) + 3 packed 32-bit
    gating/flag words.  FP64 mode: 73 double words + the 3 packed words."""
    if precision == 1:
        return 2 * (55 * 4 + 18 * 8 + 3 * 4)
    return 2 * (73 * 8 + 3 * 4)


def alg_bytes_per_kernel(precision, sched=0):
    """Algorithmic HBM bytes per interior CRU of each step kernel (scheme A),
    every word the kernel must touch counted once.
    k_cyto (sub-voxel kernel): reads 8 ci, 8 cnsr, 5x8 fast buffers, 8 slow
    troponin, cjsr, cs, cp; writes 8 ci, 8 cnsr, 5x8 fast buffers, 8 slow
    troponin.  FP32: 336 + 320 = 656 B; FP64: 536 + 512 = 1048 B.
    k_cru_step (CRU kernel): reads ci/cnsr of sub-voxels 0-1, cs, cp, cjsr,
    cjsrb, Ancx, 4 sarcolemmal buffers, 3 packed words; writes Ancx, the 4
    buffers, cp, cs, cjsr, cjsrb, the packed words.  FP32: 80 + 56 = 136 B;
    FP64: 116 + 84 = 200 B.  The two add up to the whole-step figure plus the
    40 (FP64: 56) bytes both kernels read.
    k_membrane: per cell, not per CRU (a few hundred bytes): 0 here.
    Schedule 3 (RyR gating in the sub-voxel launch): the sub-voxel launch also
    reads and writes the RyR-count and flag words (cp and cjsr it reads
    anyway), +16 B; the CRU kernel no longer writes the RyR-count word, -4 B."""
    if precision == 1:
        b = {"k_cyto": 656, "k_cru_step": 136, "k_membrane": 0}
    else:
        b = {"k_cyto": 1048, "k_cru_step": 200, "k_membrane": 0}
    if sched == 3:
        b["k_cyto"] += 16
        b["k_cru_step"] -= 4
    return b


def make_cell(prec):
    from ryr_hyperactivity_amd import CellDesc, Params
    p = Params.default(precision=prec)
    d = CellDesc.default(nx=100, ny=20, nz=10, cl=2000.0, Ku2=2.5, ikr_scale=0.0, seed=1)
    return p, [d]


def make_batch(prec, ncells=256, seed=2024):
    from ryr_hyperactivity_amd import CellDesc, Params
    rng = np.random.default_rng(seed)
    dims = [(50, 10, 10), (75, 15, 10), (100, 20, 10)]
    cells = []
    for c in range(ncells):
        nx, ny, nz = dims[rng.integers(0, 3)]
        cells.append(CellDesc.default(nx=nx, ny=ny, nz=nz, Ku2=float(rng.uniform(1, 3)),
                                      ikr_scale=float(rng.uniform(0, 1)),
                                      cl=float([1000.0, 2000.0, 3000.0][rng.integers(0, 3)]), seed=c))
    return Params.default(precision=prec, lqt=0), cells    # lqt off: the I_Kr scale applies


def run_gpu(p, cells, steps, warmup, prof_steps, shape=None):
    """Timed region: `steps` graph-replayed steps, bracketed by a device
    synchronisation on both sides (wall clock) and by HIP events on the
    engine's stream.  Then `prof_steps` more steps with every kernel stamped
    by its own start/end events (the per-kernel durations of the roofline)."""
    from ryr_hyperactivity_amd import CellEngine
    ncru = sum(c.ncru for c in cells)
    with CellEngine(p, cells, device=0) as eng:
        if shape:
            eng.set_launch_shape(*shape)
        eng.step(warmup)
        while eng.launch_shape()[2] < 0:       # the shape tuner runs inside the warm-up, never in the timed region
            eng.step(50)
        eng.sync()
        eng.kernel_time()                      # reset the event window
        t0 = time.perf_counter()
        eng.step(steps)
        eng.sync()
        wall = time.perf_counter() - t0
        ev_ms, launches = eng.kernel_time()
        v = eng.scalars()[:, 1]
        assert np.all(np.isfinite(v)), "non-finite membrane potential"
        shape = eng.launch_shape()
        prof = eng.profile_kernels(prof_steps)
    sim_ms = len(cells) * steps * p.dt
    step_s = ev_ms * 1e-3 / max(launches, 1)
    bpc = alg_bytes_per_kernel(p.precision, shape[2])
    kernels = {}
    for name, (us, cnt) in prof.items():
        d = {"avg_us": us, "launches": cnt}
        if name in bpc and bpc[name] and us > 0:
            d["bytes_per_launch"] = ncru * bpc[name]
            d["achieved_GBs"] = ncru * bpc[name] / (us * 1e-6) / 1e9
        kernels[name] = d
    # dominant kernel for the HBM roofline: of the kernels that take at least a quarter of the summed kernel
    # time, the one that moves the most bytes.  k_membrane moves no per-CRU bytes (one workgroup per cell,
    # latency-bound; in the fused schedules not a launch of its own) and k_diag runs once per output row.  On
    # the single cell the sub-voxel and CRU kernels take about the same time, and a rule by time alone would
    # flip between them from run to run.
    hbm = {k: v for k, v in kernels.items() if v.get("bytes_per_launch")}
    tot = sum(v["avg_us"] for v in hbm.values())
    big = {k: v for k, v in hbm.items() if v["avg_us"] >= 0.25 * tot}
    dom = max(big, key=lambda k: big[k]["bytes_per_launch"])
    return {
        "cell_ms_per_s": sim_ms / wall,
        "wall_s": wall,
        "ms_per_step": 1e3 * wall / steps,
        "event_ms_per_step": 1e3 * step_s,
        "ncru": ncru,
        "step_bytes": ncru * alg_bytes_per_cru_step(p.precision),
        "step_GBs": ncru * alg_bytes_per_cru_step(p.precision) / step_s / 1e9,
        "launch_shape": {"k_cyto": shape[0], "k_cru_step": shape[1], "schedule": ["sequential", "membrane-fused", "concurrent", "membrane-and-gating-fused"][shape[2]]},
        "kernels": kernels,
        "dominant": dom,
    }


def roofline(r, workload):
    k = r["dominant"]
    kd = r["kernels"][k]
    return {"bound": "hbm", "achieved": kd.get("achieved_GBs"), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": (kd.get("achieved_GBs") or 0.0) / HBM_PEAK_GBS,
            "traffic": traffic_from_profiles(k, workload),
            "kernel": k, "kernel_us": kd["avg_us"], "bytes_per_launch": kd.get("bytes_per_launch"),
            "step_GBs": r["step_GBs"], "step_frac": r["step_GBs"] / HBM_PEAK_GBS,
            "kernels": {n: {"us": v["avg_us"], "GBs": v.get("achieved_GBs")} for n, v in r["kernels"].items()},
            "launch_shape": r["launch_shape"]}


def cpu_baseline(p, cell, nsteps, nsteps_1core=None):
    """The FP64 oracle (OpenMP over CRUs) on this host, on a bounded sample:
    all threads, then (a smaller sample) one thread."""
    from oracle.binding import OracleCell, lib
    nthr = lib().orc_num_threads()
    o = OracleCell(p, cell)
    o.step(10)
    t0 = time.perf_counter()
    o.step(nsteps)
    wall = time.perf_counter() - t0
    out = {"value": nsteps * p.dt / wall, "unit": "simulated cell-ms/wall-s",
           "cores": nthr, "kind": "port",
           "sample": f"{nsteps} steps ({nsteps * p.dt:g} ms) of one {cell.nx}x{cell.ny}x{cell.nz} cell, FP64 "
                     f"oracle, {nthr} OpenMP threads, {platform.processor() or 'x86_64'}",
           "wall_s": wall}
    n1 = nsteps_1core if nsteps_1core is not None else max(1, nsteps // 16)
    lib().orc_set_num_threads(1)
    try:
        t0 = time.perf_counter()
        o.step(n1)
        w1 = time.perf_counter() - t0
    finally:
        lib().orc_set_num_threads(nthr)
    out["single_core"] = {"value": n1 * p.dt / w1, "cores": 1, "sample": f"the next {n1} steps, one thread",
                          "wall_s": w1}
    return out


def ead_map_from_profiles():
    """The metric's third part, the steered EAD map's GPU time against the
    plain map's, is a sweep of minutes, not a bench line: it is reported
    from its committed run (profiles/ead_map.json), marked as such."""
    path = os.path.join(ROOT, "profiles", "ead_map.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    d["measured_in_this_run"] = False
    return d


def traffic_from_profiles(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed PMC
    summary (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by the factor
    the same counter shows on a copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    return t.get(f"{kernel}/{workload}", {}).get("hbm_bytes_per_launch")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed time steps of the 1-cell workload (default 8000)")
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--precision", choices=["f32", "f64"], default="f32")
    ap.add_argument("--workload", choices=["cell", "batch", "both"], default="both")
    ap.add_argument("--batch-steps", type=int, default=1200)
    ap.add_argument("--prof-steps", type=int, default=None,
                    help="steps with per-kernel event stamps after the timed region (default 400 cell / 100 batch)")
    ap.add_argument("--cpu-steps", type=int, default=4000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--shape", default=None,
                    help="fix the launch shape WG_CYTO,WG_CRU,SCHED instead of tuning it (e.g. 256,64,2)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (independent cells shard trivially across GPUs; "
              "no multi-GPU path is built). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)
    prec = 1 if args.precision == "f32" else 0
    steps = args.steps if args.steps is not None else 8000
    warmup = args.warmup if args.warmup is not None else 400

    shape = tuple(int(x) for x in args.shape.split(",")) if args.shape else None
    p, cells = make_cell(prec)
    r = None
    if args.workload in ("cell", "both"):
        r = run_gpu(p, cells, steps, warmup, args.prof_steps or 400, shape)
        print(f"# cell: {json.dumps(r)}", file=sys.stderr, flush=True)
    batch = None
    if args.workload in ("batch", "both"):
        pb, bcells = make_batch(prec)
        bw = args.warmup if args.warmup is not None else 100
        rb = run_gpu(pb, bcells, args.batch_steps, bw, args.prof_steps or 100, shape)
        print(f"# batch: {json.dumps(rb)}", file=sys.stderr, flush=True)
        batch = {"value": rb["cell_ms_per_s"], "unit": "simulated cell-ms/wall-s", "ncells": len(bcells),
                 "ncru": rb["ncru"], "steps": args.batch_steps, "warmup": bw, "ms_per_step": rb["ms_per_step"],
                 "roofline": roofline(rb, "batch")}
        if r is None:
            r, steps, warmup = rb, args.batch_steps, bw
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(make_cell(0)[0], cells[0], args.cpu_steps)
        print(f"# cpu: {cpu}", file=sys.stderr, flush=True)
    wl = "cell" if args.workload != "batch" else "batch"
    out = {
        "metric": METRIC,
        "value": r["cell_ms_per_s"],
        "unit": "simulated cell-ms/wall-s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if prec == 1 else "f64",
        "data": "synthetic (model initial conditions, Philox seeds)",
        "config": {"workload": "1 cell 100x20x10 CRUs, Ku2=2.5, I_Kr=0, CL 2000 ms" if wl == "cell"
                   else "256-cell uneven batch", "dt_ms": p.dt, "precision": args.precision,
                   "ncru": r["ncru"]},
        "roofline": roofline(r, wl),
        "cpu_baseline": cpu,
        "batch": batch,
        "ead_map": ead_map_from_profiles(),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
