#!/usr/bin/env python3
"""Benchmark: simulated cell-ms per wall-second of the CRU update on one MI355X.

A "step" is one time step (dt = 0.025 ms) of the hot path over the whole
batch: the CRU kernels over every CRU, the on-device whole-cell reduction and
membrane ODE, and (every 100 steps) the output-row reduction.  Inputs live in
HBM before the timed region starts; nothing is copied to the host inside it.
The timed steps are the ones production runs: replays of captured step
graphs (ryr_prepare captures them before the timed region).

Workloads (synthetic, model ICs; no datasets involved):
  cell  : one 100x20x10-CRU cell (20,000 CRUs), RyR multiplier Ku2 = 2.5,
          I_Kr scaled to 0, CL 2000 ms, FP32 production precision, seed 1
  batch : 256 independent cells, lattices drawn from {50x10x10, 75x15x10,
          100x20x10}, Ku2 ~ U[1,3], I_Kr scale ~ U[0,1] (lqt off, so the scale
          applies), CL in {1000,2000,3000}, seed = cell index, same precision
The JSON line reports the `cell` workload as `value`, the batch beside it, and
both workloads again in FP64 (the reference's arithmetic) under `f64`.
"""
import argparse
import json
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "simulated cell-ms/wall-s (1 cell, 256-cell batch); EAD-map GPU-time vs plain"
HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes of one interior CRU-step, the unit of work of the
# roofline: every live state word read once and written once, halo re-reads,
# output-step extras and per-step scalars excluded.  FP32 production: 73 live
# FP32 words read and written (8 ci, 8 cnsr, 6x8 cytosolic buffers, 4
# sarcolemmal buffers, cs, cp, cjsr, cjsrb, Ancx) = 584 B, plus one 8-byte
# packed gating word read and written = 16 B.  FP64: 73 x 8 x 2 + 16.
UNIT_BYTES = {1: 600, 0: 1184}


def stored_bytes_per_cru_step(precision):
    """HBM bytes of one interior CRU-step as this build stores the state:
    FP32 mode keeps all 73 concentration and buffer words in float (the slow
    pools rounded stochastically) and 3 packed 32-bit gating/flag words; FP64
    mode 73 doubles and the 3 packed words.  Each read once and written once."""
    if precision == 1:
        return 2 * (73 * 4 + 3 * 4)
    return 2 * (73 * 8 + 3 * 4)


def kernel_bytes(precision, sched=0):
    """Bytes per interior CRU that each step kernel must touch (scheme A),
    every word counted once per kernel.
    k_cyto (sub-voxel kernel): reads 8 ci, 8 cnsr, 5x8 fast buffers, 8 slow
    troponin, cjsr, cs, cp; writes 8 ci, 8 cnsr, 5x8 fast buffers, 8 slow
    troponin.  FP32: 268 + 256 = 524 B; FP64: 536 + 512 = 1048 B.
    k_cru_step (CRU kernel): reads ci/cnsr of sub-voxels 0-1, cs, cp, cjsr,
    cjsrb, Ancx, 4 sarcolemmal buffers, 3 packed words; writes Ancx, the 4
    buffers, cp, cs, cjsr, cjsrb, the packed words.  FP32: 64 + 48 = 112 B;
    FP64: 116 + 84 = 200 B.
    Schedule 3 (RyR gating in the sub-voxel launch): the sub-voxel launch also
    reads and writes the RyR-count and flag words, +16 B; the CRU kernel no
    longer writes the RyR-count word, -4 B.
    Schedule 4 (one launch per step, k_one): the sum of the two."""
    if precision == 1:
        b = {"k_cyto": 524, "k_cru_step": 112}
    else:
        b = {"k_cyto": 1048, "k_cru_step": 200}
    if sched == 3:
        b["k_cyto"] += 16
        b["k_cru_step"] -= 4
    if sched == 4:
        b = {"k_one": b["k_cyto"] + b["k_cru_step"]}
    return b


SCHED_NAMES = ["sequential", "membrane-fused", "concurrent", "membrane-and-gating-fused", "one-launch"]


def make_cell(prec):
    from ryr_hyperactivity_amd import CellDesc, Params
    p = Params.default(precision=prec)
    d = CellDesc.default(nx=100, ny=20, nz=10, cl=2000.0, Ku2=2.5, ikr_scale=0.0, seed=1)
    return p, [d]


def make_batch(prec, ncells=256, seed=2024):
    """The uneven ensemble's cells (ryr_hyperactivity_amd.ensemble.uneven_cells)."""
    from ryr_hyperactivity_amd import Params
    from ryr_hyperactivity_amd.ensemble import uneven_cells
    return Params.default(precision=prec, lqt=0), uneven_cells(ncells, seed)    # lqt off: the I_Kr scale applies


def window_start(steps, spo):
    """Phase (step mod spo) at which the timed window starts, so that it
    holds round(steps/spo) output steps, the share production carries: a
    window starting on an output step holds ceil(steps/spo), one starting
    just after it floor(steps/spo)."""
    return 0 if round(steps / spo) == math.ceil(steps / spo) else 1


def run_gpu(p, cells, steps, warmup, prof_steps, shape=None, floor=False):
    """Timed region: exactly `steps` steps, all of them graph replays (the
    graphs are captured by ryr_prepare before it), bracketed by a device
    synchronisation on both sides (wall clock) and by HIP events on the
    engine's stream.  Then `prof_steps` more steps in the same schedule with
    every launch stamped by its own start/end events (the per-kernel
    durations of the roofline), and the launch floor of the step graph."""
    from ryr_hyperactivity_amd import CellEngine
    from ryr_hyperactivity_amd._abi import S
    ncru = sum(c.ncru for c in cells)
    spo = p.steps_per_output
    with CellEngine(p, cells, device=0) as eng:
        if shape:
            eng.set_launch_shape(*shape)
        eng.step(warmup)
        extra = 0
        while eng.launch_shape()[2] < 0:       # the shape tuner runs inside the warm-up, never in the timed region
            eng.step(50)
            extra += 50
        done = int(eng.scalars()[0][S["STEP"]])
        align = (window_start(steps, spo) - done) % spo
        eng.step(align)
        eng.prepare(steps)
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
        floor_us = eng.latency_floor(spo) if floor else None
    if shape[2] == 4:
        # one launch per step (k_one) carries the membrane step, the gating,
        # the CRU update and the sweep; it is stamped in the sub-voxel slot
        prof = {("k_one" if k == "k_cyto" else k): v for k, v in prof.items() if k != "k_cru_step"}
    sim_ms = len(cells) * steps * p.dt
    step_s = ev_ms * 1e-3 / max(launches, 1)
    kb = kernel_bytes(p.precision, shape[2])
    kernels = {}
    for name, (us, cnt) in prof.items():
        d = {"avg_us": us, "launches": cnt}
        if name in kb and us > 0:
            d["kernel_bytes_per_launch"] = ncru * kb[name]
            d["kernel_GBs"] = ncru * kb[name] / (us * 1e-6) / 1e9
        kernels[name] = d
    # The dominant kernel is the per-CRU kernel with the longest duration as production runs it (under the
    # concurrent schedule the two overlap; under the fused ones the sub-voxel launch carries the membrane step
    # and the gating).  The unit of work's algorithmic bytes (UNIT_BYTES per CRU-step, the whole step's state)
    # over that kernel's duration is the roofline's achieved rate.
    dom = "k_one" if "k_one" in kernels else max(("k_cyto", "k_cru_step"), key=lambda k: kernels[k]["avg_us"])
    return {
        "cell_ms_per_s": sim_ms / wall,
        "wall_s": wall,
        "ms_per_step": 1e3 * wall / steps,
        "event_ms_per_step": 1e3 * step_s,
        "graph_replayed_steps": steps,
        "window_start_phase": window_start(steps, spo),
        "warmup_total": warmup + extra + align,
        "ncru": ncru,
        "unit_bytes": UNIT_BYTES[p.precision],
        "stored_bytes_per_cru_step": stored_bytes_per_cru_step(p.precision),
        "step_GBs": ncru * UNIT_BYTES[p.precision] / step_s / 1e9,
        "launch_shape": {"k_cyto": shape[0], "k_cru_step": shape[1], "schedule": SCHED_NAMES[shape[2]]},
        "shape": list(shape),
        "kernels": kernels,
        "dominant": dom,
        "latency_floor_us": floor_us,
    }


def roofline(r, workload):
    k = r["dominant"]
    us = r["kernels"][k]["avg_us"]
    alg = r["ncru"] * r["unit_bytes"]
    ach = alg / (us * 1e-6) / 1e9
    return {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "traffic": traffic_from_profiles(workload, r["shape"]),
            "kernel": k, "kernel_us": us, "bytes_per_launch": alg, "unit_bytes": r["unit_bytes"],
            "step_GBs": r["step_GBs"], "step_frac": r["step_GBs"] / HBM_PEAK_GBS,
            "stored_bytes_per_cru_step": r["stored_bytes_per_cru_step"],
            "latency_floor_us_per_step": r["latency_floor_us"],
            "kernels": {n: {"us": v["avg_us"], "launches": v["launches"], "kernel_GBs": v.get("kernel_GBs")}
                        for n, v in r["kernels"].items()},
            "launch_shape": r["launch_shape"]}


def cpu_baseline(p, cell, nsteps, nsteps_1core=None):
    """The FP64 oracle (OpenMP over CRUs) on this host, on a bounded sample:
    all threads over `nsteps` steps from rest through a stimulus at 1 ms (the
    upstroke, the release and the start of the plateau), then one thread over
    the next steps, on the plateau."""
    from oracle.binding import OracleCell, lib
    nthr = lib().orc_num_threads()
    p.stim_offset = 1.0
    o = OracleCell(p, cell)
    t0 = time.perf_counter()
    o.step(nsteps)
    wall = time.perf_counter() - t0
    v_end = float(o.scalars()[1])
    out = {"value": nsteps * p.dt / wall, "unit": "simulated cell-ms/wall-s",
           "cores": nthr, "kind": "port",
           "sample": f"{nsteps} steps ({nsteps * p.dt:g} ms) of one {cell.nx}x{cell.ny}x{cell.nz} cell from rest, "
                     f"stimulated at 1 ms (upstroke, release, plateau; V = {v_end:.1f} mV at the end), FP64 "
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
    out["single_core"] = {"value": n1 * p.dt / w1, "cores": 1,
                          "sample": f"the next {n1} steps (on the plateau), one thread", "wall_s": w1}
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


def traffic_from_profiles(workload, shape):
    """HBM bytes of one step's per-CRU kernels (the sub-voxel sweep and the
    CRU kernel, or the one-launch kernel) on `workload` at launch shape
    `shape` (k_cyto size, k_cru_step size, schedule), from the committed PMC
    summary (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by the factor
    the same counter shows on a copy kernel of known size), or None.  Without
    an entry for the exact shape, an entry with the same schedule and the same
    sweep form (one lane per CRU or per sub-voxel) is taken: the bytes do not
    depend on the workgroup sizes.  It is the measured counterpart of the
    unit's algorithmic bytes, which also cover the whole step."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    wc, wr, sc = shape
    exact = t.get(f"{workload}@{wc},{wr},{sc}")
    if exact:
        return exact["hbm_bytes_per_step"]
    for key, e in t.items():
        if not key.startswith(workload + "@"):
            continue
        kc, kr, ks = (int(x) for x in key.split("@")[1].split(","))
        if ks == sc and (kc == 1024) == (wc == 1024):
            return e["hbm_bytes_per_step"]
    return None


def sub_record(r, workload, steps, warmup):
    return {"value": r["cell_ms_per_s"], "unit": "simulated cell-ms/wall-s", "ncru": r["ncru"], "steps": steps,
            "warmup": warmup, "ms_per_step": r["ms_per_step"], "event_ms_per_step": r["event_ms_per_step"],
            "roofline": roofline(r, workload)}


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
    ap.add_argument("--no-f64", action="store_true", help="skip the FP64 sub-records")
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
        r = run_gpu(p, cells, steps, warmup, args.prof_steps or 400, shape, floor=True)
        print(f"# cell: {json.dumps(r)}", file=sys.stderr, flush=True)
    batch = None
    bw = args.warmup if args.warmup is not None else 100
    if args.workload in ("batch", "both"):
        pb, bcells = make_batch(prec)
        rb = run_gpu(pb, bcells, args.batch_steps, bw, args.prof_steps or 100, shape)
        print(f"# batch: {json.dumps(rb)}", file=sys.stderr, flush=True)
        batch = dict(sub_record(rb, "batch", args.batch_steps, bw), ncells=len(bcells))
        if r is None:
            r, steps, warmup = rb, args.batch_steps, bw
    f64 = None
    if prec == 1 and not args.no_f64:
        f64 = {}
        if args.workload in ("cell", "both"):
            p64, c64 = make_cell(0)
            n64 = min(steps, 2000)
            r64 = run_gpu(p64, c64, n64, warmup, 200, None)
            print(f"# cell f64: {json.dumps(r64)}", file=sys.stderr, flush=True)
            f64["cell"] = sub_record(r64, "cell_f64", n64, warmup)
        if args.workload in ("batch", "both"):
            pb64, bc64 = make_batch(0)
            nb64 = min(args.batch_steps, 300)
            rb64 = run_gpu(pb64, bc64, nb64, bw, 50, None)
            print(f"# batch f64: {json.dumps(rb64)}", file=sys.stderr, flush=True)
            f64["batch"] = dict(sub_record(rb64, "batch_f64", nb64, bw), ncells=len(bc64))
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
                   "ncru": r["ncru"], "graph_replayed_steps": r["graph_replayed_steps"],
                   "window_start_phase": r["window_start_phase"], "warmup_total": r["warmup_total"],
                   "event_ms_per_step": r["event_ms_per_step"]},
        "roofline": roofline(r, wl),
        "cpu_baseline": cpu,
        "batch": batch,
        "f64": f64,
        "ead_map": ead_map_from_profiles(),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
