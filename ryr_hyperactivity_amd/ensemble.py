"""The uneven ensemble: independent cells of three lattice sizes, each with its
own RyR open-rate multiplier, I_Kr scale, cycle length and Philox seed, in
one handle, run until every cell has completed a number of beats.  The
reference runs one cell per process (cuda_Cell_Detailed.cu:49-92, its argv);
here the cells share every launch.

    python -m ryr_hyperactivity_amd.ensemble [--ncells 256] [--nbeats 5] [--precision f32|f64] [--out file.json]
"""
import argparse
import json
import time

import numpy as np

from ._abi import BEAT_NAMES, PREC_F32, PREC_F64, CellDesc, Params
from .engine import CellEngine

LATTICES = [(50, 10, 10), (75, 15, 10), (100, 20, 10)]
CYCLE_LENGTHS = [1000.0, 2000.0, 3000.0]
B_APD, B_EAD, B_DONE = BEAT_NAMES.index("apd90"), BEAT_NAMES.index("ead"), BEAT_NAMES.index("complete")


def uneven_cells(ncells: int = 256, seed: int = 2024):
    """ncells cells: lattice drawn from LATTICES, Ku2 ~ U[1, 3], I_Kr scale
    ~ U[0, 1], cycle length drawn from CYCLE_LENGTHS, Philox seed = cell
    index (the generator draws in that order, cell by cell)."""
    rng = np.random.default_rng(seed)
    cells = []
    for c in range(ncells):
        nx, ny, nz = LATTICES[rng.integers(0, 3)]
        cells.append(CellDesc.default(nx=nx, ny=ny, nz=nz, Ku2=float(rng.uniform(1, 3)),
                                      ikr_scale=float(rng.uniform(0, 1)),
                                      cl=float(CYCLE_LENGTHS[rng.integers(0, 3)]), seed=c))
    return cells


def ensemble_steps(cells, nbeats: int, p: Params) -> int:
    """Steps after which every cell has completed nbeats beats: the stimulus
    offset, nbeats of the longest cycle length, and two steps for the last
    beat's close."""
    return int(round((p.stim_offset + nbeats * max(c.cl for c in cells)) / p.dt)) + 2


def summarize(cells, beats, nbeats: int):
    """Per cycle length and I_Kr third: cells, the fewest completed beats, and
    over beats 2..nbeats of each cell (the first beat follows the initial
    conditions and is not scored) the EAD incidence, the mean APD90 of the
    beats that repolarised to 90% and the count of those that did not (the
    beat tracker's APD90 of -1)."""
    groups = {}
    for c, b in zip(cells, beats):
        done = b[b[:, B_DONE] > 0]
        scored = done[1:nbeats]
        key = (c.cl, min(int(c.ikr_scale * 3), 2))
        g = groups.setdefault(key, {"cells": 0, "min_beats": None, "apd90": [], "ead": [], "unrepolarised": 0})
        g["cells"] += 1
        g["min_beats"] = len(done) if g["min_beats"] is None else min(g["min_beats"], len(done))
        ok = scored[:, B_APD] > 0
        g["apd90"].extend(scored[ok, B_APD].tolist())
        g["unrepolarised"] += int((~ok).sum())
        g["ead"].extend((scored[:, B_EAD] > 0).astype(float).tolist())
    out = []
    for (cl, third), g in sorted(groups.items()):
        out.append({"cl": cl, "ikr_third": third, "cells": g["cells"], "min_beats": g["min_beats"],
                    "scored_beats": len(g["ead"]), "unrepolarised": g["unrepolarised"],
                    "apd90_mean": float(np.mean(g["apd90"])) if g["apd90"] else float("nan"),
                    "ead_incidence": float(np.mean(g["ead"])) if g["ead"] else float("nan")})
    return out


def run(cells, nbeats: int, precision: int = PREC_F32, chunk_steps: int = 20000, progress=None):
    """All cells in one handle until each has completed nbeats beats; the
    beats per cell, the per-group summary and the cost."""
    p = Params.default(precision=precision, lqt=0)         # lqt off: each cell's I_Kr scale applies
    nsteps = ensemble_steps(cells, nbeats, p)
    with CellEngine(p, cells) as eng:
        eng.kernel_time()
        t0 = time.perf_counter()
        done = 0
        while done < nsteps:
            n = min(chunk_steps, nsteps - done)
            eng.step(n)
            done += n
            if progress:
                progress(f"ensemble: {done}/{nsteps} steps, {len(cells)} cells")
        eng.sync()
        wall = time.perf_counter() - t0
        ms, launches = eng.kernel_time()
        beats = [eng.beats(i) for i in range(len(cells))]
        shape = eng.launch_shape()
    ncru = sum(c.ncru for c in cells)
    cost = {"steps": nsteps, "cells": len(cells), "crus": ncru, "wall_s": wall, "gpu_s": ms * 1e-3,
            "launches": launches, "cell_ms_per_wall_s": len(cells) * nsteps * p.dt / wall,
            "us_per_step": wall / nsteps * 1e6, "launch_shape": list(shape)}
    return beats, summarize(cells, beats, nbeats), cost


def main(argv=None):
    ap = argparse.ArgumentParser(description="the uneven ensemble, run to a number of completed beats per cell")
    ap.add_argument("--ncells", type=int, default=256)
    ap.add_argument("--nbeats", type=int, default=5)
    ap.add_argument("--seed", type=int, default=2024, help="the generator of the cells' parameters")
    ap.add_argument("--precision", choices=["f32", "f64"], default="f32")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    cells = uneven_cells(a.ncells, a.seed)
    prec = PREC_F32 if a.precision == "f32" else PREC_F64
    beats, groups, cost = run(cells, a.nbeats, prec, progress=lambda s: print(s, flush=True))
    res = {"ncells": a.ncells, "nbeats": a.nbeats, "seed": a.seed, "precision": a.precision, "cost": cost,
           "groups": groups,
           "cells": [{"nx": c.nx, "ny": c.ny, "nz": c.nz, "ku2": c.Ku2, "ikr": c.ikr_scale, "cl": c.cl, "seed": c.seed,
                      "beats": b[b[:, B_DONE] > 0].tolist()} for c, b in zip(cells, beats)]}
    txt = json.dumps(res)
    if a.out:
        with open(a.out, "w") as f:
            f.write(txt)
    print(json.dumps({"cost": cost, "groups": groups}, indent=1), flush=True)


if __name__ == "__main__":
    main()
