"""Error of the FP32 production mode against the FP64 mode, per output.

Both modes run the same cells on the same Philox seeds.  They draw the same
uniforms, so gating decisions agree until one comparison of a uniform with a
rate falls on the other side of the threshold: from that step on the two runs
are different stochastic realisations.  The report therefore gives, per
output-row column (the wholetest columns), the relative RMS difference before
the first divergence of the channel counts, the time of that divergence, and
the per-beat statistics of both modes (mean and 95% interval over the seeds),
which is where FP32 has to agree after divergence.
"""
from __future__ import annotations

import argparse
import json
import math

import numpy as np

from ._abi import DIAG_NAMES, PREC_F32, PREC_F64, CellDesc, Params
from .engine import CellEngine

# columns that are integer counts of channel states: the first row where any
# differs marks the divergence of the gating sequences
COUNT_COLS = [DIAG_NAMES.index(c) for c in ("nou", "nob", "ncu", "ncb", "tn1", "tn2", "tn3", "tn4", "tn7")]
TRACE_COLS = ["v", "ci", "cp", "cs", "cjsr", "cnsr", "i_cal", "i_ncx", "po", "nai", "catf", "cats", "cacal",
              "csqn_bound", "ancx", "total_ca"]


def run_traces(precision: int, cells, nrows: int, lqt: int = 1):
    p = Params.default(precision=precision, lqt=lqt)
    rows = np.zeros((len(cells), nrows, len(DIAG_NAMES)))
    with CellEngine(p, cells) as eng:
        eng.set_launch_shape(256, 256, 0)          # the shape changes no value; skip the tuner's steps
        for r in range(nrows):
            eng.step(p.steps_per_output if r else 1)
            d, st = eng.diag()
            rows[:, r] = d
            if r % 1000 == 0:
                print(f"precision {precision}: row {r} of {nrows}", flush=True)
        beats = [eng.beats(c) for c in range(len(cells))]
    return rows, beats


def beat_stats(beats, nbeats):
    out = {}
    for k, name in ((1, "apd90"), (6, "peak_ci"), (7, "jsr_load"), (4, "ead")):
        per = [np.mean(b[1:nbeats, k]) for b in beats if len(b) > 1]
        m = float(np.mean(per))
        h = float(1.96 * np.std(per, ddof=1) / math.sqrt(len(per))) if len(per) > 1 else float("nan")
        out[name] = (m, h)
    return out


def compare(lattice=(100, 20, 10), ku2s=(1.0, 2.5), seeds=range(1, 9), cl=2000.0, nbeats=10, ikr=0.0):
    cells = [CellDesc.default(nx=lattice[0], ny=lattice[1], nz=lattice[2], Ku2=k, ikr_scale=ikr, cl=cl, seed=s)
             for k in ku2s for s in seeds]
    dt, spo = 0.025, 100
    nrows = int(round((100.0 + nbeats * cl) / (dt * spo))) + 2
    r32, b32 = run_traces(PREC_F32, cells, nrows)
    r64, b64 = run_traces(PREC_F64, cells, nrows)
    res = {"lattice": lattice, "ku2": list(ku2s), "seeds": list(seeds), "cl": cl, "nbeats": nbeats, "rows": nrows,
           "runs": []}
    for c in range(len(cells)):
        diff = np.any(r32[c][:, COUNT_COLS] != r64[c][:, COUNT_COLS], axis=1)
        div = int(np.argmax(diff)) if diff.any() else nrows
        run = {"ku2": cells[c].Ku2, "seed": cells[c].seed, "diverge_row": div,
               "diverge_ms": float(r64[c][min(div, nrows - 1), 0]), "rel_rms_pre": {}}
        for name in TRACE_COLS:
            j = DIAG_NAMES.index(name)
            a, b = r32[c][1:div, j], r64[c][1:div, j]
            if len(a) == 0:
                continue
            scale = max(np.sqrt(np.mean(b * b)), 1e-30)
            run["rel_rms_pre"][name] = float(np.sqrt(np.mean((a - b) ** 2)) / scale)
        res["runs"].append(run)
    for k in ku2s:
        sel = [i for i, cd in enumerate(cells) if cd.Ku2 == k]
        res[f"beats_ku2_{k:g}"] = {"fp32": beat_stats([b32[i] for i in sel], nbeats),
                                   "fp64": beat_stats([b64[i] for i in sel], nbeats)}
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(description="FP32 production mode against the FP64 mode, per output")
    ap.add_argument("--lattice", default="100x20x10")
    ap.add_argument("--nbeats", type=int, default=10)
    ap.add_argument("--cl", type=float, default=2000.0)
    ap.add_argument("--seeds", type=int, default=8)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    lat = tuple(int(v) for v in a.lattice.split("x"))
    res = compare(lat, seeds=range(1, a.seeds + 1), cl=a.cl, nbeats=a.nbeats)
    txt = json.dumps(res, indent=1)
    print(txt, flush=True)
    if a.out:
        with open(a.out, "w") as f:
            f.write(txt)


if __name__ == "__main__":
    main()
