"""Python host mirror of the reference's run interface over the C ABI.

`CellEngine` owns one ryr_handle (one or many independent cells on one GPU).
Its methods map onto the reference's main() loop (cuda_Cell_Detailed.cu:
232-1105): construction = set-up and Init, step() = the time loop body,
diag()/linescan() = the output block.  `run_cell` reproduces the reference's
command line (positional argv plus named flags) and writes the same files as
the C++ CLI `bin/ryr_cell`.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _abi
from ._abi import CellDesc, Params

_D = ctypes.POINTER(ctypes.c_double)


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_D)


class EngineError(RuntimeError):
    pass


class CellEngine:
    """Cells advanced together by the gfx950 kernels."""

    def __init__(self, params: Params, cells, device: int = 0):
        self.lib = _abi.load_library()
        if isinstance(cells, CellDesc):
            cells = [cells]
        self.params = params
        self.cells = list(cells)
        arr = (CellDesc * len(self.cells))(*self.cells)
        h = ctypes.c_void_p()
        rc = self.lib.ryr_create(device, ctypes.byref(params), arr, len(self.cells), ctypes.byref(h))
        if rc != _abi.RYR_OK:
            raise EngineError(f"ryr_create failed ({rc}): {self.lib.ryr_last_error(None).decode()}")
        self.h = h

    # -- lifecycle --------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.lib.ryr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc, what):
        if rc != _abi.RYR_OK:
            raise EngineError(f"{what} failed ({rc}): {self.lib.ryr_last_error(self.h).decode()}")

    # -- stepping ---------------------------------------------------------
    def step(self, nsteps: int = 1):
        self._check(self.lib.ryr_step(self.h, int(nsteps)), "ryr_step")

    def prepare(self, nsteps: int):
        """Capture now the graphs that step(nsteps) called next would replay."""
        self._check(self.lib.ryr_prepare(self.h, int(nsteps)), "ryr_prepare")

    def latency_floor(self, nsteps: int) -> float:
        """Microseconds per step of the production step graph with every
        kernel an empty one of the same shape (the launch floor)."""
        us = ctypes.c_double()
        self._check(self.lib.ryr_latency_floor(self.h, int(nsteps), ctypes.byref(us)), "ryr_latency_floor")
        return us.value

    def sync(self):
        self._check(self.lib.ryr_sync(self.h), "ryr_sync")

    def kernel_time(self):
        ms, n = ctypes.c_double(), ctypes.c_int64()
        self._check(self.lib.ryr_kernel_time(self.h, ctypes.byref(ms), ctypes.byref(n)), "ryr_kernel_time")
        return ms.value, n.value

    def set_launch_shape(self, wg_cyto: int, wg_cru: int, sched: int = 0):
        """Workgroup sizes (64/128/256) of the sub-voxel and CRU kernels (CRU
        kernel also 512: two lanes per CRU) and the step schedule (0 sequential,
        1 membrane fused into the next sub-voxel launch, 2 sub-voxel and CRU
        kernels concurrent, 3 as 1 with the RyR gating in the sub-voxel launch
        too); sizes (0, 0) re-tune."""
        self._check(self.lib.ryr_set_launch_shape(self.h, int(wg_cyto), int(wg_cru), int(sched)),
                    "ryr_set_launch_shape")

    def launch_shape(self):
        """(wg_cyto, wg_cru, sched); (0, 0, -1) while tuning is pending."""
        a, b, f = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.ryr_get_launch_shape(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(f)),
                    "ryr_get_launch_shape")
        return a.value, b.value, f.value

    def profile_kernels(self, nsteps: int):
        """Advance nsteps steps with per-kernel event stamps; returns
        {kernel: (mean_us, launches)} for k_cyto, k_cru_step, k_membrane, k_diag."""
        us = (ctypes.c_double * 4)()
        cnt = (ctypes.c_int64 * 4)()
        self._check(self.lib.ryr_profile_kernels(self.h, int(nsteps), us, cnt), "ryr_profile_kernels")
        names = ("k_cyto", "k_cru_step", "k_membrane", "k_diag")
        return {nm: (us[i], cnt[i]) for i, nm in enumerate(names)}

    def cru_steps(self) -> int:
        n = ctypes.c_int64()
        self._check(self.lib.ryr_cru_steps(self.h, ctypes.byref(n)), "ryr_cru_steps")
        return n.value

    # -- reads ------------------------------------------------------------
    def scalars(self) -> np.ndarray:
        out = np.zeros((len(self.cells), _abi.S_COUNT))
        self._check(self.lib.ryr_read_scalars(self.h, _dp(out)), "ryr_read_scalars")
        return out

    def field(self, cell: int, name: str) -> np.ndarray:
        planes = _abi.PLANES[name]
        n = self.cells[cell].ncru
        out = np.zeros(planes * n)
        self._check(self.lib.ryr_read_field(self.h, cell, _abi.F[name], _dp(out), out.size), "ryr_read_field")
        return out.reshape(planes, n) if planes > 1 else out

    def set_field(self, cell: int, name: str, values):
        planes = _abi.PLANES[name]
        v = np.ascontiguousarray(np.asarray(values, dtype=np.float64).reshape(-1))
        if v.size != planes * self.cells[cell].ncru:
            raise ValueError("wrong number of values")
        self._check(self.lib.ryr_write_field(self.h, cell, _abi.F[name], _dp(v), v.size), "ryr_write_field")

    def set_scalar(self, cell: int, name: str, value: float):
        self._check(self.lib.ryr_write_scalar(self.h, cell, _abi.S[name], float(value)), "ryr_write_scalar")

    def diag(self):
        out = np.zeros((len(self.cells), _abi.DIAG_COLS))
        st = ctypes.c_int64()
        self._check(self.lib.ryr_read_diag(self.h, _dp(out), ctypes.byref(st)), "ryr_read_diag")
        return out, st.value

    def linescan(self, cell: int = 0) -> np.ndarray:
        out = np.zeros((self.cells[cell].nx, _abi.LINE_COLS))
        self._check(self.lib.ryr_read_linescan(self.h, cell, _dp(out), out.size), "ryr_read_linescan")
        return out

    def counters(self) -> np.ndarray:
        out = np.zeros((len(self.cells), _abi.C_COUNT), dtype=np.int64)
        self._check(self.lib.ryr_read_counters(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))),
                    "ryr_read_counters")
        return out

    def beats(self, cell: int = 0) -> np.ndarray:
        """Per-beat summary rows (columns _abi.BEAT_NAMES)."""
        out = np.zeros((_abi.MAX_BEATS + 1, _abi.BEAT_COLS))
        n = ctypes.c_int()
        self._check(self.lib.ryr_read_beats(self.h, cell, _dp(out), _abi.MAX_BEATS + 1, ctypes.byref(n)),
                    "ryr_read_beats")
        return out[:n.value]

    def set_vclamp(self, cell: int, v, periodic: bool = False):
        """Replace V after each step by a waveform (see clamp_waveform)."""
        v = np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(-1))
        self._check(self.lib.ryr_set_vclamp(self.h, cell, _dp(v), v.size, int(bool(periodic))), "ryr_set_vclamp")

    def save(self, path: str):
        self._check(self.lib.ryr_save(self.h, path.encode()), "ryr_save")

    def load(self, path: str):
        self._check(self.lib.ryr_load(self.h, path.encode()), "ryr_load")


def gating_probe(seed, cru, step, Ku, Ku2, Kb, dt, iso2, counts, lcc, conc, device=0):
    """Frozen-input RyR+LCC sampler on the GPU (ryr_gating_probe).

    counts (n,4) nou,ncu,nob,ncb; lcc (n,8) states (16 absent); conc (n,4)
    cp,cjsr,cs,v.  Returns new counts, new lcc, draws, open LCCs, fallbacks."""
    lib = _abi.load_library()
    cru = np.ascontiguousarray(cru, dtype=np.uint32)
    n = cru.size
    counts = np.ascontiguousarray(counts, dtype=np.int32).copy()
    lcc = np.ascontiguousarray(lcc, dtype=np.int32).copy()
    conc = np.ascontiguousarray(conc, dtype=np.float64)
    draws = np.zeros(n, np.int32)
    nopen = np.zeros(n, np.int32)
    fails = np.zeros(n, np.int32)
    I32 = ctypes.POINTER(ctypes.c_int32)
    rc = lib.ryr_gating_probe(device, n, int(seed), cru.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), int(step),
                              float(Ku), float(Ku2), float(Kb), float(dt), int(iso2),
                              counts.ctypes.data_as(I32), lcc.ctypes.data_as(I32), _dp(conc),
                              draws.ctypes.data_as(I32), nopen.ctypes.data_as(I32), fails.ctypes.data_as(I32))
    if rc != _abi.RYR_OK:
        raise EngineError(f"ryr_gating_probe failed ({rc}): {lib.ryr_last_error(None).decode()}")
    return counts, lcc, draws, nopen, fails


def clamp_waveform(name: str, nsteps: int, dt: float = 0.025, clamp_voltage: float = 0.0):
    """V after step n (n = 0..nsteps-1) for the reference's clamp protocols,
    with t = n*dt the time at the start of the step:
      vclamp        -86 mV, then clamp_voltage (argv[3]) once t > 100 ms (:733-736)
      brownvclamp   -86 / -10 / 10 / -30 / -86 mV with edges at 4000, 4200,
                    6200, 7200 ms (:741-752)
      linearapclamp piecewise-linear AP from 100 ms (:758-773)
    apclamp (:754-757) is a periodic waveform read from a file: pass the
    array itself to CellEngine.set_vclamp(..., periodic=True)."""
    t = np.arange(nsteps) * dt
    if name == "vclamp":
        return np.where(t > 100, clamp_voltage, -86.0)
    if name == "brownvclamp":
        return np.select([t < 4000, t < 4200, t < 6200, t < 7200], [-86.0, -10.0, 10.0, -30.0], -86.0)
    if name == "linearapclamp":
        tt = t - 100
        return np.select([t < 100, t < 101.865, t < 119.1, t < 138.9, t < 364.43, t < 395.43],
                         [-86.0, 68 * tt - 80, -tt + 48, 27 + 0.1 * tt, 43 - tt / 3.2, -1.5 * tt + 357], -86.0)
    raise ValueError(f"unknown clamp protocol {name}")


# -- output files (cuda_Cell_Detailed.cu:974-997, :1041-1072) --------------
def format_wholetest_row(r) -> str:
    i = [int(r[k]) for k in range(45)]
    f = list(r)
    return ("%f %f %f %f %f  %f %f %f %f %f  "
            "%f %f %f %f %f  %f %f %f %f %f  "
            "%f %f %f %f %f  %f %i %i %i %i  "
            "%i %i %i %i %i  %i %i %f %f %f  "
            "%i %f %f %f %f\n") % (
        f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7], f[8], f[9],
        f[10], f[11], f[12], f[13], f[14], f[15], f[16], f[17], f[18], f[19],
        f[20], f[21], f[22], f[23], f[24], f[25], i[26], i[27], i[28], i[29],
        i[30], i[31], i[32], i[33], i[34], i[35], i[36], f[37], f[38], f[39],
        i[40], f[41], f[42], f[43], f[44])


def format_linescan(rows) -> str:
    out = []
    for L in rows:
        out.append(("%f %f %f %i %i %i %i %i %f %f %f %f %f %f %f %f %f %f %f %f %f %f\n") % (
            L[0], L[1], L[2], int(L[3]), int(L[4]), int(L[5]), int(L[6]), int(L[7]), L[8], L[9], L[10],
            L[11], L[12], L[13], L[14], L[15], L[16], L[17], L[18], L[19], L[20], L[21]))
    out.append("\n")
    return "".join(out)


def parse_argv(argv):
    """Reference positional argv (run.sh:9) + named flags -> (Params, CellDesc, nsteps, name, device)."""
    if len(argv) < 10:
        raise ValueError("need <Device> <name> <a3> <cjsr0> <ncx> <Ku> <r1> <Kb> <ica> <Ku2>")
    p = Params.default()
    d = CellDesc.default()
    device = max(0, int(argv[0]))
    name = argv[1]
    d.cjsr0 = float(argv[3]); d.ncx_scale = float(argv[4]); d.Ku = float(argv[5])
    d.Kb = float(argv[7]); d.ica_scale = float(argv[8]); d.Ku2 = float(argv[9])
    nbeat, tail, nsteps = 7, 2100.0, None
    it = iter(argv[10:])
    for f in it:
        if f == "--lattice":
            d.nx, d.ny, d.nz = (int(x) for x in next(it).split("x"))
        elif f == "--cl": d.cl = float(next(it))
        elif f == "--nbeat": nbeat = int(next(it))
        elif f == "--tail": tail = float(next(it))
        elif f == "--nsteps": nsteps = int(next(it))
        elif f == "--dt": p.dt = float(next(it))
        elif f == "--stim-dur": p.stim_dur = float(next(it))
        elif f == "--stim-amp": p.stim_amp = float(next(it))
        elif f == "--stim-offset": p.stim_offset = float(next(it))
        elif f == "--seed": d.seed = int(next(it))
        elif f == "--ikr-scale": d.ikr_scale = float(next(it))
        elif f == "--lqt": p.lqt = 1
        elif f == "--no-lqt": p.lqt = 0
        elif f == "--iso2": p.iso2 = 1
        elif f == "--no-iso2": p.iso2 = 0
        elif f == "--tapsi": p.tapsi = 1
        elif f == "--perm":
            p.variants |= _abi.VAR_PERM
            p.ci0, d.cjsr0 = 0.3, 900.0
        elif f == "--iso": p.variants |= _abi.VAR_ISO
        elif f == "--slowncxchange": p.variants |= _abi.VAR_SLOWNCX
        elif f == "--randomlcc": p.variants |= _abi.VAR_RANDOMLCC
        elif f == "--nopacing": p.stim_amp = 0.0
        elif f == "--precision": p.precision = _abi.PREC_F32 if next(it) == "f32" else _abi.PREC_F64
        elif f == "--scheme": p.scheme = _abi.SCHEME_B if next(it).upper() == "B" else _abi.SCHEME_A
        elif f == "--ic":
            if next(it) == "baseline":
                p.ci0, p.v0, d.cjsr0 = 0.1, -86.0, 750.0
        elif f == "--cm": p.cm = float(next(it))
        elif f == "--cm-fixed": p.cm_mode = 1
        else:
            raise ValueError(f"unknown flag {f}")
    if nsteps is None:
        nsteps = int((nbeat * d.cl + tail) / p.dt + 0.5)
    return p, d, nsteps, name, device


def run_cell(argv, outdir="."):
    """Python twin of bin/ryr_cell: run one cell and write the trace files."""
    p, d, nsteps, name, device = parse_argv(argv)
    spo = p.steps_per_output
    with CellEngine(p, d, device) as eng, \
            open(os.path.join(outdir, "wholetest_" + name), "w") as fw, \
            open(os.path.join(outdir, "linescan_" + name), "w") as fl, \
            open(os.path.join(outdir, "cadist_" + name), "w"):
        done = 0
        while done < nsteps:
            chunk = min(spo - done % spo, nsteps - done)
            has_row = done % spo == 0
            eng.step(chunk)
            done += chunk
            if has_row:
                rows, _ = eng.diag()
                fw.write(format_wholetest_row(rows[0]))
                fl.write(format_linescan(eng.linescan(0)))
