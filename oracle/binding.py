"""ctypes binding of the FP64 CPU oracle (liboracle_ryr.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
benchmark's CPU-baseline leg, as the reference results are compared with.  The
product never imports it.
`OracleCell` exposes the same reads as ryr_hyperactivity_amd.CellEngine for
one cell so that parity tests compare like with like.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from ryr_hyperactivity_amd._abi import CellDesc, Params, S_COUNT, DIAG_COLS, LINE_COLS, C_COUNT, F, PLANES, S
from ryr_hyperactivity_amd._abi import BEAT_COLS

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle_ryr.so")
_D = ctypes.POINTER(ctypes.c_double)
_I32 = ctypes.POINTER(ctypes.c_int32)
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        d, i, i64, u64, vp = ctypes.c_double, ctypes.c_int, ctypes.c_int64, ctypes.c_uint64, ctypes.c_void_p
        sig = {
            "orc_create": (vp, [ctypes.POINTER(Params), ctypes.POINTER(CellDesc)]),
            "orc_destroy": (None, [vp]),
            "orc_step": (None, [vp, i64]),
            "orc_read_scalars": (i, [vp, _D]),
            "orc_read_field": (i, [vp, i, _D, i64]),
            "orc_write_field": (i, [vp, i, _D, i64]),
            "orc_write_scalar": (i, [vp, i, d]),
            "orc_read_diag": (i, [vp, _D, ctypes.POINTER(i64)]),
            "orc_read_linescan": (i, [vp, _D, i64]),
            "orc_read_counters": (i, [vp, ctypes.POINTER(i64)]),
            "orc_num_threads": (i, []),
            "orc_set_num_threads": (None, [i]),
            "orc_set_vclamp": (i, [vp, _D, i64, i]),
            "orc_read_beats": (i, [vp, _D, i, ctypes.POINTER(i)]),
            "orc_ncx": (d, [d, d, d, _D, d]),
            "orc_ncx_slow": (d, [d, d, d, _D, d]),
            "orc_read_scheme_b_lag": (i, [vp, _D, i64]),
            "orc_uptake": (d, [d, d]),
            "orc_leak": (d, [d, d]),
            "orc_luminal": (d, [d]),
            "orc_luminal_SS": (d, [d]),
            "orc_release": (d, [d, d, d]),
            "orc_lcccurrent": (d, [d, d]),
            "orc_buffer_flux": (d, [i, d, d, d, d]),
            "orc_fracbound": (d, [d]),
            "orc_lcc_markov": (i, [i, d, d, d, d, i]),
            "orc_canonical_sum": (d, [_D, i64]),
            "orc_ryrgating_replay": (i, [ctypes.POINTER(u64), i, d, d, d, d, d, d, _I32, _I32, _I32]),
            "orc_gating_probe": (i, [i64, u64, ctypes.POINTER(ctypes.c_uint32), u64, d, d, d, d, i, _I32, _I32,
                                     _D, _I32, _I32, _I32]),
            "orc_membrane_probe": (i, [_D, d, d, d, d, d, d, d, d, _D]),
            "orc_dm_exp": (d, [d]),
            "orc_dm_log": (d, [d]),
            "orc_dm_cos": (d, [d]),
            "orc_dm_pow": (d, [d, d]),
            "orc_philox": (None, [ctypes.POINTER(ctypes.c_uint32)] * 3),
            "orc_slot_uniform": (d, [ctypes.c_uint32, ctypes.c_uint32]),
            "orc_sr_bits": (ctypes.c_uint32, [u64, u64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32]),
            "orc_mix32": (ctypes.c_uint32, [ctypes.c_uint32]),
            "orc_sr_f32": (ctypes.c_float, [d, ctypes.c_uint32]),
            "orc_stream": (None, [u64, ctypes.c_uint32, u64, i, _D, ctypes.POINTER(ctypes.c_uint32)]),
            "orc_stencil_probe": (i, [ctypes.POINTER(Params), ctypes.POINTER(CellDesc), _D, d, d, _D]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(_D)


class OracleCell:
    """One cell on the CPU oracle (FP64, the reference's layout)."""

    def __init__(self, params: Params, cell: CellDesc):
        self.L = lib()
        self.params = params
        self.cell = cell
        self.h = self.L.orc_create(ctypes.byref(params), ctypes.byref(cell))
        if not self.h:
            raise ValueError("orc_create rejected the parameters")

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n=1):
        self.L.orc_step(self.h, int(n))

    def scalars(self):
        out = np.zeros(S_COUNT)
        self.L.orc_read_scalars(self.h, _dp(out))
        return out

    def field(self, name):
        planes = PLANES[name]
        n = self.cell.nx * self.cell.ny * self.cell.nz
        out = np.zeros(planes * n)
        rc = self.L.orc_read_field(self.h, F[name], _dp(out), out.size)
        assert rc == 0
        return out.reshape(planes, n) if planes > 1 else out

    def scheme_b_lag(self):
        """Per CRU, the last scheme-B step's sum over its sub-steps of
        (step-start cs - sub-step cs)."""
        out = np.zeros(self.cell.nx * self.cell.ny * self.cell.nz)
        assert self.L.orc_read_scheme_b_lag(self.h, _dp(out), out.size) == 0
        return out

    def set_field(self, name, values):
        v = np.ascontiguousarray(np.asarray(values, dtype=np.float64).reshape(-1))
        assert self.L.orc_write_field(self.h, F[name], _dp(v), v.size) == 0

    def set_scalar(self, name, value):
        assert self.L.orc_write_scalar(self.h, S[name], float(value)) == 0

    def diag(self):
        out = np.zeros(DIAG_COLS)
        st = ctypes.c_int64()
        self.L.orc_read_diag(self.h, _dp(out), ctypes.byref(st))
        return out, st.value

    def linescan(self):
        out = np.zeros((self.cell.nx, LINE_COLS))
        self.L.orc_read_linescan(self.h, _dp(out), out.size)
        return out

    def beats(self):
        out = np.zeros((65, BEAT_COLS))
        n = ctypes.c_int()
        self.L.orc_read_beats(self.h, _dp(out), 65, ctypes.byref(n))
        return out[:n.value]

    def set_vclamp(self, v, periodic=False):
        v = np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(-1))
        self.L.orc_set_vclamp(self.h, _dp(v), v.size, int(bool(periodic)))

    def counters(self):
        out = np.zeros(C_COUNT, dtype=np.int64)
        self.L.orc_read_counters(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
        return out


def gating_probe(seed, cru, step, Ku, Ku2, Kb, dt, iso2, counts, lcc, conc):
    L = lib()
    cru = np.ascontiguousarray(cru, dtype=np.uint32)
    n = cru.size
    counts = np.ascontiguousarray(counts, dtype=np.int32).copy()
    lcc = np.ascontiguousarray(lcc, dtype=np.int32).copy()
    conc = np.ascontiguousarray(conc, dtype=np.float64)
    draws, nopen, fails = (np.zeros(n, np.int32) for _ in range(3))
    L.orc_gating_probe(n, int(seed), cru.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), int(step), Ku, Ku2, Kb,
                       dt, int(iso2), counts.ctypes.data_as(_I32), lcc.ctypes.data_as(_I32), _dp(conc),
                       draws.ctypes.data_as(_I32), nopen.ctypes.data_as(_I32), fails.ctypes.data_as(_I32))
    return counts, lcc, draws, nopen, fails


def ryrgating_replay(slots, Ku, Ku2, Kb, cp, cjsr, dt, counts):
    """RyR sampler fed with raw 64-bit slots instead of Philox."""
    L = lib()
    slots = np.ascontiguousarray(slots, dtype=np.uint64)
    c = np.ascontiguousarray(counts, dtype=np.int32).copy()
    used, fails = np.zeros(1, np.int32), np.zeros(1, np.int32)
    L.orc_ryrgating_replay(slots.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), slots.size, Ku, Ku2, Kb, cp, cjsr,
                           dt, c.ctypes.data_as(_I32), used.ctypes.data_as(_I32), fails.ctypes.data_as(_I32))
    return c, int(used[0]), int(fails[0])


def canonical_sum(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().orc_canonical_sum(_dp(x), x.size)


def membrane_probe(state13, sum_icat, sum_inaca, sum_cs, ncru, cm, dt, ikr_scale, stim):
    s = np.ascontiguousarray(state13, dtype=np.float64).copy()
    cur = np.zeros(10)
    lib().orc_membrane_probe(_dp(s), sum_icat, sum_inaca, sum_cs, ncru, cm, dt, ikr_scale, stim, _dp(cur))
    return s, cur


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().orc_philox(c, k, o)
    return list(o)


def stream(seed, cru, step, n):
    u = np.zeros(n)
    w = np.zeros(n, np.uint32)
    lib().orc_stream(int(seed), int(cru), int(step), int(n), _dp(u), w.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
    return u, w


def stencil(params, cell, field, tt, tl):
    f = np.ascontiguousarray(field, dtype=np.float64).reshape(-1)
    out = np.zeros_like(f)
    assert lib().orc_stencil_probe(ctypes.byref(params), ctypes.byref(cell), _dp(f), tt, tl, _dp(out)) == 0
    return out.reshape(8, -1)
