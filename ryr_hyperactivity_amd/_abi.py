"""ctypes mirror of include/ryr_cru.h and the loader of libryr_cru.so.

The structs and enums here are field-for-field copies of the C header; the
test suite checks the defaults against ryr_params_default/ryr_cell_desc_default
of the built library.  Loading never falls back to anything: if the HIP
library is missing the import of the engine fails with a message that says how
to build it.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libryr_cru.so")

ABI_VERSION = 4
PROBE_NOALLOC_MAX = 16384      # RYR_PROBE_NOALLOC_MAX
RYR_OK = 0
RYR_E_ARG = -1
RYR_E_HIP = -2
RYR_E_NODEV = -3
RYR_E_STATE = -4

SCHEME_A, SCHEME_B = 0, 1
VAR_PERM, VAR_ISO, VAR_SLOWNCX, VAR_RANDOMLCC = 1, 2, 4, 8
PREC_F64, PREC_F32 = 0, 1

DIAG_COLS = 45
LINE_COLS = 22

SCALARS = ["T", "V", "XM", "XH", "XJ", "XR", "XS1", "XS2", "QKS", "XTOS", "YTOS", "XTOF", "YTOF",
           "XNAI", "CSUBT", "ICAL", "INCX", "INA", "IKS", "IKR", "IK1", "INAK", "ITOS", "ITOF",
           "STIM", "OUT_ICA", "OUT_NCX", "OUT_INA", "STEP"]
S = {name: i for i, name in enumerate(SCALARS)}
S_COUNT = len(SCALARS)

FIELDS = ["CI", "CNSR", "CATF", "CATS", "CACAL", "CASR", "CAMYO", "MGMYO", "CS", "CP", "CJSR",
          "CJSRB", "ANCX", "CASAR", "CASARH", "CASARJ", "CASARHJ", "NOU", "NCU", "NOB", "NCB",
          "LCC", "NL", "NSPARK"]
F = {name: i for i, name in enumerate(FIELDS)}
PLANES = {name: (8 if name in ("CI", "CNSR", "CATF", "CATS", "CACAL", "CASR", "CAMYO", "MGMYO", "LCC")
                 else 1) for name in FIELDS}

BEAT_COLS = 12
MAX_BEATS = 64
BEAT_NAMES = ["onset", "apd90", "vmax", "vrest", "ead", "ead_t", "peak_ci", "jsr_load", "sparks", "complete",
              "ead_lit", "ead_lit_t"]

COUNTERS = ["GAUSS_FALLBACK", "CLAMP_CP", "CLAMP_CS", "CI_RUNAWAY", "NONFINITE_V"]
C_COUNT = len(COUNTERS)

# wholetest_<name> columns (cuda_Cell_Detailed.cu:1041-1072)
DIAG_NAMES = ["t", "ci", "v", "i_ncx", "i_cal", "cp", "cs", "cjsr", "cnsr", "po", "nai", "iks",
              "ikr", "ik1", "inak", "itos", "itof", "ina", "xr", "ncx_fwd", "ncb_frac", "cacal",
              "catf", "leak", "uptake", "jrel", "nou", "nob", "ncu", "ncb", "tn1", "tn2", "tn3",
              "tn4", "tn5", "tn6", "tn7", "ancx", "xs1", "qks", "sparks", "cats", "csqn_bound",
              "total_ci", "total_ca"]
DIAG_INT_COLS = (26, 27, 28, 29, 30, 31, 32, 33, 34, 35, 36, 40)


class Params(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("scheme", ctypes.c_int), ("precision", ctypes.c_int),
                ("iso2", ctypes.c_int), ("tapsi", ctypes.c_int), ("lqt", ctypes.c_int),
                ("nlcc", ctypes.c_int), ("steps_per_output", ctypes.c_int),
                ("stim_amp", ctypes.c_double), ("stim_dur", ctypes.c_double),
                ("stim_offset", ctypes.c_double), ("ci0", ctypes.c_double), ("v0", ctypes.c_double),
                ("cm_mode", ctypes.c_int), ("cm", ctypes.c_double), ("variants", ctypes.c_int),
                ("reserved_", ctypes.c_int)]

    @classmethod
    def default(cls, **kw):
        """The reference's values (run.sh `make il`: iso2 + lqt)."""
        p = cls(dt=0.025, scheme=SCHEME_A, precision=PREC_F64, iso2=1, tapsi=0, lqt=1, nlcc=4,
                steps_per_output=100, stim_amp=80.0, stim_dur=1.0, stim_offset=100.0, ci0=0.0944,
                v0=-80.0, cm_mode=0, cm=45.0, variants=0, reserved_=0)
        for k, v in kw.items():
            setattr(p, k, v)
        return p


class CellDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("pad_", ctypes.c_int), ("seed", ctypes.c_uint64), ("cl", ctypes.c_double),
                ("cjsr0", ctypes.c_double), ("ncx_scale", ctypes.c_double), ("Ku", ctypes.c_double),
                ("Kb", ctypes.c_double), ("ica_scale", ctypes.c_double), ("Ku2", ctypes.c_double),
                ("ikr_scale", ctypes.c_double)]

    @classmethod
    def default(cls, **kw):
        """run.sh:9 "stab" values on the reference 62x26x10 interior lattice."""
        d = cls(nx=62, ny=26, nz=10, pad_=0, seed=1, cl=4000.0, cjsr0=740.0, ncx_scale=1.0, Ku=10.0,
                Kb=1.0, ica_scale=0.7, Ku2=1.0, ikr_scale=1.0)
        for k, v in kw.items():
            setattr(d, k, v)
        return d

    @property
    def ncru(self) -> int:
        return self.nx * self.ny * self.nz


class LibraryMissing(RuntimeError):
    pass


_lib = None


def load_library(path: str | None = None):
    """Load libryr_cru.so (built by __graft_entry__.build()); raise if absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise LibraryMissing(
            f"{p} not found: build the HIP extension first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(p)
    c_int, c_double, c_i64, c_u64 = ctypes.c_int, ctypes.c_double, ctypes.c_int64, ctypes.c_uint64
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    sig = {
        "ryr_abi_version": (c_int, []),
        "ryr_params_default": (None, [P(Params)]),
        "ryr_cell_desc_default": (None, [P(CellDesc)]),
        "ryr_create": (c_int, [c_int, P(Params), P(CellDesc), c_int, P(vp)]),
        "ryr_step": (c_int, [vp, c_i64]),
        "ryr_prepare": (c_int, [vp, c_i64]),
        "ryr_latency_floor": (c_int, [vp, c_i64, P(c_double)]),
        "ryr_sync": (c_int, [vp]),
        "ryr_read_scalars": (c_int, [vp, P(c_double)]),
        "ryr_read_field": (c_int, [vp, c_int, c_int, P(c_double), c_i64]),
        "ryr_write_field": (c_int, [vp, c_int, c_int, P(c_double), c_i64]),
        "ryr_write_scalar": (c_int, [vp, c_int, c_int, c_double]),
        "ryr_read_diag": (c_int, [vp, P(c_double), P(c_i64)]),
        "ryr_read_linescan": (c_int, [vp, c_int, P(c_double), c_i64]),
        "ryr_read_counters": (c_int, [vp, P(c_i64)]),
        "ryr_cru_steps": (c_int, [vp, P(c_i64)]),
        "ryr_kernel_time": (c_int, [vp, P(c_double), P(c_i64)]),
        "ryr_set_launch_shape": (c_int, [vp, c_int, c_int, c_int]),
        "ryr_get_launch_shape": (c_int, [vp, P(c_int), P(c_int), P(c_int)]),
        "ryr_profile_kernels": (c_int, [vp, c_i64, P(c_double), P(c_i64)]),
        "ryr_save": (c_int, [vp, ctypes.c_char_p]),
        "ryr_load": (c_int, [vp, ctypes.c_char_p]),
        "ryr_destroy": (None, [vp]),
        "ryr_last_error": (ctypes.c_char_p, [vp]),
        "ryr_read_beats": (c_int, [vp, c_int, P(c_double), c_int, P(c_int)]),
        "ryr_set_vclamp": (c_int, [vp, c_int, P(c_double), c_i64, c_int]),
        "ryr_mlp_create": (c_int, [c_int, P(vp)]),
        "ryr_mlp_destroy": (None, [vp]),
        "ryr_mlp_set_params": (c_int, [vp, P(ctypes.c_float), c_i64]),
        "ryr_mlp_get_params": (c_int, [vp, P(ctypes.c_float), c_i64]),
        "ryr_mlp_forward": (c_int, [vp, P(ctypes.c_float), c_i64, P(ctypes.c_float)]),
        "ryr_mlp_grad": (c_int, [vp, P(ctypes.c_float), P(ctypes.c_float), P(ctypes.c_float),
                                 P(ctypes.c_float), P(ctypes.c_float), c_i64, ctypes.c_float,
                                 P(ctypes.c_float), P(ctypes.c_float)]),
        "ryr_mlp_train": (c_int, [vp, P(ctypes.c_float), P(ctypes.c_float), P(ctypes.c_float),
                                  P(ctypes.c_float), c_i64, ctypes.c_float, c_i64, c_i64, ctypes.c_float,
                                  ctypes.c_float, ctypes.c_float, ctypes.c_float, c_u64, P(ctypes.c_float)]),
        "ryr_mlp_last_error": (ctypes.c_char_p, [vp]),
        "ryr_gating_probe": (c_int, [c_int, c_i64, c_u64, P(ctypes.c_uint32), c_u64, c_double, c_double,
                                     c_double, c_double, c_int, P(ctypes.c_int32), P(ctypes.c_int32),
                                     P(c_double), P(ctypes.c_int32), P(ctypes.c_int32),
                                     P(ctypes.c_int32)]),
        "ryr_gating_probe_noalloc": (c_int, [c_int, c_i64, c_u64, P(ctypes.c_uint32), c_u64, c_double, c_double,
                                             c_double, c_double, c_int, P(ctypes.c_int32), P(ctypes.c_int32),
                                             P(c_double), P(ctypes.c_int32), P(ctypes.c_int32),
                                             P(ctypes.c_int32)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.ryr_abi_version() != ABI_VERSION:
        raise LibraryMissing(f"{p}: ABI version mismatch")
    if path is None:
        _lib = lib
    return lib


# exported symbols every build must carry (checked by the CPU test suite)
EXPORTED = ["ryr_abi_version", "ryr_params_default", "ryr_cell_desc_default", "ryr_create", "ryr_step",
            "ryr_prepare", "ryr_latency_floor",
            "ryr_sync", "ryr_read_scalars", "ryr_read_field", "ryr_write_field", "ryr_write_scalar",
            "ryr_read_diag", "ryr_read_linescan", "ryr_read_counters", "ryr_cru_steps",
            "ryr_kernel_time", "ryr_save", "ryr_load", "ryr_destroy", "ryr_last_error",
            "ryr_read_beats", "ryr_set_vclamp", "ryr_gating_probe", "ryr_gating_probe_noalloc",
            "ryr_set_launch_shape",
            "ryr_get_launch_shape", "ryr_profile_kernels"]
# include/ryr_mlp.h
EXPORTED_MLP = ["ryr_mlp_create", "ryr_mlp_destroy", "ryr_mlp_set_params", "ryr_mlp_get_params",
                "ryr_mlp_forward", "ryr_mlp_grad", "ryr_mlp_train", "ryr_mlp_last_error"]
