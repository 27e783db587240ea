"""MI355X (gfx950) engine for the per-time-step CRU update of the detailed
rabbit ventricular myocyte model (reference: colinmrees/RyR_hyperactivity).

The compute path is the HIP library libryr_cru.so (C ABI: include/ryr_cru.h);
this package is its Python host mirror.  There is no CPU fallback: without the
built library, importing the engine raises.
"""
from ._abi import (CellDesc, Params, LibraryMissing, load_library, PREC_F32, PREC_F64, SCHEME_A,
                   SCHEME_B, DIAG_NAMES, FIELDS, SCALARS, COUNTERS)
from .engine import (CellEngine, EngineError, gating_probe, format_wholetest_row, format_linescan, run_cell,
                     clamp_waveform)

__all__ = ["CellDesc", "Params", "LibraryMissing", "load_library", "CellEngine", "EngineError",
           "gating_probe", "format_wholetest_row", "format_linescan", "run_cell", "clamp_waveform", "PREC_F32",
           "PREC_F64", "SCHEME_A", "SCHEME_B", "DIAG_NAMES", "FIELDS", "SCALARS", "COUNTERS"]
