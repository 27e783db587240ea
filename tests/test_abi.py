"""The C ABI library loads and exports what include/ryr_cru.h declares (CPU;
no kernel launches), and its host-side checks reject bad input."""
import ctypes
import os
import re
import subprocess

import pytest

from ryr_hyperactivity_amd import _abi
from ryr_hyperactivity_amd import CellDesc, Params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("ryr_cru.h", "ryr_mlp.h")]


def _declared():
    names = set()
    for hdr in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", open(hdr).read(), flags=re.S)
        names |= set(re.findall(r"\b(ryr_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = _abi.load_library()
    names = _declared()
    assert "ryr_create" in names and "ryr_gating_probe" in names and "ryr_mlp_train" in names
    for n in names:
        assert hasattr(lib, n), n
    assert set(_abi.EXPORTED) | set(_abi.EXPORTED_MLP) <= set(names)
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], capture_output=True, text=True).stdout
    for n in names:
        assert re.search(rf"\bT {n}\b", out), n


def test_struct_layout_and_defaults_match_the_header():
    lib = _abi.load_library()
    p = Params()
    lib.ryr_params_default(ctypes.byref(p))
    q = Params.default()
    for f, _ in Params._fields_:
        assert getattr(p, f) == getattr(q, f), f
    d = CellDesc()
    lib.ryr_cell_desc_default(ctypes.byref(d))
    e = CellDesc.default()
    for f, _ in CellDesc._fields_:
        assert getattr(d, f) == getattr(e, f), f
    assert ctypes.sizeof(Params) == 104 and ctypes.sizeof(CellDesc) == 88
    assert lib.ryr_abi_version() == _abi.ABI_VERSION == 4


def _create(p, cells):
    lib = _abi.load_library()
    arr = (CellDesc * len(cells))(*cells)
    h = ctypes.c_void_p()
    rc = lib.ryr_create(0, ctypes.byref(p), arr, len(cells), ctypes.byref(h))
    return rc, lib.ryr_last_error(None).decode()


@pytest.mark.parametrize("mutate,msg", [
    (lambda p, d: setattr(p, "dt", 0.0), "dt"),
    (lambda p, d: setattr(p, "scheme", 7), "scheme"),
    (lambda p, d: setattr(p, "nlcc", 9), "nlcc"),
    (lambda p, d: setattr(d, "nx", 0), "lattice"),
    (lambda p, d: setattr(d, "cl", 1000.01), "multiple of dt"),
    (lambda p, d: setattr(d, "Ku", 0.0), "Ku"),
])
def test_create_rejects_bad_parameters_before_touching_a_device(mutate, msg):
    p, d = Params.default(), CellDesc.default(nx=4, ny=4, nz=4)
    mutate(p, d)
    rc, err = _create(p, [d])
    assert rc == _abi.RYR_E_ARG
    assert msg in err


def test_product_path_has_no_cpu_fallback(tmp_path):
    with pytest.raises(_abi.LibraryMissing):
        _abi.load_library(str(tmp_path / "missing.so"))


def test_cli_usage():
    cli = os.path.join(ROOT, "ryr_hyperactivity_amd", "bin", "ryr_cell")
    r = subprocess.run([cli], capture_output=True, text=True)
    assert r.returncode == 2 and "usage" in r.stderr


def test_argv_parser_matches_run_sh():
    from ryr_hyperactivity_amd.engine import parse_argv
    # run.sh:11 "hyper" run
    p, d, nsteps, name, dev = parse_argv(["0", "Rabbit_lqt2_hyper", "1.0", "740", "1.0", "3.333", "0.00", "18.",
                                          "0.7", "1"])
    assert (d.cjsr0, d.ncx_scale, d.Ku, d.Kb, d.ica_scale, d.Ku2) == (740.0, 1.0, 3.333, 18.0, 0.7, 1.0)
    assert name == "Rabbit_lqt2_hyper" and dev == 0
    assert nsteps == round((7 * 4000 + 2100) / 0.025)
    assert p.iso2 == 1 and p.lqt == 1
    p, d, nsteps, _, _ = parse_argv(["0", "x", "1", "740", "1", "10", "0", "1", "0.7", "2.5", "--lattice", "100x20x10",
                                     "--cl", "2000", "--nbeat", "10", "--tail", "0", "--precision", "f32",
                                     "--scheme", "B", "--ic", "baseline"])
    assert (d.nx, d.ny, d.nz, d.cl, d.Ku2) == (100, 20, 10, 2000.0, 2.5)
    assert nsteps == 800000 and p.precision == 1 and p.scheme == 1
    assert (p.ci0, p.v0, d.cjsr0) == (0.1, -86.0, 750.0)


def test_argv_variant_flags_match_the_makefile_builds():
    from ryr_hyperactivity_amd.engine import parse_argv
    from ryr_hyperactivity_amd._abi import VAR_PERM, VAR_ISO, VAR_SLOWNCX, VAR_RANDOMLCC
    base = ["0", "x", "1.0", "740", "1.0", "10", "0.00", "1.", "0.7", "1"]
    p, d, _, _, _ = parse_argv(base + ["--perm"])
    assert p.variants == VAR_PERM and (p.ci0, d.cjsr0) == (0.3, 900.0)          # :87-89
    p, _, _, _, _ = parse_argv(base + ["--iso", "--slowncxchange", "--randomlcc", "--nopacing"])
    assert p.variants == VAR_ISO | VAR_SLOWNCX | VAR_RANDOMLCC and p.stim_amp == 0.0
    with pytest.raises(ValueError):
        parse_argv(base + ["--nosuchbuild"])
