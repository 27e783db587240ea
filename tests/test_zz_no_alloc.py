"""ryr_gating_probe_noalloc (staging in the library's fixed device array, no
allocation) against ryr_gating_probe (staging in buffers allocated per call):
the same counts, channel states, draw counts, open channels and fallbacks, bit
for bit, on two calls in a row (the second smaller, so stale scratch from the
first would show), and a call above the size limit is refused."""
import ctypes

import numpy as np
import pytest

from ryr_hyperactivity_amd import _abi

pytestmark = pytest.mark.gpu

I32 = ctypes.POINTER(ctypes.c_int32)


def _queries(n, seed):
    rng = np.random.default_rng(seed)
    cru = rng.integers(0, 2**31, n, dtype=np.uint32)
    counts = np.zeros((n, 4), np.int32)
    nou = rng.integers(0, 60, n)
    nob = rng.integers(0, 100 - nou + 1)
    ncu = rng.integers(0, 100 - nou - nob + 1)
    counts[:, 0], counts[:, 1], counts[:, 2] = nou, ncu, nob
    counts[:, 3] = 100 - nou - nob - ncu
    lcc = np.full((n, 8), 16, np.int32)
    lcc[:, :4] = rng.integers(0, 16, (n, 4))
    conc = np.stack([10 ** rng.uniform(-2, 2.5, n), rng.uniform(100, 1000, n),
                     10 ** rng.uniform(-2, 1.5, n), rng.uniform(-90, 50, n)], axis=1)
    return cru, counts, lcc, np.ascontiguousarray(conc)


def _call(fn, cru, counts, lcc, conc, dt):
    n = cru.size
    c, l = counts.copy(), lcc.copy()
    draws, nopen, fails = (np.full(n, -7, np.int32) for _ in range(3))
    rc = fn(0, n, 4242, cru.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), 77, 3.333, 2.5, 18.0, dt, 1,
            c.ctypes.data_as(I32), l.ctypes.data_as(I32), conc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
            draws.ctypes.data_as(I32), nopen.ctypes.data_as(I32), fails.ctypes.data_as(I32))
    return rc, (c, l, draws, nopen, fails)


def test_noalloc_probe_equals_the_allocating_probe_on_two_calls():
    lib = _abi.load_library()
    for n, seed, dt in ((_abi.PROBE_NOALLOC_MAX, 1, 0.4), (3001, 2, 0.025)):
        print(f"case noalloc probe n={n} dt={dt}", flush=True)
        q = _queries(n, seed)
        rc_a, ref = _call(lib.ryr_gating_probe, *q, dt)
        rc_b, got = _call(lib.ryr_gating_probe_noalloc, *q, dt)
        assert rc_a == rc_b == _abi.RYR_OK, lib.ryr_last_error(None)
        for a, b, name in zip(ref, got, ["counts", "lcc", "draws", "nopen", "fails"]):
            assert np.array_equal(a, b), name
        assert np.all(got[2] >= 0)                     # every query's outputs were written


def test_noalloc_probe_refuses_more_queries_than_its_array_holds():
    lib = _abi.load_library()
    q = _queries(_abi.PROBE_NOALLOC_MAX + 1, 3)
    rc, _ = _call(lib.ryr_gating_probe_noalloc, *q, 0.025)
    assert rc == _abi.RYR_E_ARG
    assert "at most" in lib.ryr_last_error(None).decode()
