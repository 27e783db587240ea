"""The two arithmetic primitives shared by the kernels and the oracle.

Philox4x32-10 is pinned by the published known-answer vectors of its authors'
Random123 distribution; the deterministic exp/log/cos/pow by glibc (through
Python's math module) within a few ulp over the ranges the model uses, plus
their special cases.
"""
import math

import numpy as np
import pytest

from oracle import binding as orc


# Random123 known-answer vectors for philox4x32 with 10 rounds:
# (counter[4], key[2]) -> output[4]
KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_philox_known_answers(ctr, key, expect):
    assert orc.philox(ctr, key) == list(expect)


def test_uniforms_open_interval_and_extremes():
    L = orc.lib()
    assert L.orc_slot_uniform(0, 0) == 2.0 ** -53
    assert L.orc_slot_uniform(0xFFFFFFFF, 0xFFFFFFFF) == 1.0 - 2.0 ** -53
    assert L.orc_slot_uniform(0x80000000, 0) == 0.5 + 2.0 ** -53


def test_stream_statistics_and_addressing():
    u, w = orc.stream(1, 17, 99, 200000)
    assert u.min() > 0 and u.max() < 1
    assert abs(u.mean() - 0.5) < 4 * math.sqrt(1 / 12 / u.size)
    assert abs(u.var() - 1 / 12) < 0.002
    # integer draws are the low word of the slot: fair bits
    assert abs((w & 1).mean() - 0.5) < 0.01
    # different CRU / step / seed give unrelated streams; same key repeats
    assert not np.array_equal(orc.stream(1, 18, 99, 16)[0], u[:16])
    assert not np.array_equal(orc.stream(1, 17, 100, 16)[0], u[:16])
    assert not np.array_equal(orc.stream(2, 17, 99, 16)[0], u[:16])
    assert np.array_equal(orc.stream(1, 17, 99, 16)[0], u[:16])
    # slot 2k and 2k+1 come from Philox block k of counter (k, cru, step)
    blk = orc.philox((1, 17, 99, 0), (1, 0))
    x = (blk[2] << 20) | (blk[3] >> 12)
    assert u[3] == (x + 0.5) * 2.0 ** -52


def _slot_u(blk, half):
    a, b = blk[2 * half], blk[2 * half + 1]
    return (((a << 20) | (b >> 12)) + 0.5) * 2.0 ** -52


def test_transition_draw_addressing():
    """RyR transition t draws its first uniform from static slot 8 + t (block
    4 + t//2 of counter (block, CRU, step)) and its further ones from its own
    substream, block ((t + 1) << 24) | (j >> 1).  Only OU->CU (t = 0) moves
    here, so its count is Knuth's product rule applied to those uniforms,
    recomputed from the known-answer-pinned Philox blocks."""
    n = 300
    seed, step = 0x1234_5678_9ABC, 41
    counts = np.zeros((n, 4), np.int32)
    counts[:, 0] = 100
    conc = np.tile([1.0, 740.0, 0.1, -80.0], (n, 1))
    lcc = np.full((n, 8), 16, np.int32)
    out, _, draws, _, _ = orc.gating_probe(seed, np.arange(n, dtype=np.uint32), step, 10.0, 1.0, 0.0, 0.025, 1,
                                           counts, lcc, conc)
    key = (seed & 0xFFFFFFFF, seed >> 32)
    lam = orc.lib().orc_dm_exp(-(100) * 0.025)
    for q in range(n):
        puu = _slot_u(orc.philox((4, q, step, 0), key), 0)
        kk, j = 1, 0
        while puu >= lam and kk < 195:
            kk += 1
            puu = puu * _slot_u(orc.philox(((1 << 24) | (j >> 1), q, step, 0), key), j & 1)
            j += 1
        assert out[q, 1] == kk - 1
        assert draws[q] == 8 + j


def test_knuth_early_out_bound():
    """The kernels settle a Knuth transition at count 0 without the
    exponential when the first uniform lies below 1 - x (x >= 2^-20) or
    1 - 2x (2^-40 <= x < 2^-20), x = n p.  That needs the rounded bound to lie
    at or below the rounded exp(-x) that the loop compares against."""
    L = orc.lib()
    rng = np.random.default_rng(5)
    xs = np.concatenate([2.0 ** rng.uniform(-40, 0, 60000), 2.0 ** -np.arange(0, 41, dtype=float),
                         np.nextafter(2.0 ** -20, 0) * np.ones(1), np.linspace(0.5, 1.0, 2001)])
    for x in xs:
        thr = 1.0 - x if x >= 2.0 ** -20 else 1.0 - 2.0 * x
        assert thr <= L.orc_dm_exp(-x)


def _ulps(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    ia = a.view(np.int64).astype(object)
    ib = b.view(np.int64).astype(object)
    return max(abs(int(x) - int(y)) for x, y in zip(ia, ib))


def test_detmath_exp_accuracy():
    L = orc.lib()
    xs = np.concatenate([np.linspace(-745, 709.7, 4001), np.random.default_rng(0).uniform(-50, 50, 20000),
                         np.linspace(-1e-3, 1e-3, 1001)])
    got = np.array([L.orc_dm_exp(x) for x in xs])
    ref = np.array([math.exp(x) for x in xs])
    normal = ref > 2.3e-308
    assert _ulps(got[normal], ref[normal]) <= 2
    assert L.orc_dm_exp(0.0) == 1.0
    assert L.orc_dm_exp(-800.0) == 0.0
    assert math.isinf(L.orc_dm_exp(710.0))
    assert math.isnan(L.orc_dm_exp(float("nan")))


def test_detmath_log_accuracy():
    L = orc.lib()
    rng = np.random.default_rng(1)
    xs = np.concatenate([10.0 ** rng.uniform(-300, 300, 20000), rng.uniform(0.5, 2.0, 20000),
                         [5e-324, 1e-310, 1.0, 2.0, 0.5]])
    got = np.array([L.orc_dm_log(x) for x in xs])
    ref = np.array([math.log(x) for x in xs])
    far = np.abs(ref) > 1e-3
    assert _ulps(got[far], ref[far]) <= 2
    near = ~far
    assert np.all(np.abs(got[near] - ref[near]) <= 4e-16 * np.maximum(np.abs(ref[near]), 1e-300) + 1e-300)
    assert L.orc_dm_log(1.0) == 0.0
    assert math.isinf(L.orc_dm_log(0.0)) and L.orc_dm_log(0.0) < 0
    assert math.isnan(L.orc_dm_log(-1.0))


def test_detmath_cos_accuracy():
    L = orc.lib()
    xs = np.concatenate([np.linspace(0, 2 * 3.1415926, 20001), np.random.default_rng(2).uniform(-20, 20, 5000)])
    got = np.array([L.orc_dm_cos(x) for x in xs])
    ref = np.array([math.cos(x) for x in xs])
    assert np.max(np.abs(got - ref)) <= 2.5e-16


def test_detmath_pow():
    L = orc.lib()
    rng = np.random.default_rng(3)
    # integral exponents by binary powering: pow(x,2) == x*x, pow(x,3) == x*(x*x), exactly
    for x in rng.uniform(0.01, 100, 1000):
        assert L.orc_dm_pow(x, 2.0) == x * x
        assert L.orc_dm_pow(x, 3.0) == x * (x * x)
        assert L.orc_dm_pow(x, 1.0) == x
        assert abs(L.orc_dm_pow(x, 23.0) / x ** 23 - 1) < 8 * 2.2e-16
    # non-integral (SERCA Hill exponent 1.787): exp(y log x) loses |y log x|
    # ulps of relative accuracy through the rounding of y*log(x)
    for x in np.concatenate([rng.uniform(0, 50, 2000), 10 ** rng.uniform(-8, 0, 500)]):
        ref = x ** 1.787
        bound = (abs(1.787 * math.log(x)) + 3) * 2.3e-16 * ref if x > 0 else 0.0
        assert abs(L.orc_dm_pow(x, 1.787) - ref) <= bound
    assert L.orc_dm_pow(0.0, 1.787) == 0.0


# ---- stochastic rounding of the FP32 production mode's slow pools ----------
def _lo_hi(v):
    """The floats just below/at and just above a double v > 0."""
    lo = np.float32(v)
    if float(lo) > v:
        lo = np.nextafter(lo, np.float32(0))
    return lo, np.nextafter(lo, np.float32(np.inf))


def test_stochastic_rounding_brackets_and_keeps_exact_floats():
    L = orc.lib()
    rng = np.random.default_rng(5)
    rs = [0, 8, 0xFFFFFFF8, 0xFFFFFFFF, 0x80000000] + [int(x) for x in rng.integers(0, 2**32, 20)]
    for scale in (1e-3, 0.1, 1.0, 118.0, 750.0, 1000.0, 3.0e4):
        for v in scale * (1.0 + rng.random(20)):
            lo, hi = _lo_hi(float(v))
            for r in rs:
                out = np.float32(L.orc_sr_f32(float(v), r))
                assert out == lo or out == hi, (v, r, out, lo, hi)
            f = float(lo)                      # an exact float rounds to itself whatever the bits
            assert all(np.float32(L.orc_sr_f32(f, r)) == lo for r in rs)
    assert all(L.orc_sr_f32(0.0, r) == 0.0 for r in rs)
    # rounding up across a binade boundary lands on the next binade's first float
    below2 = float(np.nextafter(np.float32(2.0), np.float32(0)))
    v = below2 + 0.75 * (2.0 - below2)
    assert {float(L.orc_sr_f32(v, r)) for r in (0, 0xFFFFFFF8)} == {below2, 2.0}


@pytest.mark.parametrize("frac", [0.001, 0.25, 0.5, 0.875, 0.999])
def test_stochastic_rounding_rounds_up_with_the_fractional_probability(frac):
    """Stratified random words (the 29 bits used step through [0, 2^29) in
    steps of 2^13): the share rounded up equals the fraction of an ulp the
    value lies above the lower float, to the stratification's 2^-16."""
    L = orc.lib()
    for base in (118.0, 750.0, 0.0944):
        lo, hi = _lo_hi(base)
        v = float(lo) + frac * (float(hi) - float(lo))
        ups = sum(L.orc_sr_f32(v, ((k << 13) | 1) << 3) == float(hi) for k in range(65536))
        assert abs(ups / 65536 - frac) <= 2.0 / 65536, (base, frac, ups / 65536)


def test_stochastic_rounding_accumulates_increments_below_an_ulp():
    """Slow troponin at rest: 2.5e-6 uM per step on 118 uM (a float ulp there
    is 7.6e-6).  Rounding to nearest never moves; stochastic rounding with the
    model's own random bits tracks the exact sum within its random-walk
    spread (5 sigma)."""
    L = orc.lib()
    x, xn, d, n = np.float32(118.0), np.float32(118.0), 2.5e-6, 20000
    for k in range(n):
        x = np.float32(L.orc_sr_f32(float(x) + d, L.orc_sr_bits(7, k, 1, 123, 5)))
        xn = np.float32(float(xn) + d)
    ulp = float(np.spacing(np.float32(118.0)))
    p = d / ulp
    sigma = ulp * math.sqrt(n * p * (1 - p))
    assert xn == np.float32(118.0)
    assert abs(float(x) - (118.0 + n * d)) < 5 * sigma, (float(x), 118.0 + n * d, sigma)


def test_stochastic_rounding_bits_are_uniform_and_independent():
    L = orc.lib()
    # the mixer is a bijection: no two of 2^18 consecutive inputs collide
    m = {L.orc_mix32(x) for x in range(1 << 18)}
    assert len(m) == 1 << 18
    # over the sub-voxels of one key the words are uniform (64-bin chi-square)
    w = np.array([L.orc_sr_bits(3, 1000, 0, n, ii) for n in range(8192) for ii in range(8)], dtype=np.float64)
    h = np.histogram(w / 2**32, bins=64, range=(0, 1))[0]
    e = len(w) / 64
    chi2 = float(((h - e) ** 2 / e).sum())
    assert chi2 < 120, chi2                    # 63 dof: p ~ 3e-5 at 120
    # one sub-voxel over successive steps: no lag-1 correlation
    s = np.array([L.orc_sr_bits(3, k, 0, 77, 2) for k in range(20000)], dtype=np.float64) / 2**32
    c = np.corrcoef(s[:-1], s[1:])[0, 1]
    assert abs(c) < 0.03, c
    # fields, cells and sub-voxels draw different words
    a = [L.orc_sr_bits(3, 9, f, 5, 1) for f in range(4)] + [L.orc_sr_bits(4, 9, 0, 5, 1), L.orc_sr_bits(3, 9, 0, 5, 2)]
    assert len(set(a)) == len(a)
