"""The reference's remaining compile-time builds as runtime variants
(ryr_params.variants): the oracle against what each variant fixes (CPU), then
the GPU bit for bit against the oracle (FP64)."""
import math

import numpy as np
import pytest

from ryr_hyperactivity_amd import CellDesc, Params, FIELDS
from ryr_hyperactivity_amd._abi import S, VAR_PERM, VAR_ISO, VAR_SLOWNCX, VAR_RANDOMLCC, DIAG_NAMES
from oracle.binding import OracleCell
from oracle import binding as orc

FRT = 96.485 / 8.314 / 308.0


def test_perm_freezes_the_sarcolemmal_channels_and_pacing():
    """perm (cuda_Cell_Detailed.cu:697-699,1249-1263,1294-1304): no stimulus,
    no LCC gating, no I_CaL and no NCX."""
    p = Params.default(variants=VAR_PERM, ci0=0.3, stim_offset=1.0)
    d = CellDesc.default(nx=5, ny=4, nz=3, cjsr0=900.0, cl=50.0, seed=4)
    o = OracleCell(p, d)
    lcc0, ancx0 = o.field("LCC").copy(), o.field("ANCX").copy()
    stims = []
    for _ in range(30):
        o.step(101)
        stims.append(o.scalars()[S["STIM"]])
    assert max(abs(s) for s in stims) == 0.0
    assert np.array_equal(o.field("LCC"), lcc0) and np.all(o.field("NL") == 0)
    assert np.array_equal(o.field("ANCX"), ancx0)
    row, _ = o.diag()
    assert row[DIAG_NAMES.index("i_cal")] == 0.0 and row[DIAG_NAMES.index("i_ncx")] == 0.0


def test_perm_bath_holds_the_lattice_near_ci0():
    """With no sarcolemmal fluxes, the cs relaxation to ci0 (tau 5 ms) and the
    ci/cs bath at the lattice faces (:1409-1411, :1662-1689) keep a small
    resting lattice at the bath value."""
    p = Params.default(variants=VAR_PERM, ci0=0.3, tapsi=1, lqt=0)
    d = CellDesc.default(nx=4, ny=3, nz=3, cjsr0=900.0, seed=2)
    o = OracleCell(p, d)
    o.set_field("CS", np.full(d.ncru, 2.0))                # perturb the sub-membrane space
    o.step(4000)                                           # 100 ms = 20 relaxation times
    assert np.allclose(o.field("CS"), 0.3, rtol=0.1)
    assert np.allclose(o.field("CI"), 0.3, rtol=0.1)


def test_iso_paces_from_phase_zero_for_twelve_beats_and_adds_channels():
    """iso (:683-693, :1213-1219): stimulus while fmod(t, CL) < 1 and t < 11.9
    CL; at t = 5.5 CL (mod 15.5 CL) the absent LCC slots 4..7 appear in
    state 3."""
    dt = 0.025
    p = Params.default(variants=VAR_ISO)
    d = CellDesc.default(nx=2, ny=2, nz=2, cl=5.0, seed=1)    # 200 steps per cycle
    o = OracleCell(p, d)
    ncl = 200
    on = []
    n_act = 11 * ncl // 2                                     # 5.5 CL
    for n in range(13 * ncl):
        if n == n_act:
            assert np.all(o.field("LCC")[4:] == 16)
        o.step(1)
        on.append(o.scalars()[S["STIM"]] > 0)
        if n == n_act:
            assert np.all(o.field("LCC")[4:] < 16)
    on = np.array(on)
    starts = np.flatnonzero(on & ~np.r_[False, on[:-1]])
    assert list(starts) == [k * ncl for k in range(12)]           # t < 11.9 CL: 12 beats from t = 0
    assert on[:40].all() and not on[40:ncl].any()                 # 1 ms = 40 steps


def test_slowncx_background_current_vanishes_at_the_calcium_reversal():
    """slowncxchange (:1266-1268): 0.003016/32 (V - ln(180000/cs)/(2 FRT)) per
    CRU.  With V at the Ca reversal of the uniform cs the first step's cs
    equals the base model's; below it the current is inward and cs rises.
    The exchanger is off (ncx scale 0) so that only the background current
    differs between the two builds."""
    cs0 = 0.0944
    eca = math.log(180000 / cs0) / 2.0 / FRT
    res = {}
    for var in (0, VAR_SLOWNCX):
        for v in (eca, -80.0):
            p = Params.default(variants=var)
            d = CellDesc.default(nx=3, ny=3, nz=3, seed=5, ncx_scale=0.0)
            o = OracleCell(p, d)
            o.set_scalar("V", v)
            o.step(1)
            res[(var, v)] = o.field("CS").copy()
    assert np.allclose(res[(VAR_SLOWNCX, eca)], res[(0, eca)], rtol=1e-12, atol=0)
    assert np.all(res[(VAR_SLOWNCX, -80.0)] > res[(0, -80.0)])


def test_slowncx_runs_the_exchanger_at_five_times_the_allosteric_factor():
    """slowncxchange (subroutine_cuda.c:59-62): Ka = 5 Ancx in the exchanger
    current, while Ancx itself relaxes as in the base model."""
    import ctypes
    L = orc.lib()
    for cs, v, nai in ((1e-4, -80.0, 8.0), (5e-4, 10.0, 10.0), (2e-3, 40.0, 12.0)):
        a0, a5 = ctypes.c_double(0.2), ctypes.c_double(0.2)
        i0 = L.orc_ncx(cs, v, nai, ctypes.byref(a0), 0.025)
        i5 = L.orc_ncx_slow(cs, v, nai, ctypes.byref(a5), 0.025)
        assert i5 == pytest.approx(5.0 * i0, rel=1e-14)
        assert a5.value == a0.value != 0.2


def test_randomlcc_draws_zero_to_eight_channels_uniformly():
    """randomlcc (:1150-1155): per CRU curand % 9 channels in slots
    0..count-1; here one 32-bit Philox draw per CRU."""
    p = Params.default(variants=VAR_RANDOMLCC)
    d = CellDesc.default(nx=30, ny=20, nz=6, seed=8)          # 3600 CRUs
    o = OracleCell(p, d)
    lcc = o.field("LCC")                                    # 8 x N, 16 = absent
    present = lcc < 16
    cnt = present.sum(0)
    # slots are filled from 0 up
    assert np.all(present == (np.arange(8)[:, None] < cnt[None, :]))
    h = np.bincount(cnt, minlength=9)
    chi2 = ((h - d.ncru / 9) ** 2 / (d.ncru / 9)).sum()
    assert h.size == 9 and chi2 < 26.1                     # chi-square, 8 dof, p = 0.001


VARIANT_CASES = [
    ("perm", dict(variants=VAR_PERM, ci0=0.3), dict(cjsr0=900.0)),
    ("iso", dict(variants=VAR_ISO, iso2=0), {}),
    ("iso+iso2", dict(variants=VAR_ISO), {}),
    ("slowncx", dict(variants=VAR_SLOWNCX), {}),
    ("randomlcc", dict(variants=VAR_RANDOMLCC), {}),
    ("randomlcc+iso+slowncx/B", dict(variants=VAR_RANDOMLCC | VAR_ISO | VAR_SLOWNCX, scheme=1), {}),
    ("tapsi", dict(tapsi=1), {}),                       # with lqt: cjsr0 10, SERCA 0.05
    ("tapsi-nolqt", dict(tapsi=1, lqt=0), {}),          # SERCA 0
    ("tapsi+iso", dict(tapsi=1, variants=VAR_ISO, iso2=0), {}),
]


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [None, (256, 256, 3), (128, 64, 1), (256, 256, 4), (256, 512, 4), (1024, 64, 2)],
                         ids=["tuned", "gating-fused", "fused", "one-launch", "one-launch-2", "lane-per-cru-sweep"])
@pytest.mark.parametrize("name,pk,dk", VARIANT_CASES, ids=[c[0] for c in VARIANT_CASES])
def test_variant_fp64_bitexact_against_the_oracle(name, pk, dk, shape):
    """Every variant, FP64 against the oracle, at the tuned shape and forced
    onto both CRU kernel families: the gating-free one of the gating-fused
    schedule (state, buffer and neighbour loads issued together) and the
    128-lane one that also gates (loads at their use)."""
    from ryr_hyperactivity_amd import CellEngine
    print(f"case variant {name} shape {shape}", flush=True)
    p = Params.default(stim_offset=1.0, **pk)
    if shape is not None and shape[2] in (2, 4) and p.scheme == 1:
        pytest.skip("the one-launch and concurrent schedules need cleft scheme A")
    d = CellDesc.default(nx=9, ny=5, nz=3, cl=10.0, seed=6, **dk)   # 400 steps per cycle: iso events by step 2200
    o = OracleCell(p, d)
    with CellEngine(p, d) as eng:
        if shape is not None:
            eng.set_launch_shape(*shape)
        for chunk in (1, 99, 700, 1500, 37):
            eng.step(chunk)
            o.step(chunk)
            for f in FIELDS:
                g, r = eng.field(0, f), o.field(f)
                assert np.array_equal(g, r), (name, chunk, f, np.argwhere(g != r)[:3])
            assert np.array_equal(eng.scalars()[0], o.scalars()), (name, chunk)
        grow, gst = eng.diag()
        orow, ost = o.diag()
        assert gst == ost and np.array_equal(grow[0], orow)
        assert np.array_equal(eng.beats(0), o.beats())


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", [0, 1])
@pytest.mark.parametrize("shape", [(256, 512, 1), (256, 256, 3)])
def test_split_shapes_bitexact_with_gapped_lcc_slots(scheme, shape):
    """Two lanes per CRU (CRU kernel size 512) split the RyR transitions and
    the LCC channels by parity; the gating schedule (3) moves the RyR gating
    into the sub-voxel launch.  randomlcc + iso leaves gaps in the slot mask
    (0..count-1 and, from 5.5 CL, 4..7): only the slots before the first gap
    gate, as in the reference's loop.  FP64 against the oracle, both cleft
    schemes, with the shape forced."""
    from ryr_hyperactivity_amd import CellEngine
    print(f"case shape {shape} scheme {scheme}", flush=True)
    p = Params.default(stim_offset=1.0, variants=VAR_RANDOMLCC | VAR_ISO, scheme=scheme)
    d = CellDesc.default(nx=9, ny=5, nz=3, cl=10.0, seed=9)
    o = OracleCell(p, d)
    with CellEngine(p, d) as eng:
        eng.set_launch_shape(*shape)
        for chunk in (1, 99, 2100, 300):
            eng.step(chunk)
            o.step(chunk)
            for f in FIELDS:
                g, r = eng.field(0, f), o.field(f)
                assert np.array_equal(g, r), (scheme, chunk, f, np.argwhere(g != r)[:3])
            assert np.array_equal(eng.scalars()[0], o.scalars()), (scheme, chunk)
