"""The FP64 oracle against closed forms, invariants and brute force.

Each check is chosen so that a plausible transcription mistake (a dropped
term, a wrong volume factor, a swapped operand) fails it; none of them retypes
the formula under test.
"""
import math

import numpy as np
import pytest

from ryr_hyperactivity_amd import CellDesc, Params, SCHEME_A
from oracle import binding as orc
from oracle.binding import OracleCell


def test_initial_buffers_sit_at_their_steady_state():
    """Init (cuda_Cell_Detailed.cu:1136-1147) puts every buffer at the fixed
    point of its own flux (buffers.c): the flux there must vanish, and the
    values are the closed forms for ci = 0.0944 uM."""
    p = Params.default()
    d = CellDesc.default(nx=3, ny=2, nz=2)
    o = OracleCell(p, d)
    expect = {"CATF": 9.5245181356, "CATS": 117.7888963838, "CACAL": 0.5060014991, "CASR": 2.5829493088,
              "CASAR": 6.0557184751, "CASARH": 71.8052738337, "CASARJ": 6.0557184751, "CASARHJ": 71.8052738337}
    for f, v in expect.items():
        assert np.allclose(o.field(f), v, rtol=1e-10), f
    L = orc.lib()
    ci = 0.0944
    for which, f in [(0, "CATF"), (1, "CATS"), (2, "CACAL"), (3, "CASR"), (6, "CASAR"), (7, "CASARH"),
                     (8, "CASARJ"), (9, "CASARHJ")]:
        b = float(o.field(f).flat[0])
        flux = L.orc_buffer_flux(which, b, 0.0, ci, 0.025)
        on_rate = L.orc_buffer_flux(which, 0.0, 0.0, ci, 0.025)
        assert abs(flux) < 1e-12 * abs(on_rate), f
    # saturation branch: k_on*c*dt > 1 caps the on-rate at 0.95/dt (buffers.c:74-75)
    big = L.orc_buffer_flux(0, 0.0, 0.0, 1e4, 0.025)
    assert big == pytest.approx(0.95 / 0.025 * 70.0)


def test_csqn_buffer_closed_forms_and_identity():
    L = orc.lib()
    assert L.orc_luminal_SS(740.0) == pytest.approx(7614.953517, rel=1e-9)
    assert L.orc_luminal(740.0) == pytest.approx(0.1755539589, rel=1e-9)
    # beta = 1/(1 + dB/dc) for the rapid-buffer CSQN pool over the physiological range
    for c in np.linspace(400.0, 1000.0, 13):
        h = 2e-5 * c
        dB = (L.orc_luminal_SS(c + h) - L.orc_luminal_SS(c - h)) / (2 * h)
        assert L.orc_luminal(c) == pytest.approx(1.0 / (1.0 + dB), rel=2e-6)


def test_initial_ryr_split():
    """ncb = int(100 * fracbound(cjsr0)) (cuda_Cell_Detailed.cu:1165-1174)."""
    L = orc.lib()
    assert L.orc_fracbound(740.0) == pytest.approx(0.339760288, rel=1e-8)
    for cj, ncb in [(740.0, 33), (750.0, 30)]:
        o = OracleCell(Params.default(), CellDesc.default(nx=2, ny=2, nz=2, cjsr0=cj))
        assert np.all(o.field("NCB") == ncb) and np.all(o.field("NCU") == 100 - ncb)
        assert np.all(o.field("NOU") == 0) and np.all(o.field("NOB") == 0)
        lcc = o.field("LCC")
        assert np.all(lcc[:4] == 3) and np.all(lcc[4:] == 16)


def test_ghk_limits_and_sign():
    L = orc.lib()
    frt = 96.485 / 8.314 / 308.0
    v_of = lambda za: za / frt
    # V = 0: the small-za branch is the za -> 0 limit 2 Pca F gamma (cp/1000 - Cext)
    assert L.orc_lcccurrent(0.0, 100.0) == pytest.approx(2 * 11.9 * 96.485 * 0.341 * (0.1 - 1.8), rel=1e-14)
    # the two branches meet continuously at |za| = 0.001 (the jump is O(za))
    lo = L.orc_lcccurrent(v_of(0.000999), 1.0)
    hi = L.orc_lcccurrent(v_of(0.001001), 1.0)
    assert abs(hi / lo - 1) < 2.5e-3
    # inward only: clipped to <= 0, zero beyond the GHK reversal
    for v in np.linspace(-100, 80, 37):
        assert L.orc_lcccurrent(v, 1.0) <= 0.0
    assert L.orc_lcccurrent(200.0, 1.0) == 0.0


def test_ncx_reversal_and_allosteric_relaxation():
    L = orc.lib()
    import ctypes
    frt = 96.485 / 8.314 / 308.0
    nai, cs = 10.0, 0.5e-3          # mM
    # I = 0 where e^{eta za} nai^3 Cext = e^{(eta-1) za} nao^3 cs
    za = math.log(140.0 ** 3 * cs / (nai ** 3 * 1.8))
    vrev = za / frt
    a = ctypes.c_double(1.0)
    i0 = L.orc_ncx(cs, vrev, nai, ctypes.byref(a), 0.0)
    a = ctypes.c_double(1.0)
    i1 = L.orc_ncx(cs, vrev + 20.0, nai, ctypes.byref(a), 0.0)
    a = ctypes.c_double(1.0)
    i2 = L.orc_ncx(cs, vrev - 20.0, nai, ctypes.byref(a), 0.0)
    assert abs(i0) < 1e-12 * abs(i1)
    assert i1 > 0 > i2                 # outward (reverse mode) above, inward below
    # the allosteric factor relaxes to the Hill function 1/(1+(0.3 uM/cs)^3), tau 150 ms
    for cs_ in (1e-4, 3e-4, 1e-3):
        a = ctypes.c_double(0.025)
        for _ in range(200):
            L.orc_ncx(cs_, -80.0, nai, ctypes.byref(a), 15.0)
        assert a.value == pytest.approx(1.0 / (1.0 + (0.0003 / cs_) ** 3), rel=1e-6)


def test_serca_leak_release_zeros():
    L = orc.lib()
    # SERCA: zero when ci/Ki == cnsr/Knsr, sign follows the imbalance
    assert abs(L.orc_uptake(0.123, 1700.0)) < 1e-16
    assert L.orc_uptake(0.5, 700.0) > 0 > L.orc_uptake(0.01, 1500.0)
    # leak: zero at equal concentrations; half-activation at cnsr = 500
    assert L.orc_leak(300.0, 300.0) == 0.0
    assert L.orc_leak(500.0, 0.0) == pytest.approx(0.00001035 * 500.0 / 2.0, rel=1e-14)
    # release: zero at cjsr == cp, linear in po
    assert L.orc_release(0.3, 500.0, 500.0) == 0.0
    assert L.orc_release(0.4, 700.0, 1.0) == pytest.approx(2 * L.orc_release(0.2, 700.0, 1.0), rel=1e-14)


def test_membrane_currents_vanish_at_their_reversal():
    frt = 96.485 / 8.314 / 308.0
    ek = math.log(5.4 / 140.0) / frt
    assert ek == pytest.approx(-86.394, abs=1e-3)
    st = np.array([ek, 0.001, 1.0, 1.0, 0.3, 0.1, 0.1, 0.2, 0.01, 1.0, 0.02, 0.8, 10.0])
    s, cur = orc.membrane_probe(st, 0.0, 0.0, 100.0, 1000.0, 45.0, 0.025, 1.0, 0.0)
    # xikr, xik1, xitos, xitof are (v - ek) times conductances
    for k in (5, 6, 8, 9):
        assert abs(cur[k]) < 1e-12, k
    ena = math.log(140.0 / 10.0) / frt
    st2 = st.copy()
    st2[0] = ena
    _, cur2 = orc.membrane_probe(st2, 0.0, 0.0, 100.0, 1000.0, 45.0, 0.025, 1.0, 0.0)
    assert abs(cur2[3]) < 1e-11            # xina
    # initial [Na]i closed form (cuda_Cell_Detailed.cu:353): 6.1 at CL = 4000
    o = OracleCell(Params.default(), CellDesc.default(nx=2, ny=2, nz=2, cl=4000.0))
    assert o.scalars()[13] == pytest.approx(6.1, rel=1e-15)


def test_membrane_gates_relax_to_boltzmann_steady_states():
    """With a huge dt the Rush-Larsen gates land on their steady states."""
    v = -20.0
    st = np.array([v, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5, 0.2, 0.5, 0.5, 0.5, 0.5, 10.0])
    s, _ = orc.membrane_probe(st, 0.0, 0.0, 100.0, 1000.0, 45.0, 1e7, 0.0, 0.0)
    assert s[4] == pytest.approx(1.0 / (1.0 + math.exp(-(v + 50.0) / 7.5)), rel=1e-12)   # xr
    assert s[5] == pytest.approx(1.0 / (1.0 + math.exp(-(v - 1.5) / 16.7)), rel=1e-12)   # xs1
    assert s[8] == pytest.approx(1.0 / (1.0 + math.exp(-(v + 3.0) / 15.0)), rel=1e-12)   # xtos
    assert s[2] == pytest.approx(0.0, abs=1e-12)          # h inactivated above -40 mV (ah = 0)


def test_stencil_equals_dense_anisotropic_laplacian():
    """Zero-flux (mirror-ghost) 7-point coupling on the 2x-refined grid of a
    3x3x3 lattice equals a dense Neumann Laplacian built from scratch."""
    p = Params.default()
    d = CellDesc.default(nx=3, ny=3, nz=3)
    X, Y, Z = 3, 3, 3
    tt, tl = 1.8 * 0.7, 6.0 * 0.7
    rng = np.random.default_rng(5)
    fine = rng.uniform(100, 900, (2 * X, 2 * Y, 2 * Z))
    field = np.zeros((8, X * Y * Z))
    for ii in range(8):
        for k in range(Z):
            for j in range(Y):
                for i in range(X):
                    field[ii, i + X * (j + Y * k)] = fine[2 * i + (ii >> 2 & 1), 2 * j + (ii >> 1 & 1), 2 * k + (ii & 1)]
    out = orc.stencil(p, d, field, tt, tl)
    n = fine.size
    idx = np.arange(n).reshape(fine.shape)
    Lm = np.zeros((n, n))
    for (fx, fy, fz), a in np.ndenumerate(idx):
        for dx, dy, dz, tau in [(1, 0, 0, tl), (-1, 0, 0, tl), (0, 1, 0, tt), (0, -1, 0, tt), (0, 0, 1, tt),
                                (0, 0, -1, tt)]:
            gx, gy, gz = fx + dx, fy + dy, fz + dz
            if 0 <= gx < 2 * X and 0 <= gy < 2 * Y and 0 <= gz < 2 * Z:
                Lm[a, idx[gx, gy, gz]] += 1.0 / tau
                Lm[a, a] -= 1.0 / tau
    lap = (Lm @ fine.reshape(-1)).reshape(fine.shape)
    for ii in range(8):
        for k in range(Z):
            for j in range(Y):
                for i in range(X):
                    got = out[ii, i + X * (j + Y * k)]
                    ref = lap[2 * i + (ii >> 2 & 1), 2 * j + (ii >> 1 & 1), 2 * k + (ii & 1)]
                    assert got == pytest.approx(ref, rel=1e-12, abs=1e-10)
    assert abs(out.sum()) < 1e-9 * np.abs(out).sum()       # mirror BCs conserve mass
    const = orc.stencil(p, d, np.full((8, 27), 740.0), tt, tl)
    assert np.all(const == 0.0)


@pytest.mark.parametrize("n", [1, 7, 255, 256, 257, 1000, 20000])
def test_canonical_sum_is_correctly_rounded_or_one_ulp(n):
    rng = np.random.default_rng(n)
    for x in (rng.normal(0, 1, n), rng.uniform(0, 1e-3, n) * 10 ** rng.uniform(-5, 5, n),
              np.concatenate([rng.normal(0, 1e8, n // 2 + 1), -rng.normal(0, 1e8, n // 2 + 1)])[:n]):
        got = orc.canonical_sum(x)
        ref = math.fsum(x)
        assert abs(got - ref) <= abs(np.spacing(ref)), (got, ref)


def test_scheme_a_conserves_calcium_exactly_up_to_sarcolemmal_flux():
    """Scheme A moves Ca between compartments with the same operands on both
    sides, so the change of Total_Ca over a step (cuda_Cell_Detailed.cu:
    1031-1035) equals dt * (Vs*sum I_NCX - (1+gamma) Vp * sum I_CaL) / N, the
    only sarcolemmal fluxes.  A dropped term or a wrong volume factor breaks
    this at once."""
    p = Params.default(scheme=SCHEME_A, steps_per_output=1, stim_offset=1.0)
    d = CellDesc.default(nx=6, ny=4, nz=3, cl=1000.0, seed=3)
    o = OracleCell(p, d)
    N = d.ncru
    cm = 45.0 * N / 16120.0
    o.step(1)
    prev = o.diag()[0][44]
    checked = 0
    for step in range(1, 160):
        c0 = o.counters().copy()
        o.step(1)
        row, _ = o.diag()
        sc = o.scalars()
        if np.array_equal(o.counters(), c0):
            flux = p.dt * (sc[16] * cm / 0.0965 - sc[15] * cm / (0.0965 * 2.0)) / N
            assert row[44] - prev == pytest.approx(flux, rel=1e-6, abs=1e-12), step
            checked += 1
        prev = row[44]
    assert checked > 100


def test_output_row_columns_recomputed_from_the_state():
    """The output row (cuda_Cell_Detailed.cu:839-962, :1031-1072) aggregates
    the state the step leaves behind: every column that is a mean, a count or
    a histogram of exported fields is recomputed here from those fields with
    numpy, and the LCC histogram from the gate bits of each channel's state
    (bit0 c closed, bit1 d closed, bit2 f inactivated, bit3 fCa inactivated;
    reesica.h:65-113) instead of the reference's 16-way table (:904-923)."""
    p = Params.default(stim_offset=1.0)
    d = CellDesc.default(nx=9, ny=6, nz=4, cl=1000.0, seed=3)
    o = OracleCell(p, d)
    o.step(301)                                   # the last step (300) is an output step
    row, st = o.diag()
    assert st == 300 and row[0] == 300 * p.dt
    fld = {f: o.field(f) for f in ("CI", "CNSR", "CATF", "CATS", "CACAL", "CASR", "CAMYO", "CS", "CP", "CJSR",
                                   "CJSRB", "ANCX", "CASAR", "CASARH", "CASARJ", "CASARHJ", "NOU", "NCU", "NOB",
                                   "NCB", "LCC", "NSPARK")}
    m8 = {f: fld[f].mean() for f in ("CI", "CNSR", "CATF", "CATS", "CACAL", "CASR", "CAMYO")}
    m1 = {f: fld[f].mean() for f in ("CS", "CP", "CJSR", "CJSRB", "ANCX", "CASAR", "CASARH", "CASARJ", "CASARHJ")}
    close = dict(rel=1e-12, abs=0)
    assert row[1] == pytest.approx(m8["CI"], **close)
    assert row[5] == pytest.approx(m1["CP"], **close)
    assert row[6] == pytest.approx(m1["CS"], **close)
    assert row[7] == pytest.approx(m1["CJSR"], **close)
    assert row[8] == pytest.approx(m8["CNSR"], **close)
    assert row[9] == pytest.approx(((fld["NOU"] + fld["NOB"]) / 100.0).mean(), **close)
    assert row[20] == pytest.approx(fld["NCB"].mean() / 100.0, **close)
    assert row[21] == pytest.approx(m8["CACAL"], **close)
    assert row[22] == pytest.approx(m8["CATF"], **close)
    assert [row[26], row[27], row[28], row[29]] == [fld[f].sum() for f in ("NOU", "NOB", "NCU", "NCB")]
    lcc = fld["LCC"].astype(np.int64).ravel()
    lcc = lcc[lcc < 16]                           # 16 marks an absent channel slot
    b = [(lcc >> q) & 1 for q in range(4)]
    hist = [b[0].sum(), b[1].sum(), b[2].sum(), b[3].sum(), (b[2] & (1 - b[3])).sum(),
            (lcc <= 1).sum(), (lcc == 0).sum()]
    assert list(row[30:37]) == hist
    assert row[37] == pytest.approx(m1["ANCX"], **close)
    assert row[40] == fld["NSPARK"].sum()
    assert row[41] == pytest.approx(m8["CATS"], **close)
    assert row[42] == pytest.approx(m1["CJSRB"], **close)
    # whole-cell calcium content: cytosol (8 sub-voxels of Vi each), sub-membrane,
    # cleft, NSR (8 of Vnsr), JSR (free + CSQN-bound), volumes of :66-69,95
    Vi, Vs, Vp, Vnsr, Vjsr = 0.5 / 8.0, 0.025, 0.00126, 0.025 / 8.0, 0.02
    cyt = (m8["CI"] + m8["CATF"] + m8["CATS"] + m8["CASR"] + m8["CAMYO"] + m8["CACAL"]) * Vi * 8
    tot_ci = cyt + (m1["CS"] + m1["CASAR"] + m1["CASARH"]) * Vs + (m1["CP"] + m1["CASARJ"] + m1["CASARHJ"]) * Vp
    assert row[43] == pytest.approx(tot_ci, **close)
    assert row[44] == pytest.approx(tot_ci + m8["CNSR"] * Vnsr * 8 + (m1["CJSR"] + m1["CJSRB"]) * Vjsr, **close)


@pytest.mark.parametrize("scheme", [0, 1])
def test_uniform_equilibrium_is_a_fixed_point_of_both_schemes(scheme):
    """With every compartment at one concentration c, every buffer at its own
    steady state for c (myosin with Mg 500 uM competing), and the membrane
    sources off (no LCC, no NCX, no SERCA, no stimulus), no calcium flux is
    left in either cleft scheme (cuda_Cell_Detailed.cu:1356-1473 and
    :1475-1597): the calcium state must stay where it is.  A wrong buffer
    constant, a mis-scaled volume ratio on a non-difference term or a flux
    applied twice moves it."""
    c = 0.3
    p = Params.default(scheme=scheme, tapsi=1, lqt=0, nlcc=0, stim_amp=0.0)   # tapsi+lqt would keep SERCA at 0.05
    d = CellDesc.default(nx=4, ny=3, nz=3, ncx_scale=0.0, seed=9)
    o = OracleCell(p, d)
    n = d.ncru
    ss = lambda kon, koff, B: kon * c * B / (kon * c + koff)      # simple buffer: kon c (B-b) = koff b
    free = 140.0 / (1 + 13.8 * c / 0.46 + 0.0157 * 500.0 / 0.057)  # myosin: Ca and Mg sites share B = 140
    values8 = {"CI": c, "CNSR": c, "CATF": ss(0.0327, 0.0196, 70.0), "CATS": ss(0.00254, 0.000033, 134.0),
               "CACAL": ss(0.0543, 0.238, 24.0), "CASR": ss(0.1, 0.06, 19.0),
               "CAMYO": 13.8 * c * free / 0.46, "MGMYO": 0.0157 * 500.0 * free / 0.057}
    vr = (0.5 / 8.0) * 8.0 / 0.025          # sarcolemmal sites scale with Vcyt/Vsub (buffers.c:118-155)
    values1 = {"CS": c, "CP": c, "CJSR": c, "CASAR": ss(0.1, 1.3, 42.0 * vr), "CASARH": ss(0.1, 0.03, 15.0 * vr),
               "CASARJ": ss(0.1, 1.3, 42.0 * vr), "CASARHJ": ss(0.1, 0.03, 15.0 * vr)}
    for f, v in values8.items():
        o.set_field(f, np.full(8 * n, v))
    for f, v in values1.items():
        o.set_field(f, np.full(n, v))
    before = {f: o.field(f).copy() for f in list(values8) + list(values1) + ["CJSRB"]}
    o.step(200)
    for f, v in before.items():
        assert np.allclose(o.field(f), v, rtol=1e-11, atol=0), f
