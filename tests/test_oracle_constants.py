"""The oracle's constants and leaf formulas against the reference's own.

tests/golden/make_reference_literals.py read the reference sources as text
and wrote tests/golden/reference_literals.json: the value of every numeric
`#define` of the published build (run.sh `make il`: iso2 + lqt) and, per
hot-path function, its numeric literals in order (numbers only).  The tests
here check, on the CPU:

  * every `#define O_*` constant of oracle/ryr_oracle.c equals the reference
    define of the same name;
  * the literals of each restated function, in order, equal the reference's,
    up to the differences listed in ALLOWED (dead assignments the oracle
    leaves out; unit scale macros it spells as 1.0), so a mistyped, dropped or
    swapped constant fails;
  * the formulas, evaluated from the fixture's values by a transcription
    of the reference written here, against the oracle's functions: the LCC
    gate rates read off orc_lcc_markov's thresholds at several voltages in
    both iso2 settings (with the taudin limit branch and the d-gate midpoint),
    the NCX, SERCA and leak magnitudes, the membrane ICs and every membrane
    current and gate update;
  * scheme B's exact calcium balance: the sarcolemmal flux plus the term its
    mixed old/new operands leave over.
"""
import difflib
import json
import math
import os
import re
import sys

import numpy as np
import pytest

from ryr_hyperactivity_amd import CellDesc, Params, SCHEME_B
from oracle import binding as orc
from oracle.binding import OracleCell

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(HERE, "golden"))
from make_reference_literals import literals, strip_comments  # noqa: E402

FIX = json.load(open(os.path.join(HERE, "golden", "reference_literals.json")))
DEF = {k: v["value"] for k, v in FIX["defines"].items()}
LIT = {k: [float(x) for x in v["literals"]] for k, v in FIX["regions"].items()}
ORACLE_SRC = strip_comments(open(os.path.join(ROOT, "oracle", "ryr_oracle.c")).read())

# oracle function(s) restating each fixture region
REGION_FUNCS = {"ncx": ["orc_ncx_v"], "uptake": ["orc_uptake"], "leak": ["orc_leak"], "luminal": ["orc_luminal"],
                "luminal_ss": ["orc_luminal_SS"], "release": ["orc_release"], "lcccurrent": ["orc_lcccurrent"],
                "lcc_markov": ["orc_lcc_markov"], "membrane": ["orc_membrane"]}
# differences from the reference that the oracle makes on purpose
ALLOWED = {
    # subroutine_cuda.c:60 `Ka = 0.025*12.` is overwritten on the next line
    "ncx": {"delete": [[0.025, 12.0]], "insert_ones": 0},
    # reesica.h:62 `betafca = .0012/2.*cpx*cpx` is overwritten on the next line
    "lcc_markov": {"delete": [[0.0012, 2.0]], "insert_ones": 0},
    # the unit scale macros sviks, sviks_iso, svk1, svtos, svtof, svnak (all 1.0 in the
    # published build) are written as 1.0 where the reference names them
    "membrane": {"delete": [], "insert_ones": 6},
}


def _body(name):
    m = re.search(r"\b" + name + r"\s*\([^;{]*\)\s*\{", ORACLE_SRC)
    assert m, name
    i, depth = m.end(), 1
    while depth:
        depth += (ORACLE_SRC[i] == "{") - (ORACLE_SRC[i] == "}")
        i += 1
    # array subscripts (state slots) are not constants of the model
    return re.sub(r"\[\s*\d+\s*\]", "[]", ORACLE_SRC[m.end():i - 1]).split("\n")


def test_fixture_was_made_for_the_published_build():
    assert {"iso2", "lqt"} <= set(FIX["switches"])
    assert DEF["svikr"] == 0.0 and DEF["sv_iup"] == 1.75 and DEF["tauu"] == 220.0


def test_oracle_defines_equal_the_reference_defines():
    defs = dict(re.findall(r"#define\s+O_(\w+)\s+(.+)", ORACLE_SRC))
    norm = {k.lower().replace("_", ""): k for k in DEF}
    vals = {}
    for name, expr in defs.items():
        e = expr.strip()
        for ref in sorted(vals, key=len, reverse=True):
            e = re.sub(rf"\bO_{ref}\b", repr(vals[ref]), e)
        vals[name] = float(eval(e, {"__builtins__": {}}))
        key = name.lower().replace("_", "")
        assert key in norm, f"O_{name} has no reference define"
        assert vals[name] == DEF[norm[key]], (name, vals[name], DEF[norm[key]])
    assert len(vals) >= 60


@pytest.mark.parametrize("region", sorted(REGION_FUNCS))
def test_oracle_literals_follow_the_reference_in_order(region):
    ref = LIT[region]
    got = [float(x) for f in REGION_FUNCS[region] for x in literals(_body(f))]
    allow = ALLOWED.get(region, {"delete": [], "insert_ones": 0})
    deletes, ones = [], 0
    for tag, i1, i2, j1, j2 in difflib.SequenceMatcher(a=ref, b=got, autojunk=False).get_opcodes():
        if tag == "equal":
            continue
        assert tag in ("delete", "insert"), (region, tag, ref[i1:i2], got[j1:j2])
        if tag == "delete":
            deletes.append(ref[i1:i2])
        else:
            assert all(x == 1.0 for x in got[j1:j2]), (region, got[j1:j2])
            ones += j2 - j1
    assert deletes == allow["delete"], (region, deletes)
    assert ones == allow["insert_ones"], (region, ones)


# ---- LCC gates (reesica.h:14-116) ------------------------------------------------------------------------
def _lcc_params(iso2):
    L = LIT["lcc_markov"]
    dv5, dvk, fv5, fvk, alphac, betac = L[0], L[1], -L[2], L[3], L[4], L[5]
    if iso2:
        betac, dv5, fv5, fvk = L[6], L[7], -L[8], L[9]
    return dv5, dvk, fv5, fvk, alphac, betac


def _lcc_rates(v, cpx, iso2):
    """reesica.h:37-63, transcribed here."""
    dv5, dvk, fv5, fvk, alphac, betac = _lcc_params(iso2)
    dinf = 1.0 / (1.0 + math.exp(-(v - dv5) / dvk))
    if -0.0001 < v < 0.0001:
        taudin = 0.035 * dvk / dinf
    else:
        taudin = 1.0 / ((1.0 - math.exp(-(v - dv5) / dvk)) / (0.035 * (v - dv5)) * dinf)
    finf = 1.0 - 1.0 / (1.0 + math.exp(-(v - fv5) / fvk)) / (1.0 + math.exp((v - 60.0) / 12.0))
    taufin = 0.02 - 0.007 * math.exp(-(0.0337 * (v + 10.5)) ** 2)
    return {"alphac": alphac, "betac": betac, "alphad": dinf * taudin, "betad": (1 - dinf) * taudin,
            "alphaf": finf * taufin, "betaf": (1 - finf) * taufin, "alphafca": 0.012 / 2.0,
            "betafca": 0.175 / (1 + (25.0 / cpx) ** 2), "dinf": dinf}


def _walk_rates(state, cpx, v, dt, iso2):
    """The four rates of the walk from `state`, in gate order c, d, f, fCa,
    from the thresholds between successive outcomes."""
    L = orc.lib()
    outs = [state ^ (1 << b) for b in range(4)] + [state]
    cum = []
    for b in range(4):
        lo, hi = 0.0, 1.0
        for _ in range(90):                 # first u whose outcome lies past gate b
            mid = 0.5 * (lo + hi)
            r = L.orc_lcc_markov(state, cpx, v, dt, mid, iso2)
            if outs.index(r) > b:
                hi = mid
            else:
                lo = mid
        cum.append(hi / dt)
    return [cum[0]] + [cum[b] - cum[b - 1] for b in range(1, 4)]


@pytest.mark.parametrize("iso2", [0, 1])
@pytest.mark.parametrize("v", [-70.0, -30.0, 0.00005, 20.0, 45.0])
def test_lcc_gate_rates_follow_the_reference_formulas(v, iso2):
    cpx, dt = 12.0, 0.05
    ref = _lcc_rates(v, cpx, iso2)
    closed = _walk_rates(15, cpx, v, dt, iso2)        # every gate set: the clearing rates
    opened = _walk_rates(0, cpx, v, dt, iso2)         # every gate clear: the setting rates
    for got, name in zip(closed, ("alphac", "alphad", "alphaf", "alphafca")):
        assert got == pytest.approx(ref[name], rel=1e-9, abs=1e-12), (name, v, iso2)
    for got, name in zip(opened, ("betac", "betad", "betaf", "betafca")):
        assert got == pytest.approx(ref[name], rel=1e-9, abs=1e-12), (name, v, iso2)


def test_lcc_d_gate_midpoint_and_limit_branch():
    cpx, dt = 12.0, 0.05
    # iso2: dv5 = 0, so the |v| < 1e-4 branch is the limit of taudin at the midpoint, taudin = 0.035 dvk / dinf:
    # alphad = dinf taudin = 0.035 dvk exactly, betad the same to first order, their sum taudin = 2 x 0.035 dvk
    a = _walk_rates(15, cpx, 0.00005, dt, 1)[1]
    b = _walk_rates(0, cpx, 0.00005, dt, 1)[1]
    assert a == pytest.approx(0.035 * 8.0, rel=1e-12) and b == pytest.approx(a, rel=1e-4)
    assert a + b == pytest.approx(2 * 0.035 * 8.0, rel=1e-4)
    # without iso2 the midpoint is dv5 = 5 mV: the stationary d occupancy a/(a+b) = dinf(dv5) = 1/2
    a = _walk_rates(15, cpx, 5.0 + 1e-6, dt, 0)[1]
    b = _walk_rates(0, cpx, 5.0 + 1e-6, dt, 0)[1]
    assert a / (a + b) == pytest.approx(0.5, abs=1e-6)
    # and there the |v| < 1e-4 branch fires at V = 0 although the formula is finite: the reference's value
    ref = _lcc_rates(0.00005, cpx, 0)
    assert ref["alphad"] != pytest.approx(ref["dinf"] / ((1.0 - math.exp(-(0.00005 - 5.0) / 8.0)) /
                                                          (0.035 * (0.00005 - 5.0))), rel=1e-3)
    assert _walk_rates(15, cpx, 0.00005, dt, 0)[1] == pytest.approx(ref["alphad"], rel=1e-9)


# ---- NCX, SERCA, leak ---------------------------------------------------------------------------------------
def test_ncx_magnitude_follows_the_reference_formula():
    import ctypes
    D = DEF
    L = orc.lib()
    nao = D["xnao"]
    for cs, v, nai, ka in ((1e-4, -80.0, 8.0, 0.3), (5e-4, 10.0, 10.0, 0.7), (2e-3, 40.0, 12.0, 0.05),
                           (1e-2, -20.0, 6.1, 1.0)):
        za = v * D["Farad"] / D["xR"] / D["Temper"]
        t1 = D["Kmcai"] * nao ** 3 * (1.0 + (nai / D["Kmnai"]) ** 3)
        t2 = D["Kmnao"] ** 3 * cs * (1.0 + cs / D["Kmcai"])
        t3 = (D["Kmcao"] + D["Cext"]) * nai ** 3 + cs * nao ** 3
        ref = ka * D["vnaca"] * (math.exp(D["eta"] * za) * nai ** 3 * D["Cext"]
                                 - math.exp((D["eta"] - 1.0) * za) * nao ** 3 * cs) \
            / ((t1 + t2 + t3) * (1.0 + D["ksat"] * math.exp((D["eta"] - 1.0) * za)))
        a = ctypes.c_double(ka)
        got = L.orc_ncx(cs, v, nai, ctypes.byref(a), 0.0)
        assert got == pytest.approx(ref, rel=1e-12)
        n = LIT["ncx"]        # allosteric relaxation: Kda 0.0003 mM, tau 150 ms
        a = ctypes.c_double(ka)
        L.orc_ncx(cs, v, nai, ctypes.byref(a), 0.4)
        assert a.value == pytest.approx(ka + 0.4 * (1.0 / (1.0 + (n[-2] / cs) ** 3) - ka) / n[-1], rel=1e-13)


def test_serca_and_leak_follow_the_reference_formulas():
    L = orc.lib()
    vup, Ki, Knsr, HH, upfactor = LIT["uptake"][:5]
    gleak, Kjsr = LIT["leak"]
    for ci, cnsr in ((0.05, 400.0), (0.1, 740.0), (0.5, 1000.0), (2.0, 300.0), (0.123, 1700.0)):
        a, b = (ci / Ki) ** HH, (cnsr / Knsr) ** HH
        assert L.orc_uptake(ci, cnsr) == pytest.approx(upfactor * vup * (a - b) / (1.0 + a + b), rel=1e-12,
                                                       abs=1e-15)
        assert L.orc_leak(cnsr, ci) == pytest.approx(gleak * (cnsr - ci) * cnsr ** 2 / (cnsr ** 2 + Kjsr ** 2),
                                                     rel=1e-13)


# ---- membrane -----------------------------------------------------------------------------------------------
def test_membrane_initial_conditions_follow_the_reference():
    nai, n9, n5, v80, xm, xh, xj, xr, xs1, qks, xtos, ytos, xtof, ytof = LIT["membrane_ics"]
    for cl in (1000.0, 4000.0):
        o = OracleCell(Params.default(v0=-v80), CellDesc.default(nx=2, ny=2, nz=2, cl=cl))
        s = o.scalars()
        assert list(s[1:14]) == pytest.approx([-v80, xm, xh, xj, xr, xs1, xs1, qks, xtos, ytos, xtof, ytof,
                                              nai * (n9 / (cl + n5))], rel=1e-15)


def _membrane_reference(st, sum_icat, sum_inaca, sum_cs, ncru, cm, dt, svikr, stim):
    """cuda_Cell_Detailed.cu:452-456 and 458-731, transcribed here (lqt off
    through svikr, no slowncxchange); returns the new state and currents."""
    D = DEF
    v, xm, xh, xj, xr, xs1, xs2, qks, xtos, ytos, xtof, ytof, xnai = st
    xki, xko, xnao, frt = 140.0, 5.40, D["xnao"], D["frt"]
    ex = math.exp
    csubt = sum_cs / ncru
    xicatto = sum_icat / cm * 0.0965 * D["Vp"] * 2.0 * (D["icagamma"] + 1.0)
    xinacato = sum_inaca / cm * 0.0965 * D["Vs"]
    ena = (1.0 / frt) * math.log(xnao / xnai)
    am = 0.32 * (v + 47.13) / (1.0 - ex(-0.1 * (v + 47.13)))
    bm = 0.08 * ex(-v / 11.0)
    if v < -40.0:
        ah = 0.135 * ex((80.0 + v) / (-6.8))
        bh = 3.56 * ex(0.079 * v) + 310000.0 * ex(0.35 * v)
        aj = (-127140.0 * ex(0.2444 * v) - 0.00003474 * ex(-0.04391 * v)) * ((v + 37.78) / (1.0 + ex(0.311 * (v + 79.23))))
        bj = (0.1212 * ex(-0.01052 * v)) / (1.0 + ex(-0.1378 * (v + 40.14)))
    else:
        ah, aj = 0.0, 0.0
        bh = 1.0 / (0.130 * (1.0 + ex((v + 10.66) / (-11.10))))
        bj = (0.3 * ex(-0.00000025350 * v)) / (1.0 + ex(-0.10 * (v + 32.00)))
    xina = 15 * xh * xj * xm ** 3 * (v - ena) + 0.3e-3 * 5 * (v - ena)
    xh_n = ah / (ah + bh) - (ah / (ah + bh) - xh) * ex(-dt * (ah + bh))
    xj_n = aj / (aj + bj) - (aj / (aj + bj) - xj) * ex(-dt * (aj + bj))
    xm_n = am / (am + bm) - (am / (am + bm) - xm) * ex(-dt * (am + bm))
    ek = (1.0 / frt) * math.log(xko / xki)
    gss = math.sqrt(xko / 5.40)
    xkrv1 = 0.001380 * (v + 7.00) / (1.0 - ex(-0.123 * (v + 7.00)))
    xkrv2 = 0.000610 * (v + 10.00) / (ex(0.1450 * (v + 10.00)) - 1.00)
    xkrinf = 1.0 / (1.0 + ex(-(v + 50.0) / 7.50))
    rg = 1.0 / (1.0 + ex((v + 33.0) / 22.40))
    xikr = svikr * 0.0078360 * gss * xr * rg * (v - ek)
    xr_n = xkrinf - (xkrinf - xr) * ex(-dt * (xkrv1 + xkrv2))
    prnak = 0.0183300
    qks_inf = 0.2 * (1 + 0.8 / (1 + (0.28 / csubt) ** 3))
    eks = (1.0 / frt) * math.log((xko + prnak * xnao) / (xki + prnak * xnai))
    xs1ss = 1.0 / (1.0 + ex(-(v - 1.500) / 16.700))
    tauxs = D["svtauxs"] / (0.0000719 * (v + 30.0) / (1.0 - ex(-0.1480 * (v + 30.0)))
                            + 0.0001310 * (v + 30.0) / (ex(0.06870 * (v + 30.0)) - 1.0))
    xiks = 0.2000 * qks * xs1 * xs2 * (v - eks)
    xs1_n = xs1ss - (xs1ss - xs1) * ex(-dt / tauxs)
    xs2_n = xs1ss - (xs1ss - xs2) * ex(-dt / tauxs)
    qks_n = qks + dt * (qks_inf - qks) / 1000.0
    gki = 0.600 * math.sqrt(xko / 5.4)
    aki = 1.02 / (1.0 + ex(0.2385 * (v - ek - 59.215)))
    bki = (0.49124 * ex(0.08032 * (v - ek + 5.476)) + ex(0.061750 * (v - ek - 594.31))) / \
        (1.0 + ex(-0.5143 * (v - ek + 4.753)))
    xik1 = gki * aki / (aki + bki) * (v - ek)
    rt1, rt2, rt3 = -(v + 3.0) / 15.0, (v + 33.5) / 10.0, (v + 60.0) / 10.0
    xtos_inf, ytos_inf = 1.0 / (1.0 + ex(rt1)), 1.0 / (1.0 + ex(rt2))
    txs = 9.0 / (1.0 + ex(-rt1)) + 0.50
    tys = 3000.0 / (1.0 + ex(rt3)) + 30.0
    xitos = 0.040 * xtos * (ytos + 0.50 * ytos_inf) * (v - ek)
    xtos_n = xtos_inf - (xtos_inf - xtos) * ex(-dt / txs)
    ytos_n = ytos_inf - (ytos_inf - ytos) * ex(-dt / tys)
    txf = 3.50 * ex(-(v / 30.0) ** 2) + 1.50
    tyf = 20.0 / (1.0 + ex((v + 33.50) / 10.0)) + 20.0
    xitof = 0.10 * xtof * ytof * (v - ek)
    xtof_n = xtos_inf - (xtos_inf - xtof) * ex(-dt / txf)
    ytof_n = ytos_inf - (ytos_inf - ytof) * ex(-dt / tyf)
    sigma = (ex(xnao / 67.30) - 1.0) / 7.0
    fnak = 1.0 / (1.0 + 0.1245 * ex(-0.1 * v * frt) + 0.0365 * sigma * ex(-v * frt))
    xinak = 1.5000 * fnak * (1.0 / (1.0 + (12.00 / xnai) ** 1.0)) * xko / (xko + 1.50)
    xrr = (1.0 / (1.0 / 50.0) / (0.18 * 12500)) / 1000.0
    xnai_n = xnai + dt * (-xrr * (xina + 3.0 * xinak + 3.0 * xinacato))
    v_n = v + dt * (-(xina + xik1 + xikr + xiks + xitos + xitof + xinacato + xicatto + xinak) + stim)
    new = [v_n, xm_n, xh_n, xj_n, xr_n, xs1_n, xs2_n, qks_n, xtos_n, ytos_n, xtof_n, ytof_n, xnai_n]
    cur = [csubt, xicatto, xinacato, xina, xiks, xikr, xik1, xinak, xitos, xitof]
    return np.array(new), np.array(cur)


STATES = [[-86.0, 0.001, 1.0, 1.0, 0.0, 0.0843, 0.0843, 0.2, 0.01, 1.0, 0.02, 0.8, 6.1],     # rest
          [-55.0, 0.05, 0.8, 0.9, 0.01, 0.1, 0.12, 0.25, 0.05, 0.9, 0.1, 0.7, 8.0],       # V < -40 branch
          [-12.0, 0.9, 0.3, 0.5, 0.2, 0.3, 0.25, 0.4, 0.6, 0.4, 0.7, 0.3, 9.5],           # plateau
          [35.0, 0.99, 0.01, 0.05, 0.4, 0.5, 0.45, 0.5, 0.9, 0.2, 0.95, 0.1, 12.0]]       # peak


@pytest.mark.parametrize("st", STATES, ids=["rest", "low", "plateau", "peak"])
def test_membrane_currents_and_gates_follow_the_reference(st):
    args = (-3.2e3, 1.5e2, 2.4e3, 20000.0, 45.0 * 20000 / 16120.0, 0.025)
    for ikr_scale, stim in ((1.0, 0.0), (0.4, 80.0)):
        s, cur = orc.membrane_probe(np.array(st), *args, ikr_scale, stim)
        rs, rcur = _membrane_reference(st, *args, 3.0 * ikr_scale, stim)
        assert s == pytest.approx(rs, rel=1e-11, abs=1e-14)
        assert cur == pytest.approx(rcur, rel=1e-11, abs=1e-14)


# ---- scheme B balance -------------------------------------------------------------------------------------------
def test_scheme_b_calcium_change_is_the_sarcolemmal_flux_plus_its_exchange_lag():
    """Scheme B (cuda_Cell_Detailed.cu:1475-1597) exchanges cp and cs with
    the step-start cs on the cp side (:1515) and the sub-step cs on the cs
    side (:1503), and diffuses cs in z from step-start neighbours against
    the sub-step value (:1496).  Every other exchange uses one operand on
    both sides.  So per step Total_Ca (:1031-1035) changes by the
    sarcolemmal flux plus (Vp/taups + 2 Vs/taust) dt/5 sum_CRU sum_m
    (cs_start - cs_m), the lag the oracle records (orc_read_scheme_b_lag).
    A wrong operand in either exchange, a dropped term or a wrong volume
    factor breaks the identity; the lag term is a material part of it."""
    D = DEF
    p = Params.default(scheme=SCHEME_B, steps_per_output=1, stim_offset=1.0)
    d = CellDesc.default(nx=6, ny=4, nz=3, cl=1000.0, seed=3)
    o = OracleCell(p, d)
    N = d.ncru
    cm = 45.0 * N / 16120.0
    kappa = D["Vp"] / D["taups"] + 2.0 * D["Vs"] / D["taust"]
    o.step(1)
    prev = o.diag()[0][44]
    checked, material = 0, 0
    for step in range(1, 200):
        c0 = o.counters().copy()
        o.step(1)
        row, _ = o.diag()
        sc = o.scalars()
        if np.array_equal(o.counters(), c0):
            flux = p.dt * (sc[16] * cm / 0.0965 - sc[15] * cm / (0.0965 * 2.0)) / N
            lag = kappa * (p.dt / 5.0) * o.scheme_b_lag().sum() / N
            resid = row[44] - prev - flux - lag
            assert abs(resid) <= 1e-12 + 1e-9 * abs(flux + lag), (step, resid, flux, lag)
            material += abs(lag) > 1e4 * abs(resid) and abs(lag) > 1e-9
            checked += 1
        prev = row[44]
    assert checked > 150 and material > 100
