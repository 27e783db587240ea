"""Stochastic gating of the oracle (subroutine_cuda.c:251-669, reesica.h:14-116)
against textbook distributions.

- one isolated transition is Knuth's Poisson sampler: its counts follow the
  Poisson pmf and it consumes exactly count + 1 uniforms;
- the dithered Gaussian branch has the binomial mean and variance (+1/3 for
  the floor and the coin);
- the rejection fallback after 200 negative tries returns 0 and is counted;
- the four-state RyR chain satisfies detailed balance (the reference picks the
  OB->OU rate abu*pku/pkb for that), so a population of 100 channels sits at
  a closed-form stationary distribution;
- the LCC chain flips at most one gate per step, with the voltage-independent
  c-gate and fCa-gate rates of reesica.h:27-28,61-63.
"""
import math

import numpy as np
import pytest
from scipy import stats

from oracle import binding as orc

DT = 0.025


def _probe(counts, conc, *, Ku=10.0, Ku2=1.0, Kb=1.0, dt=DT, step=0, seed=1, lcc=None, iso2=1):
    n = counts.shape[0]
    if lcc is None:
        lcc = np.full((n, 8), 16, np.int32)
    return orc.gating_probe(seed, np.arange(n, dtype=np.uint32), step, Ku, Ku2, Kb, dt, iso2, counts, lcc, conc)


def test_knuth_poisson_pmf_and_draw_count():
    n = 40000
    counts = np.zeros((n, 4), np.int32)
    counts[:, 0] = 100                        # all open-unbound: OU->CU at rate 1/ms
    conc = np.tile([1.0, 740.0, 0.1, -80.0], (n, 1))
    out, _, draws, _, fails = _probe(counts, conc, Kb=0.0)   # Kb = 0: pkb = 0, OU->OB off
    k = out[:, 1]                             # closed-unbound after the step = OU->CU count
    assert np.all(out[:, 0] + out[:, 1] == 100)
    lam = 100 * 1.0 * DT
    assert np.all(draws == k + 1 + 7)         # Knuth: count+1 uniforms; 7 empty transitions: 1 each
    assert not fails.any()
    edges = np.arange(0, 11)
    obs = np.array([np.sum(k == e) for e in edges[:-1]] + [np.sum(k >= 10)])
    pmf = stats.poisson.pmf(edges[:-1], lam)
    exp = np.concatenate([pmf, [1 - pmf.sum()]]) * n
    chi2 = ((obs - exp) ** 2 / exp).sum()
    assert chi2 < stats.chi2.ppf(0.999, len(obs) - 1)


def test_gaussian_branch_moments():
    n = 30000
    counts = np.zeros((n, 4), np.int32)
    counts[:, 0] = 50
    conc = np.tile([1.0, 740.0, 0.1, -80.0], (n, 1))
    dt = 0.4                                  # p = 0.4 >= 0.2: the Gaussian approximation
    out, _, draws, _, fails = _probe(counts, conc, Kb=0.0, dt=dt)
    k = out[:, 1].astype(float)
    mu, var = 50 * dt, 50 * dt * (1 - dt)
    assert abs(k.mean() - mu) < 4 * math.sqrt(var / n)
    assert k.var() == pytest.approx(var + 1.0 / 3.0, rel=0.05)
    assert np.all((draws - 7) % 3 == 0)       # u1, u2 and one integer per try
    assert not fails.any()


def test_gaussian_rejection_fallback_is_counted():
    """Feed slots that make every Gaussian try very negative: after 201 tries
    the count is 0 and one fallback is recorded (subroutine_cuda.c:358-361)."""
    near_one = (0xFFFFFFFF << 32) | 0xFFFFF000   # u1 = 1 - 2^-53: sqrt(-2 log(1-u1)) = 8.57
    half = (0x80000000 << 32) | 0x00000000        # u2 ~ 1/2: cos(2 pi u2) = -1
    coin = 0
    slots = np.array([near_one, half, coin] * 201 + [half] * 64, dtype=np.uint64)
    counts = np.array([10, 90, 0, 0], np.int32)    # nou = 10, p = 0.3 -> Gaussian
    new, used, fails = orc.ryrgating_replay(slots, 10.0, 1.0, 0.0, 1.0, 740.0, 0.3, counts)
    assert fails == 1
    assert used == 603 + 7
    assert new[0] == 10                            # no OU->CU moves (the others are tiny)


def test_ryr_population_conserved_at_reference_dt():
    rng = np.random.default_rng(7)
    n = 20000
    counts = np.zeros((n, 4), np.int32)
    nou = rng.integers(0, 60, n)
    nob = rng.integers(0, 100 - nou + 1)
    ncu = rng.integers(0, 100 - nou - nob + 1)
    counts[:, 0], counts[:, 1], counts[:, 2], counts[:, 3] = nou, ncu, nob, 100 - nou - nob - ncu
    conc = np.stack([10 ** rng.uniform(-2, 2.5, n), rng.uniform(100, 1000, n), np.full(n, 0.1),
                     np.full(n, -80.0)], axis=1)
    out, _, _, _, _ = _probe(counts, conc, Ku=3.333, Ku2=2.5, Kb=18.0)
    assert np.all(out >= 0) and np.all(out <= 100)
    assert np.all(out.sum(axis=1) == 100)


def test_ryr_chain_stationary_distribution_detailed_balance():
    """CU<->OU (pku, 1), CU<->CB (aub, abu), CB<->OB (pkb, 1), OU<->OB
    (aub, abu pku/pkb): detailed balance gives pi_OU/pi_CU = pku,
    pi_CB/pi_CU = aub/abu, pi_OB/pi_CU = aub pkb/abu."""
    cp, cjsr, Ku, Ku2, Kb = 20.0, 740.0, 10.0, 1.0, 1.0
    pku = 6.75 * Ku2 / (1 + (5000 / cjsr) ** 2) / (1 + (Ku / cp) ** 2)
    pkb = 6.75 / 36 * Kb / (1 + (5000 / cjsr) ** 2) / (1 + (Ku / cp) ** 2)
    abu = 1 / 220.0
    fb = orc.lib().orc_fracbound(cjsr)           # = 1/(1 + abu/aub)
    aub = abu * fb / (1 - fb)
    w = np.array([1.0, pku, aub / abu, aub * pkb / abu])   # CU, OU, CB, OB
    pi = w / w.sum()
    n = 1000
    start = np.round(100 * pi).astype(int)
    start[0] += 100 - start.sum()
    counts = np.tile([start[1], start[0], start[3], start[2]], (n, 1)).astype(np.int32)  # nou ncu nob ncb
    conc = np.tile([cp, cjsr, 0.1, -80.0], (n, 1))
    acc = np.zeros(4)
    nstep, burn = 3000, 500
    for s in range(nstep):
        counts, _, _, _, _ = _probe(counts, conc, Ku=Ku, Ku2=Ku2, Kb=Kb, step=s)
        if s >= burn:
            acc += counts.mean(axis=0)
    frac = acc / (nstep - burn) / 100.0            # nou ncu nob ncb
    got = np.array([frac[1], frac[0], frac[3], frac[2]])
    assert got[0] == pytest.approx(pi[0], rel=0.02)
    assert got[1] == pytest.approx(pi[1], rel=0.03)
    assert got[2] == pytest.approx(pi[2], rel=0.03)
    assert got[3] == pytest.approx(pi[3], rel=0.15)


def test_lcc_single_flip_and_gate_rates():
    n = 20000
    rng = np.random.default_rng(11)
    counts = np.zeros((n, 4), np.int32)
    conc = np.tile([5.0, 740.0, 5.0, -20.0], (n, 1))      # cpx = 5 uM
    lcc = np.full((n, 8), 16, np.int32)
    lcc[:, :4] = rng.integers(0, 16, (n, 4))
    flips = np.zeros((4, 2))      # per gate: (set->clear, clear->set) transitions
    occ = np.zeros((4, 2))        # per gate: time in (set, clear)
    for s in range(400):
        old = lcc[:, :4].copy()
        _, lcc, draws, nopen, _ = _probe(counts, conc, lcc=lcc, step=s)
        new = lcc[:, :4]
        changed = old ^ new
        assert np.all(new >= 0) and np.all(new <= 15)
        assert np.all((changed & (changed - 1)) == 0)     # at most one bit per channel per step
        assert np.array_equal(nopen, (new == 0).sum(axis=1))
        for b in range(4):
            was = (old >> b) & 1
            now = (new >> b) & 1
            flips[b, 0] += np.sum((was == 1) & (now == 0))
            flips[b, 1] += np.sum((was == 0) & (now == 1))
            occ[b, 0] += np.sum(was == 1)
            occ[b, 1] += np.sum(was == 0)
    rate = flips / occ / DT
    cpx = (5.0 + 5.0 * 0.5) / 1.5
    # c gate (bit 0): clear at alphac 0.22/ms, set at betac = 2/ms under iso2
    assert rate[0, 0] == pytest.approx(0.22, rel=0.05)
    assert rate[0, 1] == pytest.approx(2.0, rel=0.03)
    # fCa gate (bit 3): recovery 0.006/ms, inactivation 0.175/(1+(25/cpx)^2)
    assert rate[3, 0] == pytest.approx(0.006, rel=0.15)
    assert rate[3, 1] == pytest.approx(0.175 / (1 + (25 / cpx) ** 2), rel=0.1)
