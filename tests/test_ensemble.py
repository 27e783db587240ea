"""The uneven ensemble's cell generator and per-group summary (CPU), and a
short GPU run of a few cells to completed beats."""
import numpy as np
import pytest

from ryr_hyperactivity_amd import Params
from ryr_hyperactivity_amd._abi import BEAT_COLS, BEAT_NAMES
from ryr_hyperactivity_amd.ensemble import (CYCLE_LENGTHS, LATTICES, ensemble_steps, run, summarize,
                                            uneven_cells)


def test_uneven_cells_draws():
    cells = uneven_cells(256, 2024)
    assert len(cells) == 256
    assert [c.seed for c in cells] == list(range(256))
    assert {(c.nx, c.ny, c.nz) for c in cells} == set(LATTICES)
    assert {c.cl for c in cells} == set(CYCLE_LENGTHS)
    assert all(1.0 <= c.Ku2 <= 3.0 and 0.0 <= c.ikr_scale <= 1.0 for c in cells)
    again = uneven_cells(256, 2024)
    assert all(a.Ku2 == b.Ku2 and a.ikr_scale == b.ikr_scale and a.cl == b.cl and a.nx == b.nx
               for a, b in zip(cells, again))
    # the first cell's draws, in order: lattice index, Ku2, I_Kr, cycle-length index
    rng = np.random.default_rng(2024)
    li, ku2, ikr, ci = rng.integers(0, 3), rng.uniform(1, 3), rng.uniform(0, 1), rng.integers(0, 3)
    c0 = cells[0]
    assert (c0.nx, c0.ny, c0.nz) == LATTICES[li] and c0.Ku2 == ku2 and c0.ikr_scale == ikr
    assert c0.cl == CYCLE_LENGTHS[ci]


def test_ensemble_steps_cover_the_longest_cycle():
    p = Params.default()
    cells = uneven_cells(8, 3)
    n = ensemble_steps(cells, 5, p)
    assert n * p.dt >= p.stim_offset + 5 * max(c.cl for c in cells)
    assert (n - 3) * p.dt < p.stim_offset + 5 * max(c.cl for c in cells)


def _beats(apds, eads, complete=True):
    b = np.zeros((len(apds) + 1, BEAT_COLS))
    for k, (a, e) in enumerate(zip(apds, eads)):
        b[k, BEAT_NAMES.index("apd90")] = a
        b[k, BEAT_NAMES.index("ead")] = e
        b[k, BEAT_NAMES.index("complete")] = 1.0 if complete else 0.0
    return b                                   # one trailing incomplete row


def test_summary_scores_beats_after_the_first_per_group():
    cells = uneven_cells(3, 11)
    cells[0].cl, cells[0].ikr_scale = 1000.0, 0.1
    cells[1].cl, cells[1].ikr_scale = 1000.0, 0.2
    cells[2].cl, cells[2].ikr_scale = 2000.0, 0.9
    beats = [_beats([900, 400, 420, 440], [1, 0, 1, 0]),
             _beats([800, 500, -1, 540], [0, 1, 2, 1]),      # an unrepolarised beat; an EAD flag above 1
             _beats([700, 300, 310], [1, 0, 0])]
    g = summarize(cells, beats, nbeats=3)
    assert [(x["cl"], x["ikr_third"]) for x in g] == [(1000.0, 0), (2000.0, 2)]
    a, b = g
    assert a["cells"] == 2 and a["min_beats"] == 4 and a["scored_beats"] == 4 and a["unrepolarised"] == 1
    assert a["apd90_mean"] == pytest.approx(np.mean([400, 420, 500]))
    assert a["ead_incidence"] == pytest.approx(np.mean([0, 1, 1, 1]))
    assert b["cells"] == 1 and b["min_beats"] == 3 and b["scored_beats"] == 2
    assert b["ead_incidence"] == 0.0


@pytest.mark.gpu
def test_small_ensemble_completes_its_beats():
    print("case ensemble 6 cells 2 beats", flush=True)
    cells = uneven_cells(6, 5)
    for c in cells:
        c.nx, c.ny, c.nz = 16, 6, 4            # small lattices: 84,000 steps take seconds
        c.cl = 1000.0
    beats, groups, cost = run(cells, nbeats=2)
    for b in beats:
        done = b[b[:, BEAT_NAMES.index("complete")] > 0]
        assert len(done) >= 2
        apd = done[:, BEAT_NAMES.index("apd90")]
        assert np.all((apd > 0) | (apd == -1.0)) and np.all(apd < 1000.0)
    assert sum(g["cells"] for g in groups) == 6
    assert cost["steps"] == ensemble_steps(cells, 2, Params.default(lqt=0)) and cost["gpu_s"] > 0
