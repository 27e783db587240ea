"""GPU parity: the HIP CRU update against the FP64 oracle.

FP64 mode is bit-exact by construction (same Philox draws, same detmath
transcendentals, same operand order, contraction off, same canonical
reduction order), so these tests compare with `==`, not a tolerance:
state fields, membrane scalars, output rows, linescans, counters.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from ryr_hyperactivity_amd import CellDesc, CellEngine, Params, PREC_F32, SCHEME_A, SCHEME_B, FIELDS
from ryr_hyperactivity_amd import format_wholetest_row, format_linescan, gating_probe
from ryr_hyperactivity_amd._abi import S
from oracle.binding import OracleCell
from oracle import binding as orc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _assert_same_state(eng, cell, o, label):
    for f in FIELDS:
        g = eng.field(cell, f)
        r = o.field(f)
        if not np.array_equal(g, r):
            idx = np.argwhere(g != r)
            raise AssertionError(f"{label}: field {f} differs at {len(idx)} places, first {idx[:3].tolist()}: "
                                 f"gpu {g[tuple(idx[0])]!r} oracle {r[tuple(idx[0])]!r}")
    gs = eng.scalars()[cell]
    os_ = o.scalars()
    assert np.array_equal(gs, os_), f"{label}: scalars differ: {np.nonzero(gs != os_)[0]}"


@pytest.mark.parametrize("scheme", [SCHEME_A, SCHEME_B])
@pytest.mark.parametrize("dims", [(7, 5, 3), (20, 13, 5), (1, 3, 4), (6, 1, 2), (3, 2, 1)])
def test_fp64_bitexact_through_an_action_potential(scheme, dims):
    p = Params.default(scheme=scheme, stim_offset=2.0)
    d = CellDesc.default(nx=dims[0], ny=dims[1], nz=dims[2], cl=1000.0, seed=11)
    print(f"case scheme={scheme} dims={dims}", flush=True)
    o = OracleCell(p, d)
    with CellEngine(p, d) as eng:
        for chunk in (1, 99, 100, 37, 163):          # eager steps and graph replays
            eng.step(chunk)
            o.step(chunk)
            _assert_same_state(eng, 0, o, f"after +{chunk}")
            grow, gst = eng.diag()
            orow, ost = o.diag()
            assert gst == ost
            if gst >= 0:
                assert np.array_equal(grow[0], orow), np.nonzero(grow[0] != orow)
                assert format_wholetest_row(grow[0]) == format_wholetest_row(orow)
                assert np.array_equal(eng.linescan(0), o.linescan())
        assert np.array_equal(eng.counters()[0], o.counters())
        # the action potential fired (stimulus at step 80)
        assert eng.scalars()[0][S["V"]] > -60.0


def test_fp64_bitexact_at_full_size_through_upstroke_and_release():
    """The benchmark's own cell (100x20x10 CRUs, Ku2 2.5, I_Kr off) in FP64,
    stimulated at 2 ms: bit-identical to the oracle through 400 ms, i.e. the
    upstroke, the L-type trigger, the release and the plateau of the long
    action potential, with the launch-shape tuning and graph-replayed
    intervals on the way."""
    p = Params.default(stim_offset=2.0)
    d = CellDesc.default(nx=100, ny=20, nz=10, cl=2000.0, Ku2=2.5, ikr_scale=0.0, seed=1)
    print("case 100x20x10 FP64 through 16000 steps", flush=True)
    o = OracleCell(p, d)
    with CellEngine(p, d) as eng:
        done = 0
        for chunk in (1, 99, 1900, 2000, 12000):
            eng.step(chunk)
            o.step(chunk)
            done += chunk
            _assert_same_state(eng, 0, o, f"full size after {done} steps")
            print(f"  {done} steps equal", flush=True)
            if done == 4000:
                assert eng.scalars()[0][S["V"]] > -40.0           # on the plateau at 100 ms
        grow, gst = eng.diag()
        orow, ost = o.diag()
        assert gst == ost and gst > 0
        assert np.array_equal(grow[0], orow), np.nonzero(grow[0] != orow)
        assert np.array_equal(eng.linescan(0), o.linescan())
        assert np.array_equal(eng.counters()[0], o.counters())
        assert np.array_equal(eng.beats(0), o.beats())


def test_fp64_full_batch_cells_match_the_oracle():
    """The benchmark's 256-cell batch (3.07 M CRUs in three lattice sizes,
    per-cell Ku2, I_Kr and cycle length) in FP64, at the launch shape the
    tuner keeps for it: the first cell of each size is bit-identical to the
    oracle's run of that cell alone after 2,030 steps through the stimulus
    and the upstroke."""
    import bench
    pb, cells = bench.make_batch(0)
    assert pb.lqt == 0                               # each cell's I_Kr scale applies, as in the benchmark
    p = Params.default(stim_offset=2.0, lqt=0)
    pick, seen = [], set()
    for c, d in enumerate(cells):
        if (d.nx, d.ny, d.nz) not in seen:
            seen.add((d.nx, d.ny, d.nz))
            pick.append(c)
    assert len(pick) == 3
    print(f"case 256-cell batch FP64, cells {pick}", flush=True)
    with CellEngine(p, cells) as eng:
        for chunk in (1, 99, 130, 1800):
            eng.step(chunk)
        for c in pick:
            o = OracleCell(p, cells[c])
            o.step(2030)
            _assert_same_state(eng, c, o, f"batch cell {c}")
            print(f"  cell {c} equal", flush=True)


def test_config0_three_beats_from_baseline_ics_match_the_oracle():
    """The CPU-runnable configuration: 32x8x4 = 1,024 CRUs, baseline ICs
    (ci 0.1 uM, cjsr = cnsr 750 uM, V -86 mV, every channel closed), CL
    1000 ms with a 2 ms stimulus, 3 beats = 120,000 steps, FP64, seed 1.
    Every output row (1,200 of them), the final state, the counters and the
    three beats' summaries equal the oracle's."""
    p = Params.default(ci0=0.1, v0=-86.0, stim_dur=2.0)
    d = CellDesc.default(nx=32, ny=8, nz=4, cl=1000.0, cjsr0=750.0, seed=1)
    print("case config 0: 32x8x4, 120000 steps", flush=True)
    o = OracleCell(p, d)
    nrow = 0
    with CellEngine(p, d) as eng:
        for r in range(1200):
            eng.step(100 if r else 1)
            o.step(100 if r else 1)
            grow, gst = eng.diag()
            orow, ost = o.diag()
            assert gst == ost == 100 * r, r
            assert np.array_equal(grow[0], orow), (r, np.nonzero(grow[0] != orow))
            nrow += 1
            if r % 200 == 0:
                print(f"  row {r} equal (t = {orow[0]:g} ms, V = {orow[2]:.2f} mV)", flush=True)
        eng.step(99)
        o.step(99)
        _assert_same_state(eng, 0, o, "config 0 after 120000 steps")
        assert np.array_equal(eng.counters()[0], o.counters())
        gb = eng.beats(0)
        assert np.array_equal(gb, o.beats())
        # onsets at 100, 1100 and 2100 ms: two closed beats and the third in progress, each with an upstroke
        assert gb.shape[0] == 3 and list(gb[:, 9]) == [1.0, 1.0, 0.0] and np.all(gb[:, 2] > 0.0)
    assert nrow == 1200


@pytest.mark.parametrize("shape", [None, (256, 256, 4), (256, 512, 4)], ids=["tuned", "one-launch", "one-launch-2"])
def test_fp64_bitexact_on_a_cell_of_more_than_256_chunks(shape):
    """A 160x20x21 cell (67,200 CRUs, 263 chunks): the whole-cell sums and
    the output-row sums run their loops over groups of 256 chunks more than
    once, and the sums of the second group join the first in chunk order
    (under the one-launch schedule in every CRU workgroup too)."""
    p = Params.default(stim_offset=1.0)
    d = CellDesc.default(nx=160, ny=20, nz=21, cl=1000.0, seed=13)
    print(f"case 160x20x21 FP64, 263 chunks, shape {shape}", flush=True)
    o = OracleCell(p, d)
    with CellEngine(p, d) as eng:
        if shape:
            eng.set_launch_shape(*shape)
        for chunk in (1, 100, 37):                        # an eager step, then a graph-replayed interval
            eng.step(chunk)
            o.step(chunk)
        _assert_same_state(eng, 0, o, "after 138 steps")
        grow, gst = eng.diag()
        orow, ost = o.diag()
        assert gst == ost == 100
        assert np.array_equal(grow[0], orow), np.nonzero(grow[0] != orow)
        assert np.array_equal(eng.linescan(0), o.linescan())
        assert np.array_equal(eng.counters()[0], o.counters())


def test_gating_probe_bitexact_including_gaussian_branch():
    rng = np.random.default_rng(3)
    n = 4096
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
    for dt in (0.025, 0.4):   # 0.4 ms pushes p*dt past 0.2: the Gaussian branch
        for iso2 in (0, 1):
            print(f"case dt={dt} iso2={iso2}", flush=True)
            g = gating_probe(5, cru, 12345, 3.333, 2.5, 18.0, dt, iso2, counts, lcc, conc)
            r = orc.gating_probe(5, cru, 12345, 3.333, 2.5, 18.0, dt, iso2, counts, lcc, conc)
            for a, b, name in zip(g, r, ["counts", "lcc", "draws", "nopen", "fails"]):
                assert np.array_equal(a, b), name
            if dt > 0.1:
                # OU->CU with nou > 5 and p = 0.4 takes the Gaussian branch: 3 slots per try
                many = counts[:, 0] > 5
                assert np.all(g[2][many] >= 4 + 3 + 7)


@pytest.mark.parametrize("shape", [None, (256, 256, 4), (256, 512, 4), (1024, 64, 2)],
                         ids=["tuned", "one-launch", "one-launch-2", "lane-per-cru-sweep"])
def test_batch_equals_cells_run_alone(shape):
    p = Params.default(stim_offset=2.0)
    cells = [CellDesc.default(nx=9, ny=4, nz=3, cl=1000.0, seed=1, Ku2=1.0),
             CellDesc.default(nx=30, ny=10, nz=2, cl=2000.0, seed=2, Ku2=2.5, ikr_scale=0.3),
             CellDesc.default(nx=5, ny=5, nz=5, cl=1000.0, seed=3, Ku=3.333, Kb=18.0)]
    with CellEngine(p, cells) as batch:
        if shape:
            batch.set_launch_shape(*shape)
        batch.step(230)
        bs = batch.scalars()
        for c, d in enumerate(cells):
            print(f"case cell {c}", flush=True)
            with CellEngine(p, d) as one:
                one.step(230)
                assert np.array_equal(one.scalars()[0], bs[c])
                for f in ("CI", "CS", "CJSR", "NOU", "LCC"):
                    assert np.array_equal(one.field(0, f), batch.field(c, f))


@pytest.mark.parametrize("shape", [None, (256, 256, 4), (256, 512, 4), (1024, 64, 2), (128, 64, 2)],
                         ids=["tuned", "one-launch", "one-launch-2", "lane-per-cru-sweep", "concurrent"])
def test_fp32_batch_equals_cells_run_alone(shape):
    """FP32 production mode: a cell's trajectory does not depend on the other
    cells of its batch or on the launch shape (the stochastic rounding's bits
    are keyed by cell seed, step and the cell's own CRU index; the membrane
    step is contraction-free), through the upstroke and 100 ms of plateau,
    per-cell I_Kr scale applied (lqt off); every field, the scalars and the
    beat rows."""
    p = Params.default(stim_offset=2.0, precision=PREC_F32, lqt=0)
    cells = [CellDesc.default(nx=9, ny=4, nz=3, cl=1000.0, seed=1, Ku2=1.0),
             CellDesc.default(nx=30, ny=10, nz=2, cl=2000.0, seed=2, Ku2=2.5, ikr_scale=0.3),
             CellDesc.default(nx=5, ny=5, nz=5, cl=1000.0, seed=3, Ku=3.333, Kb=18.0, ikr_scale=0.0)]
    print(f"case fp32 batch shape={shape}", flush=True)
    with CellEngine(p, cells) as batch:
        if shape:
            batch.set_launch_shape(*shape)
        batch.step(4000)
        for c, d in enumerate(cells):
            with CellEngine(p, d) as one:
                one.step(4000)
                assert np.array_equal(one.scalars()[0], batch.scalars()[c]), f"cell {c}: scalars"
                for f in FIELDS:
                    assert np.array_equal(one.field(0, f), batch.field(c, f)), f"cell {c}: {f}"
                assert np.array_equal(one.beats(0), batch.beats(c)), f"cell {c}: beats"


def test_full_size_scheme_a_conserves_calcium_up_to_sarcolemmal_flux():
    """The GPU path at the full 100x20x10 size, checked by a property that
    holds at any size: under scheme A the change of Total_Ca over a step
    (cuda_Cell_Detailed.cu:1031-1035) equals dt * (Vs*sum I_NCX - (1+gamma)
    Vp*sum I_CaL) / N, the only sarcolemmal fluxes (the identity the oracle
    obeys; test_oracle_closed_forms).  The window holds the stimulus and the
    upstroke, so I_CaL and I_NCX are both active."""
    p = Params.default(scheme=SCHEME_A, steps_per_output=1, stim_offset=1.0)
    d = CellDesc.default(nx=100, ny=20, nz=10, cl=1000.0, seed=5)
    print("case full-size FP64 calcium balance 100x20x10", flush=True)
    N = d.ncru
    cm = 45.0 * N / 16120.0
    checked = 0
    with CellEngine(p, d) as eng:
        eng.step(1)
        prev = eng.diag()[0][0, 44]
        for step in range(1, 161):
            c0 = eng.counters()[0].copy()
            eng.step(1)
            row = eng.diag()[0][0]
            sc = eng.scalars()[0]
            if np.array_equal(eng.counters()[0], c0):
                flux = p.dt * (sc[16] * cm / 0.0965 - sc[15] * cm / (0.0965 * 2.0)) / N
                assert row[44] - prev == pytest.approx(flux, rel=1e-6, abs=1e-12), step
                checked += 1
            prev = row[44]
        assert abs(eng.scalars()[0][15]) > 0.0           # I_CaL flowed inside the window
    assert checked > 100


def test_fp32_production_tracks_fp64():
    """FP32 concentrations vs FP64: same draws; deviations stay small over a
    short window (per-output error of the precision pattern)."""
    p64 = Params.default(stim_offset=1.0)
    p32 = Params.default(stim_offset=1.0, precision=PREC_F32)
    d = CellDesc.default(nx=16, ny=8, nz=4, cl=1000.0, seed=9)
    with CellEngine(p64, d) as a, CellEngine(p32, d) as b:
        a.step(101)
        b.step(101)
        ra, _ = a.diag()
        rb, _ = b.diag()
        for col in (1, 2, 5, 6, 7, 8, 21, 22, 41, 43, 44):
            assert abs(ra[0, col] - rb[0, col]) <= 1e-4 * max(1.0, abs(ra[0, col])), col
        for f in ("CI", "CNSR", "CP", "CJSR", "CATS"):
            x, y = a.field(0, f), b.field(0, f)
            assert np.allclose(x, y, rtol=2e-4, atol=1e-6), f


def test_save_load_continues_bit_identically():
    p = Params.default(stim_offset=2.0)
    d = CellDesc.default(nx=10, ny=6, nz=3, cl=1000.0, seed=4)
    with tempfile.TemporaryDirectory() as td, CellEngine(p, d) as eng:
        path = os.path.join(td, "snap.bin")
        eng.step(150)
        eng.save(path)
        eng.step(120)
        ref = eng.scalars()[0].copy()
        ref_ci = eng.field(0, "CI").copy()
        eng.load(path)
        eng.step(120)
        assert np.array_equal(eng.scalars()[0], ref)
        assert np.array_equal(eng.field(0, "CI"), ref_ci)


def test_cli_writes_the_oracle_traces():
    """bin/ryr_cell with the reference argv writes wholetest/linescan files
    byte-identical to the oracle's rows formatted the same way."""
    cli = os.path.join(ROOT, "ryr_hyperactivity_amd", "bin", "ryr_cell")
    nsteps = 450
    argv = ["0", "t1", "1.0", "740", "1.0", "3.333", "0.00", "18.", "0.7", "1",
            "--lattice", "8x6x5", "--cl", "1000", "--nsteps", str(nsteps), "--stim-offset", "3"]
    with tempfile.TemporaryDirectory() as td:
        r = subprocess.run([cli] + argv + ["--outdir", td], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        whole = open(os.path.join(td, "wholetest_t1")).read()
        line = open(os.path.join(td, "linescan_t1")).read()
        assert os.path.exists(os.path.join(td, "cadist_t1"))
    p = Params.default(stim_offset=3.0)
    d = CellDesc.default(nx=8, ny=6, nz=5, cl=1000.0, Ku=3.333, Kb=18.0, seed=1)
    o = OracleCell(p, d)
    exp_w, exp_l = [], []
    done = 0
    while done < nsteps:
        o.step(1)
        row, st = o.diag()
        if st == done:
            exp_w.append(format_wholetest_row(row))
            exp_l.append(format_linescan(o.linescan()))
        done += 1
    assert whole == "".join(exp_w)
    assert line == "".join(exp_l)


def test_beats_and_clamp_match_the_oracle():
    """Beat summaries under a clamp protocol and in a free-running paced cell
    are bit-identical to the oracle's."""
    from tests.test_beats import synthetic_trace, CL
    tr = synthetic_trace(3, ead_beats=(1,))
    p = Params.default(stim_offset=100.0)
    d = CellDesc.default(nx=2, ny=2, nz=2, cl=CL)
    o = OracleCell(p, d)
    o.set_vclamp(tr)
    o.step(tr.size)
    with CellEngine(p, d) as eng:
        eng.set_vclamp(0, tr)
        eng.step(tr.size)
        assert np.array_equal(eng.beats(0), o.beats())
        assert eng.scalars()[0][S["V"]] == tr[-1]
    # free-running, paced: onsets at 2, 502 and 1002 ms on a small lattice
    p = Params.default(stim_offset=2.0, lqt=0)
    d = CellDesc.default(nx=8, ny=8, nz=4, cl=500.0, seed=5, Ku2=2.5)
    o = OracleCell(p, d)
    o.step(40400)
    with CellEngine(p, d) as eng:
        eng.step(40400)
        gb = eng.beats(0)
        assert np.array_equal(gb, o.beats())
        assert gb.shape[0] == 3 and list(gb[:, 9]) == [1.0, 1.0, 0.0]


@pytest.mark.parametrize("shape", [(64, 64, 1), (128, 256, 0), (256, 128, 1), (256, 256, 0), (64, 128, 2),
                                   (256, 64, 2), (256, 512, 1), (128, 512, 0), (64, 512, 2), (256, 256, 3),
                                   (64, 128, 3), (128, 512, 3), (256, 256, 4), (256, 512, 4)])
def test_every_launch_shape_is_bitexact(shape):
    """Workgroup sizes and step schedules change no value: FP64 against the
    oracle, and FP32 against the FP32 run at another shape (whole-cell sums
    are formed per 64-CRU group, so the reduction order is the same for every
    size, two lanes per CRU (CRU size 512) included;
    the fused membrane step reads only the CRU kernel's sums; the concurrent
    kernels read only step-start values of each other's fields).  The FP32
    runs go 4,000 steps, through the upstroke into the plateau, and compare
    the beat rows too: the membrane step, FP64 in both precisions, is formed
    in different places by the one-launch schedule and must round the same."""
    print(f"case shape={shape}", flush=True)
    p = Params.default(stim_offset=2.0)
    d = CellDesc.default(nx=20, ny=13, nz=5, cl=1000.0, seed=5)
    o = OracleCell(p, d)
    with CellEngine(p, d) as eng:
        eng.set_launch_shape(*shape)
        assert eng.launch_shape() == shape
        for chunk in (101, 149):                   # output rows at steps 100 and 200
            eng.step(chunk)
            o.step(chunk)
            _assert_same_state(eng, 0, o, f"shape {shape}")
            grow, gst = eng.diag()
            orow, ost = o.diag()
            assert gst == ost and np.array_equal(grow[0], orow), f"shape {shape}: output row"
        assert np.array_equal(eng.counters()[0], o.counters()), f"shape {shape}: counters"
    p32 = Params.default(stim_offset=2.0, precision=PREC_F32)
    with CellEngine(p32, d) as a, CellEngine(p32, d) as b:
        a.set_launch_shape(256, 256, (shape[2] + 1) % 3)
        b.set_launch_shape(*shape)
        for chunk in (250, 3750):
            a.step(chunk)
            b.step(chunk)
            for f in FIELDS:
                assert np.array_equal(a.field(0, f), b.field(0, f)), (chunk, f)
            assert np.array_equal(a.scalars(), b.scalars()), chunk
        assert np.array_equal(a.beats(0), b.beats(0))


LANE_PER_CRU_CASES = [((20, 13, 5), SCHEME_A, 0, (1024, 64, 2)), ((100, 4, 3), SCHEME_A, 0, (1024, 128, 2)),
                      ((70, 3, 2), SCHEME_B, 0, (1024, 128, 0)), ((1, 3, 4), SCHEME_A, 0, (1024, 256, 2)),
                      ((6, 1, 2), SCHEME_B, 0, (1024, 64, 0)), ((3, 2, 1), SCHEME_A, 0, (1024, 64, 2)),
                      ((33, 7, 4), SCHEME_A, 1, (1024, 64, 2))]


@pytest.mark.parametrize("dims,scheme,perm,shape", LANE_PER_CRU_CASES,
                         ids=["20x13x5", "100x4x3", "70x3x2-B", "1x3x4", "6x1x2-B", "3x2x1", "33x7x4-perm"])
def test_lane_per_cru_sweep_is_bitexact(dims, scheme, perm, shape):
    """The sweep with one lane per CRU (sub-voxel size 1024): the x partners
    come from the neighbouring lanes, and from memory at a wave's two ends
    (rows of 70 and 100 CRUs cross waves mid-row), the partners inside a CRU
    from registers.  FP64 against the oracle through an action potential,
    both cleft schemes, lattices one CRU thick, the perm bath; FP32 against
    the per-sub-voxel sweep."""
    from ryr_hyperactivity_amd._abi import VAR_PERM
    print(f"case lane-per-CRU sweep {dims} scheme {scheme} perm {perm} shape {shape}", flush=True)
    kw = dict(variants=VAR_PERM, ci0=0.3) if perm else {}
    p = Params.default(stim_offset=2.0, scheme=scheme, **kw)
    d = CellDesc.default(nx=dims[0], ny=dims[1], nz=dims[2], cl=1000.0, seed=11)
    o = OracleCell(p, d)
    with CellEngine(p, d) as eng:
        eng.set_launch_shape(*shape)
        assert eng.launch_shape() == shape
        for chunk in (101, 199):
            eng.step(chunk)
            o.step(chunk)
            _assert_same_state(eng, 0, o, f"lane-per-CRU sweep {dims}")
            grow, gst = eng.diag()
            orow, ost = o.diag()
            assert gst == ost and np.array_equal(grow[0], orow), "output row"
        assert np.array_equal(eng.counters()[0], o.counters())
    p32 = Params.default(stim_offset=2.0, scheme=scheme, precision=PREC_F32, **kw)
    with CellEngine(p32, d) as a, CellEngine(p32, d) as b:
        a.set_launch_shape(128, shape[1], shape[2])
        b.set_launch_shape(*shape)
        a.step(300)
        b.step(300)
        for f in FIELDS:
            assert np.array_equal(a.field(0, f), b.field(0, f)), f
        assert np.array_equal(a.scalars(), b.scalars())


def test_tuning_and_kernel_profile():
    p = Params.default(precision=PREC_F32)
    d = CellDesc.default(nx=30, ny=10, nz=6, seed=2)
    with CellEngine(p, d) as eng:
        assert eng.launch_shape() == (0, 0, -1)    # tuning pending
        eng.step(5)
        assert eng.launch_shape() == (0, 0, -1)
        eng.step(500)                              # sizes after 9 steps, then the schedule candidates
        wc, wr, sc = eng.launch_shape()
        assert wc in (64, 128, 256, 1024) and wr in (64, 128, 256, 512) and sc in (0, 1, 2, 3, 4)
        prof = eng.profile_kernels(120)
        assert eng.cru_steps() == 625 * d.ncru
        assert eng.scalars()[0][S["STEP"]] == 625
        for k in ("k_cyto", "k_cru_step"):
            us, n = prof[k]
            if sc == 4 and k == "k_cru_step":
                assert n == 0                      # one launch per step, stamped in the sub-voxel slot
                continue
            assert n == 120 and 0.5 < us < 5e4, (k, us, n)
        # the fused schedules run the membrane step inside the next sub-voxel launch, and on its own only at
        # output steps and at the last step
        us, n = prof["k_membrane"]
        assert (n == 120 if sc in (0, 2) else n == 2) and 0.5 < us < 5e4, (us, n, sc)
        assert prof["k_diag"][1] == 1              # steps 505..624 hold one output step (600)
        assert eng.latency_floor(100) > 0.0
        eng.set_launch_shape(0, 0, 0)
        assert eng.launch_shape() == (0, 0, -1)
