"""Beat analyser (APD90, EAD flag, peak Ca, JSR load, sparks) pinned by
synthetic voltage traces with known answers, driven through the clamp
protocol (cuda_Cell_Detailed.cu:754-757 apclamp: V replaced after each step)."""
import numpy as np
import pytest

from ryr_hyperactivity_amd import CellDesc, Params, clamp_waveform
from oracle.binding import OracleCell

DT = 0.025
CL = 500.0


def synthetic_trace(nbeats, ead_beats=(), onset=100.0):
    """V after step n (time (n+1)*dt): rest -86 mV, 2 ms linear upstroke to
    +40, linear plateau decay to 0 at +200 ms, linear repolarisation to -86 at
    +300 ms.  In EAD beats a 10 mV hump of 20 ms starts at +230 ms."""
    nsteps = int(round((onset + nbeats * CL) / DT))
    t = (np.arange(nsteps) + 1) * DT
    v = np.full(nsteps, -86.0)
    for b in range(nbeats):
        s = t - (onset + b * CL)
        up = (s >= 0) & (s < 2)
        v[up] = -86.0 + 126.0 * s[up] / 2.0
        pl = (s >= 2) & (s < 200)
        v[pl] = 40.0 - 40.0 * (s[pl] - 2.0) / 198.0
        rp = (s >= 200) & (s < 300)
        v[rp] = -86.0 * (s[rp] - 200.0) / 100.0
        if b in ead_beats:
            hump = (s >= 230) & (s < 250)
            v[hump] += 10.0 * np.sin(np.pi * (s[hump] - 230.0) / 20.0)
    return v


def run_clamped(trace, nsteps):
    p = Params.default(stim_offset=100.0)
    d = CellDesc.default(nx=2, ny=2, nz=2, cl=CL)
    o = OracleCell(p, d)
    o.set_vclamp(trace, periodic=False)
    o.step(nsteps)
    return o.beats(), o


def test_apd90_and_ead_on_synthetic_trace():
    tr = synthetic_trace(3, ead_beats=(1,))
    beats, o = run_clamped(tr, tr.size)
    assert beats.shape[0] == 3
    assert list(beats[:2, 9]) == [1.0, 1.0] and beats[2, 9] == 0.0
    assert np.allclose(beats[:, 0], [100.0, 600.0, 1100.0])
    # vrest -86, vmax 40: threshold -73.4 mV, crossed on the last ramp at 200 + 73.4/86*100 ms
    apd_exact = 200.0 + 73.4 / 86.0 * 100.0
    assert beats[0, 1] == pytest.approx(apd_exact, abs=1e-9)
    assert beats[2, 1] == pytest.approx(apd_exact, abs=1e-9)
    assert beats[1, 1] == pytest.approx(apd_exact, abs=1e-9)      # the hump ends before the crossing
    assert list(beats[:, 4]) == [0.0, 1.0, 0.0]
    assert 230.0 < beats[1, 5] < 240.0
    assert np.allclose(beats[:, 2], 40.0) and np.allclose(beats[:, 3], -86.0)
    # the literal flag (dV/dt > 0 above -40 mV once repolarisation began): only the hump, from its first step
    # (at +230 ms the ramp is at -25.8 mV and the hump's slope 10 pi / 20 mV/ms beats the ramp's -0.86)
    assert list(beats[:, 10]) == [0.0, 1.0, 0.0]
    assert beats[1, 11] == pytest.approx(230.0 + DT, abs=DT / 2) and beats[0, 11] == -1.0


def test_no_ead_for_spike_notch_dome():
    """A notch-and-dome rise during the plateau (above 40% of the amplitude)
    is not an EAD."""
    tr = synthetic_trace(2)
    t = (np.arange(tr.size) + 1) * DT
    for b in range(2):
        s = t - (100.0 + b * CL)
        notch = (s >= 2) & (s < 20)
        tr[notch] -= 15.0 * np.sin(np.pi * (s[notch] - 2.0) / 18.0)
    beats, _ = run_clamped(tr, tr.size)
    assert list(beats[:, 4]) == [0.0, 0.0]
    # read literally, the dome's rise after the notch (far above -40 mV) is dV/dt > 0 during repolarisation:
    # the literal flag counts it, the flag with the 40% / 50 ms floor does not.  The notch's minimum is where
    # 40/198 = (15 pi/18) (-cos(pi (s-2)/18)); the first rising step ends within a step after it
    s_min = 2.0 + 18.0 / np.pi * np.arccos(-(40.0 / 198.0) / (15.0 * np.pi / 18.0))
    assert list(beats[:, 10]) == [1.0, 1.0]
    assert np.all((beats[:, 11] > s_min) & (beats[:, 11] <= s_min + 2 * DT))


def test_failed_beat_has_no_apd():
    tr = np.full(int(round((100 + 2 * CL) / DT)), -86.0)
    beats, _ = run_clamped(tr, tr.size)
    assert beats.shape[0] == 2
    assert np.all(beats[:, 1] == -1.0) and np.all(beats[:, 4] == 0.0) and np.all(beats[:, 10] == 0.0)


def test_reference_clamp_protocols():
    v = clamp_waveform("vclamp", 8000, clamp_voltage=-10.0)
    assert v[4000] == -86.0 and v[4001] == -10.0       # t > 100 ms (strict)
    b = clamp_waveform("brownvclamp", 300000)
    assert b[159999] == -86.0 and b[160000] == -10.0 and b[168000] == 10.0 and b[248000] == -30.0
    assert b[288000] == -86.0
    la = clamp_waveform("linearapclamp", 20000)
    assert la[4000] == -80.0 and la[3999] == -86.0
    # the clamp drives the oracle's V exactly
    p = Params.default()
    o = OracleCell(p, CellDesc.default(nx=2, ny=2, nz=2))
    o.set_vclamp(v)
    o.step(4002)
    assert o.scalars()[1] == -10.0
