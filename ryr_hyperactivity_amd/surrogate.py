"""Beat surrogate: an 11-64-64-2 MLP trained and evaluated by the gfx950
kernels of libryr_cru.so (C ABI include/ryr_mlp.h).

It predicts, for one beat, the EAD probability and APD90 from the cell's
parameters (RyR open-rate multiplier Ku2, I_Kr scale, cycle length) and the
previous beat's summary, so that a parameter sweep can be steered: points
where the prediction is confident take the prediction, the others are
simulated.  The reference simulates every point; it has no counterpart.

Feature vector (RYR_MLP_NIN = 11):
  0 Ku2, 1 I_Kr scale, 2 cycle length (s),
  3..10 previous beat: APD90 (ms), EAD flag, EAD time (ms, 0 if none),
        Vmax (mV), Vrest (mV), peak ci (uM), JSR load (uM), sparks per 1000 CRUs.
Targets of a beat: EAD flag; APD90 (ms) where the beat repolarised.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from ._abi import load_library, RYR_OK

NIN, NHID, NOUT = 11, 64, 2
NPARAM = 7312
OFF = {"W1": 0, "b1": 2048, "W2": 2112, "b2": 6208, "W3": 6272, "b3": 7296}
FEATURE_NAMES = ["ku2", "ikr_scale", "cl_s", "apd90", "ead", "ead_t", "vmax", "vrest", "peak_ci", "jsr_load",
                 "sparks_per_1000"]

_fp = ctypes.POINTER(ctypes.c_float)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_fp)


class MLPError(RuntimeError):
    pass


def unpack(params: np.ndarray) -> dict:
    """Views of the packed parameter vector: W1 (32x64, rows >= 11 padding),
    b1, W2 (64x64), b2, W3 (64x16, columns >= 2 padding), b3 (16)."""
    p = np.asarray(params)
    return {"W1": p[0:2048].reshape(32, 64), "b1": p[2048:2112], "W2": p[2112:6208].reshape(64, 64),
            "b2": p[6208:6272], "W3": p[6272:7296].reshape(64, 16), "b3": p[7296:7312]}


def init_params(seed: int = 0) -> np.ndarray:
    """Glorot-uniform weights on the live entries, zero biases and padding."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, dtype=np.float32)
    v = unpack(p)
    for name, (fi, fo) in (("W1", (NIN, NHID)), ("W2", (NHID, NHID)), ("W3", (NHID, NOUT))):
        lim = np.sqrt(6.0 / (fi + fo))
        v[name][:fi, :fo] = rng.uniform(-lim, lim, size=(fi, fo))
    return p


class BeatMLP:
    """The network on one GPU.  Every call runs the HIP kernels; there is no
    CPU path (a missing library raises at construction)."""

    def __init__(self, device: int = 0, params: np.ndarray | None = None, seed: int = 0):
        self.lib = load_library()
        h = ctypes.c_void_p()
        rc = self.lib.ryr_mlp_create(device, ctypes.byref(h))
        if rc != RYR_OK:
            raise MLPError(f"ryr_mlp_create failed ({rc})")
        self.h = h
        self.set_params(init_params(seed) if params is None else params)

    def _check(self, rc, what):
        if rc != RYR_OK:
            raise MLPError(f"{what} failed ({rc}): {self.lib.ryr_mlp_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.ryr_mlp_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params):
        p = _f32(params)
        assert p.size == NPARAM
        self._check(self.lib.ryr_mlp_set_params(self.h, _ptr(p), NPARAM), "ryr_mlp_set_params")

    def params(self) -> np.ndarray:
        p = np.zeros(NPARAM, dtype=np.float32)
        self._check(self.lib.ryr_mlp_get_params(self.h, _ptr(p), NPARAM), "ryr_mlp_get_params")
        return p

    def forward(self, x):
        """(EAD probability, normalised APD90) for each row of x (n x 11)."""
        x = _f32(x).reshape(-1, NIN)
        out = np.zeros((x.shape[0], 2), dtype=np.float32)
        self._check(self.lib.ryr_mlp_forward(self.h, _ptr(x), x.shape[0], _ptr(out)), "ryr_mlp_forward")
        return out[:, 0], out[:, 1]

    def grad(self, x, y, t, mask, s=None, w_apd: float = 1.0):
        x = _f32(x).reshape(-1, NIN)
        n = x.shape[0]
        y, t, mask = _f32(y), _f32(t), _f32(mask)
        sw = None if s is None else _f32(s)
        g = np.zeros(NPARAM, dtype=np.float32)
        loss = ctypes.c_float()
        self._check(self.lib.ryr_mlp_grad(self.h, _ptr(x), _ptr(y), _ptr(t), _ptr(mask),
                                          _ptr(sw) if sw is not None else None, n, w_apd, ctypes.byref(loss),
                                          _ptr(g)), "ryr_mlp_grad")
        return float(loss.value), g

    def train(self, x, y, t, mask, steps: int, batch: int = 256, lr: float = 3e-3, w_apd: float = 1.0,
              beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8, seed: int = 0) -> np.ndarray:
        x = _f32(x).reshape(-1, NIN)
        y, t, mask = _f32(y), _f32(t), _f32(mask)
        trace = np.zeros(max(steps, 1), dtype=np.float32)
        self._check(self.lib.ryr_mlp_train(self.h, _ptr(x), _ptr(y), _ptr(t), _ptr(mask), x.shape[0], w_apd,
                                           steps, batch, lr, beta1, beta2, eps, seed, _ptr(trace)),
                    "ryr_mlp_train")
        return trace[:steps]


@dataclass
class Standardizer:
    mean: np.ndarray
    std: np.ndarray

    @classmethod
    def fit(cls, a: np.ndarray) -> "Standardizer":
        a = np.asarray(a, dtype=np.float64)
        sd = a.std(axis=0)
        return cls(a.mean(axis=0), np.where(sd > 1e-12, sd, 1.0))

    def __call__(self, a):
        return (np.asarray(a, dtype=np.float64) - self.mean) / self.std

    def inverse(self, z):
        return np.asarray(z, dtype=np.float64) * self.std + self.mean


def beat_features(ku2, ikr, cl_ms, prev_beat, ncru: int) -> np.ndarray:
    """Feature row from the parameters and one previous-beat record (a row of
    ryr_read_beats: onset, apd90, vmax, vrest, ead, ead_t, peak_ci, jsr_load,
    sparks, complete)."""
    b = np.asarray(prev_beat, dtype=np.float64)
    return np.array([ku2, ikr, cl_ms / 1000.0, max(b[1], 0.0), b[4], max(b[5], 0.0), b[2], b[3], b[6], b[7],
                     1000.0 * b[8] / ncru])


def training_set(records):
    """Rows (x, y, t, mask) from per-cell beat tables.  `records` holds
    (ku2, ikr, cl_ms, ncru, beats) with beats an (nbeats x 10) array of
    completed beats; every beat after the first gives one sample."""
    xs, ys, ts, ms = [], [], [], []
    for ku2, ikr, cl, ncru, beats in records:
        for b in range(1, len(beats)):
            xs.append(beat_features(ku2, ikr, cl, beats[b - 1], ncru))
            ys.append(float(beats[b][4] > 0))
            apd = beats[b][1]
            ts.append(apd if apd > 0 else 0.0)
            ms.append(1.0 if apd > 0 else 0.0)
    return np.array(xs), np.array(ys), np.array(ts), np.array(ms)


class BeatSurrogate:
    """Standardised inputs and APD target around a BeatMLP (or several, whose
    spread flags the points where the networks disagree)."""

    def __init__(self, n_models: int = 3, device: int = 0, seed: int = 0):
        self.models = [BeatMLP(device, seed=seed + k) for k in range(n_models)]
        self.xn: Standardizer | None = None
        self.tn: Standardizer | None = None

    def close(self):
        for m in self.models:
            m.close()

    def fit(self, x, y, t, mask, steps: int = 3000, batch: int = 256, lr: float = 3e-3, seed: int = 0):
        x, t, mask = np.asarray(x, dtype=np.float64), np.asarray(t, dtype=np.float64), np.asarray(mask)
        self.xn = Standardizer.fit(x)
        self.tn = Standardizer.fit(t[mask > 0][:, None] if np.any(mask > 0) else np.zeros((1, 1)))
        xz = self.xn(x)
        tz = self.tn(t[:, None])[:, 0] * (mask > 0)
        traces = []
        for k, m in enumerate(self.models):
            m.set_params(init_params(seed + 101 * k))
            traces.append(m.train(xz, y, tz, mask, steps, batch=batch, lr=lr, seed=seed + k))
        return traces

    def predict(self, x):
        """EAD probability and APD90 (ms), mean and spread over the models."""
        xz = self.xn(np.asarray(x, dtype=np.float64))
        ps, aps = [], []
        for m in self.models:
            p, a = m.forward(xz)
            ps.append(p)
            aps.append(self.tn.inverse(a[:, None])[:, 0])
        ps, aps = np.array(ps), np.array(aps)
        return ps.mean(0), aps.mean(0), ps.std(0), aps.std(0)

    def predict_point(self, ku2, ikr, cl_ms, prev_beats, ncru: int):
        """EAD incidence of a point and the networks' spread: the per-beat EAD
        probability averaged over a set of previous-beat records (rows of
        ryr_read_beats), typically the beats of the nearest simulated points,
        with this point's parameters."""
        x = np.array([beat_features(ku2, ikr, cl_ms, b, ncru) for b in prev_beats])
        p, _, sp, _ = self.predict(x)
        return float(p.mean()), float(sp.mean())
