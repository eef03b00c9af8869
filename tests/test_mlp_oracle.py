"""The FP64 CPU reference of the Riemann-guess network (oracle/mlp_oracle.py)
pinned by central finite differences of its own loss, so a slip in the
hand-written backward pass (a missing ReLU mask, a transposed weight) fails."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import mlp_oracle as M  # noqa: E402


def _rows(n, seed=0):
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 7))
    r[:, 0] = rng.normal(0, 1, n)
    r[:, 1] = rng.uniform(0.2, 2, n)
    r[:, 2] = 10 ** rng.uniform(-2, 1, n)
    r[:, 3] = rng.normal(0, 1, n)
    r[:, 4] = rng.uniform(0.2, 2, n)
    r[:, 5] = 10 ** rng.uniform(-2, 1, n)
    r[:, 6] = np.sqrt(r[:, 2] * r[:, 5]) * rng.uniform(0.5, 2, n)
    return r


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    w = rng.normal(0, 0.3, M.N_PARAMS)
    rows = _rows(50)
    loss, g = M.loss_grad(w, rows, 5.0 / 3.0)
    idx = np.concatenate([rng.choice(M.N_PARAMS, 200, replace=False), [M.OFF_B4, M.OFF_B3, M.OFF_B1, M.OFF_W4, M.OFF_W3]])
    eps = 1e-6
    for k in idx:
        wp = w.copy(); wp[k] += eps
        wm = w.copy(); wm[k] -= eps
        fd = (M.loss_grad(wp, rows, 5.0 / 3.0)[0] - M.loss_grad(wm, rows, 5.0 / 3.0)[0]) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_features_are_scale_and_galilean_invariant():
    rows = _rows(20, seed=3)
    x0, _ = M.features(rows, 1.4)
    s = rows.copy()
    s[:, [2, 5, 6]] *= 7.0         # pressure scale
    s[:, [1, 4]] *= 7.0            # density scale (c unchanged)
    s[:, [0, 3]] += 3.0            # frame velocity
    x1, _ = M.features(s, 1.4)
    np.testing.assert_allclose(x0, x1, rtol=1e-12, atol=1e-12)
