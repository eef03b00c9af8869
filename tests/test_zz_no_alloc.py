"""The allocation-free Riemann batch (sphgpu_riemann_batch_static) against
the allocating one (sphgpu_riemann_batch): same kernels, same inputs, so the
results are bit-identical.  Two calls in a row, the second smaller, show that
the shared static staging area carries nothing from one call to the next; a
size beyond SPHGPU_RIEMANN_STATIC_MAX is refused with SPHGPU_E_ARG."""
from __future__ import annotations

import numpy as np
import pytest

import ics
import sphgpu

pytestmark = pytest.mark.gpu


def _states(n: int, seed: int, gam: float):
    rng = np.random.default_rng(seed)
    L = np.stack([rng.normal(0, 1, n), rng.uniform(0.1, 2, n), 10 ** rng.uniform(-3, 1, n), np.zeros(n)], 1)
    R = np.stack([rng.normal(0, 1, n), rng.uniform(0.1, 2, n), 10 ** rng.uniform(-3, 1, n), np.zeros(n)], 1)
    # a share of strong pressure jumps, which the learned gate hands to the network
    k = rng.random(n) < 0.3
    L[k, 2] *= 50.0
    L[:, 3] = np.sqrt(gam * L[:, 2] / L[:, 1])
    R[:, 3] = np.sqrt(gam * R[:, 2] / R[:, 1])
    return L, R


def _compare(c: sphgpu.Context, solver: int, precision: int, n: int, seed: int):
    L, R = _states(n, seed, c.params["gamma"] if hasattr(c, "params") else 5.0 / 3.0)
    a = c.riemann_batch(solver, precision, L, R, with_usage=True)
    b = c.riemann_batch(solver, precision, L, R, with_usage=True, static=True)
    for x, y, name in zip(a, b, ("pstar", "vstar", "iters", "learned_used")):
        assert np.array_equal(x, y), (name, solver, precision, n)
    assert np.isfinite(b[0]).all() and np.isfinite(b[1]).all()
    if solver == ics.RIEMANN_ITERATIVE:  # floored at 1e-25 (HLL's closed form is not)
        assert (b[0] > 0).all()
    return b


@pytest.mark.parametrize("guess", [ics.GUESS_PLAIN, ics.GUESS_LEARNED])
def test_static_batch_equals_allocating_batch(guess):
    p, f = ics.sedov_3d(n1=4)
    q = dict(p)
    q["riemann_guess"] = guess
    c = sphgpu.Context(q, f.shape[1])
    try:
        if guess == ics.GUESS_LEARNED:
            c.mlp_init(7)
        n_max = sphgpu.RIEMANN_STATIC_MAX
        for solver in (ics.RIEMANN_ITERATIVE, ics.RIEMANN_HLL):
            for precision in (ics.PAIR_FP32, ics.PAIR_FP64):
                # the full area first, then a smaller call over stale contents
                big = _compare(c, solver, precision, n_max, seed=11)
                small = _compare(c, solver, precision, 1000, seed=12)
                assert not np.array_equal(big[0][:1000], small[0])
        if guess == ics.GUESS_LEARNED:
            _, _, _, used = c.riemann_batch(ics.RIEMANN_ITERATIVE, ics.PAIR_FP32, *_states(4096, 5, 5.0 / 3.0),
                                            with_usage=True, static=True)
            assert (used > 0).any()
    finally:
        c.close()


def test_static_batch_rejects_oversize():
    p, f = ics.sedov_3d(n1=4)
    c = sphgpu.Context(p, f.shape[1])
    try:
        L, R = _states(sphgpu.RIEMANN_STATIC_MAX + 1, 3, 5.0 / 3.0)
        with pytest.raises(RuntimeError, match="SPHGPU_RIEMANN_STATIC_MAX"):
            c.riemann_batch(ics.RIEMANN_ITERATIVE, ics.PAIR_FP32, L, R, static=True)
        # the context stays usable after the refusal
        _compare(c, ics.RIEMANN_ITERATIVE, ics.PAIR_FP32, 512, seed=4)
    finally:
        c.close()
