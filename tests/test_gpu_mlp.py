"""GPU tests of the learned Riemann starting guess.

* the bf16-MFMA forward/backward kernels against the FP64 CPU reference
  (oracle/mlp_oracle.py): loss within 2e-2 relative, gradient direction
  cosine > 0.999 and norm within 5e-2 (bf16 operands, FP32 accumulation);
* training on the GPU lowers the loss;
* the learned guess inside the GSPH force kernel: fewer Newton iterations
  than the plain guess on logged Sedov pair states, with accelerations equal
  to the plain-guess ones within the solver tolerance.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

import ics
from parity_util import bar

pytestmark = pytest.mark.gpu
sphgpu = pytest.importorskip("sphgpu")
sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import mlp_oracle as M  # noqa: E402


def _riemann_rows(n, seed=0, gamma=5.0 / 3.0):
    """Random pair states with the converged p* of the oracle's solver."""
    from checkers import oracle_riemann
    rng = np.random.default_rng(seed)
    rows = np.zeros((n, 7), dtype=np.float32)
    for k in range(n):
        L = np.array([rng.normal(0, 0.5), rng.uniform(0.3, 3), 10 ** rng.uniform(-3, 0.5), 0.0])
        R = np.array([rng.normal(0, 0.5), rng.uniform(0.3, 3), 10 ** rng.uniform(-3, 0.5), 0.0])
        ps, _, _ = oracle_riemann(ics.RIEMANN_ITERATIVE, gamma, L, R)
        rows[k] = [L[0], L[1], L[2], R[0], R[1], R[2], ps]
    return rows


def _ctx():
    p, f = ics.sedov_3d(n1=4)
    c = sphgpu.Context(p, f.shape[1])
    c.upload(f)
    return c, p["gamma"]


def test_mlp_loss_and_gradient_match_cpu_reference():
    c, gam = _ctx()
    c.mlp_init(seed=5)
    w = c.mlp_get_params().astype(np.float64)
    rows = _riemann_rows(1000, seed=1, gamma=gam)
    loss_g, grad_g = c.mlp_loss_grad(rows)
    loss_c, grad_c = M.loss_grad(w, rows.astype(np.float64), gam)
    cos = float(grad_g @ grad_c / (np.linalg.norm(grad_g) * np.linalg.norm(grad_c)))
    nrel = abs(np.linalg.norm(grad_g) - np.linalg.norm(grad_c)) / np.linalg.norm(grad_c)
    print(f"loss gpu {loss_g:.6e} cpu {loss_c:.6e}; grad cosine {cos:.6f} norm rel {nrel:.3e}")
    assert abs(loss_g - loss_c) / loss_c < 2e-2
    assert cos > 0.999
    assert nrel < 5e-2
    # each parameter block separately (a transposed block would fail here)
    for a, b in ((M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2), (M.OFF_B2, M.OFF_W3),
                 (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.OFF_W4), (M.OFF_W4, M.OFF_B4)):
        gg, gc = grad_g[a:b], grad_c[a:b]
        cb = float(gg @ gc / (np.linalg.norm(gg) * np.linalg.norm(gc) + 1e-30))
        assert cb > 0.995, (a, cb)
    # inference: the GPU prediction equals the CPU forward pass (FP32 weights)
    pred_g = c.mlp_predict(rows[:200])
    pred_c = M.predict(w, rows[:200].astype(np.float64), gam)
    np.testing.assert_allclose(pred_g, pred_c, rtol=1e-4)


def test_mlp_training_lowers_loss():
    c, gam = _ctx()
    c.mlp_init(seed=2)
    rows = _riemann_rows(4096, seed=3, gamma=gam)
    l0, _ = c.mlp_loss_grad(rows)
    l1 = c.mlp_train(rows, epochs=30, batch=256, lr=2e-3)
    print(f"loss {l0:.4e} -> {l1:.4e}")
    assert l1 < 0.1 * l0


def _run_sedov(p, f, steps, guess, params=None):
    q = dict(p)
    q["riemann_guess"] = guess
    q["pair_precision"] = ics.PAIR_FP32  # the benchmark's form
    c = sphgpu.Context(q, f.shape[1])
    c.upload(f)
    if params is not None:
        c.mlp_set_params(params)
    c.initialize()
    c.reset_stats()
    c.integrate(steps)
    st = c.stats()
    out = c.download()
    c.close()
    return st, out


def test_learned_guess_gains_over_plain_guess():
    """What the trained network gains: on a 24^3 Sedov blast the network is
    trained on the pair states logged from a plain-guess run, then the same
    steps run with the learned guess.  Newton iterations per Riemann call go
    down, the fallback rate stays small, and the fields equal the plain-guess
    run's to FP32 tolerance (both are converged to the solver's tolerance and
    refined by the same final FP64 Newton step).  Both runs use FP32 pairs,
    the form the benchmark times."""
    import learned
    p, f = ics.sedov_3d(n1=24)
    rows = learned.log_pairs(p, f, steps=3, capacity=4_000_000)
    w, info = learned.train(p, rows)
    print("training", info)
    assert info["loss_final"] < 0.2 * info["loss_initial"]
    st_p, out_p = _run_sedov(p, f, 3, ics.GUESS_PLAIN)
    st_l, out_l = _run_sedov(p, f, 3, ics.GUESS_LEARNED, w)
    it_p = st_p["riemann_iters"] / st_p["riemann_calls"]
    it_l = st_l["riemann_iters"] / st_l["riemann_calls"]
    used = st_l["riemann_learned"] + st_l["riemann_fallbacks"]
    print(f"iterations per call: plain {it_p:.4f} learned {it_l:.4f}; learned guesses used {used} "
          f"of {st_l['riemann_calls']} calls, fallbacks {st_l['riemann_fallbacks']}")
    assert used > 0
    assert it_l < it_p
    assert st_l["riemann_fallbacks"] <= 0.05 * used
    P, rho, h = out_p[ics.F_PRES], out_p[ics.F_DENS], out_p[ics.F_SML]
    da = np.sqrt(((out_l[ics.F_ACC:ics.F_ACC + 3] - out_p[ics.F_ACC:ics.F_ACC + 3]) ** 2).sum(axis=0))
    e_acc = da / (np.abs(P) / (rho * h) + np.sqrt((out_p[ics.F_ACC:ics.F_ACC + 3] ** 2).sum(axis=0)))
    e_rho = np.abs(out_l[ics.F_DENS] - rho) / rho
    print(f"learned vs plain: acc {e_acc.max():.3e} (p99.9 {np.quantile(e_acc, 0.999):.3e}) dens {e_rho.max():.3e}")
    bar("acc", e_acc, 5e-5, ics.PAIR_FP32)
    bar("dens", e_rho, 5e-5, ics.PAIR_FP32)


def _states(rows, gamma):
    L = np.stack([rows[:, 0], rows[:, 1], rows[:, 2], np.sqrt(gamma * rows[:, 2] / rows[:, 1])], 1)
    R = np.stack([rows[:, 3], rows[:, 4], rows[:, 5], np.sqrt(gamma * rows[:, 5] / rows[:, 4])], 1)
    return L.astype(np.float64), R.astype(np.float64)


def test_learned_guess_cuts_iterations_on_hard_pairs():
    """The gain where the network is used: Sedov 24^3 pair states logged by
    the force kernel; the network is trained as the benchmark trains it (all
    rows that need two or more plain Newton iterations, then the hard rows),
    with a fifth of the hard rows -- the force kernel's gate -- held out.  On
    the held-out rows the learned start needs fewer Newton iterations than
    the plain two-shock guess, is rarely rejected, and p* agrees with the
    plain-guess solution to the FP32 bar."""
    import learned
    p, f = ics.sedov_3d(n1=24)
    log = learned.log_pairs(p, f, steps=3, capacity=4_000_000)
    w, info = learned.train(p, log, holdout=0.2, seed=0)
    ho = info["held_out"]
    print("training", {k: v for k, v in info.items() if k != "held_out"})
    print(f"held-out hard pairs {ho['pairs']}: iterations plain {ho['iters_plain']:.3f} learned "
          f"{ho['iters_learned']:.3f}; learned start used {ho['learned_used']}, rejected {ho['learned_rejected']}, "
          f"max p* difference {ho['max_pstar_diff']:.2e}")
    assert info["pairs_trained_2plus_iterations"] > 100 * info["hard_pairs_trained"]
    assert ho["learned_used"] == ho["pairs"] > 100
    assert ho["iters_learned"] < 0.8 * ho["iters_plain"]
    assert ho["learned_rejected"] <= 0.1 * ho["learned_used"]
    assert ho["max_pstar_diff"] < 1e-5
