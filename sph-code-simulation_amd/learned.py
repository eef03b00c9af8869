"""Training the learned Riemann starting guess from the solver's own pair log.

The GSPH force kernel can log the left/right states and the converged p* of
the pairs it solves (sphgpu_riemann_log_enable / _read).  The network is then
trained on the GPU by the library's bf16 MFMA forward/backward kernels
(sphgpu_mlp_train) in two stages:

* on every logged pair that the plain guess does not solve in one Newton
  step (two or more iterations of the FP32 solver: the pairs that set the
  wavefront's iteration count), a few epochs;
* then on the pairs that the force kernel hands to the network -- a pressure
  ratio above 4 or a velocity jump above the mean sound speed (the `hard`
  gate of the force kernel) -- more epochs at a lower rate.
"""
from __future__ import annotations

import time

import numpy as np

import ics
import sphgpu


def hard_rows(rows: np.ndarray, gamma: float) -> np.ndarray:
    """Rows (uL, rhoL, pL, uR, rhoR, pR, p*) that the force kernel sends to
    the network, with finite entries and positive p*."""
    pl, pr = rows[:, 2], rows[:, 5]
    pmin, pmax = np.minimum(pl, pr), np.maximum(pl, pr)
    du = rows[:, 0] - rows[:, 3]
    r_m, p_m = 0.5 * (rows[:, 1] + rows[:, 4]), 0.5 * (pl + pr)
    with np.errstate(invalid="ignore", over="ignore"):
        hard = (pmin > 0) & (r_m > 0) & ((pmax > 4 * pmin) | (du * du * r_m > gamma * p_m))
    return rows[hard & np.isfinite(rows).all(axis=1) & (rows[:, 6] > 0)]


def log_pairs(params: dict, fields: np.ndarray, steps: int, capacity: int, stride: int = 1) -> np.ndarray:
    """Run `steps` steps of the given problem with the plain guess and return
    the logged pair rows."""
    p = dict(params)
    p["riemann_guess"] = ics.GUESS_PLAIN
    ctx = sphgpu.Context(p, fields.shape[1])
    try:
        ctx.upload(fields)
        ctx.initialize()
        ctx.riemann_log_enable(capacity, stride)
        ctx.integrate(steps)
        return ctx.riemann_log_read(capacity)
    finally:
        ctx.close()


def states(rows: np.ndarray, gamma: float) -> tuple[np.ndarray, np.ndarray]:
    """Left and right states [u, rho, p, c] of log rows."""
    r = rows.astype(np.float64)
    L = np.stack([r[:, 0], r[:, 1], r[:, 2], np.sqrt(gamma * r[:, 2] / r[:, 1])], 1)
    R = np.stack([r[:, 3], r[:, 4], r[:, 5], np.sqrt(gamma * r[:, 5] / r[:, 4])], 1)
    return L, R


def evaluate(params: dict, rows: np.ndarray, w: np.ndarray) -> dict:
    """Newton iterations of the FP32 iterative solver on `rows` from the plain
    two-shock guess and from the learned guess (network w), the share of
    learned starts rejected, and the largest p* difference on the pair's
    pressure scale."""
    L, R = states(rows, params["gamma"])
    q = dict(params)
    q["riemann_guess"] = ics.GUESS_PLAIN
    ctx = sphgpu.Context(q, 64)
    try:
        ps0, _, it0 = ctx.riemann_batch(ics.RIEMANN_ITERATIVE, ics.PAIR_FP32, L, R)
    finally:
        ctx.close()
    q["riemann_guess"] = ics.GUESS_LEARNED
    ctx = sphgpu.Context(q, 64)
    try:
        ctx.mlp_set_params(w)
        ps1, _, it1, used = ctx.riemann_batch(ics.RIEMANN_ITERATIVE, ics.PAIR_FP32, L, R, with_usage=True)
    finally:
        ctx.close()
    scale = np.maximum(np.abs(ps0), np.maximum(L[:, 2], R[:, 2]))
    n_used = int((used > 0).sum())
    return {"pairs": int(rows.shape[0]), "iters_plain": round(float(it0.mean()), 4),
            "iters_learned": round(float(it1.mean()), 4), "learned_used": n_used,
            "learned_rejected": int((used == 2).sum()),
            "max_pstar_diff": float((np.abs(ps1 - ps0) / scale).max())}


def plain_iterations(params: dict, rows: np.ndarray, chunk: int = 2_000_000) -> np.ndarray:
    """Newton iterations of the FP32 iterative solver from the plain guess for
    each log row (on the GPU, sphgpu_riemann_batch)."""
    q = dict(params)
    q["riemann_guess"] = ics.GUESS_PLAIN
    its = np.zeros(rows.shape[0], dtype=np.int32)
    ctx = sphgpu.Context(q, 64)
    try:
        for a in range(0, rows.shape[0], chunk):
            L, R = states(rows[a:a + chunk], params["gamma"])
            its[a:a + chunk] = ctx.riemann_batch(ics.RIEMANN_ITERATIVE, ics.PAIR_FP32, L, R)[2]
    finally:
        ctx.close()
    return its


def _clean(rows: np.ndarray) -> np.ndarray:
    ok = np.isfinite(rows).all(axis=1) & (rows[:, 6] > 0) & (rows[:, 2] > 0) & (rows[:, 5] > 0) \
        & (rows[:, 1] > 0) & (rows[:, 4] > 0)
    return rows[ok]


def train(params: dict, rows: np.ndarray, epochs: int = 2, batch: int = 4096, lr: float = 2e-3,
          fine_epochs: int = 20, fine_batch: int = 1024, fine_lr: float = 5e-4,
          seed: int = 7, holdout: float = 0.0) -> tuple[np.ndarray, dict]:
    """Train the network on the log; returns the FP32 parameter block and a
    summary (loss on the hard rows before and after, row counts, seconds).
    Stage 1: `epochs` over the rows that need two or more plain Newton
    iterations; stage 2: `fine_epochs` over the hard rows.  With holdout > 0
    that share of the hard rows is kept out of both stages and the summary
    gains the iteration comparison on it (`evaluate`)."""
    rows = _clean(rows)
    gamma = params["gamma"]
    pl, pr = rows[:, 2], rows[:, 5]
    du = rows[:, 0] - rows[:, 3]
    r_m, p_m = 0.5 * (rows[:, 1] + rows[:, 4]), 0.5 * (pl + pr)
    is_hard = (np.maximum(pl, pr) > 4 * np.minimum(pl, pr)) | (du * du * r_m > gamma * p_m)
    if not is_hard.any():
        raise ValueError("no hard pairs in the log")
    rng = np.random.default_rng(seed)
    hard_idx = np.flatnonzero(is_hard)
    hard_idx = hard_idx[rng.permutation(hard_idx.size)]
    n_hold = int(holdout * hard_idx.size)
    held = rows[hard_idx[:n_hold]]
    keep = np.ones(rows.shape[0], dtype=bool)
    keep[hard_idx[:n_hold]] = False
    t0 = time.perf_counter()
    its = plain_iterations(params, rows)
    multi = rows[keep & (its >= 2)]
    multi = multi[rng.permutation(multi.shape[0])]
    hard = rows[hard_idx[n_hold:]]
    t1 = time.perf_counter()
    ctx = sphgpu.Context(params, 64)  # the network needs no particle state
    try:
        ctx.mlp_init(seed)
        loss0 = ctx.mlp_loss_grad(hard[: min(hard.shape[0], 200_000)])[0]
        t2 = time.perf_counter()
        loss1 = ctx.mlp_train(multi, epochs, batch, lr) if multi.shape[0] else float("nan")
        loss = ctx.mlp_train(hard, fine_epochs, fine_batch, fine_lr)
        t3 = time.perf_counter()
        w = ctx.mlp_get_params()
    finally:
        ctx.close()
    info = {"pairs_logged": int(rows.shape[0]),
            "plain_iterations_histogram": np.bincount(its).tolist(),
            "pairs_trained_2plus_iterations": int(multi.shape[0]), "hard_pairs_trained": int(hard.shape[0]),
            "epochs": [epochs, fine_epochs], "batch": [batch, fine_batch], "lr": [lr, fine_lr],
            "loss_2plus_rows": round(float(loss1), 8),
            "loss_initial": round(float(loss0), 5), "loss_final": round(float(loss), 6),
            "select_s": round(t1 - t0, 2), "train_s": round(t3 - t2, 2)}
    if n_hold > 0:
        info["held_out"] = evaluate(params, held, w)
    return w, info
