"""The form the benchmark times, against the oracle: 3D GSPH Sedov with FP32
pair loops and the learned p* starting guess (the GSPH force kernel's
learned instantiation, in place of src/gsph/g_fluid_force.cpp:44-171 with the
guess inserted at :244-246), with a network trained in the test on the pair
states the solver itself logs.

* 32^3 (32,768 particles): the network starts some Riemann problems; every
  field after 5 steps is within the FP32 bars of the FP64 oracle run (plain
  guess, reference arithmetic); the total-energy drift over 30 steps is
  within 1e-4 of the oracle's.
* 100^3 (1,000,000 particles, the benchmark's size): after initialisation
  and 3 steps, the FluidForce pair sets the pre-interaction derived are
  exactly the brute-force sets r < max(h_i, h_j), r != 0
  (src/fluid_force.cpp:43-62) for 5,000 sampled particles.
"""
from __future__ import annotations

import numpy as np
import pytest

import ics
from checkers import Oracle
from parity_util import compare_fields, force_pairs

pytestmark = pytest.mark.gpu
sphgpu = pytest.importorskip("sphgpu")


def _network(p, n1=24, steps=3):
    import learned
    q = dict(p, pair_precision=ics.PAIR_FP32)
    _, f = ics.sedov_3d(n1=n1)
    rows = learned.log_pairs(q, f, steps=steps, capacity=4_000_000)
    w, info = learned.train(q, rows)
    print("network:", info)
    assert info["loss_final"] < 0.2 * info["loss_initial"]
    return w


def _learned_ctx(p, f, w):
    q = dict(p, pair_precision=ics.PAIR_FP32, riemann_guess=ics.GUESS_LEARNED)
    c = sphgpu.Context(q, f.shape[1])
    c.upload(f)
    c.mlp_set_params(w)
    return c


def test_learned_fp32_sedov_against_oracle():
    p, f = ics.sedov_3d(n1=32)
    w = _network(p)
    o = Oracle(p, f)
    o.initialize()
    e0 = o.energy()
    c = _learned_ctx(p, f, w)
    c.initialize()
    g0 = c.energy_momentum()
    c.reset_stats()
    o.integrate(5)
    c.integrate(5)
    st = c.stats()
    print(f"5 steps: Riemann calls {st['riemann_calls']}, learned starts {st['riemann_learned']}, "
          f"fallbacks {st['riemann_fallbacks']}, iterations per call "
          f"{st['riemann_iters'] / st['riemann_calls']:.4f}")
    assert st["riemann_learned"] > 0
    assert st["riemann_fallbacks"] <= 0.05 * (st["riemann_learned"] + st["riemann_fallbacks"])
    assert abs(c.dt - o.dt) <= 1e-4 * o.dt
    compare_fields("sedov_3d_32_learned_fp32", p, c.download(), o.get_fields(), ics.PAIR_FP32)
    o.integrate(25)
    c.integrate(25)
    eo, eg = o.energy(), c.energy_momentum()
    drift_o = (eo[3] - e0[3]) / abs(e0[3])
    drift_g = (eg[3] - g0[3]) / abs(g0[3])
    print(f"30 steps: t oracle {o.time:.6e} gpu {c.time:.6e}; drift oracle {drift_o:.6e} gpu {drift_g:.6e} "
          f"|diff| {abs(drift_g - drift_o):.3e}")
    assert abs(drift_g - drift_o) <= 1e-4
    pscale = np.sqrt(2.0 * eg[0] * f[ics.F_MASS].sum())
    assert np.all(np.abs(eg[4:7]) <= 1e-5 * pscale)
    c.close()


def _periodic_superset(pos, box, radius, idx):
    """Candidates within `radius` of the particles idx (a superset of every
    pair closer than radius), from a periodic k-d tree over the wrapped
    positions."""
    from scipy.spatial import cKDTree
    lo = box[0]
    L = box[1] - box[0]
    y = np.mod(pos.T - lo, L)
    y = np.where(y >= L, 0.0, y)
    tree = cKDTree(y, boxsize=L)
    return tree.query_ball_point(y[idx], r=radius)


def test_learned_fp32_force_sets_exact_at_1m():
    p, f = ics.sedov_3d(n1=100)
    w = _network(p)
    c = _learned_ctx(p, f, w)
    c.initialize()
    c.reset_stats()
    c.integrate(3)
    c.build_tree()
    c.pre_interaction()
    c.fluid_force()
    st = c.stats()
    assert st["riemann_learned"] > 0
    assert st["flist_searches"] == 0  # the derived lists were used
    counts, lists = c.neighbor_sets(None, 2, cap=512)
    gf = c.download()
    c.close()
    n = gf.shape[1]
    h = gf[ics.F_SML]
    idx = np.random.default_rng(4).choice(n, 5000, replace=False)
    radius = float(h.max()) * (1.0 + 1e-6)
    sup = _periodic_superset(gf[ics.F_POS:ics.F_POS + 3], (-0.5, 0.5), radius, idx)
    total = 0
    for i, cand in zip(idx, sup):
        ref = force_pairs(p, gf, int(i), np.asarray(cand))
        got = lists[i, :counts[i]]
        assert counts[i] == len(ref), (i, counts[i], len(ref))
        np.testing.assert_array_equal(np.sort(got), np.sort(ref))
        total += len(ref)
    print(f"1M learned-FP32: {len(idx)} particles, {total} pairs, all identical to brute force; "
          f"learned starts {st['riemann_learned']}")
