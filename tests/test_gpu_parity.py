"""GPU parity tests: the HIP library through its C-ABI against the oracle.

Every test runs the same seeded input through the oracle (the CPU
restatement, itself pinned to the reference) and through libsphgpu.so.

Bars:
  * tree topology, leaf membership, node centres/edges, kernel_size and
    neighbour sets: bit-exact;
  * FP64 pair mode: fields within 1e-9 relative (summation order and the
    device pow differ from the reference; nothing else does);
  * FP32 pair mode (compensated sums): densities and smoothing lengths within
    5e-5 relative, accelerations and du/dt within 5e-4 of the field's largest
    magnitude, per workload below.
"""
from __future__ import annotations

import numpy as np
import pytest

import ics
from checkers import Oracle
from parity_util import SMALL, compare_fields, force_pairs, rel_err

pytestmark = pytest.mark.gpu

sphgpu = pytest.importorskip("sphgpu")


def gpu_ctx(params, fields, **over):
    p = dict(params)
    p.update(over)
    c = sphgpu.Context(p, fields.shape[1])
    c.upload(fields)
    return c


def guess_h(params, fields):
    """A smoothing length per particle of the right size (for tree tests)."""
    d = params["dim"]
    a = {1: 2.0, 2: np.pi, 3: 4.0 * np.pi / 3.0}[d]
    h = (params["neighbor_number"] * fields[ics.F_MASS] / (fields[ics.F_DENS] * a)) ** (1.0 / d)
    rng = np.random.default_rng(7)
    return h * (1.0 + 0.1 * rng.random(h.shape))


@pytest.mark.parametrize("case", list(SMALL))
def test_tree_topology(case):
    p, f = SMALL[case]()
    f = f.copy()
    f[ics.F_SML] = guess_h(p, f)
    o = Oracle(p, f)
    o.make_tree()
    o.set_kernel()
    rec, mem = o.tree_dump()
    g = gpu_ctx(p, f, pair_precision=ics.PAIR_FP64)
    g.build_tree()
    g.neighbor_sets(f[ics.F_SML], False)  # runs set_kernel with these h
    d = g.tree_dump()
    print(f"{case}: oracle nodes {len(rec)}, gpu nodes {len(d['level'])}")
    assert len(rec) == len(d["level"])
    np.testing.assert_array_equal(rec[:, 0].astype(int), d["level"])
    # counts: the reference never counts particles on the root itself
    # (root_clear zeroes num and assign increments children only,
    # src/bhtree.cpp:46-53, 242), so its root reports 0
    assert rec[0, 1] == 0 and d["count"][0] == f.shape[1]
    np.testing.assert_array_equal(rec[1:, 1].astype(int), d["count"][1:])
    np.testing.assert_array_equal(rec[:, 2].astype(int), d["is_leaf"])
    np.testing.assert_array_equal(rec[:, 3:6], d["center"])
    np.testing.assert_array_equal(rec[:, 6], d["edge"])
    # oracle record: level, num, is_leaf, centre[3], edge, mass, com[3],
    # kernel_size, member offset
    np.testing.assert_array_equal(rec[:, 11], d["kernel_size"])
    if p["use_gravity"]:
        np.testing.assert_allclose(d["mass"], rec[:, 7], rtol=1e-12, atol=0)
        np.testing.assert_allclose(d["com"], rec[:, 8:11], rtol=0, atol=1e-12)


@pytest.mark.parametrize("cluster", [False, True])
def test_tree_large_partial_sort(cluster):
    """Above 65536 particles in 2D/3D the key sort skips its low radix passes
    (it sorts the top 8 digits, 4 more after each failed check) and checks
    that no cell of the sorted depth holds more than a leaf; the tree must be
    the reference's either way.  The clustered case packs 40 particles into a
    cell of edge 1e-7 (more than a leaf below depth 20), so the check fails
    in both builds and each sorts fully."""
    p, f = ics.sedov_3d(n1=48)
    f = f.copy()
    if cluster:
        rng = np.random.default_rng(3)
        f[ics.F_POS:ics.F_POS + 3, :40] = 0.123 + 1e-7 * rng.random((3, 40))
    f[ics.F_SML] = guess_h(p, f)
    o = Oracle(p, f)
    o.make_tree()
    rec, mem = o.tree_dump()
    g = gpu_ctx(p, f, pair_precision=ics.PAIR_FP64)
    for _ in range(2):  # a second build starts from the first one's order
        g.build_tree()
        d = g.tree_dump()
        assert len(rec) == len(d["level"])
        np.testing.assert_array_equal(rec[:, 0].astype(int), d["level"])
        np.testing.assert_array_equal(rec[1:, 1].astype(int), d["count"][1:])
        np.testing.assert_array_equal(rec[:, 2].astype(int), d["is_leaf"])
        np.testing.assert_array_equal(rec[:, 3:6], d["center"])
        np.testing.assert_array_equal(rec[:, 6], d["edge"])
        for k in np.nonzero(rec[:, 2] == 1)[0]:
            cnt, off = int(rec[k, 1]), int(rec[k, 12])
            np.testing.assert_array_equal(np.sort(mem[off:off + cnt]),
                                          np.sort(d["members"][d["first"][k]:d["first"][k] + cnt]))
    if cluster:
        assert d["level"].max() > 13


@pytest.mark.parametrize("case", list(SMALL))
def test_leaf_members(case):
    p, f = SMALL[case]()
    f = f.copy()
    f[ics.F_SML] = guess_h(p, f)
    o = Oracle(p, f)
    o.make_tree()
    rec, mem = o.tree_dump()
    g = gpu_ctx(p, f, pair_precision=ics.PAIR_FP64)
    g.build_tree()
    d = g.tree_dump()
    o_leaf = np.nonzero(rec[:, 2] == 1)[0]
    g_leaf = np.nonzero(d["is_leaf"] == 1)[0]
    np.testing.assert_array_equal(o_leaf, g_leaf)
    for k in o_leaf:
        cnt = int(rec[k, 1])
        off = int(rec[k, 12])
        s = d["first"][k]
        gm = np.sort(d["members"][s:s + cnt])
        # the reference's linked-list order: the root list is the particle
        # array (src/bhtree.cpp:103-107) and every assign() prepends to the
        # child's list (:245-246), so each level reverses its parent's order;
        # a leaf at level l (root = 1) holds ascending ids for odd l,
        # descending for even l
        want = gm if d["level"][k] % 2 == 1 else gm[::-1]
        np.testing.assert_array_equal(mem[off:off + cnt], want)


@pytest.mark.parametrize("case", list(SMALL))
@pytest.mark.parametrize("is_ij", [False, True])
def test_neighbor_sets_exact(case, is_ij):
    p, f = SMALL[case]()
    f = f.copy()
    f[ics.F_SML] = guess_h(p, f)
    o = Oracle(p, f)
    o.make_tree()
    o.set_kernel()
    g = gpu_ctx(p, f, pair_precision=ics.PAIR_FP64)
    g.build_tree()
    counts, lists = g.neighbor_sets(f[ics.F_SML], is_ij, cap=1024)
    n = f.shape[1]
    idx = np.arange(n) if n <= 5000 else np.random.default_rng(1).choice(n, 5000, replace=False)
    total = 0
    for i in idx:
        ref = o.neighbor_search(int(i), is_ij)
        if is_ij:
            ref = force_pairs(p, f, int(i), ref)
        got = lists[i, :counts[i]]
        assert counts[i] == len(ref), (case, i, counts[i], len(ref))
        np.testing.assert_array_equal(np.sort(got), np.sort(ref))
        total += len(ref)
    print(f"{case} is_ij={is_ij}: {len(idx)} particles, {total} neighbours, all sets identical")


@pytest.mark.parametrize("case", list(SMALL))
@pytest.mark.parametrize("prec", [ics.PAIR_FP64, ics.PAIR_FP32])
def test_force_lists_from_pre_exact(case, prec):
    """The force pair lists the pre-interaction derives (members r < h_i plus
    the transposed pairs r < h_j) are the FluidForce sets of
    src/fluid_force.cpp:43-62, checked by brute force over all j with the
    final h, and equal to what the is_ij tree search returns."""
    p, f = SMALL[case]()
    g = gpu_ctx(p, f, pair_precision=prec)
    g.initialize()
    g.integrate(3)
    g.build_tree()
    g.pre_interaction()
    g.fluid_force()
    # still current after the force call: the force used it (a fallback
    # search would have replaced it)
    counts, lists = g.neighbor_sets(None, 2, cap=1024)
    gf = g.download()
    n = gf.shape[1]
    allj = np.arange(n)
    idx = np.arange(n) if n <= 5000 else np.random.default_rng(2).choice(n, 5000, replace=False)
    total = extra = 0
    h = gf[ics.F_SML]
    for i in idx:
        ref = force_pairs(p, gf, int(i), allj)
        got = lists[i, :counts[i]]
        assert counts[i] == len(ref), (case, i, counts[i], len(ref))
        np.testing.assert_array_equal(np.sort(got), np.sort(ref))
        total += len(ref)
        extra += int((h[ref] > h[i]).sum())
    c2, l2 = g.neighbor_sets(h, True, cap=1024)
    np.testing.assert_array_equal(c2, counts)
    for i in idx:
        np.testing.assert_array_equal(np.sort(l2[i, :c2[i]]), np.sort(lists[i, :counts[i]]))
    with pytest.raises(sphgpu.SphGpuError):
        g.neighbor_sets(None, 2, cap=1024)  # the search replaced the derived lists
    print(f"{case} prec={prec}: {len(idx)} particles, {total} pairs ({extra} with h_j > h_i), "
          "all sets identical to brute force and to the is_ij search")


@pytest.mark.parametrize("case", ["khi_2d", "sedov_3d"])
def test_force_lists_many_candidates(case):
    """Particles whose previous density is far too low get search radii with
    hundreds of candidates (beyond the 256 the member masks cover); their
    member rows are written out directly.  The derived force pair sets must
    still be exactly the FluidForce sets."""
    p, f = SMALL[case]()
    g = gpu_ctx(p, f, pair_precision=ics.PAIR_FP32)
    g.initialize()
    g.integrate(2)
    st = g.download()
    low = np.arange(0, st.shape[1], 97)
    st[ics.F_DENS, low] *= 0.02  # h guess (nn m / (rho A))^(1/D) grows by 0.02^(-1/D)
    g.upload(st)
    g.build_tree()
    g.reset_stats()
    g.pre_interaction()
    g.fluid_force()
    stats = g.stats()
    counts, lists = g.neighbor_sets(None, 2, cap=1024)
    gf = g.download()
    n = gf.shape[1]
    allj = np.arange(n)
    for i in range(n):
        ref = force_pairs(p, gf, i, allj)
        assert counts[i] == len(ref), (case, i, counts[i], len(ref))
        np.testing.assert_array_equal(np.sort(lists[i, :counts[i]]), np.sort(ref))
    pc = stats["pre_candidates"]
    print(f"{case}: {len(low)} particles with a low density guess, {pc / n:.1f} candidates per particle, "
          f"force searches {stats['flist_searches']}; all {n} sets identical to brute force")
    assert stats["flist_searches"] == 0


@pytest.mark.parametrize("case", list(SMALL))
@pytest.mark.parametrize("prec", [ics.PAIR_FP64, ics.PAIR_FP32])
def test_initialize_fields(case, prec):
    p, f = SMALL[case]()
    o = Oracle(p, f)
    o.initialize()
    of = o.get_fields()
    g = gpu_ctx(p, f, pair_precision=prec)
    g.initialize()
    gf = g.download()
    compare_fields(case, p, gf, of, prec)
    # neighbour counts are exact whenever h is (FP64 mode)
    if prec == ics.PAIR_FP64:
        same_h = gf[ics.F_SML] == of[ics.F_SML]
        mism = (gf[ics.F_NEIGHBOR] != of[ics.F_NEIGHBOR]) & same_h
        assert mism.sum() == 0


@pytest.mark.parametrize("case", list(SMALL))
@pytest.mark.parametrize("prec", [ics.PAIR_FP64, ics.PAIR_FP32])
def test_short_run(case, prec):
    p, f = SMALL[case]()
    steps = 5
    o = Oracle(p, f)
    o.initialize()
    e0 = o.energy()
    o.integrate(steps)
    of = o.get_fields()
    g = gpu_ctx(p, f, pair_precision=prec)
    g.initialize()
    g.integrate(steps)
    gf = g.download()
    eo = o.energy()
    eg = g.energy_momentum()
    print(f"{case} prec={prec}: dt oracle {o.dt:.6e} gpu {g.dt:.6e}; t {o.time:.6e} {g.time:.6e}")
    print(f"  E oracle {eo[3]:.12e} gpu {eg[3]:.12e} (E0 {e0[3]:.12e})")
    assert abs(g.dt - o.dt) <= (1e-9 if prec == ics.PAIR_FP64 else 1e-4) * o.dt
    compare_fields(case, p, gf, of, prec)
    drift_o = (eo[3] - e0[3]) / abs(e0[3])
    drift_g = (eg[3] - e0[3]) / abs(e0[3])
    assert abs(drift_g - drift_o) < (1e-10 if prec == ics.PAIR_FP64 else 1e-5)


@pytest.mark.parametrize("prec", [ics.PAIR_FP64, ics.PAIR_FP32])
def test_short_run_unequal_masses(prec):
    """The density sums read each candidate's mass unless every uploaded
    mass is equal; a Sedov run with masses spread by +-5% (seeded) takes
    the per-candidate path in the GSPH PreInteraction and must still match
    the oracle."""
    p, f = SMALL["sedov_3d"]()
    rng = np.random.default_rng(7)
    f = f.copy()
    f[ics.F_MASS] *= 1.0 + 0.05 * rng.uniform(-1.0, 1.0, f.shape[1])
    assert np.unique(f[ics.F_MASS]).size > 1
    steps = 3
    o = Oracle(p, f)
    o.initialize()
    o.integrate(steps)
    of = o.get_fields()
    g = gpu_ctx(p, f, pair_precision=prec)
    g.initialize()
    g.integrate(steps)
    gf = g.download()
    compare_fields("sedov_3d_unequal_masses", p, gf, of, prec)


def test_riemann_batch_matches_oracle():
    from checkers import oracle_riemann
    rng = np.random.default_rng(3)
    n = 2000
    L = np.stack([rng.normal(0, 1, n), rng.uniform(0.1, 2, n), 10 ** rng.uniform(-3, 1, n), np.zeros(n)], 1)
    R = np.stack([rng.normal(0, 1, n), rng.uniform(0.1, 2, n), 10 ** rng.uniform(-3, 1, n), np.zeros(n)], 1)
    gam = 5.0 / 3.0
    L[:, 3] = np.sqrt(gam * L[:, 2] / L[:, 1])
    R[:, 3] = np.sqrt(gam * R[:, 2] / R[:, 1])
    p, f = ics.sedov_3d(n1=4)
    g = gpu_ctx(p, f)
    for solver in (ics.RIEMANN_HLL, ics.RIEMANN_ITERATIVE):
        ps64, vs64, it64 = g.riemann_batch(solver, ics.PAIR_FP64, L, R)
        ps32, vs32, it32 = g.riemann_batch(solver, ics.PAIR_FP32, L, R)
        ref = np.array([oracle_riemann(solver, gam, L[k], R[k])[:2] for k in range(n)])
        e64 = rel_err(ps64, ref[:, 0]).max()
        ev = np.abs(vs64 - ref[:, 1]).max()
        # FP32: near vacuum (p* << p_L, p_R) p* is set by the cancellation of
        # terms of size p_L, p_R, and the FP32 rounding of the input states
        # alone moves it by ~1e-7 max(p_L, p_R); errors are therefore measured
        # on the pair's pressure scale max(p*, p_L, p_R), the scale on which
        # p* enters the force sum
        pscale = np.maximum(np.abs(ref[:, 0]), np.maximum(L[:, 2], R[:, 2]))
        e32 = (np.abs(ps32 - ref[:, 0]) / pscale).max()
        vscale = np.abs(L[:, 0]) + np.abs(R[:, 0]) + L[:, 3] + R[:, 3]
        ev32 = (np.abs(vs32 - ref[:, 1]) / vscale).max()
        print(f"solver {solver}: p* fp64 {e64:.3e} fp32 {e32:.3e} v* fp64 {ev:.3e} fp32 {ev32:.3e}")
        assert e64 < 1e-13
        assert ev < 1e-12
        assert e32 < 1e-5
        assert ev32 < 1e-5


@pytest.mark.parametrize("prec", [ics.PAIR_FP64, ics.PAIR_FP32])
def test_sedov_energy_drift_gate(prec):
    """The accuracy gate of the benchmark workload at a size the oracle runs
    in seconds: 3D GSPH Sedov 32^3, 30 steps; the total-energy drift of the
    GPU run stays within 1e-4 of the oracle's (FP32 pairs) or 1e-9 (FP64
    pairs), and the total momentum (zero initially) stays at the rounding
    level of the pair precision."""
    p, f = ics.sedov_3d(n1=32)
    steps = 30
    o = Oracle(p, f)
    o.initialize()
    e0 = o.energy()
    o.integrate(steps)
    eo = o.energy()
    g = gpu_ctx(p, f, pair_precision=prec)
    g.initialize()
    g0 = g.energy_momentum()
    g.integrate(steps)
    eg = g.energy_momentum()
    drift_o = (eo[3] - e0[3]) / abs(e0[3])
    drift_g = (eg[3] - g0[3]) / abs(g0[3])
    print(f"prec={prec}: t oracle {o.time:.6e} gpu {g.time:.6e}; drift oracle {drift_o:.6e} gpu {drift_g:.6e} "
          f"|diff| {abs(drift_g - drift_o):.3e}; momentum {eg[4]:.3e} {eg[5]:.3e} {eg[6]:.3e}")
    assert abs(drift_g - drift_o) <= (1e-9 if prec == ics.PAIR_FP64 else 1e-4)
    # sum_i m_i |v_i| <= sqrt(2 K M) (total mass M = 1): momentum relative to it
    pscale = np.sqrt(2.0 * eg[0] * f[ics.F_MASS].sum())
    assert np.all(np.abs(eg[4:7]) <= (1e-12 if prec == ics.PAIR_FP64 else 1e-5) * pscale)


def test_deferred_search_overflow_reruns_pre():
    """In a time step the pre-interaction search's overflow flag is read
    together with the force-list flag, after the pre-interaction kernels ran
    on the (capacity-clipped) lists.  An overflow then grows the lists and
    runs the pre-interaction again from the same search radii, so the step
    equals the one the separate entry points take (each checks its search at
    once).  Densities cut 100x on 20 particles make their next h guess 4.6x
    larger: about 100x the usual candidate count, far above the lists'
    capacity (and below the reference's limit of nn * 1000)."""
    p, f = ics.sedov_3d(n1=24)
    ctxs = []
    for _ in range(2):
        c = gpu_ctx(p, f)
        c.initialize()
        g = c.download()
        g[ics.F_DENS, :20] *= 0.01
        c.upload(g, mask=1 << ics.F_DENS)
        c.reset_stats()
        ctxs.append(c)
    a, b = ctxs
    a.integrate(1)
    b.timestep()
    b.predict()
    b.build_tree()
    b.pre_interaction()
    b.fluid_force()
    b.gravity()
    b.correct()
    ga, gb = a.download(), b.download()
    for r in (ics.F_POS, ics.F_VEL, ics.F_ENE, ics.F_DENS, ics.F_PRES, ics.F_SML, ics.F_ACC, ics.F_NEIGHBOR):
        np.testing.assert_array_equal(ga[r], gb[r], err_msg=str(r))
    sa, sb = a.stats(), b.stats()
    # the step's pre-interaction kernel ran twice (on clipped lists searched
    # within the time step's tight radius, then on grown lists at the
    # reference's radius); the separate entry points searched at the
    # reference's radius until the lists fitted and ran it once
    assert sa["pre_candidates"] > 20 * 5000 and sb["pre_candidates"] > 20 * 5000
    assert sa["pre_newton_iters"] > sb["pre_newton_iters"]


@pytest.mark.parametrize("case", ["sedov_3d", "evrard_lattice"])
@pytest.mark.parametrize("prec", [ics.PAIR_FP64, ics.PAIR_FP32])
def test_tight_search_radius_reruns_pre(case, prec):
    """The time step searches the pre-interaction candidates within
    (1 + eps) h_guess instead of the reference's 1.2 h_guess; a particle
    whose Newton iterates reach that radius makes the step run its
    pre-interaction again at the reference's radius.  Densities raised 1.3x
    on 20 particles shrink their h guess 1.3^(-1/D) = 0.92x, so their h ends
    ~9% above the guess: past the tight radius, inside the reference's.  The
    step must equal the one the separate entry points take (they search at
    the reference's radius), field for field, and must have searched twice;
    an unperturbed step searches once.  The Evrard lattice (DISPH with the
    time-dependent AV) checks that the repeated run starts its alpha update
    from the step's alpha, not from the first run's result."""
    if case == "sedov_3d":
        p, f = ics.sedov_3d(n1=24)
    else:
        p, f = ics.evrard_lattice_3d(n_side=24)
    dim = p["dim"]
    ctxs = []
    for bump in (1.3, 1.3, 1.0):
        c = gpu_ctx(p, f, pair_precision=prec)
        c.initialize()
        g = c.download()
        g[ics.F_DENS, 100:120] *= bump
        c.upload(g, mask=1 << ics.F_DENS)
        c.enable_timing(1)
        ctxs.append(c)
        if len(ctxs) == 1:
            dens_raised = g[ics.F_DENS, 100:120].copy()
    a, b, quiet = ctxs
    a.integrate(1)
    quiet.integrate(1)
    b.timestep()
    b.predict()
    b.build_tree()
    b.pre_interaction()
    b.fluid_force()
    b.gravity()
    b.correct()
    ga, gb = a.download(), b.download()
    for r in (ics.F_POS, ics.F_VEL, ics.F_ENE, ics.F_DENS, ics.F_PRES, ics.F_SML, ics.F_ACC, ics.F_NEIGHBOR,
              ics.F_ALPHA, ics.F_BALSARA):
        np.testing.assert_array_equal(ga[r], gb[r], err_msg=str(r))
    area = {1: 2.0, 2: np.pi, 3: 4.0 * np.pi / 3.0}[dim]
    h_guess = (p["neighbor_number"] * f[ics.F_MASS, 100:120] / (dens_raised * area)) ** (1 / dim)
    ratio = ga[ics.F_SML, 100:120] / h_guess
    searches = [c.kernel_times()["search_pre"][1] for c in (a, b, quiet)]
    print(f"{case} prec={prec}: h / h_guess of the raised particles {ratio.min():.4f}..{ratio.max():.4f}; "
          f"searches: step {searches[0]}, entry points {searches[1]}, unperturbed step {searches[2]}")
    assert ratio.max() > 1.05
    assert searches == [2, 1, 1]


@pytest.mark.parametrize("n1,fits", [(33, True), (37, False)])
def test_list_grows_to_reference_limit(n1, fits):
    """A particle whose density is cut 1e6x gets a pre-interaction search
    radius that covers the whole box, so its candidate list holds every
    particle.  The reference's list takes nn * 1000 = 50,000 entries
    (include/defines.hpp:35) and throws only beyond that
    (src/bhtree.cpp:300-302): with 35,937 particles the GPU lists grow to fit
    and the fields equal the oracle's; with 50,653 the GPU returns the
    reference's overflow error."""
    p, f = ics.sedov_3d(n1=n1)
    f = f.copy()
    f[ics.F_DENS, 0] *= 1e-6
    g = gpu_ctx(p, f, pair_precision=ics.PAIR_FP64)
    g.build_tree()
    if not fits:
        with pytest.raises(sphgpu.SphGpuError, match="Neighbor list overflow"):
            g.pre_interaction()
        return
    g.reset_stats()
    g.pre_interaction()
    st = g.stats()
    n = f.shape[1]
    assert st["pre_candidates"] >= 2 * n - 1  # particle 0 sees all n
    o = Oracle(p, f)
    o.make_tree()
    o.pre_interaction()
    of, gf = o.get_fields(), g.download()
    for r in (ics.F_SML, ics.F_DENS, ics.F_PRES):
        e = rel_err(gf[r], of[r])
        assert e.max() < 1e-9, (r, float(e.max()))
    print(f"n={n}: {st['pre_candidates']} candidates, particle 0 h {gf[ics.F_SML, 0]:.6e} "
          f"(oracle {of[ics.F_SML, 0]:.6e})")
