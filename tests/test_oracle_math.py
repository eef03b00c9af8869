"""Closed forms and invariants that pin the oracle independently of the
reference build (kernels, Riemann solvers, neighbour search, Newton h,
density, momentum conservation, tree gravity).  Each check is chosen so a
plausible slip (a dropped term, a wrong sign or index, a transposed operand)
fails it."""
from __future__ import annotations

import math

import numpy as np
import pytest
from scipy.integrate import quad
from scipy.optimize import brentq

import ics
from checkers import Oracle, oracle_kernel, oracle_riemann
from parity_util import r2_ref

KERNELS = [(1, ics.CUBIC_SPLINE), (2, ics.CUBIC_SPLINE), (3, ics.CUBIC_SPLINE),
           (2, ics.WENDLAND), (3, ics.WENDLAND), (1, ics.GAUSSIAN), (3, ics.GAUSSIAN)]


def _w(d, k, r, h):
    return oracle_kernel(d, k, r, h)[0]


@pytest.mark.parametrize("dim,kernel", KERNELS)
def test_kernel_normalisation(dim, kernel):
    """integral of W over space is 1 (cubic/Wendland exactly; Gaussian
    truncated at the 3h used by quad's infinite tail being negligible)."""
    h = 0.7
    surf = {1: 2.0, 2: 2.0 * math.pi, 3: 4.0 * math.pi}[dim]
    upper = h if kernel != ics.GAUSSIAN else 12 * h
    val, _ = quad(lambda r: surf * r ** (dim - 1) * _w(dim, kernel, r, h), 0, upper, limit=400, points=[h / 2])
    assert abs(val - 1.0) < 1e-8


@pytest.mark.parametrize("dim,kernel", KERNELS)
def test_kernel_gradients_match_finite_differences(dim, kernel):
    """grad W = r_ij * dwc, so dW/dr = r * dwc; dhw = dW/dh (the property the
    reference's test/kernel_test/kernel_test.cpp prints)."""
    h = 0.9
    for q in (0.1, 0.3, 0.45, 0.6, 0.8, 0.95):
        r = q * h
        e = 1e-6
        dwdr = (_w(dim, kernel, r + e, h) - _w(dim, kernel, r - e, h)) / (2 * e)
        dwdh = (_w(dim, kernel, r, h + e) - _w(dim, kernel, r, h - e)) / (2 * e)
        _, c, dh = oracle_kernel(dim, kernel, r, h)
        scale = abs(_w(dim, kernel, 0.0, h)) / h
        assert abs(r * c - dwdr) < 1e-6 * scale, (q, r * c, dwdr)
        assert abs(dh - dwdh) < 1e-6 * scale, (q, dh, dwdh)


def test_kernel_closed_forms():
    # cubic spline, D=3, h_ = h/2: W(0) = (1/pi)/h_^3 * (0.25*8 - 1) = 1/(pi h_^3)
    h = 0.8
    assert abs(_w(3, ics.CUBIC_SPLINE, 0.0, h) - 1.0 / (math.pi * (h / 2) ** 3)) < 1e-12
    # support ends at r = h
    assert _w(3, ics.CUBIC_SPLINE, h, h) == 0.0
    assert _w(2, ics.WENDLAND, h, h) == 0.0
    # Wendland C4, D=2: W(0) = 9/(pi h^2)
    assert abs(_w(2, ics.WENDLAND, 0.0, h) - 9.0 / (math.pi * h * h)) < 1e-12
    # cubic, D=1 at q = r/h_ = 1: (2/3)/h_ * 0.25
    assert abs(_w(1, ics.CUBIC_SPLINE, h / 2, h) - (2.0 / 3.0) / (h / 2) * 0.25) < 1e-14


def _toro_two_shock_exact(gam, L, R):
    """Exact p* when both waves are shocks, from Toro (2009) eq. 4.6-4.7
    (shock branch f_K = (p - p_K) sqrt(A_K / (p + B_K))) -- an independent
    formulation of the same wave curves."""
    (uL, rL, pL), (uR, rR, pR) = L, R

    def fk(p, r_, p_):
        A = 2.0 / ((gam + 1.0) * r_)
        B = (gam - 1.0) / (gam + 1.0) * p_
        return (p - p_) * math.sqrt(A / (p + B))

    g = lambda p: fk(p, rL, pL) + fk(p, rR, pR) + (uR - uL)  # noqa: E731
    p = brentq(g, max(pL, pR), 1e6, xtol=1e-15, rtol=1e-15)
    u = 0.5 * (uL + uR) + 0.5 * (fk(p, rR, pR) - fk(p, rL, pL))
    return p, u


@pytest.mark.parametrize("case", [
    ((1.0, 1.0, 1.0), (-1.0, 1.0, 1.0)),        # symmetric collision
    ((2.0, 1.0, 0.5), (-0.5, 0.3, 0.2)),
    ((0.3, 4.0, 10.0), (-3.0, 1.0, 0.01)),
])
def test_iterative_solver_matches_exact_two_shock_solution(case):
    """When p* exceeds both pressures both waves are shocks and the
    reference's two-shock Newton iteration must land on the exact solution."""
    gam = 1.4
    L = np.array([case[0][0], case[0][1], case[0][2], math.sqrt(gam * case[0][2] / case[0][1])])
    R = np.array([case[1][0], case[1][1], case[1][2], math.sqrt(gam * case[1][2] / case[1][1])])
    p_ex, u_ex = _toro_two_shock_exact(gam, L[:3], R[:3])
    ps, vs, it = oracle_riemann(ics.RIEMANN_ITERATIVE, gam, L, R)
    assert p_ex > max(L[2], R[2])
    assert abs(ps - p_ex) / p_ex < 1e-6
    assert abs(vs - u_ex) < 1e-6 * (1 + abs(u_ex))
    assert 1 <= it <= 20


def test_iterative_solver_symmetry_and_guards():
    gam = 5.0 / 3.0
    L = np.array([0.7, 1.3, 0.4, 0.0])
    R = np.array([-0.7, 1.3, 0.4, 0.0])
    ps, vs, _ = oracle_riemann(ics.RIEMANN_ITERATIVE, gam, L, R)
    assert vs == 0.0 and ps > 0.4
    # negative input states return zeros (g_fluid_force.cpp:224-228)
    ps, vs, it = oracle_riemann(ics.RIEMANN_ITERATIVE, gam, np.array([0, -1.0, 1, 1]), R)
    assert ps == 0.0 and vs == 0.0 and it == 0
    # mirrored problem: p* identical, v* negated
    L2 = np.array([0.2, 2.0, 1.5, 0.0])
    R2 = np.array([-0.4, 0.5, 0.1, 0.0])
    a = oracle_riemann(ics.RIEMANN_ITERATIVE, gam, L2, R2)
    b = oracle_riemann(ics.RIEMANN_ITERATIVE, gam, R2 * np.array([-1, 1, 1, 1]), L2 * np.array([-1, 1, 1, 1]))
    assert abs(a[0] - b[0]) < 1e-12 * a[0] and abs(a[1] + b[1]) < 1e-12


def test_hll_solver_properties():
    """HLL: equal states give p* = p, v* = u (consistency), and the
    star state satisfies the HLL jump conditions of the left/right fans."""
    gam = 1.4
    s = np.array([0.3, 1.2, 0.8, math.sqrt(gam * 0.8 / 1.2)])
    ps, vs, _ = oracle_riemann(ics.RIEMANN_HLL, gam, s, s)
    assert abs(ps - 0.8) < 1e-14 and abs(vs - 0.3) < 1e-14
    ps, vs, _ = oracle_riemann(ics.RIEMANN_HLL, gam, np.array([1.0, 1, 1, 1.18321596]),
                               np.array([-1.0, 1, 1, 1.18321596]))
    assert abs(vs) < 1e-14 and ps > 1.0


def _lattice(n1, dim=3, jitter=0.0, seed=0):
    p, f = ics.sedov_3d(n1=n1, jitter=jitter, e_blast=0.0, seed=seed)
    f[ics.F_ENE] = 1.0
    f[ics.F_PRES] = (p["gamma"] - 1.0) * f[ics.F_ENE]
    f[ics.F_SOUND] = np.sqrt(p["gamma"] * (p["gamma"] - 1.0))
    return p, f


def test_uniform_lattice_density_and_newton_fixed_point():
    p, f = _lattice(12)
    o = Oracle(p, f)
    o.initialize()
    g = o.get_fields()
    rho = g[ics.F_DENS]
    assert np.ptp(rho) < 1e-12 * rho.mean()          # lattice symmetry
    assert abs(rho.mean() - 1.0) < 2e-2               # SPH estimate of rho0 = 1
    h = g[ics.F_SML]
    target = p["neighbor_number"] * f[ics.F_MASS] / (4.0 * math.pi / 3.0)
    assert np.all(np.abs(rho * h ** 3 - target) / target < 5e-4)   # converged f(h) = 0


def test_neighbor_search_equals_brute_force():
    p, f = ics.evrard_3d(n=800)
    f[ics.F_SML] = 0.25
    o = Oracle(p, f)
    o.make_tree()
    o.set_kernel()
    n = f.shape[1]
    for i in range(0, n, 7):
        got = np.sort(o.neighbor_search(i, False))
        r2 = r2_ref(p, f, i, np.arange(n))
        ref = np.nonzero(r2 < 0.25 * 0.25)[0]
        np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("sph", [ics.SSPH, ics.DISPH, ics.GSPH])
def test_momentum_conservation(sph):
    p, f = ics.sedov_3d(n1=10)
    p["sph_type"] = sph
    o = Oracle(p, f)
    o.initialize()
    g = o.get_fields()
    m = g[ics.F_MASS]
    a = g[ics.F_ACC:ics.F_ACC + 3]
    tot = np.abs((m * a).sum(axis=1)).max()
    scale = (m * np.sqrt((a ** 2).sum(axis=0))).sum()
    assert tot < 1e-10 * scale, (tot, scale)


def test_tree_gravity_converges_to_direct_sum():
    import ctypes
    from checkers import ORACLE_SO
    lib = ctypes.CDLL(str(ORACLE_SO))
    lib.orc_gravity_direct.argtypes = [ctypes.c_void_p]
    p, f = ics.evrard_3d(n=1500)
    f[ics.F_SML] = 0.05
    res = {}
    for theta in (1e-3, 0.5):
        q = dict(p)
        q["theta"] = theta
        o = Oracle(q, f)
        o.make_tree()
        o.gravity()
        res[theta] = o.get_fields()
        o2 = Oracle(q, f)
        lib.orc_gravity_direct(ctypes.c_void_p(o2.h))
        res["direct"] = o2.get_fields()
    d = res["direct"][ics.F_ACC:ics.F_ACC + 3]
    for theta, tol in ((1e-3, 1e-11), (0.5, 5e-2)):
        t = res[theta][ics.F_ACC:ics.F_ACC + 3]
        err = np.sqrt(((t - d) ** 2).sum(axis=0)) / np.sqrt((d ** 2).sum(axis=0))
        assert np.median(err) < tol, (theta, np.median(err))
    # the direct sum (src/gravity_force.cpp:75-81) runs over j == i too, adding
    # the self term G m f(0, h) = G m * 1.4 / (h/2); the tree walk skips j == i
    self_term = p["gravity_constant"] * f[ics.F_MASS] * 1.4 / (0.5 * f[ics.F_SML])
    np.testing.assert_allclose(res[1e-3][ics.F_PHI] - self_term, res["direct"][ics.F_PHI], rtol=1e-11)
