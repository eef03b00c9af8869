"""Closed forms that pin the oracle's integrator and energy sums
(Solver::predict / correct / compute_total_energies, src/solver.cpp:1047-1060,
1106-1111, 1149-1175; Periodic::apply, include/periodic.hpp:61-72)
independently of any restatement of them.

* Kick-drift-kick under a constant acceleration a and a constant du/dt is
  exact in real arithmetic: x1 = x0 + v0 dt + a dt^2 / 2, v1 = v0 + a dt,
  u1 = u0 + du dt, with the half-step values v0 + a dt / 2 and u0 + du dt / 2
  in vel_p and ene_p.  On dyadic states (tests/parity_util.dyadic_kdk_state)
  every rounding vanishes, so the oracle must reproduce these bit for bit; a
  dropped half, a drift with the kicked velocity or a correct step from vel
  instead of vel_p each fails.
* The sound speed satisfies the ideal-gas identity c^2 = gamma (gamma - 1) u.
* Periodic::apply wraps a coordinate that left the box by exactly one period
  (strictly beyond a bound only) and leaves a non-periodic run alone.
* Energies of designed states: a uniform velocity v0 gives K = M |v0|^2 / 2
  and P = M v0; U = sum m u; W = sum m phi / 2 only with gravity.
"""
from __future__ import annotations

import numpy as np
import pytest

import ics
from checkers import Oracle
from parity_util import dyadic_kdk_state


def _kdk(p, f, dt):
    o = Oracle(p, f)
    o.dt = dt
    o.predict()
    mid = o.get_fields()
    o.correct()  # a and du/dt unchanged: the force of a constant field
    return mid, o.get_fields()


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_kdk_exact_under_constant_acceleration(dim):
    p, f, dt = dyadic_kdk_state(n=257, dim=dim, periodic=True, seed=dim)
    mid, out = _kdk(p, f, dt)
    for k in range(dim):
        x0, v0, a = f[ics.F_POS + k], f[ics.F_VEL + k], f[ics.F_ACC + k]
        np.testing.assert_array_equal(mid[ics.F_VEL_P + k], v0 + a * dt / 2)
        np.testing.assert_array_equal(out[ics.F_POS + k], x0 + v0 * dt + a * dt * dt / 2)
        np.testing.assert_array_equal(out[ics.F_VEL + k], v0 + a * dt)
        # predict leaves the fully kicked velocity, correct rebuilds it from
        # the half step: both equal v0 + a dt here
        np.testing.assert_array_equal(mid[ics.F_VEL + k], v0 + a * dt)
    u0, du = f[ics.F_ENE], f[ics.F_DENE]
    np.testing.assert_array_equal(mid[ics.F_ENE_P], u0 + du * dt / 2)
    np.testing.assert_array_equal(out[ics.F_ENE], u0 + du * dt)
    g = p["gamma"]
    c2 = out[ics.F_SOUND] ** 2
    np.testing.assert_allclose(c2, g * (g - 1.0) * out[ics.F_ENE], rtol=4e-16, atol=0)


def test_kdk_two_steps_match_closed_form():
    """Two steps of constant a: x2 = x0 + 2 v0 dt + 2 a dt^2 (the update is
    second-order exact for constant a, so no error accumulates)."""
    p, f, dt = dyadic_kdk_state(n=128, dim=3, periodic=False, seed=11)
    f[ics.F_POS:ics.F_POS + 3] *= 0.5  # room for two drifts
    o = Oracle(p, f)
    o.dt = dt
    for _ in range(2):
        o.predict()
        o.correct()
    out = o.get_fields()
    for k in range(3):
        x0, v0, a = f[ics.F_POS + k], f[ics.F_VEL + k], f[ics.F_ACC + k]
        np.testing.assert_array_equal(out[ics.F_POS + k], x0 + 2 * v0 * dt + 2 * a * dt * dt)
        np.testing.assert_array_equal(out[ics.F_VEL + k], v0 + 2 * a * dt)


@pytest.mark.parametrize("periodic", [True, False])
def test_periodic_apply_wraps_by_one_period(periodic):
    p, f, dt = dyadic_kdk_state(n=6, dim=3, periodic=periodic)
    f[ics.F_ACC:ics.F_ACC + 3] = 0.0
    # x-axis: leaves through +0.5, through -0.5, lands exactly on +0.5 (kept:
    # the wrap is strict), and stays inside
    f[ics.F_POS, :4] = [0.5 - 2.0 ** -8, -0.5 + 2.0 ** -8, 0.5 - 2.0 ** -5, 0.0]
    f[ics.F_VEL, :4] = [1.0, -1.0, 1.0, 1.0]
    _, out = _kdk(p, f, dt)
    raw = f[ics.F_POS, :4] + f[ics.F_VEL, :4] * dt
    assert raw[2] == 0.5
    if periodic:
        np.testing.assert_array_equal(out[ics.F_POS, :4], [raw[0] - 1.0, raw[1] + 1.0, 0.5, raw[3]])
    else:
        np.testing.assert_array_equal(out[ics.F_POS, :4], raw)


def test_energies_of_designed_states():
    p, f, _ = dyadic_kdk_state(n=64, dim=3, periodic=True, seed=5)
    v0 = np.array([0.75, -1.5, 0.25])
    for k in range(3):
        f[ics.F_VEL + k] = v0[k]
    f[ics.F_PHI] = np.linspace(-2.0, -1.0, 64)
    M = f[ics.F_MASS].sum()
    assert M == 1.0
    e = Oracle(p, f).energy()
    assert e[0] == 0.5 * M * float(v0 @ v0)
    assert e[1] == pytest.approx(float((f[ics.F_MASS] * f[ics.F_ENE]).sum()), rel=1e-15)
    assert e[2] == 0.0  # no gravity: phi does not count
    np.testing.assert_array_equal(e[4:7], M * v0)
    assert e[3] == e[0] + e[1] + e[2]
    pg = dict(p, use_gravity=1)
    eg = Oracle(pg, f).energy()
    assert eg[2] == pytest.approx(0.5 * float((f[ics.F_MASS] * f[ics.F_PHI]).sum()), rel=1e-15)
    assert eg[2] < 0.0
    assert eg[3] == pytest.approx(eg[0] + eg[1] + eg[2], rel=1e-15)
    # opposite velocities cancel in P but not in K
    f[ics.F_VEL, 32:] = -f[ics.F_VEL, 32:]
    f[ics.F_VEL + 1, 32:] = -f[ics.F_VEL + 1, 32:]
    f[ics.F_VEL + 2, 32:] = -f[ics.F_VEL + 2, 32:]
    e2 = Oracle(p, f).energy()
    np.testing.assert_array_equal(e2[4:7], [0.0, 0.0, 0.0])
    assert e2[0] == e[0]
