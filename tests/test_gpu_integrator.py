"""The GPU's FP64 kick-drift-kick update, periodic wrap and energy/momentum
reductions (sphgpu_predict / sphgpu_correct / sphgpu_energy_momentum in place
of src/solver.cpp:1047-1060, 1106-1111, 1149-1175) against the closed forms
that tests/test_oracle_integrator.py holds the oracle to: exact on dyadic
states, whatever the reduction order."""
from __future__ import annotations

import numpy as np
import pytest

import ics
from parity_util import dyadic_kdk_state

pytestmark = pytest.mark.gpu
sphgpu = pytest.importorskip("sphgpu")


def _ctx(p, f, dt):
    c = sphgpu.Context(p, f.shape[1])
    c.upload(f)
    c.dt = dt
    return c


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_gpu_kdk_exact_under_constant_acceleration(dim):
    p, f, dt = dyadic_kdk_state(n=257, dim=dim, periodic=True, seed=dim)
    c = _ctx(p, f, dt)
    c.predict()
    mid = c.download()
    c.correct()
    out = c.download()
    c.close()
    for k in range(dim):
        x0, v0, a = f[ics.F_POS + k], f[ics.F_VEL + k], f[ics.F_ACC + k]
        np.testing.assert_array_equal(mid[ics.F_VEL_P + k], v0 + a * dt / 2)
        np.testing.assert_array_equal(out[ics.F_POS + k], x0 + v0 * dt + a * dt * dt / 2)
        np.testing.assert_array_equal(out[ics.F_VEL + k], v0 + a * dt)
    u0, du = f[ics.F_ENE], f[ics.F_DENE]
    np.testing.assert_array_equal(mid[ics.F_ENE_P], u0 + du * dt / 2)
    np.testing.assert_array_equal(out[ics.F_ENE], u0 + du * dt)
    g = p["gamma"]
    np.testing.assert_allclose(out[ics.F_SOUND] ** 2, g * (g - 1.0) * out[ics.F_ENE], rtol=4e-16, atol=0)


@pytest.mark.parametrize("periodic", [True, False])
def test_gpu_periodic_apply_wraps_by_one_period(periodic):
    p, f, dt = dyadic_kdk_state(n=6, dim=3, periodic=periodic)
    f[ics.F_ACC:ics.F_ACC + 3] = 0.0
    f[ics.F_POS, :4] = [0.5 - 2.0 ** -8, -0.5 + 2.0 ** -8, 0.5 - 2.0 ** -5, 0.0]
    f[ics.F_VEL, :4] = [1.0, -1.0, 1.0, 1.0]
    c = _ctx(p, f, dt)
    c.predict()
    c.correct()
    out = c.download()
    c.close()
    raw = f[ics.F_POS, :4] + f[ics.F_VEL, :4] * dt
    want = [raw[0] - 1.0, raw[1] + 1.0, 0.5, raw[3]] if periodic else raw
    np.testing.assert_array_equal(out[ics.F_POS, :4], want)


def test_gpu_energies_of_designed_states():
    p, f, _ = dyadic_kdk_state(n=4096, dim=3, periodic=True, seed=5)
    f[ics.F_MASS] = 2.0 ** -12
    v0 = np.array([0.75, -1.5, 0.25])
    for k in range(3):
        f[ics.F_VEL + k] = v0[k]
    f[ics.F_PHI] = -np.arange(1, 4097) * 2.0 ** -10
    M = f[ics.F_MASS].sum()
    c = _ctx(p, f, 0.0)
    e = c.energy_momentum()
    c.close()
    assert e[0] == 0.5 * M * float(v0 @ v0)
    assert e[1] == float((f[ics.F_MASS] * f[ics.F_ENE]).sum())  # dyadic: exact in any order
    assert e[2] == 0.0
    np.testing.assert_array_equal(e[4:7], M * v0)
    c = _ctx(dict(p, use_gravity=1), f, 0.0)
    eg = c.energy_momentum()
    c.close()
    assert eg[2] == 0.5 * float((f[ics.F_MASS] * f[ics.F_PHI]).sum())
    assert eg[3] == eg[0] + eg[1] + eg[2]
