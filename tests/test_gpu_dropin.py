"""The drop-in, end to end: the reference's own modules, Simulation and
driver sequence (oracle/ref_harness.cpp, restating Solver::initialize and
Solver::integrate), with the GPU Module adapters of
sph-code-simulation_amd/dropin in place of PreInteraction, FluidForce,
GravityForce and TimeStep -- the module substitution a reference user makes.  The
run is compared with the same driver on the reference's CPU modules.  The
adapters use the library's default FP32 pair form, so the FP32 tolerances of
tests/test_gpu_parity.py apply.

oracle/_ref/libsphdropin_dD.so is built by `make -C oracle dropin` (from
__graft_entry__.build() where the reference tree is present); without it the
test is skipped.
"""
from __future__ import annotations

import numpy as np
import pytest

import ics
from checkers import DropIn, Reference, dropin_available, ref_available
from parity_util import SMALL, bar, r2_ref, rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", list(SMALL))
def test_dropin_modules_match_reference_modules(case):
    p, f = SMALL[case]()
    d = p["dim"]
    if not (dropin_available(d) and ref_available(d)):
        pytest.skip("oracle/_ref drop-in / reference libraries not built")
    steps = 3
    r = Reference(p, f)
    r.initialize()
    r.integrate(steps)
    rf = r.get_fields()
    g = DropIn(p, f)
    g.initialize()
    g.integrate(steps)
    gf = g.get_fields()
    print(f"{case}: dt ref {r.dt:.6e} drop-in {g.dt:.6e}; t {r.time:.6e} {g.time:.6e}")
    assert abs(g.dt - r.dt) <= 1e-4 * r.dt
    for fld, name in ((ics.F_SML, "sml"), (ics.F_DENS, "dens"), (ics.F_PRES, "pres")):
        e = rel_err(gf[fld], rf[fld])
        print(f"  {name}: max {e.max():.3e} p99.9 {np.quantile(e, 0.999):.3e}")
        bar(name, e, 5e-5, ics.PAIR_FP32)
    P, rho, h = rf[ics.F_PRES], rf[ics.F_DENS], rf[ics.F_SML]
    acc_r, acc_g = rf[ics.F_ACC:ics.F_ACC + d], gf[ics.F_ACC:ics.F_ACC + d]
    scale = np.abs(P) / (rho * h) + np.sqrt((acc_r ** 2).sum(axis=0))
    e_acc = np.sqrt(((acc_g - acc_r) ** 2).sum(axis=0)) / scale
    print(f"  acc: max {e_acc.max():.3e} p99.9 {np.quantile(e_acc, 0.999):.3e}")
    bar("acc", e_acc, 5e-5, ics.PAIR_FP32)
    er, eg = r.energy(), g.energy()
    assert abs(eg[3] - er[3]) <= 1e-5 * abs(er[3])
    # neighbour counts are exact for a given h; h itself carries the FP32
    # error, so a count may differ only by neighbours sitting at r = h within
    # that error (lattice ties)
    mism = np.nonzero(gf[ics.F_NEIGHBOR] != rf[ics.F_NEIGHBOR])[0]
    allj = np.arange(f.shape[1])
    for i in mism:
        r = np.sqrt(r2_ref(p, rf, int(i), allj))
        h = rf[ics.F_SML, i]
        ties = int((np.abs(r - h) <= 1e-5 * h).sum())
        assert abs(int(gf[ics.F_NEIGHBOR, i]) - int(rf[ics.F_NEIGHBOR, i])) <= ties, (int(i), ties)
    print(f"  neighbour counts: {len(mism)} differ, all by ties at r = h")
