"""Helpers shared by the parity tests: workloads, min-image distances and
field comparisons (test infrastructure)."""
from __future__ import annotations

import numpy as np

import ics

# small instances of the five workloads (sizes the oracle finishes in seconds)
def _gdisph(builder, second_order):
    def make():
        p, f = builder()
        p.update(sph_type=ics.GDISPH, gsph_2nd_order=int(second_order))
        return p, f
    return make


def _with(make, **kw):
    def f():
        p, fl = make()
        p.update(kw)
        return p, fl
    return f


SMALL = {
    "sod_1d": lambda: ics.sod_1d(),
    "khi_2d": lambda: ics.khi_2d(nx=32),
    "sedov_3d": lambda: ics.sedov_3d(n1=16),
    "evrard_3d": lambda: ics.evrard_3d(n=3000),
    # GDISPH: MUSCL-reconstructed HLL p* on the KHI, first order on the
    # Sedov (the reference's driver reads use2ndOrderGSPH only for GSPH, so
    # its GDISPH runs first order)
    "khi_2d_gdisph": _gdisph(lambda: ics.khi_2d(nx=32), True),
    "sedov_3d_gdisph": _gdisph(lambda: ics.sedov_3d(n1=16), False),
    # tree gravity in a periodic box: the walk takes minimum-image distances
    # to node centres of mass and leaf members (src/bhtree.cpp:347-393 with
    # Periodic::calc_r_ij)
    "sedov_3d_gravity_periodic": _with(lambda: ics.sedov_3d(n1=12), use_gravity=1),
}


def min_image(d: np.ndarray, L: float) -> np.ndarray:
    """Periodic::calc_r_ij per axis, tie order (d, d+L, d-L)
    (include/periodic.hpp:34-59)."""
    r1, r2, r3 = d, d + L, d - L
    a1, a2, a3 = np.abs(r1), np.abs(r2), np.abs(r3)
    out = np.where((a1 <= a2) & (a1 <= a3), r1,
                   np.where((a2 <= a3) & (a2 <= a1), r2, r3))
    return out


def rij(params: dict, fields: np.ndarray, i: int, js: np.ndarray) -> np.ndarray:
    dim = params["dim"]
    out = np.zeros((len(js), dim))
    for k in range(dim):
        d = fields[ics.F_POS + k, i] - fields[ics.F_POS + k, js]
        if params["periodic"]:
            L = params["range_max"][k] - params["range_min"][k]
            d = min_image(d, L)
        out[:, k] = d
    return out


def r2_ref(params, fields, i, js):
    """inner_product(r_ij, r_ij) with the reference's association."""
    d = rij(params, fields, i, js)
    s = d[:, 0] * d[:, 0]
    for k in range(1, d.shape[1]):
        s = s + d[:, k] * d[:, k]
    return s


def force_pairs(params, fields, i, cand):
    """Filter a FluidForce candidate list with the pair test of
    src/fluid_force.cpp:62 (r >= max(h_i, h_j) || r == 0 -> skip)."""
    cand = np.asarray(cand, dtype=np.int64)
    if len(cand) == 0:
        return cand
    r = np.sqrt(r2_ref(params, fields, i, cand))
    h = fields[ics.F_SML]
    H = np.maximum(h[i], h[cand])
    keep = ~((r >= H) | (r == 0.0))
    return cand[keep]


def scale_err(a: np.ndarray, b: np.ndarray, floor: float = 0.0) -> float:
    """max |a-b| / max(max|b|, floor)"""
    s = max(float(np.max(np.abs(b))), floor)
    if s == 0.0:
        return float(np.max(np.abs(a - b)))
    return float(np.max(np.abs(a - b))) / s


def rel_err(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)


def vec_err(g: np.ndarray, o: np.ndarray, row: int, dim: int) -> float:
    """Per-particle vector error relative to the largest vector magnitude."""
    gv = g[row:row + dim]
    ov = o[row:row + dim]
    mag = np.sqrt((ov ** 2).sum(axis=0))
    err = np.sqrt(((gv - ov) ** 2).sum(axis=0))
    return float(err.max() / max(mag.max(), 1e-300))


def dyadic_kdk_state(n: int = 64, dim: int = 3, periodic: bool = True, seed: int = 0):
    """A state whose kick-drift-kick update is exact in FP64: positions,
    velocities, accelerations, energies and du/dt are small multiples of
    powers of two and dt = 2^-5, so every product and sum of the update is
    representable.  Returns (params, fields, dt).  The box is [-0.5, 0.5]^dim;
    |x0| <= 1/4, |v0| <= 2, |a| <= 8, so x1 stays inside the box."""
    p = ics.default_params(dim)
    p.update(periodic=int(periodic), gamma=5.0 / 3.0,
             range_min=[-0.5 if k < dim else 0.0 for k in range(3)],
             range_max=[0.5 if k < dim else 0.0 for k in range(3)])
    rng = np.random.default_rng(seed)
    f = ics.empty_fields(n)
    for k in range(dim):
        f[ics.F_POS + k] = rng.integers(-256, 257, n) * 2.0 ** -10
        f[ics.F_VEL + k] = rng.integers(-128, 129, n) * 2.0 ** -6
        f[ics.F_ACC + k] = rng.integers(-128, 129, n) * 2.0 ** -4
    f[ics.F_MASS] = 2.0 ** -6
    f[ics.F_ENE] = rng.integers(256, 1024, n) * 2.0 ** -8
    f[ics.F_DENE] = rng.integers(-256, 257, n) * 2.0 ** -6
    f[ics.F_DENS] = 1.0
    f[ics.F_SML] = 0.1
    f[ics.F_PRES] = (p["gamma"] - 1.0) * f[ics.F_ENE]
    f[ics.F_SOUND] = np.sqrt(p["gamma"] * (p["gamma"] - 1.0) * f[ics.F_ENE])
    return p, f, 2.0 ** -5


def force_errors(gf, of, d):
    """Acceleration and du/dt errors per particle, normalised by the size of
    one pair's pressure term, P/(rho h) (and P c/(rho h) for du/dt), plus the
    oracle value itself.  Net forces of equilibrium regions are roundoff-level
    sums of such terms, so their relative error is not meaningful."""
    P, rho, h, c = of[ics.F_PRES], of[ics.F_DENS], of[ics.F_SML], of[ics.F_SOUND]
    ga, oa = gf[ics.F_ACC:ics.F_ACC + d], of[ics.F_ACC:ics.F_ACC + d]
    amag = np.sqrt((oa ** 2).sum(axis=0))
    scale_a = np.abs(P) / (rho * h) + amag
    e_acc = np.sqrt(((ga - oa) ** 2).sum(axis=0)) / scale_a
    scale_u = np.abs(P) * c / (rho * h) + np.abs(of[ics.F_DENE])
    e_dene = np.abs(gf[ics.F_DENE] - of[ics.F_DENE]) / scale_u
    return e_acc, e_dene


def bar(name, e, tol, prec):
    """p99.9 below the bar, and no particle beyond it: the maximum stays
    within the bar in the FP64 form and within 10x the bar in the FP32 form
    (a dropped neighbour, a stale pair list or a bad Riemann state shows up
    as one particle far beyond either)."""
    cap = tol if prec == ics.PAIR_FP64 else 10.0 * tol
    q = float(np.quantile(e, 0.999))
    worst = int(np.argmax(e))
    n_out = int((e > cap).sum())
    assert q < tol, f"{name}: p99.9 {q:.3e} >= {tol:.1e}"
    assert n_out == 0, f"{name}: {n_out} particles above {cap:.1e} (worst {worst}: {e[worst]:.3e})"


def gradient_errors(gf, of, d):
    """MUSCL gradient rows (grad rho, grad p, grad v_k; src/gsph/
    g_pre_interaction.cpp:103-132) on the scale of one kernel's worth of
    variation: rho/h, P/h and c/h, plus the oracle gradient itself."""
    h = of[ics.F_SML]
    out = {}
    for name, row, s in (("grad_d", ics.F_GRAD_D, of[ics.F_DENS] / h), ("grad_p", ics.F_GRAD_P, of[ics.F_PRES] / h)):
        g, o = gf[row:row + d], of[row:row + d]
        out[name] = np.sqrt(((g - o) ** 2).sum(axis=0)) / (np.abs(s) + np.sqrt((o ** 2).sum(axis=0)))
    for k in range(d):
        row = ics.F_GRAD_V + 3 * k
        g, o = gf[row:row + d], of[row:row + d]
        s = of[ics.F_SOUND] / h
        out[f"grad_v{k}"] = np.sqrt(((g - o) ** 2).sum(axis=0)) / (s + np.sqrt((o ** 2).sum(axis=0)))
    return out


def compare_fields(case, p, gf, of, prec):
    d = p["dim"]
    t = p["sph_type"]
    if prec == ics.PAIR_FP64:
        tol_rho, tol_f = 1e-9, 1e-9
    else:
        tol_rho, tol_f = 5e-5, 5e-5
    e_sml = rel_err(gf[ics.F_SML], of[ics.F_SML])
    e_rho = rel_err(gf[ics.F_DENS], of[ics.F_DENS])
    e_pres = rel_err(gf[ics.F_PRES], of[ics.F_PRES])
    e_acc, e_dene = force_errors(gf, of, d)
    print(f"{case} prec={prec}: sml {e_sml.max():.3e} (p99.9 {np.quantile(e_sml, 0.999):.3e}) "
          f"dens {e_rho.max():.3e} pres {e_pres.max():.3e} acc {e_acc.max():.3e} "
          f"(p99.9 {np.quantile(e_acc, 0.999):.3e}) dene {e_dene.max():.3e}")
    bar("sml", e_sml, tol_rho, prec)
    bar("dens", e_rho, tol_rho, prec)
    bar("pres", e_pres, tol_rho, prec)
    bar("acc", e_acc, tol_f, prec)
    bar("dene", e_dene, tol_f, prec)
    if t == ics.SSPH:
        e_gh = rel_err(gf[ics.F_GRADH], of[ics.F_GRADH])
        bar("gradh", e_gh, tol_rho * 10, prec)
    elif t in (ics.DISPH, ics.GDISPH):
        # DISPH grad-h (d_pre_interaction.cpp) is a ratio whose denominator
        # 1 + h/(D n) sum dW/dh nearly cancels on a uniform lattice, so its
        # relative error is not meaningful; it enters the force only through
        # f_ij = 1 - gradh_i / (m_j u_j), so it is compared on the m u scale
        mu = of[ics.F_MASS] * of[ics.F_ENE]
        e_gh = np.abs(gf[ics.F_GRADH] - of[ics.F_GRADH]) / (np.abs(of[ics.F_GRADH]) + mu)
        print(f"  gradh on the m u scale {e_gh.max():.3e}")
        bar("gradh", e_gh, tol_rho * 10, prec)
    if t in (ics.GSPH, ics.GDISPH) and p["gsph_2nd_order"]:
        ge = gradient_errors(gf, of, d)
        print("  " + " ".join(f"{k} {v.max():.3e}" for k, v in ge.items()))
        for k, v in ge.items():
            bar(k, v, tol_f, prec)
    if p["use_gravity"]:
        e_phi = rel_err(gf[ics.F_PHI], of[ics.F_PHI])
        print(f"  phi {e_phi.max():.3e}")
        bar("phi", e_phi, 1e-9 if prec == ics.PAIR_FP64 else 1e-5, prec)
    np.testing.assert_array_equal(gf[ics.F_ID], of[ics.F_ID])
