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
