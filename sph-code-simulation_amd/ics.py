"""Initial conditions and parameter sets for the five benchmark workloads.

Each builder returns ``(params, fields)`` where ``params`` is a dict of
``sphgpu_params`` members and ``fields`` is the FP64 host block
``[SPHGPU_NFIELDS, n]`` (index == particle id) that ``sphgpu_upload`` takes.

The reference's own sample generators do not produce these workloads (its
Sedov sample is 2D only, src/sample/sedov.cpp:16-18; its KHI, Evrard and
shock-tube samples use other layouts, src/sample/{khi,evrard,shock_tube}.cpp),
so the ICs are built here from the workload definitions.  Parameters not named
by a workload take the reference's code defaults (src/solver.cpp:299-438);
the Evrard workload takes the rest of sample/evrard/evrard.json.
"""
from __future__ import annotations

import math

import numpy as np

# field rows, include/sphgpu.h
F_POS, F_VEL, F_VEL_P, F_ACC = 0, 3, 6, 9
F_MASS, F_DENS, F_PRES, F_ENE, F_ENE_P, F_DENE = 12, 13, 14, 15, 16, 17
F_SML, F_SOUND, F_BALSARA, F_ALPHA, F_GRADH, F_PHI = 18, 19, 20, 21, 22, 23
F_ID, F_NEIGHBOR, F_GRAD_D, F_GRAD_P, F_GRAD_V = 24, 25, 26, 29, 32
NFIELDS = 41

SSPH, DISPH, GSPH, GDISPH = 0, 1, 2, 3
CUBIC_SPLINE, WENDLAND, GAUSSIAN = 0, 1, 2
RIEMANN_HLL, RIEMANN_ITERATIVE = 0, 1
PAIR_FP32, PAIR_FP64 = 0, 1
GUESS_PLAIN, GUESS_LEARNED = 0, 1

DEFAULT_SEED = 20261202


def default_params(dim: int) -> dict:
    """Reference code defaults, src/solver.cpp:299-438."""
    return dict(
        dim=dim, sph_type=SSPH, kernel=CUBIC_SPLINE, riemann_solver=RIEMANN_HLL,
        neighbor_number=50, leaf_particle_num=1, max_tree_level=20,
        iterative_sml=1, use_balsara_switch=1, use_time_dependent_av=0,
        use_ac=0, periodic=0, use_gravity=0, gsph_2nd_order=1,
        pair_precision=PAIR_FP64, riemann_guess=GUESS_PLAIN,
        gamma=1.6666666666666667, cfl_sound=0.3, cfl_force=0.125,
        av_alpha=1.0, alpha_max=2.0, alpha_min=0.1, epsilon_av=0.2,
        alpha_ac=1.0, gravity_constant=1.0, theta=0.5,
        range_min=[0.0, 0.0, 0.0], range_max=[0.0, 0.0, 0.0],
    )


def empty_fields(n: int) -> np.ndarray:
    f = np.zeros((NFIELDS, n), dtype=np.float64)
    f[F_ID] = np.arange(n, dtype=np.float64)
    return f


def _finish(f: np.ndarray, params: dict) -> np.ndarray:
    """alpha, balsara and sound as Solver::initialize sets them
    (src/solver.cpp:947-953); the library recomputes them too."""
    g = params["gamma"]
    f[F_ALPHA] = params["av_alpha"]
    f[F_BALSARA] = 1.0
    f[F_SOUND] = np.sqrt(g * (g - 1.0) * f[F_ENE])
    return f


def sod_1d(n_left: int = 400, n_right: int = 100) -> tuple[dict, np.ndarray]:
    """1D Sod-type tube, SSPH, cubic spline, 32 neighbours, x in [-0.5, 0.5],
    equal-mass particles: 400 at rho=1, p=1 on the left half, 100 at
    rho=0.25, p=0.1795 on the right half, v=0, gamma=1.4, non-periodic."""
    p = default_params(1)
    p.update(sph_type=SSPH, kernel=CUBIC_SPLINE, neighbor_number=32, gamma=1.4,
             periodic=0)
    n = n_left + n_right
    f = empty_fields(n)
    dxl = 0.5 / n_left
    dxr = 0.5 / n_right
    xl = -0.5 + (np.arange(n_left) + 0.5) * dxl
    xr = (np.arange(n_right) + 0.5) * dxr
    f[F_POS] = np.concatenate([xl, xr])
    rho = np.concatenate([np.full(n_left, 1.0), np.full(n_right, 0.25)])
    pres = np.concatenate([np.full(n_left, 1.0), np.full(n_right, 0.1795)])
    f[F_MASS] = 1.0 * dxl  # equal masses: 1*0.5/400 == 0.25*0.5/100
    f[F_DENS] = rho
    f[F_PRES] = pres
    f[F_ENE] = pres / ((p["gamma"] - 1.0) * rho)
    return p, _finish(f, p)


def khi_2d(nx: int = 256) -> tuple[dict, np.ndarray]:
    """2D Kelvin-Helmholtz, DISPH, Wendland C4, 64 neighbours, periodic unit
    box, nx*nx lattice.  |y-0.5| < 0.25: rho=2, vx=-0.5; else rho=1, vx=+0.5;
    p=2.5; vy = 0.025 sin(4 pi x); gamma = 5/3."""
    p = default_params(2)
    p.update(sph_type=DISPH, kernel=WENDLAND, neighbor_number=64, periodic=1,
             gamma=5.0 / 3.0, range_min=[0.0, 0.0, 0.0], range_max=[1.0, 1.0, 0.0])
    n = nx * nx
    f = empty_fields(n)
    dx = 1.0 / nx
    c = (np.arange(nx) + 0.5) * dx
    x, y = np.meshgrid(c, c, indexing="ij")
    x = x.ravel()
    y = y.ravel()
    band = np.abs(y - 0.5) < 0.25
    rho = np.where(band, 2.0, 1.0)
    f[F_POS] = x
    f[F_POS + 1] = y
    f[F_VEL] = np.where(band, -0.5, 0.5)
    f[F_VEL + 1] = 0.025 * np.sin(4.0 * math.pi * x)
    f[F_MASS] = rho * dx * dx
    f[F_DENS] = rho
    f[F_PRES] = 2.5
    f[F_ENE] = 2.5 / ((p["gamma"] - 1.0) * rho)
    return p, _finish(f, p)


def _cubic_w3(r: np.ndarray, h: float) -> np.ndarray:
    """3D cubic spline with support r < h (include/kernel/cubic_spline.hpp:27-32)."""
    h_ = 0.5 * h
    q = r / h_
    a = np.maximum(2.0 - q, 0.0)
    b = np.maximum(1.0 - q, 0.0)
    return (1.0 / math.pi) / h_ ** 3 * (0.25 * a ** 3 - b ** 3)


SEDOV_P_BACKGROUND = 0.1


def sedov_3d(n1: int = 64, seed: int = DEFAULT_SEED, jitter: float = 0.05,
             e_blast: float = 1.0, n_deposit: int = 32,
             p_background: float = SEDOV_P_BACKGROUND,
             riemann_guess: int = GUESS_PLAIN) -> tuple[dict, np.ndarray]:
    """3D Sedov blast, GSPH 2nd order with the iterative Riemann solver, cubic
    spline, 50 neighbours, periodic box [-0.5, 0.5]^3, n1^3 lattice jittered by
    a seeded uniform +-jitter*dx.  rho=1, background pressure p_background;
    energy e_blast deposited kernel-weighted (h = distance of the
    (n_deposit+1)-th nearest particle) into the n_deposit particles nearest the
    origin.  gamma = 5/3, leaf 32.

    The background pressure defaults to 0.1, not 1e-5: the reference's own
    second-order GSPH (either Riemann solver) drives cold particles next to
    the blast to negative internal energy and NaN sound speed in the first
    step when the deposit-to-background contrast is too strong.  With the
    32-particle deposit it fails at p = 1e-5 (16^3), p = 1e-3 (32^3) and
    p = 1e-2 (64^3); p = 0.1 runs at 32^3, 64^3 and 100^3
    (tests/test_oracle_reference.py keeps the failing cases)."""
    p = default_params(3)
    p.update(sph_type=GSPH, kernel=CUBIC_SPLINE, neighbor_number=50, periodic=1,
             gamma=5.0 / 3.0, riemann_solver=RIEMANN_ITERATIVE, gsph_2nd_order=1,
             leaf_particle_num=32, use_balsara_switch=0,
             range_min=[-0.5, -0.5, -0.5], range_max=[0.5, 0.5, 0.5],
             riemann_guess=riemann_guess)
    n = n1 ** 3
    f = empty_fields(n)
    dx = 1.0 / n1
    c = -0.5 + (np.arange(n1) + 0.5) * dx
    x, y, z = np.meshgrid(c, c, c, indexing="ij")
    pos = np.stack([x.ravel(), y.ravel(), z.ravel()])
    rng = np.random.Generator(np.random.PCG64(seed))
    pos += rng.uniform(-jitter * dx, jitter * dx, size=pos.shape)
    f[F_POS:F_POS + 3] = pos
    rho0, p0 = 1.0, p_background
    m = rho0 / n
    f[F_MASS] = m
    f[F_DENS] = rho0
    u0 = p0 / ((p["gamma"] - 1.0) * rho0)
    ene = np.full(n, u0)
    r = np.sqrt((pos ** 2).sum(axis=0))
    order = np.argsort(r, kind="stable")
    near = order[:n_deposit]
    h_dep = r[order[n_deposit]]
    w = _cubic_w3(r[near], h_dep)
    ene[near] += e_blast * w / (w.sum() * m)
    f[F_ENE] = ene
    f[F_PRES] = (p["gamma"] - 1.0) * rho0 * ene
    return p, _finish(f, p)


def evrard_3d(n: int = 1_000_000, seed: int = DEFAULT_SEED) -> tuple[dict, np.ndarray]:
    """3D Evrard collapse: M=1, R=1, G=1, rho ~ 1/r, u=0.05, gamma=5/3; radii
    r = R U^(1/2) (M(r) ~ r^2) with isotropic directions from a seeded draw.
    DISPH with Balsara switch and time-dependent AV, tree gravity theta=0.5;
    neighbour number 32, Wendland C4, leaf 32, CFL 0.2/0.1 from
    sample/evrard/evrard.json."""
    p = default_params(3)
    p.update(sph_type=DISPH, kernel=WENDLAND, neighbor_number=32, periodic=0,
             gamma=5.0 / 3.0, leaf_particle_num=32, use_balsara_switch=1,
             use_time_dependent_av=1, use_gravity=1, gravity_constant=1.0,
             theta=0.5, cfl_sound=0.2, cfl_force=0.1)
    f = empty_fields(n)
    rng = np.random.Generator(np.random.PCG64(seed))
    u = rng.random(n)
    r = np.sqrt(u)
    cost = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * math.pi, n)
    sint = np.sqrt(1.0 - cost ** 2)
    f[F_POS] = r * sint * np.cos(phi)
    f[F_POS + 1] = r * sint * np.sin(phi)
    f[F_POS + 2] = r * cost
    f[F_MASS] = 1.0 / n
    f[F_DENS] = 1.0 / (2.0 * math.pi * np.maximum(r, 1e-6))
    f[F_ENE] = 0.05
    f[F_PRES] = (p["gamma"] - 1.0) * f[F_DENS] * f[F_ENE]
    return p, _finish(f, p)


def evrard_lattice_3d(n_side: int = 124) -> tuple[dict, np.ndarray]:
    """3D Evrard collapse on the reference's own initial lattice
    (src/sample/evrard.cpp:12-86): cell centres of an n_side^3 grid on
    [-1, 1]^3 inside the unit sphere, stretched radially r -> r^1.5 so that
    rho ~ 1/r; equal masses 1/n, u = 0.05 G, rho = 1/(2 pi r).  Vectorised
    with the generator's arithmetic (tests/test_samples.py compares it with
    the loop form).  n_side = 124 gives 998,024 particles.  Parameters as in
    evrard_3d (sample/evrard/evrard.json)."""
    p, _ = evrard_3d(n=8)
    dx = 2.0 / n_side
    c = (np.arange(n_side, dtype=np.float64) + 0.5) * dx - 1.0
    x, y, z = np.meshgrid(c, c, c, indexing="ij")
    x, y, z = x.ravel(), y.ravel(), z.ravel()
    r0 = np.sqrt(x * x + y * y + z * z)
    keep = ~(r0 > 1.0)
    x, y, z, r0 = x[keep], y[keep], z[keep], r0[keep]
    # libm's pow, as the generator calls it (numpy's vectorised power may
    # round differently in the last place)
    s = np.array([math.pow(v, 1.5) / v if v > 0.0 else 1.0 for v in r0.tolist()])
    x, y, z = x * s, y * s, z * s
    n = x.size
    f = empty_fields(n)
    f[F_POS], f[F_POS + 1], f[F_POS + 2] = x, y, z
    f[F_MASS] = 1.0 / n
    u = 0.05 * p["gravity_constant"]
    dens = 1.0 / (2.0 * math.pi * np.sqrt(x * x + y * y + z * z))
    f[F_DENS] = dens
    f[F_ENE] = u
    f[F_PRES] = (p["gamma"] - 1.0) * dens * u
    return p, _finish(f, p)


WORKLOADS = {
    "sod_1d_ssph_500": sod_1d,
    "khi_2d_disph_65k": khi_2d,
    "sedov_3d_gsph_262k": sedov_3d,
    "sedov_3d_gsph_1m": lambda **kw: sedov_3d(n1=100, **kw),
    "evrard_3d_disph_1m": evrard_lattice_3d,
}
