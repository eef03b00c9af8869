"""The reference's parameter files and sample problems for the device-resident
driver.

* ``params_from_json``: the parameter JSON as Solver::read_parameterfile
  reads it (src/solver.cpp:299-438): the same keys, the same defaults, the
  same keys read only under their switches (alphaMax/alphaMin/epsilonAV with
  useTimeDependentAV, alphaAC with useArtificialConductivity, theta with
  useGravity, use2ndOrderGSPH/riemannSolver with SPHType "gsph" only),
  rangeMin/rangeMax only when periodic.
* ``make_sample``: the initial conditions of src/sample/*.cpp, restated with
  the same loops, so positions accumulate with the same floating-point
  additions (shock_tube, shock_tube_2d, vacuum, strong_shock, khi, sedov,
  gresho_chan_vortex, pairing_instability, hydrostatic,
  pressure_equilibrium, evrard, evrard_cold_collapse).  The sample's default
  N is the reference's (src/solver.cpp:205-260).
* ``load_sample``: the pair (parameters, fields) for a sample name and its
  JSON, ready for ``sphgpu.Context(params, n).upload(fields)``.

Fields are the library's [NFIELDS][n] block indexed by particle id.  These
are host-side generators of small inputs; nothing here runs on the device.
"""
from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np

import ics

# Solver::read_parameterfile: sample name -> (DIM, N key defaults)
SAMPLES = {
    "shock_tube": (1, {"N": 100}),
    "shock_tube_2d": (2, {"Nx": 200, "Ny": 40}),
    "vacuum": (1, {"N": 800}),
    "strong_shock": (1, {"N": 800}),
    "gresho_chan_vortex": (2, {"N": 64}),
    "pairing_instability": (2, {"N": 64}),
    "hydrostatic": (2, {"N": 32}),
    "khi": (2, {"N": 128}),
    "evrard": (3, {"N": 20}),
    "evrard_cold_collapse": (3, {"N": 30}),
    "sedov": (2, {"N": 100}),
    "pressure_equilibrium": (2, {"N": 64}),
}

_TYPES = {"ssph": ics.SSPH, "disph": ics.DISPH, "gsph": ics.GSPH, "gdisph": ics.GDISPH}
_KERNELS = {"cubic_spline": ics.CUBIC_SPLINE, "wendland": ics.WENDLAND, "gaussian": ics.GAUSSIAN}


# boost::property_tree keeps every JSON value as its text, and get<T>(key,
# default) returns the default when that text does not translate to T as a
# whole: an int key given 5.2 or 400.0 reads as its default, a bool key
# accepts only true/false/1/0.
def _int(j: dict, key: str, default: int) -> int:
    v = j.get(key)
    if v is None or isinstance(v, (bool, float)):
        return default
    if isinstance(v, int):
        return v
    try:
        return int(str(v).strip())
    except ValueError:
        return default


def _real(j: dict, key: str, default: float) -> float:
    v = j.get(key)
    if v is None or isinstance(v, bool):
        return default
    try:
        return float(v)
    except (TypeError, ValueError):
        return default


def _flag(j: dict, key: str, default: bool) -> bool:
    v = j.get(key)
    if v is None:
        return default
    if isinstance(v, bool):
        return v
    if isinstance(v, int) and v in (0, 1):
        return bool(v)
    if isinstance(v, str) and v.strip() in ("true", "false", "1", "0"):
        return v.strip() in ("true", "1")
    return default


def params_from_json(src, dim: int) -> dict:
    """Solver::read_parameterfile (src/solver.cpp:299-438) -> parameters."""
    j = json.loads(Path(src).read_text()) if not isinstance(src, dict) else dict(src)
    p = ics.default_params(dim)
    st = str(j.get("SPHType", "ssph"))
    if st not in _TYPES:
        raise ValueError(f"Unknown SPH type: {st}")  # srgsph is outside this build
    p["sph_type"] = _TYPES[st]
    p["cfl_sound"] = _real(j, "cflSound", 0.3)
    p["cfl_force"] = _real(j, "cflForce", 0.125)
    p["av_alpha"] = _real(j, "avAlpha", 1.0)
    p["use_balsara_switch"] = int(_flag(j, "useBalsaraSwitch", True))
    p["use_time_dependent_av"] = int(_flag(j, "useTimeDependentAV", False))
    if p["use_time_dependent_av"]:
        p["alpha_max"] = _real(j, "alphaMax", 2.0)
        p["alpha_min"] = _real(j, "alphaMin", 0.1)
        if p["alpha_max"] < p["alpha_min"]:
            raise ValueError("alphaMax < alphaMin")
        p["epsilon_av"] = _real(j, "epsilonAV", 0.2)
    p["use_ac"] = int(_flag(j, "useArtificialConductivity", False))
    if p["use_ac"]:
        p["alpha_ac"] = _real(j, "alphaAC", 1.0)
    p["max_tree_level"] = _int(j, "maxTreeLevel", 20)
    p["leaf_particle_num"] = _int(j, "leafParticleNumber", 1)
    p["neighbor_number"] = _int(j, "neighborNumber", 50)
    p["gamma"] = _real(j, "gamma", 1.6666666666666667)
    kn = str(j.get("kernel", "cubic_spline"))
    if kn not in _KERNELS:
        raise ValueError("kernel is unknown.")
    p["kernel"] = _KERNELS[kn]
    p["iterative_sml"] = int(_flag(j, "iterativeSmoothingLength", True))
    p["periodic"] = int(_flag(j, "periodic", False))
    if p["periodic"]:
        rmax, rmin = list(j["rangeMax"]), list(j["rangeMin"])
        if len(rmax) != dim or len(rmin) != dim:
            raise ValueError("rangeMax != DIM")
        p["range_max"] = [float(v) for v in rmax] + [0.0] * (3 - dim)
        p["range_min"] = [float(v) for v in rmin] + [0.0] * (3 - dim)
    p["use_gravity"] = int(_flag(j, "useGravity", False))
    p["gravity_constant"] = _real(j, "G", 1.0)
    if p["use_gravity"]:
        p["theta"] = _real(j, "theta", 0.5)
    if p["sph_type"] == ics.GSPH:
        p["gsph_2nd_order"] = int(_flag(j, "use2ndOrderGSPH", True))
        p["riemann_solver"] = ics.RIEMANN_ITERATIVE if str(j.get("riemannSolver", "hll")) == "iterative" \
            else ics.RIEMANN_HLL
    else:
        # SPHParameters is value-initialised: gsph.is_2nd_order = false and
        # the HLL solver for every other type (GDISPH runs first order)
        p["gsph_2nd_order"] = 0
        p["riemann_solver"] = ics.RIEMANN_HLL
    return p


class _Block:
    def __init__(self, n: int):
        self.f = np.zeros((ics.NFIELDS, n))
        self.f[ics.F_ID] = np.arange(n)

    def set(self, i, pos, vel, mass, dens, pres, ene):
        f = self.f
        for k, v in enumerate(pos):
            f[ics.F_POS + k, i] = v
        for k, v in enumerate(vel):
            f[ics.F_VEL + k, i] = v
        f[ics.F_MASS, i] = mass
        f[ics.F_DENS, i] = dens
        f[ics.F_PRES, i] = pres
        f[ics.F_ENE, i] = ene


def _shock_tube(N, gamma, **_):  # src/sample/shock_tube.cpp:12-54
    dx_r = 0.5 / N
    dx_l = dx_r * 0.25
    num = N * 10
    b = _Block(num)
    x, dx, dens, pres = -0.5 + dx_l * 0.5, dx_l, 1.0, 1.0
    mass = 0.5 / N * 0.25
    left = True
    for i in range(num):
        b.set(i, [x], [0.0], mass, dens, pres, pres / ((gamma - 1.0) * dens))
        x += dx
        if x > 0.5 and left:
            x = 0.5 + dx_r * 0.5
            dx, dens, pres, left = dx_r, 0.25, 0.1795, False
    return b.f


def _shock_tube_2d(Nx, Ny, gamma, **_):  # src/sample/shock_tube_2d.cpp:12-81
    Lx, Ly = 1.0, 0.2
    dx, dy = Lx / Nx, Ly / Ny
    num = Nx * Ny
    b = _Block(num)
    x, y = 0.0 + dx * 0.5, 0.0 + dy * 0.5
    for i in range(num):
        if x < 0.5:
            dens, pres, mass = 1.0, 1.0, 1.0 * dx * dy
        else:
            dens, pres, mass = 0.125, 0.1, 0.125 * dx * dy
        b.set(i, [x, y], [0.0, 0.0], mass, dens, pres, pres / ((gamma - 1.0) * dens))
        x += dx
        if x >= Lx:
            x = 0.0 + dx * 0.5
            y += dy
    return b.f


def _vacuum(N, gamma, **_):  # src/sample/vacuum.cpp:12-90
    nl = nr = N // 2
    dx_l, dx_r = 0.5 / nl, 0.5 / nr
    b = _Block(nl + nr)
    mass = 1.0 * 1.0 / (nl + nr)
    ene = 0.4 / ((gamma - 1.0) * 1.0)
    x = -0.5 + dx_l * 0.5
    for i in range(nl):
        b.set(i, [x], [-2.0], mass, 1.0, 0.4, ene)
        x += dx_l
    x = 0.0 + dx_r * 0.5
    for i in range(nr):
        b.set(nl + i, [x], [2.0], mass, 1.0, 0.4, ene)
        x += dx_r
    return b.f


def _strong_shock(N, gamma, **_):  # src/sample/strong_shock.cpp:12-77
    dx = 1.0 / N
    b = _Block(N)
    mass = 1.0 / N
    x = -0.5 + dx * 0.5
    for i in range(N):
        pres = 1000.0 if x < 0.0 else 0.1
        b.set(i, [x], [0.0], mass, 1.0, pres, pres / ((gamma - 1.0) * 1.0))
        x += dx
    return b.f


def _norm(v) -> float:
    s = v[0] * v[0]
    for c in v[1:]:
        s = s + c * c
    return math.sqrt(s)


def _vortex_velocity(r):  # src/sample/gresho_chan_vortex.cpp:10-19
    if r < 0.2:
        return 5.0 * r
    if r < 0.4:
        return 2.0 - 5.0 * r
    return 0.0


def _vortex_pressure(r):  # src/sample/gresho_chan_vortex.cpp:21-30
    if r < 0.2:
        return 5.0 + 12.5 * r * r
    if r < 0.4:
        return 9.0 + 12.5 * r * r - 20.0 * r + 4.0 * math.log(5.0 * r)
    return 3.0 + 4.0 * math.log(2.0)


def _gresho(N, gamma, **_):  # src/sample/gresho_chan_vortex.cpp:33-75
    dx = 1.0 / N
    num = N * N
    b = _Block(num)
    x = y = -0.5 + dx * 0.5
    mass = 1.0 / num
    for i in range(num):
        r = _norm([x, y])
        vel = _vortex_velocity(r)
        dirx, diry = -y / r, x / r
        pres = _vortex_pressure(r)
        b.set(i, [x, y], [dirx * vel, diry * vel], mass, 1.0, pres, pres / ((gamma - 1.0) * 1.0))
        x += dx
        if x > 0.5:
            x = -0.5 + dx * 0.5
            y += dx
    return b.f


class _MT19937:
    """std::mt19937 (Matsumoto & Nishimura 1998), seeded as std::mt19937(s)."""

    def __init__(self, seed: int):
        self.mt = [0] * 624
        self.mt[0] = seed & 0xffffffff
        for i in range(1, 624):
            self.mt[i] = (1812433253 * (self.mt[i - 1] ^ (self.mt[i - 1] >> 30)) + i) & 0xffffffff
        self.k = 624

    def __call__(self) -> int:
        mt = self.mt
        if self.k >= 624:
            for i in range(624):
                y = (mt[i] & 0x80000000) | (mt[(i + 1) % 624] & 0x7fffffff)
                mt[i] = mt[(i + 397) % 624] ^ (y >> 1) ^ (0x9908b0df if y & 1 else 0)
            self.k = 0
        y = mt[self.k]
        self.k += 1
        y ^= y >> 11
        y ^= (y << 7) & 0x9d2c5680
        y ^= (y << 15) & 0xefc60000
        y ^= y >> 18
        return y & 0xffffffff


def _uniform(eng: _MT19937, a: float, b: float) -> float:
    """std::uniform_real_distribution<double>(a, b) with libstdc++'s
    generate_canonical<double, 53>: two 32-bit draws, (d0 + d1 2^32) / 2^64,
    kept below 1."""
    s = float(eng()) + float(eng()) * 4294967296.0
    c = s / 18446744073709551616.0
    if c >= 1.0:
        c = math.nextafter(1.0, 0.0)
    return c * (b - a) + a


def _pairing(N, gamma, **_):  # src/sample/pairing_instability.cpp:12-54
    dx = 1.0 / N
    num = N * N
    b = _Block(num)
    x = y = -0.5 + dx * 0.5
    mass = 1.0 / num
    eng = _MT19937(1)
    for i in range(num):
        px = x + _uniform(eng, -dx * 0.05, dx * 0.05)
        py = y + _uniform(eng, -dx * 0.05, dx * 0.05)
        b.set(i, [px, py], [0.0, 0.0], mass, 1.0, 1.0, 1.0 / ((gamma - 1.0) * 1.0))
        x += dx
        if x > 0.5:
            x = -0.5 + dx * 0.5
            y += dx
    return b.f


def _hydrostatic(N, gamma, **_):  # src/sample/hydrostatic.cpp:12-75
    dx1 = 0.5 / N
    dx2 = dx1 * 2.0
    mass = 1.0 / (N * N)
    rows = []
    x = y = -0.25 + dx1 * 0.5
    while y < 0.25:
        rows.append((x, y, 4.0))
        x += dx1
        if x > 0.25:
            x = -0.25 + dx1 * 0.5
            y += dx1
    x = y = -0.5 + dx2 * 0.5
    while y < 0.5:
        rows.append((x, y, 1.0))
        while True:
            x += dx2
            if x > 0.5:
                x = -0.5 + dx2 * 0.5
                y += dx2
            if not (-0.25 < x < 0.25 and -0.25 < y < 0.25):
                break
    b = _Block(len(rows))
    for i, (px, py, dens) in enumerate(rows):
        b.set(i, [px, py], [0.0, 0.0], mass, dens, 2.5, 2.5 / ((gamma - 1.0) * dens))
    return b.f


def _khi(N, gamma, **_):  # src/sample/khi.cpp:12-77
    num = N * N * 3 // 4
    dx = 1.0 / N
    mass = 1.5 / num
    b = _Block(num)
    x = y = dx * 0.5
    region, odd = 1, True
    sigma2_inv = 2 / (0.05 * 0.05)

    def vy(x, y):
        return 0.1 * math.sin(4.0 * math.pi * x) * (math.exp(-((y - 0.25) * (y - 0.25)) * 0.5 * sigma2_inv) +
                                                     math.exp(-((y - 0.75) * (y - 0.75)) * 0.5 * sigma2_inv))
    for i in range(num):
        dens = float(region)
        b.set(i, [x, y], [-0.5 if region == 1 else 0.5, vy(x, y)], mass, dens, 2.5, 2.5 / ((gamma - 1.0) * dens))
        x += 2.0 * dx if region == 1 else dx
        if x > 1.0:
            y += dx
            region = 2 if (0.25 < y < 0.75) else 1
            if region == 1:
                if odd:
                    odd, x = False, dx * 1.5
                else:
                    odd, x = True, dx * 0.5
            else:
                x = dx * 0.5
    return b.f


def _sedov_2d(N, gamma, **_):  # src/sample/sedov.cpp:12-104
    dx = 1.0 / N
    num = N * N
    bg_p = 1.0e-6
    bg_e = bg_p / ((gamma - 1.0) * 1.0)
    blast_radius = 2.0 * dx
    mass = 1.0 * 1.0 * 1.0 / num
    x = y = -0.5 + dx * 0.5
    count = 0
    for _ in range(num):
        if math.sqrt(x * x + y * y) < blast_radius:
            count += 1
        x += dx
        if x > 0.5:
            x = -0.5 + dx * 0.5
            y += dx
    count = max(count, 1)
    e_blast = 1.0 / (count * mass)
    b = _Block(num)
    x = y = -0.5 + dx * 0.5
    for i in range(num):
        if math.sqrt(x * x + y * y) < blast_radius:
            ene, pres = e_blast, (gamma - 1.0) * 1.0 * e_blast
        else:
            ene, pres = bg_e, bg_p
        b.set(i, [x, y], [0.0, 0.0], mass, 1.0, pres, ene)
        x += dx
        if x > 0.5:
            x = -0.5 + dx * 0.5
            y += dx
    return b.f


def _pressure_equilibrium(N, gamma, **_):  # src/sample/pressure_equilibrium.cpp:12-82
    dx = 1.0 / N
    num = N * N
    b = _Block(num)
    x = y = 0.0 + dx * 0.5
    for i in range(num):
        inside = 0.25 <= x <= 0.75 and 0.25 <= y <= 0.75
        dens = 4.0 if inside else 1.0
        b.set(i, [x, y], [0.0, 0.0], dens * dx * dx, dens, 2.5, 2.5 / ((gamma - 1.0) * dens))
        x += dx
        if x >= 1.0:
            x = 0.0 + dx * 0.5
            y += dx
    return b.f


def _evrard(N, gamma, G, **_):  # src/sample/evrard.cpp:12-86 (and evrard_cold_collapse.cpp)
    dx = 2.0 / N
    pos = []
    for i in range(N):
        for j in range(N):
            for k in range(N):
                r = [(i + 0.5) * dx - 1.0, (j + 0.5) * dx - 1.0, (k + 0.5) * dx - 1.0]
                r0 = _norm(r)
                if r0 > 1.0:
                    continue
                if r0 > 0.0:
                    s = math.pow(r0, 1.5) / r0
                    r = [c * s for c in r]
                pos.append(r)
    n = len(pos)
    mass = 1.0 / n
    u = 0.05 * G
    b = _Block(n)
    for i, r in enumerate(pos):
        dens = 1.0 / (2.0 * math.pi * _norm(r))
        b.set(i, r, [0.0, 0.0, 0.0], mass, dens, (gamma - 1.0) * dens * u, u)
    return b.f


_MAKERS = {
    "shock_tube": _shock_tube, "shock_tube_2d": _shock_tube_2d, "vacuum": _vacuum,
    "strong_shock": _strong_shock, "gresho_chan_vortex": _gresho, "pairing_instability": _pairing,
    "hydrostatic": _hydrostatic, "khi": _khi, "evrard": _evrard, "evrard_cold_collapse": _evrard,
    "sedov": _sedov_2d, "pressure_equilibrium": _pressure_equilibrium,
}


def make_sample(name: str, params: dict, **sample_params) -> np.ndarray:
    """Initial fields of a reference sample (sample_params: N, or Nx/Ny)."""
    dim, defaults = SAMPLES[name]
    if params["dim"] != dim:
        raise ValueError(f"{name} needs DIM == {dim}")
    sp = dict(defaults)
    sp.update(sample_params)
    return _MAKERS[name](gamma=params["gamma"], G=params["gravity_constant"], **sp)


def load_sample(name: str | None, config=None) -> tuple[dict, np.ndarray]:
    """(parameters, fields) of a sample from its parameter JSON (a path or a
    dict; the sample's N keys are read from it with the reference's
    defaults).  With name None the JSON's "sample" key names it, as for
    parameter files given by path (src/solver.cpp:261-296)."""
    cfg = {} if config is None else (json.loads(Path(config).read_text()) if not isinstance(config, dict)
                                     else dict(config))
    if name is None:
        name = cfg.get("sample")
        if name not in SAMPLES:
            raise ValueError(f"no sample named in the parameter file: {name!r}")
    dim, defaults = SAMPLES[name]
    p = params_from_json(cfg, dim)
    sp = {k: _int(cfg, k, v) for k, v in defaults.items()}
    return p, make_sample(name, p, **sp)
