"""Snapshot CSV and energy-log files (CPU): the library's writer and reader
(sphgpu_write_csv_fields / sphgpu_read_csv_fields / sphgpu_energy_file_*)
against the reference's own fixtures (test/checkpoint_csv_resume, copied to
tests/golden/checkpoint_csv_resume) and against the iostream restatement of
CSVWriter / OutputManager in oracle/csv_oracle.cpp (the checker).

The fixtures hold headers and energy rows but no particle rows, so the rows
are pinned by the oracle, which streams every value the way the reference
does (std::scientific, setprecision(16), one operator<< per value), and by
exact round trips: 17 significant digits identify every double.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np
import pytest

import ics
import sphgpu

GOLD = Path(__file__).resolve().parent / "golden" / "checkpoint_csv_resume"
ORACLE_SO = Path(__file__).resolve().parents[1] / "oracle" / "liboracle.so"
NF = sphgpu.NFIELDS


def oracle():
    lib = C.CDLL(str(ORACLE_SO))
    lib.orc_csv_write.argtypes = [C.c_char_p, C.POINTER(sphgpu.OutputMeta), C.c_int, C.c_int,
                                  C.POINTER(C.c_double)]
    lib.orc_csv_read.argtypes = [C.c_char_p, C.POINTER(sphgpu.OutputMeta), C.c_int, C.POINTER(C.c_double),
                                 C.POINTER(C.c_int)]
    lib.orc_energy_begin.argtypes = [C.c_char_p, C.POINTER(sphgpu.OutputMeta)]
    lib.orc_energy_write.argtypes = [C.c_char_p] + [C.c_double] * 4 + [C.c_int]
    return lib


def dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


KERNELS = {"Cubic Spline": ics.CUBIC_SPLINE, "Wendland C4": ics.WENDLAND}
TYPES = {"Standard SPH": ics.SSPH, "Density Independent SPH": ics.DISPH, "Godunov SPH": ics.GSPH}


def meta_from_header(path: Path) -> sphgpu.OutputMeta:
    """Every header value of a reference snapshot file (the test's own parse,
    independent of the library's reader)."""
    kv = {}
    for line in path.read_text().splitlines()[:50]:
        if line.startswith("# ") and ": " in line:
            k, v = line[2:].split(": ", 1)
            kv[k] = v
    unit = lambda s: (s.split(" (")[0], float(s.split(" (")[1].split(" ")[0]))  # noqa: E731
    lu, lc = unit(kv["Length"])
    mu, mc = unit(kv["Mass"])
    tu, tc = unit(kv["Time"])
    vu, vc = unit(kv["Velocity"])
    en = lambda k: float(kv[k].split(" [")[0])  # noqa: E731
    first = path.read_text().splitlines()[0]
    return sphgpu.output_meta(
        format_version=int(first.rsplit("v", 1)[1]), timestamp=kv["Timestamp"], simulation_name=kv["Simulation"],
        unit_type=kv["Type"], length_unit=lu, mass_unit=mu, time_unit=tu, velocity_unit=vu,
        length_to_cgs=lc, mass_to_cgs=mc, time_to_cgs=tc, velocity_to_cgs=vc,
        step=int(kv["Step"]), time_code=float(kv["Time (code)"]),
        time_physical=float(kv["Time (physical)"].split(" ")[0]), particle_count=int(kv["Particle Count"]),
        gamma=float(kv["Gamma"]), gravitational_constant=float(kv["G"]),
        neighbor_number=int(kv["Neighbor Number"]), sph_type=TYPES[kv["SPH Type"]], kernel=KERNELS[kv["Kernel"]],
        use_gravity=int(kv["Gravity"] == "enabled"), use_balsara=int(kv["Balsara Switch"] == "enabled"),
        use_time_dependent_av=int(kv["Time-Dependent AV"] == "enabled"),
        kinetic_energy=en("Kinetic"), thermal_energy=en("Thermal"), potential_energy=en("Potential"),
        total_energy=en("Total"))


@pytest.mark.parametrize("name", ["snapshot_0000.csv", "snapshot_0001.csv"])
def test_header_matches_reference_fixture(tmp_path, name):
    gold = (GOLD / name).read_bytes()
    m = meta_from_header(GOLD / name)
    empty = np.zeros((NF, 0))
    sphgpu.write_csv_fields(tmp_path / "lib.csv", m, 3, empty)
    assert (tmp_path / "lib.csv").read_bytes() == gold
    assert oracle().orc_csv_write(str(tmp_path / "orc.csv").encode(), C.byref(m), 3, 0, dp(np.zeros(NF))) == 0
    assert (tmp_path / "orc.csv").read_bytes() == gold
    # the reference's reader: header only, no particle -> read_particles fails
    assert sphgpu.csv_row_count(GOLD / name) == 0
    with pytest.raises(sphgpu.SphGpuError):
        sphgpu.read_csv_fields(GOLD / name, 3, 1)


def test_energy_file_matches_reference_fixture(tmp_path):
    gold = (GOLD / "energy.dat").read_text()
    lines = gold.splitlines()
    m = sphgpu.output_meta(unit_type=lines[1].split(": ")[1], time_unit=lines[2].split(": ")[1],
                           energy_unit=lines[3].split(": ")[1])
    out = tmp_path / "energy.dat"
    sphgpu.energy_file_begin(out, m)
    for row in lines[6:]:
        t, k, th, p, _ = (float(x) for x in row.split())
        sphgpu.energy_file_write(out, t, k, th, p, 16)
    assert out.read_text() == gold


def random_fields(n: int, dim: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((NF, n)) * 10.0 ** rng.integers(-12, 12, size=(NF, n))
    f[ics.F_ID] = np.arange(n)  # the block is indexed by id
    f[ics.F_NEIGHBOR] = rng.integers(0, 200, n)
    # values whose printing is easy to get wrong: signed zero, ties and
    # near-ties in the 17th digit, extremes of the normal range
    special = [0.0, -0.0, 1.0, -1.0, 0.1, 1e-5, 9.9999999999999995e-1, 1e22, 1.7976931348623157e308,
               2.2250738585072014e-308, 1e-300, 123456789012345678.0, 0.5, 2.5e-16,
               1 / 3, 2 / 3, np.pi, -np.e]
    for k, v in enumerate(special):
        f[ics.F_MASS + (k % 10), k % n] = v
    return f


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_rows_match_iostream_oracle_and_round_trip(tmp_path, dim):
    n = 3000
    f = random_fields(n, dim, 11 + dim)
    m = sphgpu.output_meta(step=7, time_code=0.125, time_physical=0.125, particle_count=n, gamma=5 / 3,
                           timestamp="2026-01-01T00:00:00Z", sph_type=ics.GSPH, kernel=ics.CUBIC_SPLINE)
    lib_p, orc_p = tmp_path / "lib.csv", tmp_path / "orc.csv"
    sphgpu.write_csv_fields(lib_p, m, dim, f)
    fo = np.ascontiguousarray(f)
    assert oracle().orc_csv_write(str(orc_p).encode(), C.byref(m), dim, n, dp(fo)) == 0
    assert lib_p.read_bytes() == orc_p.read_bytes()
    # read back: every written value exactly, axes beyond dim as 0
    meta, g = sphgpu.read_csv_fields(lib_p, dim, n)
    cols = [ics.F_ID] + [b + a for b in (ics.F_POS, ics.F_VEL, ics.F_ACC) for a in range(dim)] + \
        [ics.F_MASS, ics.F_DENS, ics.F_PRES, ics.F_ENE, ics.F_SML, ics.F_SOUND, ics.F_ALPHA, ics.F_BALSARA,
         ics.F_GRADH, ics.F_PHI, ics.F_NEIGHBOR]
    for r in cols:
        np.testing.assert_array_equal(g[r], f[r], err_msg=str(r))
    others = sorted(set(range(NF)) - set(cols))
    assert not g[others].any()
    assert meta.step == 7 and meta.time_code == 0.125 and meta.particle_count == n and meta.gamma == 5 / 3
    # and the reference's reader (restated) agrees value for value
    fr = np.zeros((NF, n))
    rows = C.c_int(0)
    mo = sphgpu.output_meta()
    assert oracle().orc_csv_read(str(lib_p).encode(), C.byref(mo), n, dp(fr), C.byref(rows)) == 0
    assert rows.value == n
    for r in cols:
        np.testing.assert_array_equal(fr[r], fo[r])
    assert mo.as_dict()["time_code"] == meta.time_code


def test_relaxation_header_and_reader(tmp_path):
    m = sphgpu.output_meta(is_relaxation=1, relaxation_step=12, relaxation_total_steps=500,
                           accumulated_time=0.5, alpha_scaling=0.25, rho_center=2.0, K=0.4, R=1.5, M_total=3.0,
                           timestamp="2026-01-01T00:00:00Z")
    f = random_fields(4, 3, 3)
    p = tmp_path / "relax.csv"
    sphgpu.write_csv_fields(p, m, 3, f)
    assert oracle().orc_csv_write(str(tmp_path / "o.csv").encode(), C.byref(m), 3, 4,
                                  dp(np.ascontiguousarray(f))) == 0
    assert p.read_bytes() == (tmp_path / "o.csv").read_bytes()
    # 59 header lines: the reference's reader still skips 51 lines, so it
    # reads the last 8 comment lines as rows and fails; so does this one
    assert len(p.read_text().splitlines()) == 59 + 1 + 4
    with pytest.raises(sphgpu.SphGpuError):
        sphgpu.read_csv_fields(p, 3, 4)


def test_reader_rejects_what_the_reference_rejects(tmp_path):
    n = 5
    f = random_fields(n, 3, 5)
    m = sphgpu.output_meta(timestamp="2026-01-01T00:00:00Z")
    good = tmp_path / "good.csv"
    sphgpu.write_csv_fields(good, m, 3, f)
    text = good.read_text().splitlines(keepends=True)
    orc = oracle()

    def both_fail(lines):
        p = tmp_path / "bad.csv"
        p.write_text("".join(lines))
        with pytest.raises(sphgpu.SphGpuError):
            sphgpu.read_csv_fields(p, 3, n)
        fr = np.zeros((NF, n))
        rows = C.c_int(0)
        assert orc.orc_csv_read(str(p).encode(), C.byref(sphgpu.output_meta()), n, dp(fr), C.byref(rows)) != 0

    both_fail(text[:40])                                              # fewer than 51 lines
    both_fail(text[:51] + [text[51].rsplit(",", 1)[0] + "\n"] + text[52:])   # 20 columns
    both_fail(text[:51] + [text[51].replace(",", ",x", 1)] + text[52:])      # a field that does not parse
    both_fail(text[:51] + [text[51].rstrip("\n") + ",7\n"] + text[52:])      # 22 columns
    sub = text[51].split(",")
    sub[11] = "1.0000000000000000e-310"                                     # ERANGE: std::stod throws
    both_fail(text[:51] + [",".join(sub)] + text[52:])
    # blank lines and a trailing comma are accepted by both (std::getline)
    ok = text[:51] + ["\n"] + [text[51].rstrip("\n") + ",\n"] + text[52:]
    p = tmp_path / "ok.csv"
    p.write_text("".join(ok))
    _, g = sphgpu.read_csv_fields(p, 3, n)
    np.testing.assert_array_equal(g[ics.F_MASS], f[ics.F_MASS])


def test_snapshot_filename():
    assert sphgpu.snapshot_filename("out", "snapshot", 7, ".csv") == "out/snapshot_0007.csv"
    assert sphgpu.snapshot_filename("o", "snapshot", 12345, ".h5") == "o/snapshot_12345.h5"
