"""VTK and HDF5 snapshot files and the metadata JSON (CPU): the library's
writers and readers (sphgpu_write_vtk_fields, sphgpu_output_meta_json /
_from_json in libsphgpu.so; sphgpu_write_hdf5_fields / sphgpu_read_hdf5_fields
in libsphgpu_hdf5.so) against oracle/io_oracle.cpp, which restates VTKWriter
and HDF5Writer (src/writers/vtk_writer.cpp, src/writers/hdf5_writer.cpp) with
the reference's own machinery: std::ofstream for VTK, the HDF5 C API, and
nlohmann::json for the metadata (to_json().dump(2), json::parse, value()).

The reference holds no VTK or HDF5 fixture, so the formats are pinned by the
oracle: VTK byte for byte, HDF5 object by object (groups, attributes,
datasets with type, extent, chunking, deflate level and values; the bytes of
two HDF5 files differ by object timestamps), the JSON text byte for byte over
150,000 doubles, where a shortest-round-trip printer would
differ from nlohmann's Grisu2 in about one value in 1600.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import re
from pathlib import Path

import numpy as np
import pytest

import ics
import sphgpu

ROOT = Path(__file__).resolve().parents[1]
IO_ORACLE = ROOT / "oracle" / "liboracle_io.so"
NF = sphgpu.NFIELDS
MP = C.POINTER(sphgpu.OutputMeta)

needs_io_oracle = pytest.mark.skipif(not IO_ORACLE.exists(), reason="oracle/liboracle_io.so not built (no HDF5)")


def oracle():
    sphgpu.load_library()
    lib = C.CDLL(str(IO_ORACLE))
    dp = C.POINTER(C.c_double)
    lib.orc_meta_json.argtypes = [MP, C.c_int, C.c_char_p, C.c_int]
    lib.orc_meta_from_json.argtypes = [C.c_char_p, MP]
    lib.orc_vtk_write.argtypes = [C.c_char_p, MP, C.c_int, C.c_int, dp, C.c_int]
    lib.orc_hdf5_write.argtypes = [C.c_char_p, MP, C.c_int, C.c_int, dp, C.c_int]
    lib.orc_hdf5_read.argtypes = [C.c_char_p, MP, C.c_int, dp, C.POINTER(C.c_int)]
    lib.orc_hdf5_describe.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
    return lib


def dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def orc_json(lib, m, which=0) -> str:
    buf = C.create_string_buffer(1 << 16)
    assert lib.orc_meta_json(C.byref(m), which, buf, 1 << 16) == 0
    return buf.value.decode()


def describe(lib, path) -> str:
    buf = C.create_string_buffer(1 << 24)
    assert lib.orc_hdf5_describe(str(path).encode(), buf, 1 << 24) == 0
    return buf.value.decode()


DOUBLE_FIELDS = ("length_to_cgs", "mass_to_cgs", "time_to_cgs", "velocity_to_cgs", "time_code", "time_physical",
                 "gamma", "gravitational_constant", "accumulated_time", "alpha_scaling", "rho_center", "K", "R",
                 "M_total", "kinetic_energy", "thermal_energy", "potential_energy", "total_energy")


def random_doubles(rng, k):
    """Finite doubles across the whole exponent range, decimal-looking values
    and integers (the cases Grisu2's digit loop and layout branch on)."""
    bits = rng.integers(0, 2 ** 63, size=k, dtype=np.uint64) | (rng.integers(0, 2, size=k, dtype=np.uint64) << 63)
    a = bits.view(np.float64)
    a = np.where(np.isfinite(a), a, 1.0)
    b = rng.uniform(-1e3, 1e3, k)
    c = np.round(rng.uniform(-1e5, 1e5, k)) / 1000.0
    d = np.round(rng.uniform(-1e17, 1e17, k))
    with np.errstate(over="ignore"):
        e = 10.0 ** rng.integers(-330, 310, k).astype(np.float64) * rng.choice([1.0, 2.5, 7.0], k)
    e = np.where(np.isfinite(e), e, 2.0)
    return np.concatenate([a, b, c, d, e])


SPECIAL = [0.0, -0.0, 1.0, -1.0, 0.1, 1e-4, 1e-5, 0.000123, 1e15, 1e16, 123456789012345678.0, 999999999999999.9,
           1e21, 5e-324, 2.2250738585072014e-308, 1.7976931348623157e308, 1 / 3, 2 / 3, 5 / 3, math.pi, 1e-300,
           4.35, 0.3, 2.0 ** 52, 2.0 ** 53 + 2, 9007199254740993.0, 1e22, 1e23]


@needs_io_oracle
def test_metadata_json_matches_nlohmann_over_many_doubles():
    lib = oracle()
    rng = np.random.default_rng(7)
    vals = np.concatenate([np.array(SPECIAL), random_doubles(rng, 30000)])
    per = len(DOUBLE_FIELDS)
    vals = np.concatenate([vals, np.ones((-len(vals)) % per)])
    mism = []
    for row in vals.reshape(-1, per):
        m = sphgpu.output_meta(timestamp="2026-01-01T00:00:00Z", is_relaxation=1, relaxation_step=2,
                               relaxation_total_steps=9)
        for k, v in zip(DOUBLE_FIELDS, row):
            setattr(m, k, float(v))
        mine, ref = sphgpu.output_meta_json(m, 0), orc_json(lib, m, 0)
        if mine != ref:
            mism.append((row.tolist(), mine, ref))
        if sphgpu.output_meta_json(m, 1) != orc_json(lib, m, 1):
            mism.append(("units", row.tolist()))
    assert not mism, mism[:2]
    # the same text parses back to the same doubles (a JSON reader's view)
    j = json.loads(sphgpu.output_meta_json(m, 0))
    assert j["units"]["velocity_to_cgs"] == m.velocity_to_cgs and j["M_total"] == m.M_total


@needs_io_oracle
@pytest.mark.parametrize("case", ["plain", "relaxation", "units", "strings"])
def test_metadata_json_layout_and_round_trip(case):
    lib = oracle()
    kw = dict(timestamp="2026-02-03T04:05:06Z", step=12, time_code=0.125, time_physical=4.5, particle_count=300,
              gamma=5 / 3, gravitational_constant=6.674e-8, neighbor_number=50, sph_type=ics.GSPH,
              kernel=ics.WENDLAND, use_gravity=1, use_balsara=0, use_time_dependent_av=1,
              kinetic_energy=0.5, thermal_energy=0.25, potential_energy=-0.75, total_energy=0.0)
    if case == "relaxation":
        kw.update(is_relaxation=1, relaxation_step=40, relaxation_total_steps=500, accumulated_time=2.5,
                  alpha_scaling=0.3, rho_center=3.0, K=0.4, R=1.5, M_total=2.0)
    if case == "units":
        L, M, T = 3.0857e21, 1.989e33, 3.0857e21 / 1e5
        kw.update(unit_type="Galactic", length_unit="kpc", mass_unit="M_sun", time_unit="Myr",
                  velocity_unit="km/s", energy_unit="erg", length_to_cgs=L, mass_to_cgs=M, time_to_cgs=T,
                  velocity_to_cgs=L / T)
    if case == "strings":
        kw.update(simulation_name='quote " back\\slash tab\t nl\n ctl\x01 end', sph_type=ics.GDISPH)
    m = sphgpu.output_meta(**kw)
    text = sphgpu.output_meta_json(m, 0)
    assert text == orc_json(lib, m, 0)
    assert sphgpu.output_meta_json(m, 1) == orc_json(lib, m, 1)
    j = json.loads(text)
    assert list(j) == sorted(j) and list(j["units"]) == sorted(j["units"])
    assert ("K" in j) == (case == "relaxation")
    assert j["sph_type"] == ("Unknown" if case == "strings" else "Godunov SPH")
    # from_json: the library's reader and the oracle's agree field by field
    back = sphgpu.output_meta_from_json(text)
    ob = sphgpu.OutputMeta()
    assert lib.orc_meta_from_json(text.encode(), C.byref(ob)) == 0
    assert back.as_dict() == ob.as_dict()
    for k in ("step", "time_code", "gamma", "sph_type", "kernel", "use_balsara", "kinetic_energy",
              "length_to_cgs", "velocity_to_cgs", "is_relaxation", "K", "M_total", "simulation_name", "unit_type",
              "time_unit"):
        assert back.as_dict()[k] == m.as_dict()[k], k


@needs_io_oracle
def test_metadata_from_json_defaults_and_errors():
    lib = oracle()
    for text in ('{"step": 5, "units": {"type_enum": 2, "length_to_cgs": 100.0}}',
                 '{"is_relaxation": true, "relaxation_step": 3}',
                 '{"timestamp": "x", "gamma": 2, "use_gravity": true, "units": {}}'):
        a = sphgpu.output_meta_from_json(text)
        b = sphgpu.OutputMeta()
        assert lib.orc_meta_from_json(text.encode(), C.byref(b)) == 0
        da, db = a.as_dict(), b.as_dict()
        if '"timestamp"' not in text:  # both take the current UTC time
            assert re.fullmatch(r"\d{4}-\d\d-\d\dT\d\d:\d\d:\d\dZ", da.pop("timestamp"))
            assert re.fullmatch(r"\d{4}-\d\d-\d\dT\d\d:\d\d:\d\dZ", db.pop("timestamp"))
        assert da == db, text
    a = sphgpu.output_meta_from_json('{"is_relaxation": true}')
    assert a.rho_center == 0.0 and a.K == 0.0 and a.alpha_scaling == 1.0  # from_json's own defaults
    assert sphgpu.output_meta_from_json("{}").rho_center == 1.0           # OutputMetadata()'s
    for bad in ("[1, 2]", "{", '{"step": "five"}', '{"use_gravity": 1}', "{} x"):
        with pytest.raises(sphgpu.SphGpuError):
            sphgpu.output_meta_from_json(bad)
        if bad != "{} x" and bad != "[1, 2]":
            assert lib.orc_meta_from_json(bad.encode(), C.byref(sphgpu.OutputMeta())) != 0


def random_fields(n, seed):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal((NF, n)) * 10.0 ** rng.integers(-8, 8, size=(NF, n))
    f[ics.F_ID] = np.arange(n)
    f[ics.F_NEIGHBOR] = rng.integers(0, 200, n)
    special = [0.0, -0.0, 1e-5, 123456.5, 1e-45, 3.4e38, 1e300, -2.5e-310, 0.1, 1 / 3]
    for k, v in enumerate(special):
        f[ics.F_DENS + (k % 6), k % n] = v
    return f


@needs_io_oracle
@pytest.mark.parametrize("binary", [1, 0])
@pytest.mark.parametrize("dim", [1, 2, 3])
def test_vtk_matches_oracle_bytes(tmp_path, dim, binary):
    lib = oracle()
    n = 1500
    f = random_fields(n, 3 + dim)
    m = sphgpu.output_meta(simulation_name="sedov run", timestamp="2026-01-01T00:00:00Z")
    a, b = tmp_path / "lib.vtk", tmp_path / "orc.vtk"
    sphgpu.write_vtk_fields(a, m, dim, f, binary=bool(binary))
    assert lib.orc_vtk_write(str(b).encode(), C.byref(m), dim, n, dp(np.ascontiguousarray(f)), binary) == 0
    assert a.read_bytes() == b.read_bytes()
    data = a.read_bytes()
    assert data.startswith(b"# vtk DataFile Version 3.0\nSPH Simulation - sedov run\n")
    if binary:
        # positions: big-endian float32 right after the POINTS line
        k = data.index(b"POINTS 1500 float\n") + len(b"POINTS 1500 float\n")
        pts = np.frombuffer(data[k:k + 12 * n], dtype=">f4").reshape(n, 3)
        for ax in range(3):
            want = f[ics.F_POS + ax].astype(np.float32) if ax < dim else np.zeros(n, np.float32)
            np.testing.assert_array_equal(pts[:, ax], want)
    else:
        lines = data.decode().splitlines()
        assert lines[5] == "%g" % f[ics.F_POS, 0] + ("" if dim < 2 else " %g" % f[ics.F_POS + 1, 0]) + \
            (" 0.0" if dim < 2 else "") + (" 0.0" if dim < 3 else " %g" % f[ics.F_POS + 2, 0])


@needs_io_oracle
def test_vtk_empty_set_fails_like_the_reference(tmp_path):
    lib = oracle()
    m = sphgpu.output_meta()
    a, b = tmp_path / "a.vtk", tmp_path / "b.vtk"
    with pytest.raises(sphgpu.SphGpuError):
        sphgpu.write_vtk_fields(a, m, 3, np.zeros((NF, 0)), binary=True)
    assert lib.orc_vtk_write(str(b).encode(), C.byref(m), 3, 0, dp(np.zeros(NF)), 1) != 0
    assert a.read_bytes() == b.read_bytes() == b"# vtk DataFile Version 3.0\nSPH Simulation - SPH Simulation\n" \
        b"BINARY\nDATASET UNSTRUCTURED_GRID\n"


@needs_io_oracle
@pytest.mark.parametrize("dim,n,comp", [(3, 2000, 6), (2, 700, 0), (1, 50, 12), (3, 1024, -3), (3, 1025, 1)])
def test_hdf5_matches_oracle_and_round_trips(tmp_path, dim, n, comp):
    lib = oracle()
    f = random_fields(n, 11 + n)
    m = sphgpu.output_meta(timestamp="2026-01-01T00:00:00Z", step=4, time_code=0.03125, time_physical=0.03125,
                           particle_count=n, gamma=5 / 3, sph_type=ics.GSPH, kernel=ics.CUBIC_SPLINE,
                           kinetic_energy=0.5, thermal_energy=0.5, total_energy=1.0)
    a, b = tmp_path / "lib.h5", tmp_path / "orc.h5"
    sphgpu.write_hdf5_fields(a, m, dim, f, compression=comp)
    fo = np.ascontiguousarray(f)
    assert lib.orc_hdf5_write(str(b).encode(), C.byref(m), dim, n, dp(fo), comp) == 0
    da, db = describe(lib, a), describe(lib, b)
    assert da == db
    level = min(9, max(0, comp))
    assert ("filter=1(%d,)" % level in da) == (level > 0)
    assert "dataset /particles/pos_x class=1 size=8" in da and "dataset /particles/id class=0 size=4" in da
    # the library reads the oracle's file and the oracle reads the library's
    assert sphgpu.hdf5_row_count(b) == n
    mr, g = sphgpu.read_hdf5_fields(b, n)
    fr = np.zeros((NF, n))
    rows = C.c_int(0)
    mo = sphgpu.OutputMeta()
    assert lib.orc_hdf5_read(str(a).encode(), C.byref(mo), n, dp(fr), C.byref(rows)) == 0 and rows.value == n
    cols = [ics.F_ID, ics.F_MASS, ics.F_DENS, ics.F_PRES, ics.F_ENE, ics.F_SML, ics.F_SOUND, ics.F_ALPHA,
            ics.F_BALSARA, ics.F_GRADH, ics.F_PHI, ics.F_NEIGHBOR]
    cols += [base + ax for base in (ics.F_POS, ics.F_VEL, ics.F_ACC) for ax in range(dim)]
    for r in cols:
        np.testing.assert_array_equal(g[r], f[r], err_msg=str(r))
        np.testing.assert_array_equal(fr[r], f[r], err_msg=str(r))
    absent = sorted(set(range(NF)) - set(cols))
    assert not g[absent].any() and not fr[absent].any()
    assert mr.as_dict() == mo.as_dict()
    assert mr.step == 4 and mr.time_code == 0.03125 and mr.gamma == 5 / 3 and mr.sph_type == ics.GSPH


@needs_io_oracle
def test_hdf5_empty_set_and_bad_files(tmp_path):
    lib = oracle()
    m = sphgpu.output_meta(timestamp="2026-01-01T00:00:00Z")
    a, b = tmp_path / "a.h5", tmp_path / "b.h5"
    with pytest.raises(sphgpu.SphGpuError):
        sphgpu.write_hdf5_fields(a, m, 3, np.zeros((NF, 0)))
    assert lib.orc_hdf5_write(str(b).encode(), C.byref(m), 3, 0, dp(np.zeros(NF)), 6) != 0
    # both leave /metadata and /energy and no /particles
    assert describe(lib, a) == describe(lib, b)
    assert "group /particles" not in describe(lib, a)
    with pytest.raises(sphgpu.SphGpuError):
        sphgpu.hdf5_row_count(a)
    with pytest.raises(sphgpu.SphGpuError):
        sphgpu.read_hdf5_fields(a, 0)
    txt = tmp_path / "not.h5"
    txt.write_text("not an HDF5 file")
    with pytest.raises(sphgpu.SphGpuError):
        sphgpu.read_hdf5_fields(txt, 1)
    # a row count other than the block's is refused
    f = random_fields(10, 1)
    sphgpu.write_hdf5_fields(a, m, 3, f)
    with pytest.raises(sphgpu.SphGpuError):
        sphgpu.read_hdf5_fields(a, 11)
