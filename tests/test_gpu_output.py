"""Snapshot and energy files from a device-resident context (GPU): the
context entry points (sphgpu_output_meta_init, sphgpu_write_snapshot_csv,
sphgpu_read_snapshot_csv, sphgpu_energy_file_append) round-trip the state
exactly, and the writer's throughput at 1M particles is compared with the
iostream restatement of the reference's writer (oracle/csv_oracle.cpp)."""
from __future__ import annotations

import ctypes as C
import time

import numpy as np
import pytest

import ics
import sphgpu
from test_output_files import NF, oracle, dp

pytestmark = pytest.mark.gpu

COLS = [ics.F_ID] + [b + a for b in (ics.F_POS, ics.F_VEL, ics.F_ACC) for a in range(3)] + \
    [ics.F_MASS, ics.F_DENS, ics.F_PRES, ics.F_ENE, ics.F_SML, ics.F_SOUND, ics.F_ALPHA, ics.F_BALSARA,
     ics.F_GRADH, ics.F_PHI, ics.F_NEIGHBOR]


def test_snapshot_round_trip_through_context(tmp_path):
    p, f = ics.sedov_3d(n1=16)
    a = sphgpu.Context(p, f.shape[1])
    a.upload(f)
    a.initialize()
    a.integrate(3)
    meta = a.output_meta(3)
    e = a.energy_momentum()
    assert meta.step == 3 and meta.particle_count == f.shape[1] and meta.time_code == a.time
    assert meta.kinetic_energy == e[0] and meta.thermal_energy == e[1] and meta.potential_energy == e[2]
    assert meta.sph_type == ics.GSPH and meta.gamma == p["gamma"] and meta.neighbor_number == p["neighbor_number"]
    path = tmp_path / sphgpu.snapshot_filename(".", "snapshot", 3, ".csv")
    a.write_snapshot_csv(path, meta)
    b = sphgpu.Context(p, f.shape[1])
    m2 = b.read_snapshot_csv(path)
    assert m2.time_code == a.time and b.time == a.time and m2.step == 3
    ga, gb = a.download(), b.download()
    for r in COLS:
        np.testing.assert_array_equal(gb[r], ga[r], err_msg=str(r))
    # the file is what the reference's writer (restated) prints for this state
    orc = tmp_path / "orc.csv"
    assert oracle().orc_csv_write(str(orc).encode(), C.byref(meta), 3, ga.shape[1], dp(ga)) == 0
    assert orc.read_bytes() == path.read_bytes()
    # energy log: header + one row for the device state
    log = tmp_path / "energy.dat"
    sphgpu.energy_file_begin(log, meta)
    a.energy_file_append(log)
    row = [float(x) for x in log.read_text().splitlines()[-1].split()]
    assert row == [a.time, e[0], e[1], e[2], e[0] + e[1] + e[2]]


def test_snapshot_throughput_1m(tmp_path):
    """1M-particle Sedov snapshot: the library's threaded writer and reader
    against the single-threaded iostream writer (the reference's way)."""
    p, f = ics.sedov_3d(n1=100)
    c = sphgpu.Context(p, f.shape[1])
    c.upload(f)
    c.initialize()
    meta = c.output_meta(0)
    path = tmp_path / "snap.csv"
    t0 = time.perf_counter()
    c.write_snapshot_csv(path, meta)
    t1 = time.perf_counter()
    size = path.stat().st_size
    m2 = c.read_snapshot_csv(path)
    t2 = time.perf_counter()
    g = c.download()
    orc = tmp_path / "orc.csv"
    t3 = time.perf_counter()
    assert oracle().orc_csv_write(str(orc).encode(), C.byref(meta), 3, g.shape[1], dp(g)) == 0
    t4 = time.perf_counter()
    assert orc.read_bytes() == path.read_bytes()
    assert m2.step == 0
    n = f.shape[1]
    print(f"snapshot 1M: {size / 1e6:.0f} MB; write {t1 - t0:.2f} s ({n / (t1 - t0) / 1e6:.2f} M rows/s, "
          f"download included), read {t2 - t1:.2f} s (upload included); iostream writer {t4 - t3:.2f} s "
          f"({n / (t4 - t3) / 1e6:.2f} M rows/s)")


def test_vtk_and_hdf5_snapshots_through_context(tmp_path):
    """The context entry points sphgpu_write_snapshot_vtk /
    sphgpu_write_snapshot_hdf5 write what the host-block writers write for the
    downloaded state, and sphgpu_read_snapshot_hdf5 resumes a second context
    to the same state and time (OutputManager::load_for_resume for .h5)."""
    if not sphgpu.HDF5_LIB_PATH.exists():
        pytest.skip("libsphgpu_hdf5.so not built (no HDF5 C library at build time)")
    p, f = ics.sedov_3d(n1=16)
    a = sphgpu.Context(p, f.shape[1])
    a.upload(f)
    a.initialize()
    a.integrate(2)
    meta = a.output_meta(2)
    g = a.download()
    for binary in (True, False):
        a.write_snapshot_vtk(tmp_path / "ctx.vtk", meta, binary=binary)
        sphgpu.write_vtk_fields(tmp_path / "host.vtk", meta, 3, g, binary=binary)
        assert (tmp_path / "ctx.vtk").read_bytes() == (tmp_path / "host.vtk").read_bytes()
    a.write_snapshot_hdf5(tmp_path / "snap.h5", meta, compression=6)
    assert sphgpu.hdf5_row_count(tmp_path / "snap.h5") == f.shape[1]
    b = sphgpu.Context(p, f.shape[1])
    m2 = b.read_snapshot_hdf5(tmp_path / "snap.h5")
    assert m2.step == 2 and m2.time_code == a.time and b.time == a.time
    assert sphgpu.output_meta_json(m2) == sphgpu.output_meta_json(meta)
    gb = b.download()
    for r in COLS:
        np.testing.assert_array_equal(gb[r], g[r], err_msg=str(r))
