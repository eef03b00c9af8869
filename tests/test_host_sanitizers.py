"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer
(CPU only).

tests/sanitize/Makefile compiles the library's host sources (snapshot CSV,
VTK, energy log, metadata JSON, HDF5) and the oracle with
-fsanitize=address,undefined; tests/sanitize/host_check.cpp runs the oracle
on small instances of the workloads and sends each resulting state through
every writer and reader, including the inputs the readers reject.  The
reference offers the same kind of build for its own code (ENABLE_SANITIZERS in
its CMakeLists.txt)."""
from __future__ import annotations

import ctypes
import os
import shutil
import struct
import subprocess
from pathlib import Path

import pytest

import ics
import sphgpu

HERE = Path(__file__).resolve().parent / "sanitize"
BIN = HERE / "_build" / "host_check"


def _write_case(path: Path, params: dict, fields) -> None:
    p = sphgpu.params_struct(params)
    n = fields.shape[1]
    with open(path, "wb") as fh:
        fh.write(struct.pack("<ii", 0x53504843, ctypes.sizeof(p)))
        fh.write(bytes(p))
        fh.write(struct.pack("<i", n))
        fh.write(fields.astype("<f8").tobytes())


def _cases():
    out = {
        "sod_1d_ssph": ics.sod_1d(n_left=80, n_right=20),
        "khi_2d_disph": ics.khi_2d(nx=16),
        "sedov_3d_gsph_iter": ics.sedov_3d(n1=8),
        "evrard_3d_grav": ics.evrard_3d(n=600),
    }
    p, f = ics.sedov_3d(n1=8)
    p.update(sph_type=ics.SSPH, use_time_dependent_av=1, use_ac=1)
    out["sedov_3d_ssph_tdav_ac"] = (p, f)
    p, f = ics.khi_2d(nx=16)
    p.update(sph_type=ics.GDISPH)
    out["khi_2d_gdisph"] = (p, f)
    p, f = ics.sedov_3d(n1=8)
    p.update(riemann_solver=ics.RIEMANN_HLL, gsph_2nd_order=0, periodic=0, leaf_particle_num=1)
    out["sedov_3d_gsph_hll_open"] = (p, f)
    return out


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host compiler")
def test_host_code_under_asan_ubsan(tmp_path):
    subprocess.run(["make", "-C", str(HERE)], check=True, capture_output=True)
    files = []
    for name, (p, f) in _cases().items():
        fp = tmp_path / f"{name}.bin"
        _write_case(fp, p, f)
        files.append(str(fp))
    out_dir = tmp_path / "out"
    out_dir.mkdir()
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "abort_on_error=0:halt_on_error=1:exitcode=97:detect_leaks=1"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1:exitcode=98"
    env["OMP_NUM_THREADS"] = "2"
    r = subprocess.run([str(BIN), str(out_dir), *files], capture_output=True, text=True, env=env, timeout=900)
    log = r.stdout + r.stderr
    print(log[-4000:])
    assert "ERROR: AddressSanitizer" not in log
    assert "ERROR: LeakSanitizer" not in log
    assert "runtime error:" not in log
    assert r.returncode == 0, log[-4000:]
    assert "host_check: ok" in r.stdout
    assert r.stdout.count(": done") == len(files)
