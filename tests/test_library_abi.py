"""The C-ABI libraries load and export every entry point include/sphgpu.h
and include/sphgpu_hdf5.h declare (no compute calls: this runs without a
GPU)."""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

import ics
import sphgpu

HEADER = Path(__file__).resolve().parents[1] / "include" / "sphgpu.h"
HDF5_HEADER = Path(__file__).resolve().parents[1] / "include" / "sphgpu_hdf5.h"


def declared_entry_points(header: Path = HEADER) -> set[str]:
    text = header.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(sphgpu_[a-z0-9_]+)\s*\(", text))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(sphgpu.LIB_PATH))
    names = declared_entry_points()
    assert len(names) >= 35
    missing = sorted(n for n in names if not hasattr(lib, n))
    assert not missing, missing
    assert set(sphgpu.EXPORTED) == names


@pytest.mark.skipif(not sphgpu.HDF5_LIB_PATH.exists(), reason="libsphgpu_hdf5.so not built (no HDF5 C library)")
def test_hdf5_library_exports_every_declared_symbol():
    lib = sphgpu.load_hdf5_library()
    names = declared_entry_points(HDF5_HEADER)
    missing = sorted(n for n in names if not hasattr(lib, n))
    assert not missing, missing
    assert set(sphgpu.EXPORTED_HDF5) == names


def test_params_struct_layout_matches_header():
    """sizeof(sphgpu_params) as compiled into the library equals the ctypes
    mirror: 16 int32 + 10 double + 6 double."""
    assert ctypes.sizeof(sphgpu.Params) == 16 * 4 + 16 * 8
    lib = sphgpu.load_library()
    p = sphgpu.Params()
    lib.sphgpu_params_default(ctypes.byref(p), 3)
    d = ics.default_params(3)
    for k in ("neighbor_number", "leaf_particle_num", "max_tree_level", "use_balsara_switch", "iterative_sml"):
        assert getattr(p, k) == d[k], k
    for k in ("gamma", "cfl_sound", "cfl_force", "av_alpha", "alpha_max", "alpha_min", "epsilon_av", "theta"):
        assert getattr(p, k) == pytest.approx(d[k]), k


def test_mlp_param_count():
    lib = sphgpu.load_library()
    assert lib.sphgpu_mlp_param_count() == 6 * 64 + 64 + 64 * 64 + 64 + 64 * 64 + 64 + 64 + 1  # 6-64-64-64-1


def test_create_rejects_bad_parameters_without_touching_device():
    lib = sphgpu.load_library()
    p = sphgpu.params_struct(dict(ics.default_params(1), kernel=ics.WENDLAND))
    h = ctypes.c_void_p()
    assert lib.sphgpu_create(ctypes.byref(p), 10, ctypes.byref(h)) == -1   # Wendland in 1D
    p = sphgpu.params_struct(dict(ics.default_params(3), max_tree_level=22))
    assert lib.sphgpu_create(ctypes.byref(p), 10, ctypes.byref(h)) == -1   # 66 key bits
    assert lib.sphgpu_create(ctypes.byref(p), 0, ctypes.byref(h)) == -1
