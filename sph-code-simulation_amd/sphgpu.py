"""ctypes binding of libsphgpu.so, the C-ABI declared in include/sphgpu.h.

This is plumbing for Python callers (tests, bench, smoke): every method calls
straight into the HIP library, and loading fails loudly when the library has
not been built.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libsphgpu.so"
NFIELDS = 41
# problems per call of sphgpu_riemann_batch_static (SPHGPU_RIEMANN_STATIC_MAX)
RIEMANN_STATIC_MAX = 32768


class Params(C.Structure):
    """sphgpu_params (include/sphgpu.h)."""
    _fields_ = [
        ("dim", C.c_int32), ("sph_type", C.c_int32), ("kernel", C.c_int32),
        ("riemann_solver", C.c_int32), ("neighbor_number", C.c_int32),
        ("leaf_particle_num", C.c_int32), ("max_tree_level", C.c_int32),
        ("iterative_sml", C.c_int32), ("use_balsara_switch", C.c_int32),
        ("use_time_dependent_av", C.c_int32), ("use_ac", C.c_int32),
        ("periodic", C.c_int32), ("use_gravity", C.c_int32),
        ("gsph_2nd_order", C.c_int32), ("pair_precision", C.c_int32),
        ("riemann_guess", C.c_int32),
        ("gamma", C.c_double), ("cfl_sound", C.c_double), ("cfl_force", C.c_double),
        ("av_alpha", C.c_double), ("alpha_max", C.c_double), ("alpha_min", C.c_double),
        ("epsilon_av", C.c_double), ("alpha_ac", C.c_double),
        ("gravity_constant", C.c_double), ("theta", C.c_double),
        ("range_min", C.c_double * 3), ("range_max", C.c_double * 3),
    ]


class Stats(C.Structure):
    """sphgpu_stats (include/sphgpu.h)."""
    _fields_ = [(n, C.c_int64) for n in (
        "pre_candidates", "pre_newton_iters", "pre_newton_fail", "force_pairs",
        "force_candidates", "riemann_iters", "riemann_wave_max_iters",
        "riemann_calls", "riemann_fallbacks", "riemann_learned",
        "gravity_interactions", "nodes", "leaves", "search_node_visits",
        "search_member_tests")] + [
        ("max_depth", C.c_int32), ("pad", C.c_int32),
        ("search_windows", C.c_int64), ("search_leaves", C.c_int64), ("search_exact", C.c_int64),
        ("flist_h_beyond", C.c_int64), ("flist_other", C.c_int64), ("flist_searches", C.c_int64),
        ("search_reach_tests", C.c_int64)]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_ if n != "pad"}


class OutputMeta(C.Structure):
    """sphgpu_output_meta (include/sphgpu.h): the CSV header's metadata."""
    _fields_ = [
        ("format_version", C.c_int32), ("step", C.c_int32),
        ("timestamp", C.c_char * 32), ("simulation_name", C.c_char * 64), ("unit_type", C.c_char * 16),
        ("length_unit", C.c_char * 32), ("mass_unit", C.c_char * 32), ("time_unit", C.c_char * 32),
        ("velocity_unit", C.c_char * 32), ("energy_unit", C.c_char * 32),
        ("length_to_cgs", C.c_double), ("mass_to_cgs", C.c_double), ("time_to_cgs", C.c_double),
        ("velocity_to_cgs", C.c_double), ("time_code", C.c_double), ("time_physical", C.c_double),
        ("particle_count", C.c_int32), ("neighbor_number", C.c_int32), ("sph_type", C.c_int32),
        ("kernel", C.c_int32), ("use_gravity", C.c_int32), ("use_balsara", C.c_int32),
        ("use_time_dependent_av", C.c_int32), ("is_relaxation", C.c_int32), ("relaxation_step", C.c_int32),
        ("relaxation_total_steps", C.c_int32),
        ("gamma", C.c_double), ("gravitational_constant", C.c_double),
        ("accumulated_time", C.c_double), ("alpha_scaling", C.c_double), ("rho_center", C.c_double),
        ("K", C.c_double), ("R", C.c_double), ("M_total", C.c_double),
        ("kinetic_energy", C.c_double), ("thermal_energy", C.c_double), ("potential_energy", C.c_double),
        ("total_energy", C.c_double), ("precision", C.c_int32), ("pad", C.c_int32)]

    def as_dict(self) -> dict:
        out = {}
        for k, _ in self._fields_:
            v = getattr(self, k)
            out[k] = v.decode() if isinstance(v, bytes) else v
        return out


def output_meta(**kw) -> OutputMeta:
    """An OutputMeta with the reference's defaults (OutputMetadata(), code
    units, precision 16), overridden by kw."""
    m = OutputMeta()
    d = dict(format_version=1, timestamp="", simulation_name="SPH Simulation", unit_type="Code",
             length_unit="code_length", mass_unit="code_mass", time_unit="code_time",
             velocity_unit="code_velocity", energy_unit="code_energy", length_to_cgs=1.0, mass_to_cgs=1.0,
             time_to_cgs=1.0, velocity_to_cgs=1.0, gamma=1.4, gravitational_constant=1.0, neighbor_number=32,
             use_balsara=1, alpha_scaling=1.0, rho_center=1.0, K=1.0, R=1.0, M_total=1.0, precision=16)
    d.update(kw)
    for k, v in d.items():
        setattr(m, k, v.encode() if isinstance(v, str) else v)
    return m


def write_csv_fields(path, meta: OutputMeta, dim: int, fields: np.ndarray):
    """sphgpu_write_csv_fields: a snapshot CSV from a host [NFIELDS][n] block."""
    f = np.ascontiguousarray(fields, dtype=np.float64)
    rc = load_library().sphgpu_write_csv_fields(str(path).encode(), C.byref(meta), int(dim), f.shape[1], _dp(f))
    if rc:
        raise SphGpuError(f"sphgpu_write_csv_fields failed ({rc})")


def csv_row_count(path) -> int:
    r = C.c_int32(0)
    rc = load_library().sphgpu_csv_row_count(str(path).encode(), C.byref(r))
    if rc:
        raise SphGpuError(f"sphgpu_csv_row_count failed ({rc})")
    return r.value


def read_csv_fields(path, dim: int, n: int):
    """sphgpu_read_csv_fields -> (OutputMeta, [NFIELDS][n] block)."""
    f = np.zeros((NFIELDS, n), dtype=np.float64)
    m = OutputMeta()
    rc = load_library().sphgpu_read_csv_fields(str(path).encode(), C.byref(m), int(dim), int(n), _dp(f))
    if rc:
        raise SphGpuError(f"sphgpu_read_csv_fields failed ({rc})")
    return m, f


def energy_file_begin(path, meta: OutputMeta):
    rc = load_library().sphgpu_energy_file_begin(str(path).encode(), C.byref(meta))
    if rc:
        raise SphGpuError(f"sphgpu_energy_file_begin failed ({rc})")


def energy_file_write(path, time: float, kinetic: float, thermal: float, potential: float, precision: int = 16):
    rc = load_library().sphgpu_energy_file_write(str(path).encode(), time, kinetic, thermal, potential, precision)
    if rc:
        raise SphGpuError(f"sphgpu_energy_file_write failed ({rc})")


def snapshot_filename(directory, prefix: str, count: int, ext: str) -> str:
    buf = C.create_string_buffer(4096)
    rc = load_library().sphgpu_snapshot_filename(str(directory).encode(), prefix.encode(), count, ext.encode(),
                                                 buf, 4096)
    if rc:
        raise SphGpuError(f"sphgpu_snapshot_filename failed ({rc})")
    return buf.value.decode()


def output_meta_json(meta: OutputMeta, which: int = 0) -> str:
    """sphgpu_output_meta_json: OutputMetadata::to_json (which=0) or
    UnitSystem::to_json (which=1) as nlohmann's dump(2) prints it."""
    lib = load_library()
    need = C.c_int32(0)
    rc = lib.sphgpu_output_meta_json(C.byref(meta), int(which), None, 0, C.byref(need))
    if rc:
        raise SphGpuError(f"sphgpu_output_meta_json failed ({rc})")
    buf = C.create_string_buffer(need.value)
    rc = lib.sphgpu_output_meta_json(C.byref(meta), int(which), buf, need.value, None)
    if rc:
        raise SphGpuError(f"sphgpu_output_meta_json failed ({rc})")
    return buf.value.decode()


def output_meta_from_json(text: str) -> OutputMeta:
    """sphgpu_output_meta_from_json: OutputMetadata::from_json."""
    m = OutputMeta()
    rc = load_library().sphgpu_output_meta_from_json(text.encode(), C.byref(m))
    if rc:
        raise SphGpuError(f"sphgpu_output_meta_from_json failed ({rc})")
    return m


def write_vtk_fields(path, meta: OutputMeta, dim: int, fields: np.ndarray, binary: bool = True):
    """sphgpu_write_vtk_fields: a legacy-VTK snapshot from a host block."""
    f = np.ascontiguousarray(fields, dtype=np.float64)
    rc = load_library().sphgpu_write_vtk_fields(str(path).encode(), C.byref(meta), int(dim), f.shape[1], _dp(f),
                                                int(bool(binary)))
    if rc:
        raise SphGpuError(f"sphgpu_write_vtk_fields failed ({rc})")


HDF5_LIB_PATH = PKG / "libsphgpu_hdf5.so"
_h5 = None


def load_hdf5_library() -> C.CDLL:
    """Load libsphgpu_hdf5.so (include/sphgpu_hdf5.h), the HDF5 snapshot
    library; it is built only where the HDF5 C library is installed."""
    global _h5
    if _h5 is not None:
        return _h5
    load_library()  # its dependency, and the C++ runtime it shares
    if not HDF5_LIB_PATH.exists():
        raise SphGpuError(f"{HDF5_LIB_PATH} is not built (no HDF5 C library at build time)")
    lib = C.CDLL(str(HDF5_LIB_PATH))
    vp, ip, dp, i32p = C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int32)
    mp = C.POINTER(OutputMeta)
    sigs = {
        "sphgpu_write_hdf5_fields": (ip, [C.c_char_p, mp, C.c_int32, C.c_int32, dp, C.c_int32]),
        "sphgpu_write_snapshot_hdf5": (ip, [vp, C.c_char_p, mp, C.c_int32]),
        "sphgpu_hdf5_row_count": (ip, [C.c_char_p, i32p]),
        "sphgpu_read_hdf5_fields": (ip, [C.c_char_p, mp, C.c_int32, dp]),
        "sphgpu_read_snapshot_hdf5": (ip, [vp, C.c_char_p, mp]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _h5 = lib
    return lib


EXPORTED_HDF5 = ("sphgpu_write_hdf5_fields", "sphgpu_write_snapshot_hdf5", "sphgpu_hdf5_row_count",
                 "sphgpu_read_hdf5_fields", "sphgpu_read_snapshot_hdf5")


def write_hdf5_fields(path, meta: OutputMeta, dim: int, fields: np.ndarray, compression: int = 6):
    """sphgpu_write_hdf5_fields: an HDF5 snapshot from a host block."""
    f = np.ascontiguousarray(fields, dtype=np.float64)
    rc = load_hdf5_library().sphgpu_write_hdf5_fields(str(path).encode(), C.byref(meta), int(dim), f.shape[1],
                                                      _dp(f), int(compression))
    if rc:
        raise SphGpuError(f"sphgpu_write_hdf5_fields failed ({rc})")


def hdf5_row_count(path) -> int:
    r = C.c_int32(0)
    rc = load_hdf5_library().sphgpu_hdf5_row_count(str(path).encode(), C.byref(r))
    if rc:
        raise SphGpuError(f"sphgpu_hdf5_row_count failed ({rc})")
    return r.value


def read_hdf5_fields(path, n: int):
    """sphgpu_read_hdf5_fields -> (OutputMeta, [NFIELDS][n] block)."""
    f = np.zeros((NFIELDS, n), dtype=np.float64)
    m = OutputMeta()
    rc = load_hdf5_library().sphgpu_read_hdf5_fields(str(path).encode(), C.byref(m), int(n), _dp(f))
    if rc:
        raise SphGpuError(f"sphgpu_read_hdf5_fields failed ({rc})")
    return m, f


def params_struct(d: dict) -> Params:
    p = Params()
    for name, _ in Params._fields_:
        if name in ("range_min", "range_max"):
            v = list(d.get(name, [0.0, 0.0, 0.0])) + [0.0, 0.0, 0.0]
            setattr(p, name, (C.c_double * 3)(*v[:3]))
        else:
            setattr(p, name, type(getattr(p, name))(d[name]))
    return p


class SphGpuError(RuntimeError):
    pass


_lib = None


def load_library() -> C.CDLL:
    """Load libsphgpu.so from the package directory (built in-tree)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise SphGpuError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
    lib = C.CDLL(str(LIB_PATH), mode=os.RTLD_GLOBAL)
    vp, ip, dp, i32p, i64p = C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int64)
    fp = C.POINTER(C.c_float)
    sigs = {
        "sphgpu_params_default": (None, [C.POINTER(Params), ip]),
        "sphgpu_create": (ip, [C.POINTER(Params), ip, C.POINTER(vp)]),
        "sphgpu_destroy": (ip, [vp]),
        "sphgpu_last_error": (C.c_char_p, [vp]),
        "sphgpu_device_info": (ip, [C.c_char_p, ip, C.POINTER(C.c_int), dp]),
        "sphgpu_upload": (ip, [vp, dp, C.c_uint64]),
        "sphgpu_download": (ip, [vp, dp, C.c_uint64]),
        "sphgpu_set_dt": (ip, [vp, C.c_double]),
        "sphgpu_get_dt": (ip, [vp, dp]),
        "sphgpu_set_time": (ip, [vp, C.c_double]),
        "sphgpu_get_time": (ip, [vp, dp]),
        "sphgpu_get_h_per_v_sig": (ip, [vp, dp]),
        "sphgpu_set_h_per_v_sig": (ip, [vp, C.c_double]),
        "sphgpu_set_kernel": (ip, [vp]),
        "sphgpu_build_tree": (ip, [vp]),
        "sphgpu_pre_interaction": (ip, [vp]),
        "sphgpu_fluid_force": (ip, [vp]),
        "sphgpu_gravity": (ip, [vp]),
        "sphgpu_timestep": (ip, [vp]),
        "sphgpu_predict": (ip, [vp]),
        "sphgpu_correct": (ip, [vp]),
        "sphgpu_initialize": (ip, [vp]),
        "sphgpu_integrate": (ip, [vp, ip]),
        "sphgpu_energy_momentum": (ip, [vp, dp]),
        "sphgpu_get_stats": (ip, [vp, C.POINTER(Stats)]),
        "sphgpu_reset_stats": (ip, [vp]),
        "sphgpu_neighbor_sets": (ip, [vp, dp, ip, ip, i32p, i32p]),
        "sphgpu_tree_dump": (ip, [vp, i64p, i32p, i32p, i32p, i32p, dp, dp, dp, dp, dp, i32p]),
        "sphgpu_riemann_batch": (ip, [vp, ip, ip, ip, dp, dp, dp, dp, i32p, i32p]),
        "sphgpu_riemann_batch_static": (ip, [vp, ip, ip, ip, dp, dp, dp, dp, i32p, i32p]),
        "sphgpu_riemann_log_enable": (ip, [vp, C.c_int64, ip]),
        "sphgpu_riemann_log_read": (ip, [vp, fp, C.c_int64, i64p]),
        "sphgpu_mlp_init": (ip, [vp, C.c_uint64]),
        "sphgpu_mlp_set_params": (ip, [vp, fp, C.c_int64]),
        "sphgpu_mlp_get_params": (ip, [vp, fp, C.c_int64]),
        "sphgpu_mlp_param_count": (C.c_int64, []),
        "sphgpu_mlp_train": (ip, [vp, fp, C.c_int64, ip, ip, C.c_float, fp]),
        "sphgpu_mlp_loss_grad": (ip, [vp, fp, C.c_int64, fp, fp]),
        "sphgpu_mlp_predict": (ip, [vp, fp, C.c_int64, fp]),
        "sphgpu_kernel_times": (ip, [vp, C.c_char_p, ip, dp, i64p, ip, C.POINTER(C.c_int)]),
        "sphgpu_copy_probe": (ip, [vp, C.c_int64, dp]),
        "sphgpu_enable_timing": (ip, [vp, ip]),
        "sphgpu_get_params": (ip, [vp, C.POINTER(Params), i32p]),
        "sphgpu_output_meta_init": (ip, [vp, C.c_int32, C.POINTER(OutputMeta)]),
        "sphgpu_snapshot_filename": (ip, [C.c_char_p, C.c_char_p, C.c_int32, C.c_char_p, C.c_char_p, C.c_int32]),
        "sphgpu_write_snapshot_csv": (ip, [vp, C.c_char_p, C.POINTER(OutputMeta)]),
        "sphgpu_write_csv_fields": (ip, [C.c_char_p, C.POINTER(OutputMeta), C.c_int32, C.c_int32, dp]),
        "sphgpu_read_snapshot_csv": (ip, [vp, C.c_char_p, C.POINTER(OutputMeta)]),
        "sphgpu_read_csv_fields": (ip, [C.c_char_p, C.POINTER(OutputMeta), C.c_int32, C.c_int32, dp]),
        "sphgpu_csv_row_count": (ip, [C.c_char_p, i32p]),
        "sphgpu_energy_file_begin": (ip, [C.c_char_p, C.POINTER(OutputMeta)]),
        "sphgpu_energy_file_write": (ip, [C.c_char_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int32]),
        "sphgpu_energy_file_append": (ip, [vp, C.c_char_p, C.c_int32]),
        "sphgpu_write_vtk_fields": (ip, [C.c_char_p, C.POINTER(OutputMeta), C.c_int32, C.c_int32, dp, C.c_int32]),
        "sphgpu_write_snapshot_vtk": (ip, [vp, C.c_char_p, C.POINTER(OutputMeta), C.c_int32]),
        "sphgpu_output_meta_json": (ip, [C.POINTER(OutputMeta), C.c_int32, C.c_char_p, C.c_int32, i32p]),
        "sphgpu_output_meta_from_json": (ip, [C.c_char_p, C.POINTER(OutputMeta)]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


EXPORTED = (
    "sphgpu_params_default", "sphgpu_create", "sphgpu_destroy", "sphgpu_last_error",
    "sphgpu_device_info", "sphgpu_upload", "sphgpu_download", "sphgpu_set_dt",
    "sphgpu_get_dt", "sphgpu_set_time", "sphgpu_get_time", "sphgpu_get_h_per_v_sig",
    "sphgpu_set_h_per_v_sig", "sphgpu_set_kernel",
    "sphgpu_build_tree", "sphgpu_pre_interaction", "sphgpu_fluid_force",
    "sphgpu_gravity", "sphgpu_timestep", "sphgpu_predict", "sphgpu_correct",
    "sphgpu_initialize", "sphgpu_integrate", "sphgpu_energy_momentum",
    "sphgpu_get_stats", "sphgpu_reset_stats", "sphgpu_neighbor_sets",
    "sphgpu_tree_dump", "sphgpu_riemann_batch", "sphgpu_riemann_batch_static", "sphgpu_riemann_log_enable",
    "sphgpu_riemann_log_read", "sphgpu_mlp_init", "sphgpu_mlp_set_params",
    "sphgpu_mlp_get_params", "sphgpu_mlp_param_count", "sphgpu_mlp_train",
    "sphgpu_mlp_loss_grad", "sphgpu_mlp_predict", "sphgpu_kernel_times", "sphgpu_copy_probe",
    "sphgpu_enable_timing", "sphgpu_get_params", "sphgpu_output_meta_init", "sphgpu_snapshot_filename",
    "sphgpu_write_snapshot_csv", "sphgpu_write_csv_fields", "sphgpu_read_snapshot_csv",
    "sphgpu_read_csv_fields", "sphgpu_csv_row_count", "sphgpu_energy_file_begin",
    "sphgpu_energy_file_write", "sphgpu_energy_file_append", "sphgpu_write_vtk_fields",
    "sphgpu_write_snapshot_vtk", "sphgpu_output_meta_json", "sphgpu_output_meta_from_json",
)


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Context:
    """One device-resident simulation (sphgpu_ctx)."""

    def __init__(self, params: dict, n: int):
        self.lib = load_library()
        self.params = dict(params)
        self.n = int(n)
        self._p = params_struct(params)
        h = C.c_void_p()
        rc = self.lib.sphgpu_create(C.byref(self._p), self.n, C.byref(h))
        if rc != 0:
            raise SphGpuError(f"sphgpu_create failed ({rc})")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.sphgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            msg = self.lib.sphgpu_last_error(self.h)
            raise SphGpuError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def upload(self, fields: np.ndarray, mask: int = (1 << NFIELDS) - 1):
        f = np.ascontiguousarray(fields, dtype=np.float64)
        assert f.shape == (NFIELDS, self.n), f.shape
        self._check(self.lib.sphgpu_upload(self.h, _dp(f), mask), "upload")

    def download(self, mask: int = (1 << NFIELDS) - 1) -> np.ndarray:
        f = np.zeros((NFIELDS, self.n), dtype=np.float64)
        self._check(self.lib.sphgpu_download(self.h, _dp(f), mask), "download")
        return f

    def _scalar(self, fn):
        v = C.c_double(0)
        self._check(fn(self.h, C.byref(v)), fn.__name__)
        return v.value

    @property
    def dt(self): return self._scalar(self.lib.sphgpu_get_dt)

    @dt.setter
    def dt(self, v): self._check(self.lib.sphgpu_set_dt(self.h, float(v)), "set_dt")

    @property
    def time(self): return self._scalar(self.lib.sphgpu_get_time)

    @time.setter
    def time(self, v): self._check(self.lib.sphgpu_set_time(self.h, float(v)), "set_time")

    @property
    def h_per_v_sig(self): return self._scalar(self.lib.sphgpu_get_h_per_v_sig)

    @h_per_v_sig.setter
    def h_per_v_sig(self, v): self._check(self.lib.sphgpu_set_h_per_v_sig(self.h, float(v)), "set_h_per_v_sig")

    def set_kernel(self): self._check(self.lib.sphgpu_set_kernel(self.h), "set_kernel")

    def build_tree(self): self._check(self.lib.sphgpu_build_tree(self.h), "build_tree")
    def pre_interaction(self): self._check(self.lib.sphgpu_pre_interaction(self.h), "pre_interaction")
    def fluid_force(self): self._check(self.lib.sphgpu_fluid_force(self.h), "fluid_force")
    def gravity(self): self._check(self.lib.sphgpu_gravity(self.h), "gravity")
    def timestep(self): self._check(self.lib.sphgpu_timestep(self.h), "timestep")
    def predict(self): self._check(self.lib.sphgpu_predict(self.h), "predict")
    def correct(self): self._check(self.lib.sphgpu_correct(self.h), "correct")
    def initialize(self): self._check(self.lib.sphgpu_initialize(self.h), "initialize")
    def integrate(self, n: int = 1): self._check(self.lib.sphgpu_integrate(self.h, int(n)), "integrate")

    def energy_momentum(self) -> np.ndarray:
        out = np.zeros(7)
        self._check(self.lib.sphgpu_energy_momentum(self.h, _dp(out)), "energy_momentum")
        return out

    def stats(self) -> dict:
        s = Stats()
        self._check(self.lib.sphgpu_get_stats(self.h, C.byref(s)), "get_stats")
        return s.as_dict()

    def reset_stats(self): self._check(self.lib.sphgpu_reset_stats(self.h), "reset_stats")

    def neighbor_sets(self, h, is_ij, cap: int = 512):
        """is_ij 0/1: search with the given h; is_ij 2: the force pair sets
        derived by the last pre_interaction (h unused, may be None)."""
        hp = None
        if h is not None:
            h = np.ascontiguousarray(h, dtype=np.float64)
            hp = _dp(h)
        counts = np.zeros(self.n, dtype=np.int32)
        lists = np.zeros(self.n * cap, dtype=np.int32)
        self._check(self.lib.sphgpu_neighbor_sets(
            self.h, hp, int(is_ij), int(cap),
            counts.ctypes.data_as(C.POINTER(C.c_int32)),
            lists.ctypes.data_as(C.POINTER(C.c_int32))), "neighbor_sets")
        return counts, lists.reshape(self.n, cap)

    def tree_dump(self) -> dict:
        nn = C.c_int64(0)
        self._check(self.lib.sphgpu_tree_dump(self.h, C.byref(nn), *([None] * 11)), "tree_dump")
        m = nn.value
        i32 = lambda k: np.zeros(k, dtype=np.int32)  # noqa: E731
        out = dict(level=i32(m), count=i32(m), first=i32(m), is_leaf=i32(m),
                   center=np.zeros(3 * m), edge=np.zeros(m), mass=np.zeros(m),
                   com=np.zeros(3 * m), kernel_size=np.zeros(m), members=i32(self.n))
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))  # noqa: E731
        self._check(self.lib.sphgpu_tree_dump(
            self.h, C.byref(nn), ip(out["level"]), ip(out["count"]), ip(out["first"]),
            ip(out["is_leaf"]), _dp(out["center"]), _dp(out["edge"]), _dp(out["mass"]),
            _dp(out["com"]), _dp(out["kernel_size"]), ip(out["members"])), "tree_dump")
        out["center"] = out["center"].reshape(m, 3)
        out["com"] = out["com"].reshape(m, 3)
        return out

    def riemann_batch(self, solver: int, precision: int, left: np.ndarray, right: np.ndarray,
                      with_usage: bool = False, static: bool = False):
        """Batched Riemann problems; static=True calls the allocation-free
        entry point (sphgpu_riemann_batch_static, at most RIEMANN_STATIC_MAX
        problems)."""
        left = np.ascontiguousarray(left, dtype=np.float64)
        right = np.ascontiguousarray(right, dtype=np.float64)
        n = left.shape[0]
        assert left.shape == (n, 4) and right.shape == (n, 4)
        ps = np.zeros(n)
        vs = np.zeros(n)
        it = np.zeros(n, dtype=np.int32)
        used = np.zeros(n, dtype=np.int32)
        fn = self.lib.sphgpu_riemann_batch_static if static else self.lib.sphgpu_riemann_batch
        self._check(fn(
            self.h, int(solver), int(precision), n, _dp(left), _dp(right), _dp(ps), _dp(vs),
            it.ctypes.data_as(C.POINTER(C.c_int32)), used.ctypes.data_as(C.POINTER(C.c_int32))),
            "riemann_batch")
        return (ps, vs, it, used) if with_usage else (ps, vs, it)

    # ---- learned Riemann guess -------------------------------------------
    def riemann_log_enable(self, capacity: int, stride: int = 1):
        self._check(self.lib.sphgpu_riemann_log_enable(self.h, int(capacity), int(stride)), "log_enable")

    def riemann_log_read(self, max_rows: int) -> np.ndarray:
        rows = np.zeros((max_rows, 7), dtype=np.float32)
        n = C.c_int64(0)
        self._check(self.lib.sphgpu_riemann_log_read(
            self.h, rows.ctypes.data_as(C.POINTER(C.c_float)), int(max_rows), C.byref(n)), "log_read")
        return rows[:n.value].copy()

    def mlp_init(self, seed: int = 1):
        self._check(self.lib.sphgpu_mlp_init(self.h, int(seed)), "mlp_init")

    def mlp_param_count(self) -> int:
        return int(self.lib.sphgpu_mlp_param_count())

    def mlp_get_params(self) -> np.ndarray:
        n = self.mlp_param_count()
        a = np.zeros(n, dtype=np.float32)
        self._check(self.lib.sphgpu_mlp_get_params(self.h, a.ctypes.data_as(C.POINTER(C.c_float)), n), "mlp_get")
        return a

    def mlp_set_params(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float32)
        self._check(self.lib.sphgpu_mlp_set_params(self.h, a.ctypes.data_as(C.POINTER(C.c_float)), a.size), "mlp_set")

    def mlp_train(self, rows: np.ndarray, epochs: int, batch: int, lr: float) -> float:
        rows = np.ascontiguousarray(rows, dtype=np.float32)
        loss = C.c_float(0)
        self._check(self.lib.sphgpu_mlp_train(
            self.h, rows.ctypes.data_as(C.POINTER(C.c_float)), rows.shape[0], int(epochs),
            int(batch), float(lr), C.byref(loss)), "mlp_train")
        return loss.value

    def mlp_loss_grad(self, rows: np.ndarray):
        rows = np.ascontiguousarray(rows, dtype=np.float32)
        loss = C.c_float(0)
        g = np.zeros(self.mlp_param_count(), dtype=np.float32)
        self._check(self.lib.sphgpu_mlp_loss_grad(
            self.h, rows.ctypes.data_as(C.POINTER(C.c_float)), rows.shape[0], C.byref(loss),
            g.ctypes.data_as(C.POINTER(C.c_float))), "mlp_loss_grad")
        return loss.value, g

    def mlp_predict(self, rows: np.ndarray) -> np.ndarray:
        rows = np.ascontiguousarray(rows, dtype=np.float32)
        out = np.zeros(rows.shape[0], dtype=np.float32)
        self._check(self.lib.sphgpu_mlp_predict(
            self.h, rows.ctypes.data_as(C.POINTER(C.c_float)), rows.shape[0],
            out.ctypes.data_as(C.POINTER(C.c_float))), "mlp_predict")
        return out

    # ---- timing ------------------------------------------------------------
    def enable_timing(self, on: int = 1):
        """0 off, 1 events around every kernel group, 2 around the pair
        kernels only (searches, PreInteraction, FluidForce, gravity)."""
        self._check(self.lib.sphgpu_enable_timing(self.h, int(on)), "enable_timing")

    def copy_probe(self, nbytes: int) -> float:
        """Duration (ms) of a plain device copy of nbytes (k_copy_probe)."""
        ms = C.c_double(0)
        self._check(self.lib.sphgpu_copy_probe(self.h, int(nbytes), C.byref(ms)), "copy_probe")
        return ms.value

    def output_meta(self, step: int) -> OutputMeta:
        m = OutputMeta()
        self._check(self.lib.sphgpu_output_meta_init(self.h, int(step), C.byref(m)), "output_meta_init")
        return m

    def write_snapshot_csv(self, path, meta: OutputMeta):
        self._check(self.lib.sphgpu_write_snapshot_csv(self.h, str(path).encode(), C.byref(meta)),
                    "write_snapshot_csv")

    def read_snapshot_csv(self, path) -> OutputMeta:
        m = OutputMeta()
        self._check(self.lib.sphgpu_read_snapshot_csv(self.h, str(path).encode(), C.byref(m)), "read_snapshot_csv")
        return m

    def energy_file_append(self, path, precision: int = 16):
        self._check(self.lib.sphgpu_energy_file_append(self.h, str(path).encode(), int(precision)),
                    "energy_file_append")

    def write_snapshot_vtk(self, path, meta: OutputMeta, binary: bool = True):
        self._check(self.lib.sphgpu_write_snapshot_vtk(self.h, str(path).encode(), C.byref(meta),
                                                       int(bool(binary))), "write_snapshot_vtk")

    def write_snapshot_hdf5(self, path, meta: OutputMeta, compression: int = 6):
        h5 = load_hdf5_library()
        self._check(h5.sphgpu_write_snapshot_hdf5(self.h, str(path).encode(), C.byref(meta), int(compression)),
                    "write_snapshot_hdf5")

    def read_snapshot_hdf5(self, path) -> OutputMeta:
        h5 = load_hdf5_library()
        m = OutputMeta()
        self._check(h5.sphgpu_read_snapshot_hdf5(self.h, str(path).encode(), C.byref(m)), "read_snapshot_hdf5")
        return m

    def kernel_times(self) -> dict:
        buf = C.create_string_buffer(8192)
        avg = np.zeros(128)
        cnt = np.zeros(128, dtype=np.int64)
        nk = C.c_int(0)
        self._check(self.lib.sphgpu_kernel_times(
            self.h, buf, 8192, _dp(avg), cnt.ctypes.data_as(C.POINTER(C.c_int64)), 128, C.byref(nk)),
            "kernel_times")
        names = buf.value.decode().split(";")[:nk.value]
        return {nm: (float(avg[i]), int(cnt[i])) for i, nm in enumerate(names)}


def device_info() -> dict:
    lib = load_library()
    name = C.create_string_buffer(256)
    cu = C.c_int(0)
    hbm = C.c_double(0)
    rc = lib.sphgpu_device_info(name, 256, C.byref(cu), C.byref(hbm))
    if rc != 0:
        raise SphGpuError(f"device_info failed ({rc})")
    return {"name": name.value.decode(), "cu": cu.value, "hbm_gb": hbm.value}
