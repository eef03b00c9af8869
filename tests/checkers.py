"""ctypes bindings of the CPU checkers (test infrastructure only).

``Oracle`` wraps oracle/liboracle.so (the restatement, all DIMs) and
``Reference`` wraps oracle/_ref/libsphref_dD.so (the reference's own modules,
present only where the reference tree was available at build time).  Both
expose the same methods so tests can drive them side by side.
"""
from __future__ import annotations

import ctypes as C
import os
import sys
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(REPO / "sph-code-simulation_amd"))

import ics  # noqa: E402
from sphgpu import Params, params_struct  # noqa: E402

ORACLE_SO = REPO / "oracle" / "liboracle.so"
REF_DIR = REPO / "oracle" / "_ref"


def ref_available(dim: int) -> bool:
    return (REF_DIR / f"libsphref_d{dim}.so").exists()


class _Checker:
    prefix = ""

    def _bind(self, lib):
        p = self.prefix
        vp, dp, ip, i32p = C.c_void_p, C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_int32)
        sig = {
            "create": (vp, [C.POINTER(Params), ip]),
            "destroy": (None, [vp]),
            "set_fields": (None, [vp, dp, C.c_uint64]),
            "get_fields": (None, [vp, dp, C.c_uint64]),
            "make_tree": (ip, [vp]),
            "set_kernel": (ip, [vp]),
            "pre_interaction": (ip, [vp]),
            "fluid_force": (ip, [vp]),
            "gravity": (ip, [vp]),
            "timestep": (ip, [vp]),
            "predict": (ip, [vp]),
            "correct": (ip, [vp]),
            "initialize": (ip, [vp]),
            "integrate": (ip, [vp, ip]),
            "energy": (None, [vp, dp]),
            "get_dt": (C.c_double, [vp]),
            "set_dt": (None, [vp, C.c_double]),
            "get_time": (C.c_double, [vp]),
            "get_h_per_v_sig": (C.c_double, [vp]),
            "set_pre_first": (None, [vp, ip]),
            "neighbor_search": (ip, [vp, ip, ip, i32p, ip]),
            "tree_dump": (C.c_int64, [vp, dp, i32p, C.POINTER(C.c_int64)]),
            "riemann": (ip, [ip, C.c_double, dp, dp, dp, dp]),
        }
        self._f = {}
        for name, (res, args) in sig.items():
            fn = getattr(lib, p + name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
            self._f[name] = fn

    def __init__(self, lib, params: dict, fields: np.ndarray):
        self._bind(lib)
        self.lib = lib
        self.params = dict(params)
        self.n = fields.shape[1]
        self._p = params_struct(params)
        self.h = self._f["create"](C.byref(self._p), self.n)
        if not self.h:
            raise RuntimeError("checker create failed")
        self.set_fields(fields)

    def close(self):
        if self.h:
            self._f["destroy"](self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _dp(a):
        return a.ctypes.data_as(C.POINTER(C.c_double))

    def set_fields(self, f: np.ndarray, mask: int = (1 << ics.NFIELDS) - 1):
        f = np.ascontiguousarray(f, dtype=np.float64)
        assert f.shape == (ics.NFIELDS, self.n)
        self._f["set_fields"](self.h, self._dp(f), mask)

    def get_fields(self, mask: int = (1 << ics.NFIELDS) - 1) -> np.ndarray:
        f = np.zeros((ics.NFIELDS, self.n), dtype=np.float64)
        self._f["get_fields"](self.h, self._dp(f), mask)
        return f

    def _call(self, name, *a):
        r = self._f[name](self.h, *a)
        if isinstance(r, int) and r < 0:
            raise RuntimeError(f"{self.prefix}{name} failed: {r}")
        return r

    def make_tree(self): return self._call("make_tree")
    def set_kernel(self): return self._call("set_kernel")
    def pre_interaction(self): return self._call("pre_interaction")
    def fluid_force(self): return self._call("fluid_force")
    def gravity(self): return self._call("gravity")
    def timestep(self): return self._call("timestep")
    def predict(self): return self._call("predict")
    def correct(self): return self._call("correct")
    def initialize(self): return self._call("initialize")
    def integrate(self, n=1): return self._call("integrate", int(n))
    def set_pre_first(self, v: bool): self._f["set_pre_first"](self.h, int(v))

    @property
    def dt(self): return self._f["get_dt"](self.h)

    @dt.setter
    def dt(self, v): self._f["set_dt"](self.h, float(v))

    @property
    def time(self): return self._f["get_time"](self.h)

    @property
    def h_per_v_sig(self): return self._f["get_h_per_v_sig"](self.h)

    def energy(self) -> np.ndarray:
        out = np.zeros(7)
        self._f["energy"](self.h, self._dp(out))
        return out

    def neighbor_search(self, i: int, is_ij: bool) -> np.ndarray:
        cap = self.params["neighbor_number"] * 1000
        out = np.zeros(cap, dtype=np.int32)
        r = self._f["neighbor_search"](self.h, int(i), int(is_ij),
                                       out.ctypes.data_as(C.POINTER(C.c_int32)), cap)
        if r < 0:
            raise RuntimeError(f"neighbor_search failed: {r}")
        return out[:r].copy()

    def tree_dump(self):
        nm = C.c_int64(0)
        nn = self._f["tree_dump"](self.h, None, None, C.byref(nm))
        rec = np.zeros(nn * 13)
        mem = np.zeros(nm.value, dtype=np.int32)
        self._f["tree_dump"](self.h, self._dp(rec), mem.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(nm))
        return rec.reshape(nn, 13), mem

    def riemann(self, solver: int, gamma: float, left, right):
        lft = np.ascontiguousarray(left, dtype=np.float64)
        rgt = np.ascontiguousarray(right, dtype=np.float64)
        ps = np.zeros(1)
        vs = np.zeros(1)
        it = self._f["riemann"](int(solver), float(gamma), self._dp(lft), self._dp(rgt),
                                self._dp(ps), self._dp(vs))
        return ps[0], vs[0], it


_oracle_lib = None


class Oracle(_Checker):
    prefix = "orc_"

    def __init__(self, params, fields):
        global _oracle_lib
        if _oracle_lib is None:
            _oracle_lib = C.CDLL(str(ORACLE_SO))
        super().__init__(_oracle_lib, params, fields)


_ref_libs: dict = {}


class Reference(_Checker):
    prefix = "ref_"

    def __init__(self, params, fields):
        d = params["dim"]
        if d not in _ref_libs:
            _ref_libs[d] = C.CDLL(str(REF_DIR / f"libsphref_d{d}.so"), mode=os.RTLD_LOCAL)
        super().__init__(_ref_libs[d], params, fields)


_dropin_libs: dict = {}


def dropin_available(dim: int) -> bool:
    return (REF_DIR / f"libsphdropin_d{dim}.so").exists()


class DropIn(_Checker):
    """The reference's modules and driver with the GPU Module adapters of
    sph-code-simulation_amd/dropin wired in (oracle/_ref/libsphdropin_dD.so);
    needs a GPU."""
    prefix = "ref_"

    def __init__(self, params, fields):
        d = params["dim"]
        if d not in _dropin_libs:
            _dropin_libs[d] = C.CDLL(str(REF_DIR / f"libsphdropin_d{d}.so"), mode=os.RTLD_LOCAL)
        super().__init__(_dropin_libs[d], params, fields)


def oracle_riemann(solver: int, gamma: float, left, right):
    """One Riemann problem through the oracle's restated solvers."""
    global _oracle_lib
    if _oracle_lib is None:
        _oracle_lib = C.CDLL(str(ORACLE_SO))
    fn = _oracle_lib.orc_riemann
    fn.restype = C.c_int
    fn.argtypes = [C.c_int, C.c_double] + [C.POINTER(C.c_double)] * 4
    lft = np.ascontiguousarray(left, dtype=np.float64)
    rgt = np.ascontiguousarray(right, dtype=np.float64)
    ps = np.zeros(1)
    vs = np.zeros(1)
    it = fn(int(solver), float(gamma), lft.ctypes.data_as(C.POINTER(C.c_double)),
            rgt.ctypes.data_as(C.POINTER(C.c_double)), ps.ctypes.data_as(C.POINTER(C.c_double)),
            vs.ctypes.data_as(C.POINTER(C.c_double)))
    return ps[0], vs[0], it


def oracle_kernel(dim: int, kernel: int, r: float, h: float) -> np.ndarray:
    global _oracle_lib
    if _oracle_lib is None:
        _oracle_lib = C.CDLL(str(ORACLE_SO))
    fn = _oracle_lib.orc_kernel_eval
    fn.restype = None
    fn.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.POINTER(C.c_double)]
    out = np.zeros(3)
    fn(dim, kernel, float(r), float(h), out.ctypes.data_as(C.POINTER(C.c_double)))
    return out
