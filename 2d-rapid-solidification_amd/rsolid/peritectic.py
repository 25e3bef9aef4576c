"""Python handle on the 3D peritectic Ti-Ag stepper of librsolid.so
(C ABI in include/rs_peritectic.h): one z-slab of the reference's
multi/peritectic_multiGPU/TiAg.cu model on one GPU. Every step runs in the
library's HIP kernel; there is no CPU fallback."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import RsError, lib
from .abi import TP_FIELD_C, TP_FIELD_MU, TP_FIELD_PHI1, TP_FIELD_PHI2, TpParams  # noqa: F401

_P = C.POINTER
_vp = C.c_void_p
for _name, (_res, _args) in {
    "tp_abi_info": (C.c_int, [_P(C.c_int32)]),
    "tp_params_reference": (C.c_int, [_P(TpParams), C.c_int32, C.c_int32, C.c_int32]),
    "tp_derived": (C.c_int, [_P(TpParams), _P(C.c_double), C.c_int32]),
    "tp_create": (C.c_int, [_P(TpParams), _P(_vp)]),
    "tp_destroy": (C.c_int, [_vp]),
    "tp_last_error": (C.c_char_p, [_vp]),
    "tp_init_reference": (C.c_int, [_vp]),
    "tp_upload": (C.c_int, [_vp, C.c_int32, _P(C.c_double), C.c_int64]),
    "tp_download": (C.c_int, [_vp, C.c_int32, _P(C.c_double), C.c_int64]),
    "tp_step": (C.c_int, [_vp, C.c_int64]),
    "tp_get_step": (C.c_int, [_vp, _P(C.c_int64)]),
    "tp_set_step": (C.c_int, [_vp, C.c_int64]),
    "tp_report": (C.c_int, [_vp, C.c_int32, _P(C.c_double), C.c_int32]),
    "tp_sync": (C.c_int, [_vp]),
    "tp_timing": (C.c_int, [_vp, C.c_int32, _P(C.c_double), C.c_int32]),
    "tp_write_cross_section": (C.c_int, [C.c_char_p, _P(C.c_double), C.c_int32, C.c_int32,
                                         C.c_int32, C.c_int32, C.c_int32]),
    "tp_write_break_point": (C.c_int, [C.c_char_p, _P(C.c_double), C.c_int32, C.c_int32,
                                       C.c_int32]),
    "tp_read_break_point": (C.c_int, [C.c_char_p, _P(C.c_double), C.c_int32, C.c_int32,
                                      C.c_int32]),
}.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(_P(C.c_double))


def reference_params(Nx: int = 512, Ny: int = 512, Nz: int = 128) -> TpParams:
    """The TiAg.cu constants; 512 x 512 x 128 is one GPU's slab of the
    reference's 4-GPU run (devNz = 128)."""
    p = TpParams()
    rc = lib.tp_params_reference(C.byref(p), Nx, Ny, Nz)
    if rc:
        raise RsError(rc, "tp_params_reference")
    return p


def derived(p: TpParams) -> dict:
    v = (C.c_double * 6)()
    lib.tp_derived(C.byref(p), v, 6)
    return dict(zip(("ca", "cb", "cla", "clb", "B1", "B2"), list(v)))


def write_cross_section(path: str, data: np.ndarray, plane: int, index: int) -> None:
    a = np.ascontiguousarray(data, dtype=np.float64)
    Nz, Ny, Nx = a.shape
    rc = lib.tp_write_cross_section(path.encode(), _dptr(a), Nx, Ny, Nz, plane, index)
    if rc:
        raise RsError(rc, f"tp_write_cross_section {path}")


def write_break_point(path: str, data: np.ndarray) -> None:
    a = np.ascontiguousarray(data, dtype=np.float64)
    Nz, Ny, Nx = a.shape
    rc = lib.tp_write_break_point(path.encode(), _dptr(a), Nx, Ny, Nz)
    if rc:
        raise RsError(rc, f"tp_write_break_point {path}")


def read_break_point(path: str, shape) -> np.ndarray:
    Nz, Ny, Nx = shape
    out = np.empty(shape, dtype=np.float64)
    rc = lib.tp_read_break_point(path.encode(), _dptr(out), Nx, Ny, Nz)
    if rc:
        raise RsError(rc, f"tp_read_break_point {path}")
    return out


class Slab:
    """One GPU's Ti-Ag peritectic run (fields of shape (Nz, Ny, Nx))."""

    def __init__(self, p: TpParams):
        self.p = p.copy()
        h = _vp()
        rc = lib.tp_create(C.byref(self.p), C.byref(h))
        if rc:
            raise RsError(rc, "tp_create: " + lib.tp_last_error(None).decode())
        self._h = h
        self.shape = (p.Nz, p.Ny, p.Nx)

    def _chk(self, rc: int, what: str) -> None:
        if rc:
            raise RsError(rc, f"{what}: {lib.tp_last_error(self._h).decode()}")

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib.tp_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def init_reference(self) -> None:
        self._chk(lib.tp_init_reference(self._h), "tp_init_reference")

    def upload(self, field: int, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64)
        self._chk(lib.tp_upload(self._h, field, _dptr(a), a.size), "tp_upload")

    def download(self, field: int) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64)
        self._chk(lib.tp_download(self._h, field, _dptr(out), out.size), "tp_download")
        return out

    def step(self, n: int = 1) -> None:
        self._chk(lib.tp_step(self._h, n), "tp_step")

    def get_step(self) -> int:
        s = C.c_int64()
        self._chk(lib.tp_get_step(self._h, C.byref(s)), "tp_get_step")
        return s.value

    def report(self, k_check: int = 20) -> dict:
        v = (C.c_double * 2)()
        self._chk(lib.tp_report(self._h, k_check, v, 2), "tp_report")
        return {"mean_c": v[0], "max_abs_c_plane": v[1]}

    def sync(self) -> None:
        self._chk(lib.tp_sync(self._h), "tp_sync")

    def timing(self, reset: bool = False):
        v = (C.c_double * 2)()
        self._chk(lib.tp_timing(self._h, 1 if reset else 0, v, 2), "tp_timing")
        return v[0], int(v[1])
