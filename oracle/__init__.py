"""ctypes handle on the CPU FP64 oracle (liboracle.so, oracle/rs_oracle.h).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker or the timed CPU
baseline, never as part of the product path.
"""
from __future__ import annotations

import ctypes as C
import importlib.util
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# struct layouts only (include/rsolid.h mirror); loading them this way keeps
# the oracle independent of the HIP library
_ABI = os.path.join(os.path.dirname(_HERE), "2d-rapid-solidification_amd", "rsolid", "abi.py")
_spec = importlib.util.spec_from_file_location("rs_oracle_abi", _ABI)
abi = importlib.util.module_from_spec(_spec)
sys.modules["rs_oracle_abi"] = abi
_spec.loader.exec_module(abi)
RS_FIELD_T, RsConstants, RsParams, RsTipRow = abi.RS_FIELD_T, abi.RsConstants, abi.RsParams, abi.RsTipRow
TpParams = abi.TpParams

LIB_PATH = os.path.join(_HERE, "liboracle.so")


def _load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"oracle not built: {LIB_PATH} missing")
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER
    vp = C.c_void_p
    d = C.c_double
    sig = {
        "or_derive": (C.c_int, [P(RsParams), P(RsConstants)]),
        "or_tau0": (d, [P(RsParams)]),
        "or_G_Ti_BCC": (d, [d]), "or_G_Ti_Liq": (d, [d]),
        "or_G_Nb_BCC": (d, [d]), "or_G_Nb_Liq": (d, [d]),
        "or_Gl": (d, [P(RsParams), d, d]), "or_Gs": (d, [P(RsParams), d, d]),
        "or_Glc": (d, [P(RsParams), d, d]), "or_Gsc": (d, [P(RsParams), d, d]),
        "or_anisotropy": (d, [P(RsParams), P(d), P(d), P(d), P(d)]),
        "or_orientation": (None, [P(d), d, d, d]),
        "or_create": (vp, [P(RsParams)]),
        "or_destroy": (None, [vp]),
        "or_init_reference": (C.c_int, [vp]),
        "or_init_config": (C.c_int, [vp, C.c_uint64]),
        "or_get": (C.c_int, [vp, C.c_int, P(d), C.c_int64]),
        "or_put": (C.c_int, [vp, C.c_int, P(d), C.c_int64]),
        "or_get_constants": (C.c_int, [vp, P(RsConstants)]),
        "or_set_constants": (C.c_int, [vp, P(RsConstants)]),
        "or_step": (C.c_int, [vp, C.c_int64]),
        "or_tip_column": (C.c_int, [vp]),
        "or_pull_back": (C.c_int, [vp, C.c_int]),
        "or_solute_mass": (d, [vp]),
        "or_precise_tip": (C.c_int, [P(RsParams), P(d), P(C.c_int32), P(C.c_int32), P(d), P(d),
                                      P(d)]),
        "or_tip": (C.c_int, [vp, P(RsTipRow)]),
        "or_common_tangent": (C.c_int, [P(RsParams), d, P(d), P(d)]),
        "or_at_flux": (d, [P(RsParams), d, d, d, d, d, d, d, d, d, d, d, d]),
        "or_splitmix64": (C.c_uint64, [P(C.c_uint64)]),
        "or_interpolate_at": (C.c_int, [vp, C.c_int]),
        "tpo_g": (d, [d, d, d]), "tpo_gp1": (d, [d, d, d]), "tpo_gp2": (d, [d, d, d]),
        "tpo_mu": (d, [P(TpParams), d, d, d]),
        "tpo_bulk_force": (None, [P(TpParams), d, d, d, P(d)]),
        "tpo_create": (vp, [P(TpParams)]),
        "tpo_destroy": (None, [vp]),
        "tpo_init_reference": (C.c_int, [vp]),
        "tpo_get": (C.c_int, [vp, C.c_int, P(d), C.c_int64]),
        "tpo_put": (C.c_int, [vp, C.c_int, P(d), C.c_int64]),
        "tpo_step": (C.c_int, [vp, C.c_int64]),
        "tpo_get_step": (C.c_int64, [vp]),
        "tpo_report": (C.c_int, [vp, C.c_int, P(d)]),
        "or_integrate_phi_at": (C.c_int, [vp, P(d), P(d), C.c_int, P(d)]),
        "or_set_threads": (C.c_int, [C.c_int]),
        "or_solute_ledger": (C.c_int, [vp, P(d), C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _cast(obj, cls):
    """Same-layout struct from another module (e.g. rsolid.abi) -> cls."""
    if isinstance(obj, cls):
        return obj
    out = cls()
    if C.sizeof(obj) != C.sizeof(cls):
        raise TypeError(f"struct size mismatch for {cls.__name__}")
    C.memmove(C.byref(out), C.byref(obj), C.sizeof(cls))
    return out


def _P(p):
    return C.byref(_cast(p, RsParams))


def set_threads(n: int) -> int:
    return lib.or_set_threads(n)


def derive(p: RsParams) -> RsConstants:
    k = RsConstants()
    lib.or_derive(_P(p), C.byref(k))
    return k


def tau0(p: RsParams) -> float:
    return lib.or_tau0(_P(p))


def G(name: str, T: float) -> float:
    return getattr(lib, "or_" + name)(T)


def Gl(p, T, c): return lib.or_Gl(_P(p), T, c)
def Gs(p, T, c): return lib.or_Gs(_P(p), T, c)
def Glc(p, T, c): return lib.or_Glc(_P(p), T, c)
def Gsc(p, T, c): return lib.or_Gsc(_P(p), T, c)


def anisotropy(p: RsParams, ph3x3: np.ndarray, orient=None):
    ph = np.ascontiguousarray(ph3x3, dtype=np.float64).reshape(9)
    o = np.ascontiguousarray(orient if orient is not None else np.eye(3), dtype=np.float64).reshape(9)
    a_s = C.c_double()
    a_k = C.c_double()
    v = lib.or_anisotropy(_P(p), _dptr(ph), _dptr(o), C.byref(a_s), C.byref(a_k))
    return v, a_s.value, a_k.value


def orientation(alpha: float, beta: float, gamma: float) -> np.ndarray:
    """set_orientation (fun.cu:750-771): 3x3 rotation from Euler angles (deg)."""
    o = np.zeros(9)
    lib.or_orientation(_dptr(o), alpha, beta, gamma)
    return o.reshape(3, 3)


def at_flux(p, vdn, dt, pP, pQ, nP, nQ, cP, cQ, t1=0.0, t2=0.0, t3=0.0, t4=0.0) -> float:
    return lib.or_at_flux(_P(p), vdn, dt, pP, pQ, nP, nQ, cP, cQ, t1, t2, t3, t4)


def common_tangent(p: RsParams, T: float, cl0: float, cs0: float):
    a = C.c_double(cl0)
    b = C.c_double(cs0)
    st = lib.or_common_tangent(_P(p), T, C.byref(a), C.byref(b))
    return st, a.value, b.value


def precise_tip(p: RsParams, psi: np.ndarray):
    a = np.ascontiguousarray(psi, dtype=np.float64)
    ti = C.c_int32()
    tj = C.c_int32()
    x = C.c_double()
    y = C.c_double()
    R = C.c_double()
    st = lib.or_precise_tip(_P(p), _dptr(a), C.byref(ti), C.byref(tj), C.byref(x),
                            C.byref(y), C.byref(R))
    return st, dict(i=ti.value, j=tj.value, x=x.value, y=y.value, R=R.value)


class State:
    """The oracle's whole-run state (fields, coarse T, run constants)."""

    def __init__(self, p: RsParams):
        self.p = _cast(p, RsParams).copy()
        self._h = lib.or_create(C.byref(self.p))
        if not self._h:
            raise RuntimeError("or_create failed")
        self.k0 = derive(self.p)

    def close(self):
        if self._h:
            lib.or_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *e):
        self.close()

    def _chk(self, rc, what):
        if rc:
            raise RuntimeError(f"oracle {what} failed: {rc}")

    def init_reference(self):
        self._chk(lib.or_init_reference(self._h), "init_reference")

    def init_config(self, seed: int):
        self._chk(lib.or_init_config(self._h, C.c_uint64(seed)), "init_config")

    def get(self, field: int) -> np.ndarray:
        shp = (self.k0.NyT_all, self.k0.NxT_all) if field == RS_FIELD_T else (self.p.Ny, self.p.Nx)
        out = np.empty(shp, dtype=np.float64)
        self._chk(lib.or_get(self._h, field, _dptr(out), out.size), "get")
        return out

    def put(self, field: int, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self._chk(lib.or_put(self._h, field, _dptr(a), a.size), "put")

    def constants(self) -> RsConstants:
        k = RsConstants()
        lib.or_get_constants(self._h, C.byref(k))
        return k

    def set_constants(self, k: RsConstants):
        lib.or_set_constants(self._h, C.byref(_cast(k, RsConstants)))

    def step(self, n: int = 1):
        self._chk(lib.or_step(self._h, n), "step")

    def tip_column(self) -> int:
        return lib.or_tip_column(self._h)

    def pull_back(self, off: int):
        self._chk(lib.or_pull_back(self._h, off), "pull_back")

    def solute_mass(self) -> float:
        return lib.or_solute_mass(self._h)

    def interpolate_at(self, i_left_end: int) -> np.ndarray:
        """T_fine (Ny x Nx) interpolated from the current coarse T with the
        given i_left_end (interpolate_temperature)."""
        self._chk(lib.or_interpolate_at(self._h, i_left_end), "interpolate_at")
        return self.get(abi.RS_FIELD_TFINE)

    def integrate_phi_at(self, psi: np.ndarray, psi_next: np.ndarray, i_left_end: int) -> np.ndarray:
        """Latent-heat block sums (NyT_all x NxT_all) of psi_next - psi
        (integrate_phi); replaces the state's psi and psi_next."""
        a = np.ascontiguousarray(psi, dtype=np.float64)
        b = np.ascontiguousarray(psi_next, dtype=np.float64)
        k = self.constants()
        out = np.zeros((k.NyT_all, k.NxT_all))
        self._chk(lib.or_integrate_phi_at(self._h, _dptr(a), _dptr(b), i_left_end, _dptr(out)),
                  "integrate_phi_at")
        return out

    def solute_ledger(self, reset: bool = False) -> dict:
        out = np.zeros(7)
        lib.or_solute_ledger(self._h, _dptr(out), 1 if reset else 0)
        return dict(zip(("mass", "mass0", "pull_back", "edge_x", "edge_y", "clamped",
                         "unaccounted"), out.tolist()))

    def tip(self) -> dict:
        t = RsTipRow()
        lib.or_tip(self._h, C.byref(t))
        return t.as_dict()


# ---- 3D peritectic Ti-Ag restatement (tp_oracle.c) -------------------------
def tp_g(pi, pj, pk): return lib.tpo_g(pi, pj, pk)
def tp_gp1(pi, pj, pk): return lib.tpo_gp1(pi, pj, pk)
def tp_gp2(pi, pj, pk): return lib.tpo_gp2(pi, pj, pk)
def tp_mu(p, p1, p2, c): return lib.tpo_mu(C.byref(_cast(p, TpParams)), p1, p2, c)


def tp_bulk_force(p, p1, p2, muu) -> np.ndarray:
    out = np.zeros(3)
    lib.tpo_bulk_force(C.byref(_cast(p, TpParams)), p1, p2, muu, _dptr(out))
    return out


class TpState:
    """CPU restatement of one Ti-Ag slab (fields of shape (Nz, Ny, Nx))."""

    def __init__(self, p):
        self.p = _cast(p, TpParams)
        self._h = lib.tpo_create(C.byref(self.p))
        if not self._h:
            raise RuntimeError("tpo_create failed")
        self.shape = (self.p.Nz, self.p.Ny, self.p.Nx)

    def close(self):
        if getattr(self, "_h", None):
            lib.tpo_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def init_reference(self):
        lib.tpo_init_reference(self._h)

    def get(self, field: int) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64)
        if lib.tpo_get(self._h, field, _dptr(out), out.size):
            raise RuntimeError("tpo_get failed")
        return out

    def put(self, field: int, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if lib.tpo_put(self._h, field, _dptr(a), a.size):
            raise RuntimeError("tpo_put failed")

    def step(self, n: int = 1):
        lib.tpo_step(self._h, n)

    def get_step(self) -> int:
        return lib.tpo_get_step(self._h)

    def report(self, k_check: int = 20) -> dict:
        out = np.zeros(2)
        lib.tpo_report(self._h, k_check, _dptr(out))
        return {"mean_c": out[0], "max_abs_c_plane": out[1]}
