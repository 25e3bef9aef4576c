"""Python handle on the HIP time-stepper library (librsolid.so, C ABI in
include/rsolid.h).

The library is the product: every field update runs in its HIP kernels.
Importing this module loads the prebuilt shared library from this folder and
fails loudly when it is missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .abi import (  # noqa: F401  (re-exported)
    EXPORTED, RS_ACT_BF16, RS_ACT_BF16_SPLIT, RS_ACT_FP16, RS_ACT_FP16_PACKED, RS_ERR_ARG, RS_ERR_DOMAIN, RS_ERR_HIP, RS_ERR_NOMEM, RS_ERR_NONFINITE, RS_ERR_STATE, RS_FIELD_C,
    RS_FIELD_PHI, RS_FIELD_PSI, RS_FIELD_T, RS_FIELD_TFINE, RS_OK, RS_PREC_FP64, RS_PREC_MIXED,
    RS_PREC_SURROGATE, RS_SURR_NMETA, RS_SURR_NPARAM, RS_SURR_NREPORT, RsConstants, RsParams,
    RsSurrogateOpts, RsTipRow,
)

LEDGER_KEYS = ("mass", "mass0", "pull_back", "edge_x", "edge_y", "unaccounted")
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "librsolid.so")


class RsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"rsolid error {code}: {msg}")
        self.code = code


def _load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"HIP library not built: {LIB_PATH} is missing "
            "(run python 2d-rapid-solidification_amd/build.py)")
    lib = C.CDLL(LIB_PATH)
    P = C.POINTER
    vp = C.c_void_p
    sig = {
        "rs_abi_info": (C.c_int, [P(C.c_int32)]),
        "rs_params_reference": (C.c_int, [P(RsParams), C.c_int32, C.c_int32]),
        "rs_params_config": (C.c_int, [P(RsParams), C.c_int32, C.c_int32, C.c_double]),
        "rs_derive": (C.c_int, [P(RsParams), P(RsConstants)]),
        "rs_create": (C.c_int, [P(RsParams), P(vp)]),
        "rs_destroy": (C.c_int, [vp]),
        "rs_last_error": (C.c_char_p, [vp]),
        "rs_device_name": (C.c_int, [C.c_char_p, C.c_int32]),
        "rs_init_reference": (C.c_int, [vp]),
        "rs_init_config": (C.c_int, [vp, C.c_uint64]),
        "rs_upload": (C.c_int, [vp, C.c_int32, P(C.c_double), C.c_int64]),
        "rs_download": (C.c_int, [vp, C.c_int32, P(C.c_double), C.c_int64]),
        "rs_get_constants": (C.c_int, [vp, P(RsConstants)]),
        "rs_set_constants": (C.c_int, [vp, P(RsConstants)]),
        "rs_step": (C.c_int, [vp, C.c_int64]),
        "rs_tip_column": (C.c_int, [vp, P(C.c_int32)]),
        "rs_pull_back": (C.c_int, [vp, C.c_int32]),
        "rs_solute_mass": (C.c_int, [vp, P(C.c_double)]),
        "rs_solute_mass_noalloc": (C.c_int, [vp, C.c_int32, C.c_int32, vp, vp]),
        "rs_solute_ledger": (C.c_int, [vp, P(C.c_double), C.c_int32, C.c_int32]),
        "rs_tip": (C.c_int, [vp, P(RsTipRow)]),
        "rs_sync": (C.c_int, [vp]),
        "rs_timing_enable": (C.c_int, [vp, C.c_int32]),
        "rs_set_finite_check": (C.c_int, [vp, C.c_int32]),
        "rs_timing": (C.c_int, [vp, C.c_int32, P(C.c_double), P(C.c_int64)]),
        "rs_calibrate_copy": (C.c_int, [vp, C.c_int64, C.c_int32, P(C.c_double)]),
        "rs_context_info": (C.c_int, [vp, P(C.c_double), C.c_int32]),
        "rs_surrogate_defaults": (C.c_int, [P(RsSurrogateOpts)]),
        "rs_surrogate_train": (C.c_int, [vp, P(RsSurrogateOpts), P(C.c_double), C.c_int32]),
        "rs_surrogate_get": (C.c_int, [vp, P(C.c_float), C.c_int32, P(C.c_double), C.c_int32]),
        "rs_surrogate_set": (C.c_int, [vp, P(C.c_float), C.c_int32, P(C.c_double), C.c_int32]),
        "rs_surrogate_grad": (C.c_int, [vp, P(C.c_float), P(C.c_float), C.c_int32, P(C.c_float),
                                        P(C.c_double)]),
        "rs_surrogate_eval": (C.c_int, [vp, P(C.c_double), C.c_int64, P(C.c_double)]),
        "rs_surrogate_time": (C.c_int, [vp, C.c_int64, C.c_int32, P(C.c_double)]),
        "rs_surrogate_stats": (C.c_int, [vp, P(C.c_double), C.c_int32]),
        "rs_batch_create": (C.c_int, [P(RsParams), C.c_int32, P(vp)]),
        "rs_batch_destroy": (C.c_int, [vp]),
        "rs_batch_last_error": (C.c_char_p, [vp]),
        "rs_batch_run": (C.c_int, [vp, C.c_int32, P(vp)]),
        "rs_batch_step": (C.c_int, [vp, C.c_int64]),
        "rs_batch_surrogate_train": (C.c_int, [vp, P(RsSurrogateOpts), P(C.c_double), C.c_int32]),
        "rs_batch_tune": (C.c_int, [vp, C.c_int32, P(C.c_double), C.c_int32]),
        "rs_batch_info": (C.c_int, [vp, P(C.c_double), C.c_int32]),
        "rs_write_field_txt": (C.c_int, [C.c_char_p, P(C.c_double), C.c_int32, C.c_int32, C.c_int32,
                                         C.c_int32]),
        "rs_write_tip_row": (C.c_int, [C.c_char_p, P(RsTipRow), C.c_int32]),
        "rs_append_offset": (C.c_int, [C.c_char_p, C.c_int64, C.c_int32, C.c_int32, C.c_int32]),
        "rs_append_history": (C.c_int, [C.c_char_p, P(C.c_double), C.c_int32, C.c_int32, C.c_int32,
                                        C.c_int32, C.c_int32, P(C.c_int32)]),
        "rs_write_checkpoint": (C.c_int, [C.c_char_p, P(C.c_double), C.c_int64]),
        "rs_read_checkpoint": (C.c_int, [C.c_char_p, P(C.c_double), C.c_int64]),
        "rs_write_parameters": (C.c_int, [C.c_char_p, P(RsParams), P(RsConstants), C.c_int32]),
        "rs_read_parameters": (C.c_int, [C.c_char_p, P(RsParams), P(RsConstants)]),
        "rs_download_columns": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, P(C.c_double)]),
        "rs_save_state": (C.c_int, [vp, C.c_char_p]),
        "rs_load_state": (C.c_int, [vp, C.c_char_p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def reference_params(Nx: int = 768, Ny: int = 832) -> RsParams:
    """fun.cu values with the run.sh flags (fun.cu:20-174, run.sh:6-23)."""
    p = RsParams()
    rc = lib.rs_params_reference(C.byref(p), Nx, Ny)
    if rc:
        raise RsError(rc, "rs_params_reference")
    return p


def config_params(N: int, tfc: bool = False, alpha_ratio: float = 100.0,
                  precision: int = RS_PREC_FP64) -> RsParams:
    """Ti-10at%Nb configuration on an N x N grid (frozen or coupled T)."""
    p = RsParams()
    rc = lib.rs_params_config(C.byref(p), N, 1 if tfc else 0, alpha_ratio)
    if rc:
        raise RsError(rc, "rs_params_config")
    p.precision = precision
    return p


def ensemble_params(N: int, deltaT: float, Vp: float, precision: int = RS_PREC_MIXED,
                    tfc: bool = True) -> RsParams:
    """Ti-10at%Nb configuration on an N x N grid with undercooling deltaT (K)
    and pulling speed Vp (m/s): the coarse T sample is shifted so the initial
    front sits deltaT below the liquidus, as in the preset."""
    p = config_params(N, tfc, precision=precision)
    p.TCOLD += p.deltaT0 - deltaT
    p.deltaT0 = deltaT
    p.Vp = Vp
    return p


def draw_ensemble(nruns: int = 64, seed: int = 2024):
    """The seeded batched-ensemble draw: grid sizes from {256, 512, 1024,
    2048} with probabilities 0.4/0.3/0.2/0.1 (uneven counts), undercooling
    ~ U[25, 60] K, pulling speed ~ U[0.03, 0.3] m/s and one init_config
    seed per run. Returns (sizes, deltaT, Vp, seeds)."""
    rng = np.random.default_rng(seed)
    sizes = rng.choice([256, 512, 1024, 2048], size=nruns, p=[0.4, 0.3, 0.2, 0.1])
    dts = rng.uniform(25.0, 60.0, nruns)
    vps = rng.uniform(0.03, 0.3, nruns)
    seeds = [int(x) for x in rng.integers(1, 2**31, nruns)]
    return sizes, dts, vps, seeds


def write_field_txt(path: str, data: np.ndarray, di: int = 1, dj: int = 1) -> None:
    a = np.ascontiguousarray(data, dtype=np.float64)
    rc = lib.rs_write_field_txt(path.encode(), _dptr(a), a.shape[1], a.shape[0], di, dj)
    if rc:
        raise RsError(rc, f"rs_write_field_txt {path}")


def write_parameters(path: str, p: RsParams, k: RsConstants, i_tip: int = 0) -> None:
    rc = lib.rs_write_parameters(path.encode(), C.byref(p), C.byref(k), i_tip)
    if rc:
        raise RsError(rc, f"rs_write_parameters {path}")


def read_parameters(path: str, p: RsParams) -> RsConstants:
    k = derive(p)
    rc = lib.rs_read_parameters(path.encode(), C.byref(p), C.byref(k))
    if rc:
        raise RsError(rc, f"rs_read_parameters {path}")
    return k


def derive(p: RsParams) -> RsConstants:
    k = RsConstants()
    rc = lib.rs_derive(C.byref(p), C.byref(k))
    if rc:
        raise RsError(rc, "rs_derive")
    return k


def device_name() -> str:
    buf = C.create_string_buffer(256)
    rc = lib.rs_device_name(buf, 256)
    if rc:
        raise RsError(rc, "rs_device_name")
    return buf.value.decode()


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def surrogate_defaults(**kw) -> RsSurrogateOpts:
    """Default training options of the CALPHAD surrogate, overridden by kw."""
    o = RsSurrogateOpts()
    rc = lib.rs_surrogate_defaults(C.byref(o))
    if rc:
        raise RsError(rc, "rs_surrogate_defaults")
    fields = {f for f, _ in o._fields_}
    for key, val in kw.items():
        if key not in fields:
            raise TypeError(f"unknown surrogate option {key!r}")
        setattr(o, key, val)
    return o


SURR_REPORT_KEYS = ("mse", "trusted_frac", "err_dG", "err_fS", "err_muL", "err_dmu",
                    "max_res_trusted", "train_s", "steps", "T_lo", "T_hi", "mse0")


class Context:
    """One time-stepper on the current HIP device (own stream and buffers)."""

    def __init__(self, p: RsParams):
        self.p = p.copy()
        h = C.c_void_p()
        rc = lib.rs_create(C.byref(self.p), C.byref(h))
        if rc:
            raise RsError(rc, f"rs_create: {lib.rs_last_error(None).decode()}")
        self._h = h
        self.k0 = derive(self.p)

    # -- lifetime -----------------------------------------------------
    def close(self) -> None:
        if self._h:
            lib.rs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _chk(self, rc: int, what: str) -> None:
        if rc:
            raise RsError(rc, f"{what}: {lib.rs_last_error(self._h).decode()}")

    # -- state ----------------------------------------------------------
    @property
    def shape(self) -> tuple[int, int]:
        return (self.p.Ny, self.p.Nx)

    @property
    def shape_T(self) -> tuple[int, int]:
        return (self.k0.NyT_all, self.k0.NxT_all)

    def init_reference(self) -> None:
        self._chk(lib.rs_init_reference(self._h), "rs_init_reference")

    def init_config(self, seed: int) -> None:
        self._chk(lib.rs_init_config(self._h, C.c_uint64(seed)), "rs_init_config")

    def upload(self, field: int, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr, dtype=np.float64)
        self._chk(lib.rs_upload(self._h, field, _dptr(a), a.size), "rs_upload")

    def download(self, field: int) -> np.ndarray:
        shp = self.shape_T if field == RS_FIELD_T else self.shape
        out = np.empty(shp, dtype=np.float64)
        self._chk(lib.rs_download(self._h, field, _dptr(out), out.size), "rs_download")
        return out

    def constants(self) -> RsConstants:
        k = RsConstants()
        self._chk(lib.rs_get_constants(self._h, C.byref(k)), "rs_get_constants")
        return k

    def set_constants(self, k: RsConstants) -> None:
        self._chk(lib.rs_set_constants(self._h, C.byref(k)), "rs_set_constants")

    # -- stepping ---------------------------------------------------------
    def step(self, n: int = 1) -> None:
        self._chk(lib.rs_step(self._h, n), "rs_step")

    def tip_column(self) -> int:
        v = C.c_int32()
        self._chk(lib.rs_tip_column(self._h, C.byref(v)), "rs_tip_column")
        return v.value

    def pull_back(self, i_offset: int) -> None:
        self._chk(lib.rs_pull_back(self._h, i_offset), "rs_pull_back")

    def solute_mass(self) -> float:
        v = C.c_double()
        self._chk(lib.rs_solute_mass(self._h, C.byref(v)), "rs_solute_mass")
        return v.value

    def solute_ledger(self, reset: bool = False) -> dict:
        """Interior solute mass now and at the ledger's reset, the mass the
        pull-backs changed, the solute gained through the x and through the
        y boundary faces, and the unaccounted remainder (rs_solute_ledger).
        The ledger is off until a call with reset=True starts it at the
        current fields; reset=False before that raises (RS_ERR_STATE)."""
        v = (C.c_double * 6)()
        self._chk(lib.rs_solute_ledger(self._h, v, 6, 1 if reset else 0), "rs_solute_ledger")
        return dict(zip(LEDGER_KEYS, list(v)))

    def tip(self) -> dict:
        t = RsTipRow()
        self._chk(lib.rs_tip(self._h, C.byref(t)), "rs_tip")
        return t.as_dict()

    def sync(self) -> None:
        self._chk(lib.rs_sync(self._h), "rs_sync")

    def timing_enable(self, on: bool = True, every: int = 1) -> None:
        """HIP-event timing of the kernels of every `every`-th step."""
        self._chk(lib.rs_timing_enable(self._h, every if on else 0), "rs_timing_enable")

    def finite_check(self, every: int = 1) -> None:
        """Check psi, c and the coarse T for NaN / Inf after every `every`-th
        step (0: off); a step that finds one raises RsError
        (RS_ERR_NONFINITE) naming the step, field and first cell."""
        self._chk(lib.rs_set_finite_check(self._h, every), "rs_set_finite_check")

    def info(self) -> dict:
        v = (C.c_double * 9)()
        self._chk(lib.rs_context_info(self._h, v, 9), "rs_context_info")
        return {"poly_err": v[0], "Tref": v[1], "T_lo": v[2], "T_hi": v[3],
                "band_h": int(v[4]), "waves_per_block": int(v[5]),
                "step_variant": int(v[6]), "step_ident": int(v[7]),
                "step_variant_run": int(v[8])}

    def calibrate_copy(self, n: int, reps: int = 5) -> float:
        """Average ms of a plain FP64 copy of n doubles (destroys the fields)."""
        ms = C.c_double()
        self._chk(lib.rs_calibrate_copy(self._h, n, reps, C.byref(ms)), "rs_calibrate_copy")
        return ms.value

    # -- learned CALPHAD surrogate ----------------------------------------
    def surrogate_train(self, opts: RsSurrogateOpts | None = None, **kw) -> dict:
        o = opts if opts is not None else surrogate_defaults(**kw)
        rep = (C.c_double * RS_SURR_NREPORT)()
        self._chk(lib.rs_surrogate_train(self._h, C.byref(o), rep, RS_SURR_NREPORT),
                  "rs_surrogate_train")
        return dict(zip(SURR_REPORT_KEYS, list(rep)))

    def surrogate_get(self) -> tuple[np.ndarray, np.ndarray]:
        P = np.empty(RS_SURR_NPARAM, dtype=np.float32)
        meta = np.empty(RS_SURR_NMETA, dtype=np.float64)
        self._chk(lib.rs_surrogate_get(self._h, _fptr(P), P.size, _dptr(meta), meta.size),
                  "rs_surrogate_get")
        return P, meta

    def surrogate_set(self, params: np.ndarray, meta: np.ndarray) -> None:
        P = np.ascontiguousarray(params, dtype=np.float32)
        m = np.ascontiguousarray(meta, dtype=np.float64)
        self._chk(lib.rs_surrogate_set(self._h, _fptr(P), P.size, _dptr(m), m.size),
                  "rs_surrogate_set")

    def surrogate_grad(self, X: np.ndarray, Yn: np.ndarray) -> tuple[np.ndarray, float]:
        x = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(Yn, dtype=np.float32)
        n = x.shape[0]
        g = np.empty(RS_SURR_NPARAM, dtype=np.float32)
        loss = C.c_double()
        self._chk(lib.rs_surrogate_grad(self._h, _fptr(x), _fptr(y), n, _fptr(g), C.byref(loss)),
                  "rs_surrogate_grad")
        return g, loss.value

    def surrogate_eval(self, T: np.ndarray, c: np.ndarray) -> np.ndarray:
        """(f_L, f_S, mu_L, mu_S, trusted) per (T, c) point, J/mol."""
        tc = np.ascontiguousarray(np.stack([np.ravel(T), np.ravel(c)], axis=1), dtype=np.float64)
        out = np.empty((tc.shape[0], 5), dtype=np.float64)
        self._chk(lib.rs_surrogate_eval(self._h, _dptr(tc), tc.shape[0], _dptr(out)),
                  "rs_surrogate_eval")
        return out

    def surrogate_time(self, n: int, reps: int = 5) -> float:
        """ms per evaluation of the network alone over n device-resident
        inputs (rs_surrogate_time)."""
        ms = C.c_double()
        self._chk(lib.rs_surrogate_time(self._h, n, reps, C.byref(ms)), "rs_surrogate_time")
        return ms.value

    def surrogate_stats(self, reset: bool = False) -> dict:
        v = (C.c_double * 5)()
        self._chk(lib.rs_surrogate_stats(self._h, v, 1 if reset else 0), "rs_surrogate_stats")
        return {"fallback": v[0], "evaluated": v[1], "fallback_frac": v[2],
                "rows_without_network": v[3], "rows": v[4]}

    # -- formats -------------------------------------------------------------
    def download_columns(self, field: int, i0: int, ncols: int) -> np.ndarray:
        out = np.empty((self.p.Ny, ncols), dtype=np.float64)
        self._chk(lib.rs_download_columns(self._h, field, i0, ncols, _dptr(out)),
                  "rs_download_columns")
        return out

    def save_state(self, path: str) -> None:
        self._chk(lib.rs_save_state(self._h, path.encode()), "rs_save_state")

    def load_state(self, path: str) -> None:
        self._chk(lib.rs_load_state(self._h, path.encode()), "rs_load_state")

    def timing(self, which: int) -> tuple[float, int]:
        ms = C.c_double()
        n = C.c_int64()
        self._chk(lib.rs_timing(self._h, which, C.byref(ms), C.byref(n)), "rs_timing")
        return ms.value, n.value


class _RunView(Context):
    """A run of a Batch: the Context interface on a context the batch owns."""

    def __init__(self, h, p: RsParams):  # noqa: D401  (no rs_create)
        self.p = p.copy()
        self._h = h
        self.k0 = derive(self.p)

    def close(self) -> None:  # owned by the batch
        self._h = None


class Batch:
    """Many independent runs stepped by one persistent launch per step,
    replayed from a captured HIP graph (rs_batch_* in include/rsolid.h)."""

    def __init__(self, params: list):
        arr = (RsParams * len(params))()
        for q, p in enumerate(params):
            C.memmove(C.byref(arr[q]), C.byref(p), C.sizeof(RsParams))
        h = C.c_void_p()
        rc = lib.rs_batch_create(arr, len(params), C.byref(h))
        if rc:
            raise RsError(rc, "rs_batch_create")
        self._h = h
        self.runs = []
        for q, p in enumerate(params):
            c = C.c_void_p()
            self._chk(lib.rs_batch_run(self._h, q, C.byref(c)), "rs_batch_run")
            self.runs.append(_RunView(c, p))

    def _chk(self, rc: int, what: str) -> None:
        if rc:
            raise RsError(rc, f"{what}: {lib.rs_batch_last_error(self._h).decode()}")

    def close(self) -> None:
        if self._h:
            for r in self.runs:
                r._h = None
            lib.rs_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def step(self, n: int = 1) -> None:
        self._chk(lib.rs_batch_step(self._h, n), "rs_batch_step")

    def surrogate_train(self, opts: RsSurrogateOpts | None = None, **kw) -> dict:
        o = opts if opts is not None else surrogate_defaults(**kw)
        rep = (C.c_double * RS_SURR_NREPORT)()
        self._chk(lib.rs_batch_surrogate_train(self._h, C.byref(o), rep, RS_SURR_NREPORT),
                  "rs_batch_surrogate_train")
        return dict(zip(SURR_REPORT_KEYS, list(rep)))

    def tune(self, steps: int = 10) -> dict:
        v = (C.c_double * 32)()
        self._chk(lib.rs_batch_tune(self._h, steps, v, 32), "rs_batch_tune")
        grids = {}
        q = 2
        while q + 1 < 32 and v[q] > 0:
            grids[int(v[q])] = int(v[q + 1])
            q += 2
        return {"ms_per_step": v[0], "grid_mult": int(v[1]), "band_h": grids}

    def info(self) -> dict:
        v = (C.c_double * 10)()
        self._chk(lib.rs_batch_info(self._h, v, 10), "rs_batch_info")
        keys = ("runs", "cells", "items", "blocks", "waves_per_block", "graph_steps", "last_ms",
                "last_steps", "grid_mult", "step_variant")
        return dict(zip(keys, list(v)))
