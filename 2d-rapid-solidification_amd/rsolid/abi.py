"""ctypes mirror of include/rsolid.h (structs, enums, status codes).

Field order and types must match the C header exactly; tests check the
struct sizes against the values the library reports.
"""
from __future__ import annotations

import ctypes as C

RS_OK = 0
RS_ERR_ARG = -1
RS_ERR_HIP = -2
RS_ERR_DOMAIN = -3
RS_ERR_NOMEM = -4
RS_ERR_STATE = -5
RS_ERR_NONFINITE = -6

RS_PREC_FP64 = 0
RS_PREC_MIXED = 1
RS_PREC_SURROGATE = 2
RS_ABI_VERSION = 2  # include/rsolid.h: the layouts below

RS_ACT_BF16, RS_ACT_BF16_SPLIT, RS_ACT_FP16, RS_ACT_FP16_PACKED = 0, 1, 2, 3
RS_SURR_NPARAM = 4612
RS_SURR_NMETA = 14
RS_SURR_NREPORT = 12

RS_FIELD_C = 0
RS_FIELD_PSI = 1
RS_FIELD_PHI = 2
RS_FIELD_T = 3
RS_FIELD_TFINE = 4

_I = C.c_int32
_D = C.c_double


class RsParams(C.Structure):
    _fields_ = [
        ("Nx", _I), ("Ny", _I), ("i_tip_target", _I), ("S", _I), ("ratio_dxT_dx", _I),
        ("ifTFC", _I), ("if_Vp_ramp", _I), ("dt_cap", _I), ("num_fields_output", _I),
        ("num_tip_output", _I), ("num_check_point_output", _I), ("precision", _I),
        ("reserved0", _I),
        ("W0", _D), ("Gamma", _D), ("muk0", _D), ("Tmelt", _D), ("deltaT0", _D), ("Dl", _D),
        ("DT", _D), ("GG", _D), ("Vp", _D), ("Vp_ramp_rate", _D), ("eps1", _D), ("epk1", _D),
        ("cinf", _D), ("latentHeat", _D), ("num_moles", _D), ("v0", _D), ("cp", _D), ("RR", _D),
        ("dx", _D), ("AA", _D), ("TCOLD", _D), ("L0_bcc", _D), ("L0_liq", _D), ("L1_bcc", _D),
        ("L1_liq", _D), ("dt_coef", _D), ("dtT_coef", _D), ("length_sample_um", _D),
        ("lap_w_axis", _D), ("lap_w_diag", _D), ("q_solid", _D), ("total_time", _D),
        ("delta_initial", _D), ("phi_interface", _D), ("c_liq0", _D), ("c_sol0", _D),
        ("at_coef", _D), ("V_D", _D), ("k_part", _D), ("alpha0", _D), ("beta0", _D),
        ("gamma0", _D),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}

    def copy(self) -> "RsParams":
        q = RsParams()
        C.memmove(C.byref(q), C.byref(self), C.sizeof(RsParams))
        return q


class RsConstants(C.Structure):
    _fields_ = [
        ("step", C.c_int64), ("initial_step", C.c_int64), ("total_step", C.c_int64),
        ("step_fields", _I), ("step_tip", _I), ("step_check_point", _I), ("step_pull_back", _I),
        ("if_begin_tip", _I), ("if_begin_field", _I), ("i_offset", _I),
        ("i_offset_history_initial", _I), ("i_offset_history", _I), ("i_ignore", _I),
        ("dt", _D), ("x_tip_history", _D), ("delta_initial", _D), ("delta", _D),
        ("l_offset_Vp_initial", _D), ("l_offset_Vp", _D),
        ("i_left_end_initial", _I), ("i_left_end", _I),
        ("dtT", _D), ("ratio_dt_dtT", _I), ("NxT", _I), ("NyT", _I), ("NxT_all", _I),
        ("NyT_all", _I), ("i_left_end_interp", _I),
        ("Tcold", _D), ("Thot", _D), ("length_sample", _D),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class RsTipRow(C.Structure):
    _fields_ = [
        ("step", C.c_int64), ("t_us", _D), ("delta", _D), ("T_tip", _D), ("V_tip", _D),
        ("Vp", _D), ("R", _D), ("i", _I), ("j", _I), ("x", _D), ("y", _D),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class RsSurrogateOpts(C.Structure):
    _fields_ = [
        ("n_uniform", C.c_int64), ("n_near", C.c_int64), ("seed", C.c_uint64),
        ("epochs", _I), ("batch", _I), ("T_lo", _D), ("T_hi", _D), ("lr", _D),
        ("lr_final", _D), ("residual_tol", _D), ("n_curve", _I), ("act_format", _I),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


# every extern "C" symbol include/rsolid.h declares
EXPORTED = [
    "rs_abi_info", "rs_params_reference", "rs_params_config", "rs_derive", "rs_create", "rs_destroy",
    "rs_last_error", "rs_device_name", "rs_init_reference", "rs_init_config", "rs_upload",
    "rs_download", "rs_get_constants", "rs_set_constants", "rs_step", "rs_tip_column",
    "rs_pull_back", "rs_solute_mass", "rs_tip", "rs_sync", "rs_timing_enable", "rs_timing",
    "rs_set_finite_check",
    "rs_calibrate_copy", "rs_context_info", "rs_surrogate_defaults", "rs_surrogate_train",
    "rs_surrogate_get", "rs_surrogate_set", "rs_surrogate_grad", "rs_surrogate_eval",
    "rs_surrogate_time",
    "rs_surrogate_stats", "rs_batch_create", "rs_batch_destroy", "rs_batch_last_error",
    "rs_batch_run", "rs_batch_step", "rs_batch_surrogate_train", "rs_batch_tune", "rs_batch_info",
    "rs_write_field_txt", "rs_write_tip_row", "rs_append_offset", "rs_append_history",
    "rs_write_checkpoint", "rs_read_checkpoint", "rs_write_parameters", "rs_read_parameters",
    "rs_download_columns", "rs_save_state", "rs_load_state", "rs_solute_mass_noalloc",
    "rs_solute_ledger",
]


# ---- include/rs_peritectic.h (3D peritectic Ti-Ag stepper) ----------------
TP_FIELD_PHI1, TP_FIELD_PHI2, TP_FIELD_C, TP_FIELD_MU = 0, 1, 2, 3


class TpParams(C.Structure):
    _fields_ = [
        ("Nx", _I), ("Ny", _I), ("Nz", _I), ("zcut_from_top", _I),
        ("DT", _D), ("DXin", _D), ("DYin", _D), ("DZin", _D),
        ("a1", _D), ("a2", _D), ("a3", _D), ("b", _D), ("lambda_", _D), ("alpha", _D),
        ("Da", _D), ("Db", _D), ("A1", _D), ("A2", _D), ("B11", _D), ("B21", _D), ("B12", _D),
        ("B22", _D), ("T", _D), ("R", _D), ("theta", _D), ("ncaps", _I), ("reserved", _I),
        ("cap_x", _D * 4), ("cap_y", _D * 4),
    ]

    def copy(self) -> "TpParams":
        q = TpParams()
        C.memmove(C.byref(q), C.byref(self), C.sizeof(TpParams))
        return q


TP_EXPORTED = [
    "tp_abi_info", "tp_params_reference", "tp_derived", "tp_create", "tp_destroy", "tp_last_error",
    "tp_init_reference", "tp_upload", "tp_download", "tp_step", "tp_get_step", "tp_set_step",
    "tp_report", "tp_sync", "tp_timing", "tp_write_cross_section", "tp_write_break_point",
    "tp_read_break_point",
]
