"""C-ABI checks that need no GPU: the library loads, exports every symbol
include/rsolid.h declares, its struct layouts match the Python mirror, and
its host-side parameter code agrees with the oracle's restatement."""
import ctypes as C
import os
import re

import pytest

from rsolid.abi import (EXPORTED, RS_ABI_VERSION, TP_EXPORTED, RsConstants, RsParams, RsTipRow,
                        TpParams)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "rsolid.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(rs_[a-z_0-9]+)\s*\(", txt)))


def test_every_declared_symbol_is_exported(rs):
    decl = _declared()
    assert sorted(EXPORTED) == decl
    for name in decl:
        assert hasattr(rs.lib, name), name


def test_peritectic_header_symbols_are_exported(rs):
    """include/rs_peritectic.h (the 3D Ti-Ag stepper): every declared
    function is in the library and in the Python mirror; tp_params has the
    mirror's layout."""
    import rsolid.peritectic  # noqa: F401  (binds the signatures)
    txt = open(os.path.join(ROOT, "include", "rs_peritectic.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    decl = sorted(set(re.findall(r"\b(tp_[a-z_0-9]+)\s*\(", txt)))
    assert sorted(TP_EXPORTED) == decl
    for name in decl:
        assert hasattr(rs.lib, name), name
    n = C.c_int32()
    assert rs.lib.tp_abi_info(C.byref(n)) == 0
    assert n.value == C.sizeof(TpParams)


def test_struct_layouts(rs):
    s = (C.c_int32 * 4)()
    assert rs.lib.rs_abi_info(s) == 0
    assert s[0] == C.sizeof(RsParams)
    assert s[1] == C.sizeof(RsConstants)
    assert s[2] == C.sizeof(RsTipRow)
    assert s[3] == RS_ABI_VERSION


@pytest.mark.parametrize("maker", ["ref768", "ref512", "cfg256", "cfg4096t", "cfg2048t"])
def test_library_derive_matches_oracle(rs, oracle_mod, maker):
    p = {"ref768": lambda: rs.reference_params(768, 832),
         "ref512": lambda: rs.reference_params(512, 272),
         "cfg256": lambda: rs.config_params(256, False),
         "cfg4096t": lambda: rs.config_params(4096, True),
         "cfg2048t": lambda: rs.config_params(2048, True)}[maker]()
    a = rs.derive(p).as_dict()
    b = oracle_mod.derive(p).as_dict()
    assert a == b


def test_config_preset_sits_on_the_liquidus(rs, oracle_mod):
    p = rs.config_params(256, False)
    st, cl, cs = oracle_mod.common_tangent(p, p.Tmelt, 0.9, 0.88)
    assert st == 0
    assert cl == pytest.approx(0.90, abs=1e-9)
    assert cs == pytest.approx(p.c_sol0, abs=1e-9)
    # the SGTE breakpoint 1941 K lies inside the frozen-T window of the run
    assert p.Tmelt - p.deltaT0 < 1941.0 < p.Tmelt


def test_bad_params_rejected_without_gpu(rs):
    p = rs.reference_params(768, 832)
    p.ratio_dxT_dx = 7
    h = C.c_void_p()
    assert rs.lib.rs_create(C.byref(p), C.byref(h)) == rs.RS_ERR_ARG
    assert rs.lib.rs_params_reference(C.byref(p), 100, 100) == rs.RS_ERR_ARG


def _cfg(rs, **kw):
    p = rs.config_params(256, True)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


@pytest.mark.parametrize("case", ["grid_7", "coupled_Ny_21", "ratio_5", "precision_9",
                                  "tip_outside", "antitrap_k0", "mixed_dx_fraction", "S_0"])
def test_create_validates_every_parameter_rule(rs, case):
    """rs_create rejects each parameter set its validation rules out with
    RS_ERR_ARG before any device call (so this runs without a GPU)."""
    make = {
        "grid_7": lambda: _cfg(rs, Nx=7),
        "coupled_Ny_21": lambda: _cfg(rs, Ny=21),
        "ratio_5": lambda: _cfg(rs, ratio_dxT_dx=5),
        "precision_9": lambda: _cfg(rs, precision=9),
        "tip_outside": lambda: _cfg(rs, i_tip_target=255),
        "antitrap_k0": lambda: _cfg(rs, at_coef=0.35, k_part=0.0),
        "mixed_dx_fraction": lambda: _cfg(rs, precision=rs.RS_PREC_MIXED, dx=0.75),
        "S_0": lambda: _cfg(rs, S=0),
    }
    h = C.c_void_p()
    assert rs.lib.rs_create(C.byref(make[case]()), C.byref(h)) == rs.RS_ERR_ARG
    assert not h.value


def test_create_control_config_passes_validation(rs):
    """The unmodified configuration the rejection cases start from passes
    validation: without a GPU rs_create then fails at the device
    (RS_ERR_HIP), with one it succeeds."""
    h = C.c_void_p()
    rc = rs.lib.rs_create(C.byref(_cfg(rs)), C.byref(h))
    assert rc != rs.RS_ERR_ARG
    if rc == rs.RS_OK:
        rs.lib.rs_destroy(h)


def test_create_failure_reason_is_kept(rs):
    """rs_last_error(NULL) gives the reason of the calling thread's last
    rs_create failure: the broken rule."""
    h = C.c_void_p()
    assert rs.lib.rs_create(C.byref(_cfg(rs, ratio_dxT_dx=5)), C.byref(h)) == rs.RS_ERR_ARG
    assert b"ratio_dxT_dx must be 10" in rs.lib.rs_last_error(None)
    assert rs.lib.rs_create(C.byref(_cfg(rs, Ny=21)), C.byref(h)) == rs.RS_ERR_ARG
    assert b"coupled T needs Ny >= 22" in rs.lib.rs_last_error(None)
