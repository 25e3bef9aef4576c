"""Host-side error paths of the C ABI on a live context: each bad call
returns its RS_ERR_* code with a message in rs_last_error, launches nothing
and leaves the context's state as it was, so the run can go on."""
import ctypes as C
import sys

import numpy as np
import pytest

from rsolid import RsError
from rsolid.abi import (RS_ERR_ARG, RS_ERR_DOMAIN, RS_ERR_STATE, RS_FIELD_C, RS_FIELD_PSI,
                        RS_FIELD_T, RS_PREC_SURROGATE)

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-errors]", *a, flush=True)
    sys.stdout.flush()


def _state(ctx):
    k = ctx.constants()
    return ([ctx.download(f) for f in (RS_FIELD_PSI, RS_FIELD_C, RS_FIELD_T)],
            (k.step, k.i_offset, k.i_offset_history, k.i_left_end))


def _same(a, b):
    assert a[1] == b[1]
    for x, y in zip(a[0], b[0]):
        assert np.array_equal(x, y)


def test_argument_errors_leave_the_context_usable(rs):
    p = rs.config_params(256, True)
    say("argument errors")
    ctx = rs.Context(p)
    ctx.init_config(3)
    ctx.step(5)
    before = _state(ctx)
    with pytest.raises(RsError) as e:
        ctx.upload(RS_FIELD_PSI, np.zeros(p.Nx * p.Ny - 1))
    assert e.value.code == RS_ERR_ARG and "size mismatch" in str(e.value)
    out = np.empty(p.Nx * p.Ny)
    assert rs.lib.rs_download(ctx._h, 99, out.ctypes.data_as(C.POINTER(C.c_double)),
                              out.size) == RS_ERR_ARG
    assert rs.lib.rs_step(ctx._h, -1) == RS_ERR_ARG
    k = ctx.constants()
    k.NxT_all += 64
    with pytest.raises(RsError) as e:
        ctx.set_constants(k)
    assert e.value.code == RS_ERR_ARG and "geometry" in str(e.value)
    _same(before, _state(ctx))
    ctx.step(10)
    assert ctx.constants().step == 15
    assert np.isfinite(ctx.download(RS_FIELD_PSI)).all()
    ctx.close()


@pytest.mark.parametrize("off", [100000, -100000])
def test_pull_back_off_the_temperature_sample_is_rejected(rs, off):
    """A shift that would move the fine window off the coarse T sample
    returns RS_ERR_DOMAIN before anything is shifted; a valid shift then
    works as usual."""
    p = rs.config_params(256, True)
    say("pull-back", off)
    ctx = rs.Context(p)
    ctx.init_config(3)
    ctx.step(5)
    before = _state(ctx)
    with pytest.raises(RsError) as e:
        ctx.pull_back(off)
    say(str(e.value))
    assert e.value.code == RS_ERR_DOMAIN and "temperature sample" in str(e.value)
    _same(before, _state(ctx))
    ctx.pull_back(3)
    k = ctx.constants()
    assert k.i_offset == 3 and k.i_offset_history == before[1][2] + 3
    psi0, psi1 = before[0][0], ctx.download(RS_FIELD_PSI)
    assert np.array_equal(psi1[:, 1:p.Nx - 5], psi0[:, 4:p.Nx - 2])
    ctx.close()


def _bad_constants(k0, rule):
    k = type(k0).from_buffer_copy(k0)
    if rule == "NyT":
        k.NyT += 1
    elif rule == "NxT":
        k.NxT -= 1
    elif rule == "step_pull_back":
        k.step_pull_back = 0
    elif rule == "step_tip":
        k.step_tip = -1
    elif rule == "step_fields":
        k.step_fields = 0
    elif rule == "step_check_point":
        k.step_check_point = 0
    elif rule == "ratio_dt_dtT":
        k.ratio_dt_dtT = 0
    elif rule == "dt":
        k.dt = float("nan")
    elif rule == "dtT":
        k.dtT = -1.0
    elif rule == "step":
        k.step = -3
    elif rule == "i_left_end":
        k.i_left_end = 0
    elif rule == "i_left_end_interp":
        k.i_left_end_interp = k.NxT * 10
    return k


RULES = ["NyT", "NxT", "step_pull_back", "step_tip", "step_fields", "step_check_point",
         "ratio_dt_dtT", "dt", "dtT", "step", "i_left_end", "i_left_end_interp"]


def test_set_constants_rejects_each_broken_rule(rs):
    """rs_set_constants checks every rule before it changes anything: the
    coarse grid the buffers were sized for, positive cadences (the loop
    takes step % cadence), ratio_dt_dtT >= 1, positive finite dt and dtT, a
    non-negative step and a fine window inside the temperature sample. Each
    rejection is host-side (no kernel runs) and leaves the state as it was."""
    p = rs.config_params(256, True)
    say("set_constants rules")
    ctx = rs.Context(p)
    ctx.init_config(3)
    ctx.step(5)
    before = _state(ctx)
    k0 = ctx.constants()
    for rule in RULES:
        with pytest.raises(RsError) as e:
            ctx.set_constants(_bad_constants(k0, rule))
        say(rule, "->", str(e.value))
        assert e.value.code == RS_ERR_ARG, rule
        _same(before, _state(ctx))
    ctx.set_constants(k0)
    ctx.step(3)
    assert ctx.constants().step == 8
    ctx.close()


def test_surrogate_step_without_weights_is_a_state_error(rs):
    p = rs.config_params(256, True, precision=RS_PREC_SURROGATE)
    say("surrogate without weights")
    ctx = rs.Context(p)
    ctx.init_config(3)
    with pytest.raises(RsError) as e:
        ctx.step(1)
    assert e.value.code == RS_ERR_STATE and "weights" in str(e.value)
    ctx.close()
