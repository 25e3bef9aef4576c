"""The opt-in finite-value sentinel (rs_set_finite_check), which has no
counterpart in the reference: a NaN planted in c stops rs_step at the first
checked step with RS_ERR_NONFINITE naming the field and the first cell, and
the check itself leaves the results bitwise unchanged.

A NaN in c at one cell makes that cell's driving force (Gl - Gs) NaN, so
after one step psi is non-finite at exactly that cell (its neighbours' psi
updates read only their own c); c itself turns NaN over the 3x3 block
through the face fluxes. psi is checked first, so the message names psi.
"""
import sys

import numpy as np
import pytest

from rsolid import RsError
from rsolid.abi import RS_ERR_NONFINITE, RS_FIELD_C, RS_FIELD_PSI, RS_PREC_FP64, RS_PREC_MIXED

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-sentinel]", *a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("prec", [RS_PREC_FP64, RS_PREC_MIXED])
def test_finite_check_stops_at_the_first_nan(rs, prec):
    p = rs.config_params(256, True, precision=prec)
    say("planted NaN in c, precision", prec)
    ctx = rs.Context(p)
    ctx.init_config(3)
    ctx.step(5)
    c = ctx.download(RS_FIELD_C)
    c[100, 37] = np.nan
    ctx.upload(RS_FIELD_C, c)
    ctx.finite_check(1)
    with pytest.raises(RsError) as e:
        ctx.step(3)
    msg = str(e.value)
    say(msg)
    assert e.value.code == RS_ERR_NONFINITE
    assert "after step 6: psi has 1 such cells, the first at (i=37, j=100)" in msg
    assert ctx.constants().step == 6
    psi = ctx.download(RS_FIELD_PSI)
    assert np.isnan(psi[100, 37]) and np.isfinite(np.delete(psi.ravel(), 100 * 256 + 37)).all()
    ctx.close()


@pytest.mark.parametrize("prec", [RS_PREC_FP64, RS_PREC_MIXED])
def test_finite_check_leaves_results_unchanged(rs, prec):
    p = rs.config_params(256, True, precision=prec)
    say("sentinel on vs off, precision", prec)
    out = []
    for every in (0, 7):
        ctx = rs.Context(p)
        ctx.init_config(4)
        ctx.finite_check(every)
        ctx.step(300)
        out.append((ctx.download(RS_FIELD_PSI), ctx.download(RS_FIELD_C)))
        ctx.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
