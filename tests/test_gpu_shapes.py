"""Ragged and minimal grids on every path against the FP64 oracle: the
smallest grid the library accepts (8 x 8), widths that leave a partial
62-column tile (63, 130, 191), heights that leave a partial band and a
partial 10-row latent-heat block (9, 22, 75, 97: rows beyond 10 (NyT - 2)
get no latent heat, fun.cu:1526), frozen and coupled T. 300 steps each.

Tolerances: FP64 path 1e-10 relative per field (the run parity bound);
FP32 path psi 2e-6 absolute, c 1e-7 relative, solute-mass trajectory
within 1e-10 plus twice the largest relative c deviation (the mass is the
sum of c: on a grid of a few dozen cells, where every cell is a boundary
cell, the c deviation itself sets the mass difference); the surrogate
kernel with an untrained network (every cell fails the trust check and
takes the exact routine) bitwise equal to the FP32 path; a trained network
within the surrogate step test's bounds.
"""
import sys

import numpy as np
import pytest

from oracle import mlp
from rsolid.abi import (RS_ACT_FP16, RS_FIELD_C, RS_FIELD_PSI, RS_FIELD_T, RS_PREC_MIXED,
                        RS_PREC_SURROGATE)

pytestmark = pytest.mark.gpu

STEPS = 300
# (N of the configuration preset, Nx, Ny, coupled T, i_tip_target): the
# preset's coarse-T sample is sized for N x N, so the smaller window stays
# inside it
SHAPES = {
    "8x8_frozen": (64, 8, 8, False, 3),
    "64x9_frozen": (64, 64, 9, False, 25),
    "63x22_coupled": (64, 63, 22, True, 25),
    "130x97_coupled": (130, 130, 97, True, 25),
    "191x75_frozen": (192, 191, 75, False, 25),
}


def say(*a):
    print("[gpu-shapes]", *a, flush=True)
    sys.stdout.flush()


def _params(rs, name, precision=0):
    N, Nx, Ny, tfc, itip = SHAPES[name]
    p = rs.config_params(N, tfc, precision=precision)
    p.Nx, p.Ny, p.i_tip_target = Nx, Ny, itip
    return p


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def oracle_runs(oracle_mod):
    """The oracle's fields after STEPS steps, per shape (computed once)."""
    out = {}

    def get(rs, name):
        if name not in out:
            p = _params(rs, name)
            st = oracle_mod.State(p)
            st.init_config(3)
            m0 = st.solute_mass()
            st.step(STEPS)
            out[name] = {"psi": st.get(RS_FIELD_PSI), "c": st.get(RS_FIELD_C),
                         "T": st.get(RS_FIELD_T) if p.ifTFC else None,
                         "drift": (st.solute_mass() - m0) / m0, "k": st.constants(),
                         "tip": st.tip_column()}
            st.close()
        return out[name]
    return get


@pytest.mark.parametrize("name", list(SHAPES))
def test_fp64_shape_matches_oracle(rs, oracle_runs, name):
    o = oracle_runs(rs, name)
    say("FP64", name)
    ctx = rs.Context(_params(rs, name))
    ctx.init_config(3)
    ctx.step(STEPS)
    for f, key in ((RS_FIELD_PSI, "psi"), (RS_FIELD_C, "c")):
        e = _rel(ctx.download(f), o[key])
        say(name, key, "rel err", e)
        assert e <= 1e-10, (name, key, e)
    if o["T"] is not None:
        assert _rel(ctx.download(RS_FIELD_T), o["T"]) <= 1e-10
    k = ctx.constants()
    for f in ("step", "i_offset_history", "i_left_end"):
        assert getattr(k, f) == getattr(o["k"], f), f
    assert ctx.tip_column() == o["tip"]
    ctx.close()


@pytest.mark.parametrize("name", list(SHAPES))
def test_mixed_shape_tracks_oracle(rs, oracle_runs, name):
    o = oracle_runs(rs, name)
    say("FP32", name)
    ctx = rs.Context(_params(rs, name, RS_PREC_MIXED))
    ctx.init_config(3)
    m0 = ctx.solute_mass()
    ctx.step(STEPS)
    epsi = float(np.max(np.abs(ctx.download(RS_FIELD_PSI) - o["psi"])))
    ec = _rel(ctx.download(RS_FIELD_C), o["c"])
    dm = (ctx.solute_mass() - m0) / m0 - o["drift"]
    say(name, "psi Linf", epsi, "c rel", ec, "mass-drift diff", dm)
    assert epsi <= 2e-6
    assert ec <= 1e-7
    assert abs(dm) <= 1e-10 + 2.0 * ec
    assert ctx.tip_column() == o["tip"]
    ctx.close()


@pytest.mark.parametrize("name", ["8x8_frozen", "63x22_coupled", "130x97_coupled",
                                  "191x75_frozen"])
def test_surrogate_all_fallback_equals_fp32_path(rs, name):
    say("surrogate untrained vs FP32 path", name)
    meta = np.array([1930.0, 20.0, 0, 0, 0, 0, 1, 1, 1, 1, 1910.0, 1950.0, 1e-4,
                     float(RS_ACT_FP16)])
    out = []
    for prec in (RS_PREC_MIXED, RS_PREC_SURROGATE):
        ctx = rs.Context(_params(rs, name, prec))
        ctx.init_config(3)
        if prec == RS_PREC_SURROGATE:
            ctx.surrogate_set(mlp.init_params(3), meta)
            ctx.surrogate_stats(reset=True)
        ctx.step(STEPS)
        if prec == RS_PREC_SURROGATE:
            st = ctx.surrogate_stats()
            say(name, "fallback", st)
            assert st["fallback_frac"] == 1.0
        out.append((ctx.download(RS_FIELD_PSI), ctx.download(RS_FIELD_C)))
        ctx.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


def test_surrogate_trained_ragged_shape(rs, oracle_runs):
    name = "130x97_coupled"
    o = oracle_runs(rs, name)
    say("surrogate trained", name)
    ctx = rs.Context(_params(rs, name, RS_PREC_SURROGATE))
    ctx.init_config(3)
    # default training: 8 epochs leave the networks too few trusted bins
    # to serve this run's cells
    rep = ctx.surrogate_train()
    say("trusted bins", rep["trusted_frac"], "window", rep["T_lo"], rep["T_hi"])
    ctx.surrogate_stats(reset=True)
    ctx.step(STEPS)
    st = ctx.surrogate_stats()
    epsi = float(np.max(np.abs(ctx.download(RS_FIELD_PSI) - o["psi"])))
    ec = _rel(ctx.download(RS_FIELD_C), o["c"])
    say(name, "fallback", st, "psi Linf", epsi, "c rel", ec)
    assert st["fallback_frac"] < 1.0
    assert epsi <= 2e-2
    assert ec <= 2e-3
    assert ctx.tip_column() == o["tip"]
    ctx.close()
