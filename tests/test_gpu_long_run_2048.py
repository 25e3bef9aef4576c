"""The 2048^2 coupled-T Ti-Nb configuration at its full length: 20 000 steps
of the FP64 GPU path from the seeded planar front (init_config(21)) against
the CPU oracle's state at step 20 000, held in
tests/golden/coupled_2048_20000.npz (written by make_coupled_2048_20000.py
beside it; the oracle needs ~2 h on 8 cores for this run, the GPU a few
seconds). The fixture keeps 8 full rows and a 12-column window around the
tip column of psi and c, every 8th coarse-T row, integral figures and the
tip row; tolerance 1e-10 relative to each field's largest magnitude for
the FP64 path (the fitted tip position within 1e-6 cells). The FP32 path is held to the precision targets over the same
run (tip advance within 1%, solute mass within 1e-10) and to stated field
bounds.
"""
import os
import sys

import numpy as np
import pytest

from rsolid.abi import RS_FIELD_C, RS_FIELD_PSI, RS_FIELD_T, RS_PREC_MIXED, RS_PREC_SURROGATE

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                       "coupled_2048_20000.npz")
REL = 1e-10


def say(*a):
    print("[gpu-long-run]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def golden():
    if not os.path.exists(FIXTURE):
        pytest.skip("tests/golden/coupled_2048_20000.npz not generated")
    with np.load(FIXTURE, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_fp64_path_2048_coupled_20000_steps_vs_oracle(rs, golden):
    g = golden
    N, seed, steps = int(g["N"]), int(g["seed"]), int(g["steps"])
    p = rs.config_params(N, True)
    say(f"{N}^2 coupled T, FP64, {steps} steps, seed {seed}")
    ctx = rs.Context(p)
    ctx.init_config(seed)
    ctx.step(steps)
    k = ctx.constants()
    assert k.step == int(g["step"])
    assert k.i_left_end == int(g["i_left_end"])
    assert k.i_offset_history == int(g["i_offset_history"])
    assert k.l_offset_Vp == pytest.approx(float(g["l_offset_Vp"]), rel=1e-14)
    psi = ctx.download(RS_FIELD_PSI)
    c = ctx.download(RS_FIELD_C)
    T = ctx.download(RS_FIELD_T)
    rows = g["rows"]
    i0 = int(g["win_i0"])
    w = g["win_psi"].shape[1]
    checks = (
        ("psi rows", psi[rows, :], g["rows_psi"], float(g["max_abs_psi"])),
        ("c rows", c[rows, :], g["rows_c"], float(g["max_abs_c"])),
        ("psi tip window", psi[:, i0:i0 + w], g["win_psi"], float(g["max_abs_psi"])),
        ("c tip window", c[:, i0:i0 + w], g["win_c"], float(g["max_abs_c"])),
        ("T rows", T[::8, :], g["T_rows"], float(np.max(np.abs(g["T_rows"])))),
    )
    for name, got, ref, scale in checks:
        assert got.shape == ref.shape, name
        e = float(np.max(np.abs(got - ref))) / scale
        say(name, "max rel err", e)
        assert e <= REL, (name, e)
    say("max |psi|", float(np.max(np.abs(psi))), "oracle", float(g["max_abs_psi"]))
    assert float(np.max(np.abs(psi))) == pytest.approx(float(g["max_abs_psi"]), rel=REL)
    for name, got, ref in (("sumsq psi", float(np.sum(psi * psi)), float(g["sumsq_psi"])),
                           ("sumsq c", float(np.sum(c * c)), float(g["sumsq_c"])),
                           ("sum c", float(np.sum(c)), float(g["sum_c"]))):
        say(name, got, "oracle", ref)
        assert got == pytest.approx(ref, rel=REL), name
    m = ctx.solute_mass()
    say("solute mass", m, "oracle", float(g["solute_mass"]))
    assert m == pytest.approx(float(g["solute_mass"]), rel=1e-12)
    assert ctx.tip_column() == int(g["tip_column"])
    t = ctx.tip()
    assert (t["i"], t["j"]) == (int(g["tip_i"]), int(g["tip_j"]))
    # the tip's quartic / parabola fits (normal equations in powers of psi up
    # to the 8th) amplify the ~1e-13 field differences: 1e-6 cells
    say("tip x", t["x"], "oracle", float(g["tip_x"]), "y", t["y"], "oracle", float(g["tip_y"]))
    assert t["x"] == pytest.approx(float(g["tip_x"]), abs=1e-6)
    assert t["y"] == pytest.approx(float(g["tip_y"]), abs=1e-6)
    ctx.close()


def test_fp32_path_2048_coupled_20000_steps(rs, golden):
    """The FP32 path (FP32 psi/phi and fine T, FP64 c) over the same 20 000
    steps against the oracle's state: the tip advance within 1% of the
    oracle's and the solute mass within 1e-10 (the precision targets), the
    psi / c / T deviations reported and held to 1e-4 absolute (psi on the
    sampled rows and the tip window), 1e-6 relative (c) and 1e-7 relative
    (T)."""
    g = golden
    N, seed, steps = int(g["N"]), int(g["seed"]), int(g["steps"])
    p = rs.config_params(N, True, precision=RS_PREC_MIXED)
    say(f"{N}^2 coupled T, FP32 path, {steps} steps, seed {seed}")
    ctx = rs.Context(p)
    ctx.init_config(seed)
    x0 = ctx.constants().i_offset_history + ctx.tip()["x"]
    ctx.step(steps)
    k = ctx.constants()
    assert k.step == int(g["step"])
    psi = ctx.download(RS_FIELD_PSI)
    c = ctx.download(RS_FIELD_C)
    T = ctx.download(RS_FIELD_T)
    rows = g["rows"]
    i0 = int(g["win_i0"])
    w = g["win_psi"].shape[1]
    epsi = max(float(np.max(np.abs(psi[rows, :] - g["rows_psi"]))),
               float(np.max(np.abs(psi[:, i0:i0 + w] - g["win_psi"]))))
    ec = max(float(np.max(np.abs(c[rows, :] - g["rows_c"]))),
             float(np.max(np.abs(c[:, i0:i0 + w] - g["win_c"])))) / float(g["max_abs_c"])
    eT = float(np.max(np.abs(T[::8, :] - g["T_rows"]))) / float(np.max(np.abs(g["T_rows"])))
    m = ctx.solute_mass()
    dm = m / float(g["solute_mass"]) - 1.0
    t = ctx.tip()
    x1 = k.i_offset_history + t["x"]
    xo = int(g["i_offset_history"]) + float(g["tip_x"])
    say("psi Linf (sampled)", epsi, "c rel", ec, "T rel", eT, "mass rel", dm,
        "tip advance", x1 - x0, "oracle", xo - x0)
    assert abs((x1 - x0) - (xo - x0)) <= 0.01 * abs(xo - x0)
    assert abs(dm) <= 1e-10
    assert epsi <= 1e-4
    assert ec <= 1e-6
    assert eT <= 1e-7
    ctx.close()


def test_surrogate_path_2048_coupled_20000_steps(rs, golden):
    """The surrogate path (network trained with the default options) over
    the same 20 000 steps: the tip advance within 1% of the oracle's (the
    tip-velocity target) and the fields within the surrogate step test's
    bounds (psi 2e-2 absolute, c 2e-3 relative). The path's own scheme
    conserves solute: its ledger (pull-backs and boundary faces booked)
    leaves at most 1e-13 of the mass unaccounted. Its raw solute-mass
    difference from the oracle is held to 1e-7, a regression bound: the
    pull-backs and the edge fluxes of the reference's ghost scheme respond
    to the network's field perturbations (2.1e-8 measured)."""
    g = golden
    N, seed, steps = int(g["N"]), int(g["seed"]), int(g["steps"])
    p = rs.config_params(N, True, precision=RS_PREC_SURROGATE)
    say(f"{N}^2 coupled T, surrogate path, {steps} steps, seed {seed}")
    ctx = rs.Context(p)
    ctx.init_config(seed)
    rep = ctx.surrogate_train()
    say("trusted bins", rep["trusted_frac"], "window", rep["T_lo"], rep["T_hi"])
    x0 = ctx.constants().i_offset_history + ctx.tip()["x"]
    ctx.surrogate_stats(reset=True)
    ctx.solute_ledger(reset=True)
    ctx.step(steps)
    st = ctx.surrogate_stats()
    led = ctx.solute_ledger()
    k = ctx.constants()
    psi = ctx.download(RS_FIELD_PSI)
    c = ctx.download(RS_FIELD_C)
    rows = g["rows"]
    i0 = int(g["win_i0"])
    w = g["win_psi"].shape[1]
    epsi = max(float(np.max(np.abs(psi[rows, :] - g["rows_psi"]))),
               float(np.max(np.abs(psi[:, i0:i0 + w] - g["win_psi"]))))
    ec = max(float(np.max(np.abs(c[rows, :] - g["rows_c"]))),
             float(np.max(np.abs(c[:, i0:i0 + w] - g["win_c"])))) / float(g["max_abs_c"])
    dm = ctx.solute_mass() / float(g["solute_mass"]) - 1.0
    x1 = k.i_offset_history + ctx.tip()["x"]
    xo = int(g["i_offset_history"]) + float(g["tip_x"])
    say("fallback", st["fallback_frac"], "psi Linf (sampled)", epsi, "c rel", ec, "mass rel", dm,
        "unaccounted", led["unaccounted"] / led["mass"], "tip advance", x1 - x0,
        "oracle", xo - x0)
    assert abs((x1 - x0) - (xo - x0)) <= 0.01 * abs(xo - x0)
    assert epsi <= 2e-2
    assert ec <= 2e-3
    assert abs(led["unaccounted"]) <= 1e-13 * led["mass"]
    assert abs(dm) <= 1e-7
    ctx.close()
