"""The 2048^2 coupled-T Ti-Nb configuration at its full length with the
solute-trapping kinetics on: 20 000 steps from the seeded planar front
(init_config(21)) with the antitrapping current and the trapping factor
(a_t = 1/(2 sqrt 2), V_p / V_D = 2, the fast end of the [0.1, 2] range)
against the CPU oracle's state at step 20 000, held in
tests/golden/antitrap_2048_20000_vd2.npz (written by
make_antitrap_2048_20000.py beside it; the oracle needs ~2 h on 8 cores,
the GPU a few seconds). The fixture is reduced like coupled_2048_20000.npz:
8 full rows and a 12-column window around the tip column of psi and c,
every 8th coarse-T row, integral figures and the tip row.

Tolerances: the FP64 path 1e-10 relative to each field's largest magnitude
(the fitted tip position within 1e-6 cells, the solute mass 1e-12); the
FP32 path (FP32 psi/phi and fine T, FP64 c, the correction on FP32 psi) the
tip advance within 1% of the oracle's and the solute mass within 1e-10 (the
precision targets), psi within 1e-4 absolute on the sampled rows and the
tip window, c 1e-6 and T 1e-7 relative (measured: 2.5e-6, 4.6e-8, 2.4e-9).
"""
import os
import sys

import numpy as np
import pytest

from rsolid.abi import RS_FIELD_C, RS_FIELD_PSI, RS_FIELD_T, RS_PREC_MIXED

pytestmark = pytest.mark.gpu

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                       "antitrap_2048_20000_vd2.npz")
REL = 1e-10


def say(*a):
    print("[gpu-antitrap-2048]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def golden():
    if not os.path.exists(FIXTURE):
        pytest.skip("tests/golden/antitrap_2048_20000_vd2.npz not generated")
    with np.load(FIXTURE, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _params(rs, g, precision=0):
    p = rs.config_params(int(g["N"]), True, precision=precision)
    p.at_coef = float(g["at_coef"])
    p.V_D = float(g["V_D"])
    return p


def _sampled(g, psi, c, T):
    rows = g["rows"]
    i0 = int(g["win_i0"])
    w = g["win_psi"].shape[1]
    return (
        ("psi rows", psi[rows, :], g["rows_psi"], float(g["max_abs_psi"])),
        ("c rows", c[rows, :], g["rows_c"], float(g["max_abs_c"])),
        ("psi tip window", psi[:, i0:i0 + w], g["win_psi"], float(g["max_abs_psi"])),
        ("c tip window", c[:, i0:i0 + w], g["win_c"], float(g["max_abs_c"])),
        ("T rows", T[::8, :], g["T_rows"], float(np.max(np.abs(g["T_rows"])))),
    )


def test_fp64_path_2048_antitrap_20000_steps_vs_oracle(rs, golden):
    g = golden
    seed, steps = int(g["seed"]), int(g["steps"])
    p = _params(rs, g)
    say(f"{int(g['N'])}^2 coupled T, antitrapping + trapping (V/V_D = {float(g['ratio'])}), "
        f"FP64, {steps} steps, seed {seed}")
    ctx = rs.Context(p)
    ctx.init_config(seed)
    ctx.step(steps)
    k = ctx.constants()
    assert k.step == int(g["step"])
    assert k.i_left_end == int(g["i_left_end"])
    assert k.i_offset_history == int(g["i_offset_history"])
    psi, c, T = (ctx.download(f) for f in (RS_FIELD_PSI, RS_FIELD_C, RS_FIELD_T))
    for name, got, ref, scale in _sampled(g, psi, c, T):
        assert got.shape == ref.shape, name
        e = float(np.max(np.abs(got - ref))) / scale
        say(name, "max rel err", e)
        assert e <= REL, (name, e)
    for name, got, ref in (("sumsq psi", float(np.sum(psi * psi)), float(g["sumsq_psi"])),
                           ("sumsq c", float(np.sum(c * c)), float(g["sumsq_c"])),
                           ("sum c", float(np.sum(c)), float(g["sum_c"]))):
        say(name, got, "oracle", ref)
        assert got == pytest.approx(ref, rel=REL), name
    m = ctx.solute_mass()
    say("solute mass", m, "oracle", float(g["solute_mass"]), "initial", float(g["solute_mass0"]))
    assert m == pytest.approx(float(g["solute_mass"]), rel=1e-12)
    assert ctx.tip_column() == int(g["tip_column"])
    t = ctx.tip()
    assert (t["i"], t["j"]) == (int(g["tip_i"]), int(g["tip_j"]))
    say("tip x", t["x"], "oracle", float(g["tip_x"]), "y", t["y"], "oracle", float(g["tip_y"]))
    assert t["x"] == pytest.approx(float(g["tip_x"]), abs=1e-6)
    assert t["y"] == pytest.approx(float(g["tip_y"]), abs=1e-6)
    ctx.close()


def test_fp32_path_2048_antitrap_20000_steps(rs, golden):
    g = golden
    seed, steps = int(g["seed"]), int(g["steps"])
    p = _params(rs, g, RS_PREC_MIXED)
    say(f"FP32 path, antitrapping + trapping, {steps} steps")
    ctx = rs.Context(p)
    ctx.init_config(seed)
    x0 = ctx.constants().i_offset_history + ctx.tip()["x"]
    ctx.step(steps)
    k = ctx.constants()
    assert k.step == int(g["step"])
    psi, c, T = (ctx.download(f) for f in (RS_FIELD_PSI, RS_FIELD_C, RS_FIELD_T))
    errs = {}
    for name, got, ref, scale in _sampled(g, psi, c, T):
        errs[name] = float(np.max(np.abs(got - ref))) / (1.0 if name.startswith("psi") else scale)
    dm = ctx.solute_mass() / float(g["solute_mass"]) - 1.0
    x1 = k.i_offset_history + ctx.tip()["x"]
    xo = int(g["i_offset_history"]) + float(g["tip_x"])
    say("errors", errs, "mass rel", dm, "tip advance", x1 - x0, "oracle", xo - x0)
    assert abs((x1 - x0) - (xo - x0)) <= 0.01 * abs(xo - x0)
    assert abs(dm) <= 1e-10
    assert max(errs["psi rows"], errs["psi tip window"]) <= 1e-4
    assert max(errs["c rows"], errs["c tip window"]) <= 1e-6
    assert errs["T rows"] <= 1e-7
    ctx.close()
