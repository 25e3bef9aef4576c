"""The learned-surrogate workload (4096^2, FP32 fields, coupled T) and the
batched 64-run ensemble against the FP64 oracle, and the batch with large
grids and a non-zero pull-back shift against the runs stepped alone.

Stated tolerances (the surrogate path's bounds, as in test_gpu_surrogate):
  surrogate step vs the FP64 oracle: psi L_inf <= 2e-2 absolute, c L_inf
    <= 2e-3 relative to max |c|, the tip advance within 1% of the oracle's
    plus 2e-3 cells (the tip fit's resolution of a short advance)
  FP32 exact-table step vs the FP64 oracle: psi L_inf <= 2e-6 absolute,
    c <= 1e-7 relative (the FP32 path's bounds)
  long horizon (no oracle: 4 000 steps of 4096^2 take the CPU restatement
    about an hour): the surrogate path against the FP64 GPU path, itself
    held to 1e-10 of the oracle over 20 000 steps at 2048^2: tip velocity
    within 1%, solute-mass trajectories within 1e-8 relative
  batched runs vs the same runs stepped alone: bitwise
"""
import sys

import numpy as np
import pytest

from rsolid.abi import RS_FIELD_C, RS_FIELD_PSI, RS_FIELD_T, RS_PREC_MIXED, RS_PREC_SURROGATE

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-configs]", *a, flush=True)
    sys.stdout.flush()


def _tip_x(ctx):
    return ctx.constants().i_offset_history + ctx.tip()["x"]


def _cmp_surrogate(name, psi, c, psi_o, c_o, psi_tol, c_tol):
    epsi = float(np.max(np.abs(psi - psi_o)))
    ec = float(np.max(np.abs(c - c_o))) / float(np.max(np.abs(c_o)))
    say(name, "psi Linf", epsi, "c rel", ec)
    assert np.isfinite(psi).all() and np.isfinite(c).all(), name
    assert epsi <= psi_tol, (name, epsi)
    assert ec <= c_tol, (name, ec)


def test_surrogate_4096_coupled_vs_oracle(rs, oracle_mod):
    """4096^2 coupled T, the network trained with the default options, 50
    steps from init_config(5) (the bench's workload), against the FP64
    oracle: fields within the surrogate bounds, tip advance within 1%, and
    the fallback fraction reported."""
    N, seed, steps = 4096, 5, 50
    p = rs.config_params(N, True, precision=RS_PREC_SURROGATE)
    say(f"{N}^2 coupled T, surrogate path, {steps} steps vs oracle")
    ctx = rs.Context(p)
    ctx.init_config(seed)
    rep = ctx.surrogate_train()
    say("trusted bins", rep["trusted_frac"], "window", rep["T_lo"], rep["T_hi"])
    x0 = _tip_x(ctx)
    ctx.surrogate_stats(reset=True)
    ctx.step(steps)
    st = ctx.surrogate_stats()
    x1 = _tip_x(ctx)
    psi, c = ctx.download(RS_FIELD_PSI), ctx.download(RS_FIELD_C)
    m = ctx.solute_mass()
    ctx.close()
    po = rs.config_params(N, True)
    o = oracle_mod.State(po)
    o.init_config(seed)
    xo0 = o.constants().i_offset_history + o.tip()["x"]
    o.step(steps)
    xo1 = o.constants().i_offset_history + o.tip()["x"]
    psi_o, c_o = o.get(RS_FIELD_PSI), o.get(RS_FIELD_C)
    mo = o.solute_mass()
    o.close()
    say("fallback", st["fallback_frac"], "rows without network",
        st["rows_without_network"] / max(1.0, st["rows"]))
    _cmp_surrogate("4096 surrogate", psi, c, psi_o, c_o, 2e-2, 2e-3)
    say("tip advance", x1 - x0, "oracle", xo1 - xo0, "mass rel", m / mo - 1.0)
    assert abs((x1 - x0) - (xo1 - xo0)) <= 0.01 * abs(xo1 - xo0) + 2e-3
    assert 0.0 <= st["fallback_frac"] < 0.1


def test_surrogate_4096_long_horizon_vs_fp64_path(rs):
    """The same workload over 4 000 steps: tip velocity between steps 2 000
    and 4 000 within 1% of the FP64 GPU path's, and the solute-mass
    trajectories within 1e-8 relative of each other."""
    N, seed = 4096, 5
    res = {}
    for name, prec in (("fp64", None), ("surrogate", RS_PREC_SURROGATE)):
        p = rs.config_params(N, True) if prec is None else rs.config_params(N, True, precision=prec)
        ctx = rs.Context(p)
        ctx.init_config(seed)
        if prec is not None:
            ctx.surrogate_train()
            ctx.surrogate_stats(reset=True)
        m0 = ctx.solute_mass()
        ctx.step(2000)
        xa = _tip_x(ctx)
        ctx.step(2000)
        xb = _tip_x(ctx)
        fb = ctx.surrogate_stats()["fallback_frac"] if prec is not None else None
        res[name] = (xb - xa, ctx.solute_mass() / m0 - 1.0, fb)
        say(name, "tip advance 2000->4000", xb - xa, "mass drift", res[name][1], "fallback", fb)
        ctx.close()
    (a64, d64, _), (asg, dsg, fb) = res["fp64"], res["surrogate"]
    assert abs(asg - a64) <= 0.01 * abs(a64)
    assert abs(dsg - d64) <= 1e-8
    assert fb < 0.1


def test_precision_4096_vs_fp64_path(rs):
    """The precision variant (FP32 psi/phi and T, FP64 c) and the surrogate
    path at 4096^2 beside the plain FP64 path over 6 000 steps, compared at
    3 000 and 6 000 by the precision report (rsolid.report; its 50 000-step
    run is the profiles' precision_4096.md): per-field errors, tip and front
    velocity over each interval, and each path's solute ledger. Bounds: FP32
    path psi 1e-5 absolute, c 1e-6 and T 1e-7 relative, front velocity
    1e-5, drift within 1e-10 of the FP64 path's; surrogate path psi 2e-2, c
    2e-3, front velocity 1e-2; every path's accounted drift (the rest of its
    ledger: its scheme's conservation error) within 1e-12, and the drift
    difference equal to the ledger terms' difference (pull-backs, boundary
    faces, rest) to 1e-13."""
    from rsolid.report import run
    res = run(4096, 6000, 3000, ["mixed", "surrogate"], seed=5, log=lambda m: say(m[:220]))
    for r in res["rows"]:
        assert abs(r["unacc"]) <= 1e-12 and abs(r["unacc_fp64"]) <= 1e-12, r
        parts = r["d_pull"] + r["d_edge"] + (r["unacc"] - r["unacc_fp64"])
        assert abs(r["drift_diff"] - parts) <= 1e-13, r
        if r["path"] == "mixed":
            assert r["psi_linf"] <= 1e-5 and r["c_linf"] <= 1e-6 and r["T_linf"] <= 1e-7, r
            assert r["front_vel_err"] <= 1e-5, r
            assert abs(r["drift_diff"]) <= 1e-10, r
        else:
            assert r["psi_linf"] <= 2e-2 and r["c_linf"] <= 2e-3, r
            assert r["front_vel_err"] <= 1e-2, r
            assert r["fallback_frac"] < 0.1, r


@pytest.fixture(scope="module")
def ensemble_oracle(rs, oracle_mod):
    """The bench's seeded 64-run ensemble; for its 256^2 and 512^2 runs the
    FP64 oracle's psi and c after STEPS steps."""
    sizes, dts, vps, seeds = rs.draw_ensemble(64, 2024)
    out = {}
    say("oracle for the 256^2 / 512^2 runs of the ensemble")
    for q, (n, d, v, s) in enumerate(zip(sizes, dts, vps, seeds)):
        if n > 512:
            continue
        o = oracle_mod.State(rs.ensemble_params(int(n), float(d), float(v), 0))
        o.init_config(s)
        o.step(ENSEMBLE_STEPS)
        out[q] = (o.get(RS_FIELD_PSI), o.get(RS_FIELD_C))
        o.close()
    return (sizes, dts, vps, seeds), out


ENSEMBLE_STEPS = 100


@pytest.mark.parametrize("prec", [RS_PREC_SURROGATE, RS_PREC_MIXED])
def test_ensemble_64_runs_vs_oracle(rs, ensemble_oracle, prec):
    """The 64-run ensemble (32 x 256^2, 19 x 512^2, 8 x 1024^2, 5 x 2048^2,
    uneven undercoolings and pulling speeds) in one persistent launch per
    step with autotuned tiles and graph replay: every 256^2 / 512^2 run
    against the FP64 oracle (surrogate or FP32 bounds), every 1024^2 /
    2048^2 run bitwise equal to the same run stepped alone and held to the
    field invariants (finite, psi <= 1, c >= 1e-6)."""
    (sizes, dts, vps, seeds), ref = ensemble_oracle
    params = [rs.ensemble_params(int(n), float(d), float(v), prec)
              for n, d, v in zip(sizes, dts, vps)]
    surr = prec == RS_PREC_SURROGATE
    say("64-run ensemble,", "surrogate" if surr else "FP32 exact table")
    b = rs.Batch(params)
    for run, s in zip(b.runs, seeds):
        run.init_config(s)
    weights = None
    if surr:
        rep = b.surrogate_train()
        say("shared network: trusted bins", rep["trusted_frac"], "window", rep["T_lo"], rep["T_hi"])
        weights = b.runs[0].surrogate_get()
    say("tuned", b.tune(steps=3))
    b.step(ENSEMBLE_STEPS)
    psi_tol, c_tol = (2e-2, 2e-3) if surr else (2e-6, 1e-7)
    worst = [0.0, 0.0]
    for q, run in enumerate(b.runs):
        psi, c = run.download(RS_FIELD_PSI), run.download(RS_FIELD_C)
        assert run.constants().step == ENSEMBLE_STEPS
        assert np.isfinite(psi).all() and np.isfinite(c).all(), q
        assert psi.max() <= 1.0 and c.min() >= 1e-6, q
        if q in ref:
            psi_o, c_o = ref[q]
            epsi = float(np.max(np.abs(psi - psi_o)))
            ec = float(np.max(np.abs(c - c_o))) / float(np.max(np.abs(c_o)))
            worst = [max(worst[0], epsi), max(worst[1], ec)]
            assert epsi <= psi_tol, (q, int(sizes[q]), epsi)
            assert ec <= c_tol, (q, int(sizes[q]), ec)
    say("worst over the 256^2 / 512^2 runs: psi", worst[0], "c rel", worst[1])
    big = [q for q in range(len(params)) if int(sizes[q]) > 512]
    for q in big:
        c1 = rs.Context(params[q])
        c1.init_config(seeds[q])
        if surr:
            c1.surrogate_set(*weights)
        c1.step(ENSEMBLE_STEPS)
        for f in (RS_FIELD_PSI, RS_FIELD_C, RS_FIELD_T):
            assert np.array_equal(c1.download(f), b.runs[q].download(f)), (q, f)
        c1.close()
    say("bitwise equal to single runs:", len(big), "runs of 1024^2 / 2048^2")
    b.close()


def test_batch_large_grids_with_pull_back_shift_matches_single_runs(rs):
    """1024^2 and 2048^2 runs batched with a 256^2 run, every run's front
    moved 5 cells ahead of its target before the first step so the step-0
    pull-back shifts the fields by 5 columns on the run's own stream, and
    further pull-backs every 40 steps: bitwise equal to the runs stepped
    alone (the batch's launches must wait for the shifts)."""
    specs = [(1024, 40.0, 0.1), (2048, 50.0, 0.2), (256, 30.0, 0.06), (1024, 28.0, 0.25)]
    params = [rs.ensemble_params(n, d, v, RS_PREC_MIXED) for n, d, v in specs]
    seeds = [51, 52, 53, 54]

    def prepare(ctx, s):
        ctx.init_config(s)
        ctx.pull_back(-5)
        k = ctx.constants()
        k.step = 0
        k.step_pull_back = 40
        ctx.set_constants(k)

    b = rs.Batch(params)
    for run, s in zip(b.runs, seeds):
        prepare(run, s)
    say("large-grid batch with pull-back shifts")
    b.step(120)
    offs = [run.constants().i_offset_history for run in b.runs]
    say("i_offset_history per run", offs)
    assert all(o != -5 for o in offs)
    for q, (p, s) in enumerate(zip(params, seeds)):
        c = rs.Context(p)
        prepare(c, s)
        c.step(120)
        k1, k2 = c.constants(), b.runs[q].constants()
        assert (k1.step, k1.i_offset_history) == (k2.step, k2.i_offset_history), q
        for f in (RS_FIELD_PSI, RS_FIELD_C, RS_FIELD_T):
            assert np.array_equal(c.download(f), b.runs[q].download(f)), (q, f)
        c.close()
    b.close()
