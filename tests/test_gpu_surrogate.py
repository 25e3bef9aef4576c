"""Learned CALPHAD surrogate: HIP training/inference kernels against the CPU
restatement (oracle/mlp.py) and the exact CALPHAD oracle, the ablation
against the untrained network, and the surrogate step against the FP64
oracle.

Stated tolerances:
  gradient (random weights, 4096 rows): relative L2 error <= 2e-3 overall
    and <= 1e-2 per parameter block; loss relative error <= 1e-5
  forward, each activation format (BF16, split BF16, FP16, packed FP16):
    |y_gpu - y_oracle| <= 1e-3 max|y| (activations may round the other way
    for rare elements; accumulation is FP32 vs float64)
  trusted outputs vs exact CALPHAD: residual <= residual_tol (1e-4, the
    nondimensional bound the trust table enforces; the default packed-FP16
    format judges its bins against 0.8 of it, so points between the
    validation grid nodes stay within it)
  surrogate step (256^2, 2000 steps): psi L_inf <= 2e-2, c L_inf <= 2e-3
    relative, tip velocity within 1% of the FP64 oracle
  solute mass: exact interior conservation (drift <= 1e-13 with uniform T,
    where no flux crosses the edges); in the coupled run the only mass change
    is the edge flux of the reference's ghost scheme (a ghost copies c but
    its T is interpolated at its own position, fun.cu:1220-1230,1585-1606),
    which depends on the fields near the edges: the drift must agree with the
    oracle's to 1e-10 + 1e-2 of the oracle's drift
"""
import sys

import numpy as np
import pytest

from oracle import mlp
from rsolid.abi import (RS_ACT_BF16, RS_ACT_BF16_SPLIT, RS_ACT_FP16, RS_ACT_FP16_PACKED,
                        RS_FIELD_C, RS_FIELD_PSI, RS_PREC_SURROGATE)

pytestmark = pytest.mark.gpu

def meta_test(fmt):
    return np.array([1930.0, 20.0, 0, 0, 0, 0, 1, 1, 1, 1, 1910.0, 1950.0, 1e-4, float(fmt)])
BLOCKS = {"W1": (mlp.OFF_W1, mlp.OFF_B1), "b1": (mlp.OFF_B1, mlp.OFF_W2),
          "W2": (mlp.OFF_W2, mlp.OFF_B2), "b2": (mlp.OFF_B2, mlp.OFF_W3),
          "W3": (mlp.OFF_W3, mlp.OFF_B3), "b3": (mlp.OFF_B3, mlp.NPARAM)}


def say(*a):
    print("[gpu-surrogate]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def trained(rs):
    """A 256^2 coupled-T surrogate context trained with the default options
    (2^20 uniform + 2^18 near-boundary samples)."""
    p = rs.config_params(256, True, precision=RS_PREC_SURROGATE)
    ctx = rs.Context(p)
    ctx.init_config(7)
    say("training (default options)")
    rep = ctx.surrogate_train()
    say("report", rep)
    yield ctx, rep
    ctx.close()


@pytest.mark.parametrize("split", [False, True])
def test_gradient_matches_oracle(rs, split):
    p = rs.config_params(256, True, precision=RS_PREC_SURROGATE)
    ctx = rs.Context(p)
    P = mlp.init_params(7)
    ctx.surrogate_set(P, meta_test(RS_ACT_BF16_SPLIT if split else RS_ACT_BF16))
    rng = np.random.default_rng(11)
    X = rng.uniform(-1, 1, (4096, 2)).astype(np.float32)
    Yn = rng.normal(size=(4096, 4)).astype(np.float32)
    say("gradient check, 4096 rows, split", split)
    g, loss = ctx.surrogate_grad(X, Yn)
    lo, go = mlp.loss_grad(P, X, Yn, quantized=True, split=split)
    say("loss", loss, lo)
    assert loss == pytest.approx(lo, rel=1e-5)
    err = np.linalg.norm(g - go) / np.linalg.norm(go)
    say("grad rel L2", err)
    assert err <= 2e-3
    for name, (a, b) in BLOCKS.items():
        e = np.linalg.norm(g[a:b] - go[a:b]) / np.linalg.norm(go[a:b])
        say(" block", name, e)
        assert e <= 1e-2, name
    ctx.close()


@pytest.mark.parametrize("fmt", [RS_ACT_BF16, RS_ACT_BF16_SPLIT, RS_ACT_FP16, RS_ACT_FP16_PACKED])
def test_forward_matches_oracle(rs, fmt):
    p = rs.config_params(256, True, precision=RS_PREC_SURROGATE)
    ctx = rs.Context(p)
    P = mlp.init_params(9)
    ctx.surrogate_set(P, meta_test(fmt))
    rng = np.random.default_rng(12)
    n = 5000
    T = rng.uniform(1910.0, 1950.0, n)
    c = 1.0 - rng.uniform(0.0, 0.4, n)
    say("forward check", n, "format", fmt)
    out = ctx.surrogate_eval(T, c)
    RR = p.RR
    idl = RR * T * (c * np.log(c) + (1 - c) * np.log(1 - c))
    lr = RR * T * np.log(c / (1 - c))
    y = np.stack([out[:, 0] - out[:, 1], out[:, 1] - idl, out[:, 2] - lr, out[:, 3] - out[:, 2]],
                 axis=1)
    X = np.stack([(T - 1930.0) / 20.0, ((1 - c) - 0.2) / 0.2], axis=1).astype(np.float32)
    yo = mlp.forward(P, X, quantized=True, split=fmt == RS_ACT_BF16_SPLIT,
                     fp16_acts=fmt in (RS_ACT_FP16, RS_ACT_FP16_PACKED),
                     fp16_packed=fmt == RS_ACT_FP16_PACKED)["y"]
    err = np.max(np.abs(y - yo))
    say("forward max err", err, "scale", np.max(np.abs(yo)))
    assert err <= 1e-3 * np.max(np.abs(yo))
    ctx.close()


def test_training_reduces_error_and_trust_holds(rs, oracle_mod, trained):
    ctx, rep = trained
    assert rep["steps"] > 0
    assert rep["mse"] < 1e-3 * rep["mse0"]
    assert rep["trusted_frac"] > 0.5
    assert rep["max_res_trusted"] <= 1e-4
    # trusted points against the exact CALPHAD oracle
    p = ctx.p
    rng = np.random.default_rng(5)
    n = 4000
    T = rng.uniform(rep["T_lo"], rep["T_hi"], n)
    c = 1.0 - rng.uniform(0.0, 0.4, n)
    out = ctx.surrogate_eval(T, c)
    ex = np.array([[oracle_mod.Gl(p, t, x), oracle_mod.Gs(p, t, x), oracle_mod.Glc(p, t, x),
                    oracle_mod.Gsc(p, t, x)] for t, x in zip(T, c)])
    k = rs.derive(p)
    import math
    a1 = 2.0 * math.sqrt(2.0) / 3.0
    hh = p.Gamma * 1e-9 * p.latentHeat * p.num_moles / p.Tmelt / (a1 * p.S * p.W0 * 1e-9)
    dG_den = 2.0 * hh * p.v0
    mu_den = 1.0 / (p.RR * p.Tmelt)
    res = np.maximum.reduce([
        np.abs((out[:, 0] - out[:, 1]) - (ex[:, 0] - ex[:, 1])) / dG_den,
        np.abs(out[:, 2] - ex[:, 2]) * mu_den,
        np.abs(out[:, 3] - ex[:, 3]) * mu_den])
    tr = out[:, 4] > 0.5
    say("trusted", int(tr.sum()), "of", n, "max residual trusted", res[tr].max(),
        "untrusted median", np.median(res[~tr]) if (~tr).any() else 0.0, "dt", k.dt)
    assert tr.sum() > 0.3 * n
    # the default (packed FP16) format judges its bins against 0.8 of the
    # bound, which keeps the points between the validation nodes within it
    assert res[tr].max() <= 1e-4


def _tip_pos(ctx):
    t = ctx.tip()
    k = ctx.constants()
    return (k.i_offset_history + t["x"]) * ctx.p.dx


def test_surrogate_step_vs_fp64_and_ablation(rs, oracle_mod, trained):
    ctx_t, rep = trained
    P, meta = ctx_t.surrogate_get()
    pd = rs.config_params(256, True)
    st = oracle_mod.State(pd)
    st.init_config(7)
    ps = rs.config_params(256, True, precision=RS_PREC_SURROGATE)
    assert meta[13] == RS_ACT_FP16_PACKED  # the default inference format
    meta_split = meta.copy()
    meta_split[13] = RS_ACT_BF16_SPLIT
    meta_fp16 = meta.copy()
    meta_fp16[13] = RS_ACT_FP16
    runs = {}
    for name, weights, m in (("trained", P, meta), ("trained-split", P, meta_split),
                             ("trained-fp16", P, meta_fp16),
                             ("untrained", mlp.init_params(3), meta)):
        ctx = rs.Context(ps)
        ctx.init_config(7)
        ctx.surrogate_set(weights, m)
        runs[name] = ctx
    # tip positions after 500 and 2000 steps
    x0 = {}
    m0 = {name: c.solute_mass() for name, c in runs.items()}
    mo0 = st.solute_mass()
    st.step(500)
    xo0 = (st.constants().i_offset_history + st.tip()["x"]) * pd.dx
    for name, c in runs.items():
        c.step(500)
        x0[name] = _tip_pos(c)
        c.surrogate_stats(reset=True)
    st.step(1500)
    xo1 = (st.constants().i_offset_history + st.tip()["x"]) * pd.dx
    vo = xo1 - xo0
    psi_o, c_o = st.get(RS_FIELD_PSI), st.get(RS_FIELD_C)
    dmo = (st.solute_mass() - mo0) / mo0
    for name, c in runs.items():
        c.step(1500)
        v = _tip_pos(c) - x0[name]
        stats = c.surrogate_stats()
        epsi = float(np.max(np.abs(c.download(RS_FIELD_PSI) - psi_o)))
        ec = float(np.max(np.abs(c.download(RS_FIELD_C) - c_o)) / np.max(np.abs(c_o)))
        dm = (c.solute_mass() - m0[name]) / m0[name] - dmo
        say(name, "fallback", stats, "tip advance", v, "oracle", vo, "psi Linf", epsi, "c rel",
            ec, "mass drift diff", dm)
        assert abs(v - vo) <= 0.01 * abs(vo)
        assert epsi <= 2e-2
        assert ec <= 2e-3
        say(name, "oracle drift", dmo)
        assert abs(dm) <= 1e-10 + 1e-2 * abs(dmo)
        if name.startswith("trained"):
            # the edge rows and columns (3% at 256^2) and little else: the
            # window spans the 1941 K breakpoint the run's field crosses
            assert stats["fallback_frac"] < 0.05
        else:
            assert stats["fallback_frac"] > 0.99
    for c in runs.values():
        c.close()


def test_surrogate_conserves_solute_exactly(rs):
    p = rs.config_params(256, False, precision=RS_PREC_SURROGATE)
    p.GG = 0.0
    p.q_solid = 0.0
    say("surrogate conservation, uniform T")
    ctx = rs.Context(p)
    ctx.init_config(3)
    # default training: at 8 epochs the network met the bound in too few
    # bins to serve this run's cells
    rep = ctx.surrogate_train()
    say("trusted", rep["trusted_frac"], "window", rep["T_lo"], rep["T_hi"])
    m0 = ctx.solute_mass()
    ctx.surrogate_stats(reset=True)
    ctx.step(2000)
    st = ctx.surrogate_stats()
    drift = (ctx.solute_mass() - m0) / m0
    say("drift", drift, "fallback", st)
    assert st["fallback_frac"] < 0.5
    assert abs(drift) < 1e-13
    ctx.close()


def test_network_in_every_row_gives_the_same_fields(rs, trained, monkeypatch):
    """Rows with no cell in the network's domain skip it behind a branch;
    evaluating it in every row instead (RS_SURR_ALWAYS=1, or the per-run
    choice once such rows are rare) discards it there, so psi and c stay
    bitwise equal to the branched steps."""
    ctx_t, _ = trained
    P, meta = ctx_t.surrogate_get()
    ps = rs.config_params(256, True, precision=RS_PREC_SURROGATE)
    out = {}
    for mode in ("0", "1", None):
        if mode is None:
            monkeypatch.delenv("RS_SURR_ALWAYS", raising=False)
        else:
            monkeypatch.setenv("RS_SURR_ALWAYS", mode)
        say("network in every row:", mode)
        ctx = rs.Context(ps)
        ctx.init_config(7)
        ctx.surrogate_set(P, meta)
        ctx.step(800)
        out[mode] = (ctx.download(RS_FIELD_PSI), ctx.download(RS_FIELD_C))
        st = ctx.surrogate_stats()
        say("stats", st)
        # a wavefront row evaluates 1-64 cells (its 62 columns and the two
        # halo columns, cut at the grid's edge), and every cell of a row
        # without a trusted cell takes the exact routine
        assert st["rows"] > 0 and 0 <= st["rows_without_network"] <= st["rows"]
        assert st["rows"] <= st["evaluated"] <= 64 * st["rows"]
        assert st["fallback"] >= st["rows_without_network"]
        ctx.close()
    for mode in ("1", None):
        assert np.array_equal(out[mode][0], out["0"][0])
        assert np.array_equal(out[mode][1], out["0"][1])


def test_window_spans_the_ti_breakpoint(rs, oracle_mod, trained):
    """The 256^2 coupled run's field reaches above 1941 K, where the G_Ti
    descriptions change piece (fun.cu:1628-1646): the default window spans
    the breakpoint, and the network is trusted on both sides of it at the
    run's compositions and within the residual bound there."""
    ctx, rep = trained
    say("window", rep["T_lo"], rep["T_hi"])
    assert rep["T_lo"] < 1941.0 < rep["T_hi"]
    T = np.repeat([1939.0, 1940.5, 1941.0, 1941.5, 1943.0], 3)
    c = np.tile([0.88, 0.9, 0.92], 5)
    out = ctx.surrogate_eval(T, c)
    assert out[:, 4].all()
    p = ctx.p
    ex = np.array([[oracle_mod.Gl(p, t, x), oracle_mod.Gs(p, t, x), oracle_mod.Glc(p, t, x),
                    oracle_mod.Gsc(p, t, x)] for t, x in zip(T, c)])
    err = np.abs(out[:, :4] - ex)
    say("largest |network - exact| (J/mol) per output", err.max(axis=0))
    # the residual bound, 1e-4 of the scales the step divides by, is
    # 2.05 J/mol (f_L - f_S) and 1.64 J/mol (mu) for this preset
    assert np.all(np.abs((out[:, 0] - out[:, 1]) - (ex[:, 0] - ex[:, 1])) <= 2.5)
    assert np.all(err[:, 2:] <= 2.0)


def test_network_alone_timing(rs, trained):
    """rs_surrogate_time: the network evaluated alone over device-resident
    inputs takes a positive, finite time that grows with the number of
    cells; its argument rules and the untrained state are refused."""
    from rsolid import RsError
    from rsolid.abi import RS_ERR_ARG, RS_ERR_STATE
    ctx, _ = trained
    t_small = ctx.surrogate_time(1 << 16, 3)
    t_large = ctx.surrogate_time(1 << 22, 3)
    say("network alone: 2^16 inputs", t_small, "ms; 2^22 inputs", t_large, "ms")
    assert 0.0 < t_small < t_large < 1e3
    for n, reps in ((63, 1), ((1 << 26) + 1, 1), (1 << 16, 0)):
        with pytest.raises(RsError) as e:
            ctx.surrogate_time(n, reps)
        assert e.value.code == RS_ERR_ARG
    fresh = rs.Context(rs.config_params(64, True, precision=RS_PREC_SURROGATE))
    with pytest.raises(RsError) as e:
        fresh.surrogate_time(1 << 16, 1)
    assert e.value.code == RS_ERR_STATE
    fresh.close()
