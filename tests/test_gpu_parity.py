"""GPU parity: the HIP FP64 path (through the C ABI) against the CPU oracle
on the same seeded inputs, plus size-independent properties at the full
4096^2 size.

Tolerances: FP64 fields, coarse T and solute mass within 1e-10 relative of
the oracle (the north-star tolerance for the FP64 path); integer run state
(pull-back offsets, tip column, i_left_end) exactly equal.
"""
import sys

import numpy as np
import pytest

from rsolid.abi import (RS_FIELD_C, RS_FIELD_PHI, RS_FIELD_PSI, RS_FIELD_T, RS_FIELD_TFINE,
                        RS_PREC_MIXED)

pytestmark = pytest.mark.gpu

REL = 1e-10


def say(*a):
    print("[gpu-parity]", *a, flush=True)
    sys.stdout.flush()


def _cases(rs):
    return {
        "ref512x272": (rs.reference_params(512, 272), "ref", 0),
        "cfg256_frozen": (rs.config_params(256, False), "cfg", 7),
        "cfg256_coupled": (rs.config_params(256, True), "cfg", 7),
        "cfg256_rotated": (_rotated(rs.config_params(256, True)), "cfg", 7),
    }


def _rotated(p):
    # a grain turned away from the axes (fun.cu:129-131 Euler angles)
    p.alpha0, p.beta0, p.gamma0 = 10.0, 15.0, 25.0
    return p


def _init(obj, kind, seed):
    if kind == "ref":
        obj.init_reference()
    else:
        obj.init_config(seed)


def _relerr(a, b):
    scale = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)


def _compare(ctx, st, tfc, tag, rel=REL):
    for f, name in ((RS_FIELD_PSI, "psi"), (RS_FIELD_C, "c")):
        e = _relerr(ctx.download(f), st.get(f))
        say(tag, name, "max rel err", e)
        assert e <= rel, (tag, name, e)
    e = _relerr(ctx.download(RS_FIELD_PHI), st.get(RS_FIELD_PHI))
    assert e <= rel, (tag, "phi", e)
    if tfc:
        e = _relerr(ctx.download(RS_FIELD_T), st.get(RS_FIELD_T))
        say(tag, "T max rel err", e)
        assert e <= rel, (tag, "T", e)
    kg, ko = ctx.constants(), st.constants()
    for f in ("step", "i_offset", "i_offset_history", "i_left_end", "i_left_end_interp"):
        assert getattr(kg, f) == getattr(ko, f), (tag, f)
    assert kg.l_offset_Vp == ko.l_offset_Vp


@pytest.mark.parametrize("case", ["ref512x272", "cfg256_frozen", "cfg256_coupled"])
def test_init_matches_oracle(rs, oracle_mod, case):
    p, kind, seed = _cases(rs)[case]
    say("init", case, p.Nx, "x", p.Ny)
    ctx = rs.Context(p)
    _init(ctx, kind, seed)
    st = oracle_mod.State(p)
    _init(st, kind, seed)
    # tanh of the initial profile: device vs host libm, a few ulp
    assert _relerr(ctx.download(RS_FIELD_PSI), st.get(RS_FIELD_PSI)) < 1e-15
    assert _relerr(ctx.download(RS_FIELD_C), st.get(RS_FIELD_C)) < 1e-15
    if p.ifTFC:
        assert np.array_equal(ctx.download(RS_FIELD_T), st.get(RS_FIELD_T))
        assert np.array_equal(ctx.download(RS_FIELD_TFINE), st.get(RS_FIELD_TFINE))
    ctx.close()


@pytest.mark.parametrize("case", ["ref512x272", "cfg256_frozen", "cfg256_coupled"])
def test_single_step_from_identical_state(rs, oracle_mod, case):
    p, kind, seed = _cases(rs)[case]
    say("one step", case)
    st = oracle_mod.State(p)
    _init(st, kind, seed)
    ctx = rs.Context(p)
    _init(ctx, kind, seed)
    ctx.upload(RS_FIELD_PSI, st.get(RS_FIELD_PSI))
    ctx.upload(RS_FIELD_C, st.get(RS_FIELD_C))
    for obj in (st, ctx):
        k = obj.constants()
        k.step = 1
        obj.set_constants(k)
    st.step(1)
    ctx.step(1)
    _compare(ctx, st, p.ifTFC, case + "/1", rel=1e-13)
    ctx.close()


@pytest.mark.parametrize("case,steps", [("cfg256_frozen", 2000), ("cfg256_coupled", 2000),
                                        ("ref512x272", 2100), ("cfg256_rotated", 2000)])
def test_run_matches_oracle(rs, oracle_mod, case, steps):
    """The 256^2 frozen-T run (2000 steps), the same grid with the coupled
    heat equation (also with a rotated grain), and the reference parameter
    set with its step-0, 1025 and 2050 pull-backs."""
    p, kind, seed = _cases(rs)[case]
    say("run", case, steps, "steps")
    st = oracle_mod.State(p)
    _init(st, kind, seed)
    ctx = rs.Context(p)
    _init(ctx, kind, seed)
    m0g, m0o = ctx.solute_mass(), st.solute_mass()
    assert m0g == pytest.approx(m0o, rel=1e-13)
    done = 0
    for chunk in (1, 499, steps - 500):
        ctx.step(chunk)
        st.step(chunk)
        done += chunk
        _compare(ctx, st, p.ifTFC, f"{case}@{done}")
    assert ctx.tip_column() == st.tip_column()
    mg, mo = ctx.solute_mass(), st.solute_mass()
    say(case, "mass gpu", mg, "oracle", mo, "drift", (mg - m0g) / m0g)
    assert mg == pytest.approx(mo, rel=1e-12)
    tg, to = ctx.tip(), st.tip()
    say(case, "tip gpu", tg)
    say(case, "tip ora", to)
    for f in ("i", "j", "step"):
        assert tg[f] == to[f]
    for f in ("x", "T_tip", "delta"):
        assert tg[f] == pytest.approx(to[f], rel=1e-8, abs=1e-9)
    # the parabola's vertex row and curvature are ill-conditioned on a nearly
    # planar front (R ~ 1e3 nm): 1e-15 field differences move them by ~1e-7
    assert tg["y"] == pytest.approx(to["y"], abs=1e-5)
    assert tg["R"] == pytest.approx(to["R"], rel=1e-4)
    ctx.close()


@pytest.mark.parametrize("off", [-3, 0, 5, 40])
def test_pull_back_is_exact(rs, oracle_mod, off):
    p = rs.reference_params(512, 272)
    say("pull_back", off)
    rng = np.random.default_rng(off + 10)
    psi = rng.uniform(-1.2, 1.0, (p.Ny, p.Nx))
    c = rng.uniform(0.3, 0.7, (p.Ny, p.Nx))
    st = oracle_mod.State(p)
    st.init_reference()
    ctx = rs.Context(p)
    ctx.init_reference()
    for obj in (st, ctx):
        obj.upload(RS_FIELD_PSI, psi) if obj is ctx else obj.put(RS_FIELD_PSI, psi)
        obj.upload(RS_FIELD_C, c) if obj is ctx else obj.put(RS_FIELD_C, c)
    st.put(RS_FIELD_PHI, np.maximum(psi, -1.0))
    st.pull_back(off)
    ctx.pull_back(off)
    assert np.array_equal(ctx.download(RS_FIELD_PSI), st.get(RS_FIELD_PSI))
    assert np.array_equal(ctx.download(RS_FIELD_C), st.get(RS_FIELD_C))
    assert ctx.constants().i_left_end == st.constants().i_left_end
    assert ctx.tip_column() == st.tip_column()
    ctx.close()


def test_2048_coupled_short_run(rs, oracle_mod):
    """The 2048^2 coupled-T configuration for 60 steps."""
    p = rs.config_params(2048, True)
    say("2048^2 coupled, 60 steps")
    st = oracle_mod.State(p)
    st.init_config(21)
    ctx = rs.Context(p)
    ctx.init_config(21)
    ctx.step(60)
    st.step(60)
    _compare(ctx, st, True, "2048@60")
    assert ctx.solute_mass() == pytest.approx(st.solute_mass(), rel=1e-12)
    ctx.close()


def test_full_size_properties(rs):
    """4096^2 coupled T: finite fields, bounded phase field, bitwise
    determinism of two identical runs, and solute mass drift bounded by the
    boundary flux the reference scheme allows."""
    p = rs.config_params(4096, True)
    say("4096^2 coupled, 2 x 200 steps")
    outs = []
    for _ in range(2):
        ctx = rs.Context(p)
        ctx.init_config(5)
        m0 = ctx.solute_mass()
        ctx.step(200)
        psi = ctx.download(RS_FIELD_PSI)
        c = ctx.download(RS_FIELD_C)
        m1 = ctx.solute_mass()
        outs.append((psi, c, m1))
        ctx.close()
        assert np.isfinite(psi).all() and np.isfinite(c).all()
        assert psi.max() <= 1.0 and c.min() >= 1e-6
        say("4096 drift", (m1 - m0) / m0)
        assert abs(m1 - m0) / m0 < 1e-6
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
    assert outs[0][2] == outs[1][2]


def test_full_size_matches_oracle(rs, oracle_mod):
    """The 4096^2 coupled-T workload of the bench line itself, 20 steps from
    the seeded front, against the oracle at the run parity bound (1e-10
    relative per field, integer state equal), and the FP32 path from the
    same state within its stated bounds (psi 2e-6 absolute, c 1e-7
    relative)."""
    p = rs.config_params(4096, True)
    say("4096^2 coupled, 20 steps vs oracle")
    st = oracle_mod.State(p)
    st.init_config(5)
    st.step(20)
    ctx = rs.Context(p)
    ctx.init_config(5)
    ctx.step(20)
    _compare(ctx, st, True, "4096@20")
    ctx.close()
    pm = rs.config_params(4096, True, precision=RS_PREC_MIXED)
    cm = rs.Context(pm)
    cm.init_config(5)
    cm.step(20)
    epsi = float(np.max(np.abs(cm.download(RS_FIELD_PSI) - st.get(RS_FIELD_PSI))))
    ec = _relerr(cm.download(RS_FIELD_C), st.get(RS_FIELD_C))
    say("4096 FP32 path psi Linf", epsi, "c rel", ec)
    assert epsi <= 2e-6
    assert ec <= 1e-7
    cm.close()
    st.close()


@pytest.mark.parametrize("v_ratio", [0.1, 2.0])
def test_antitrapping_trapping_run_matches_oracle(rs, oracle_mod, v_ratio):
    """Antitrapping current with the solute-trapping factor (an extension,
    off in the reference presets) at V_p / V_D = 0.1 and 2, coupled T,
    2000 steps: FP64 path within 1e-10 of the oracle, and the solute-mass
    trajectory too."""
    p = rs.config_params(256, True)
    p.at_coef = 1.0 / (2.0 * np.sqrt(2.0))
    p.V_D = p.Vp / v_ratio
    say("antitrapping V/V_D", v_ratio)
    ctx = rs.Context(p)
    st = oracle_mod.State(p)
    ctx.init_config(7)
    st.init_config(7)
    m0, mo0 = ctx.solute_mass(), st.solute_mass()
    ctx.step(2000)
    st.step(2000)
    _compare(ctx, st, True, f"at{v_ratio}")
    dm = (ctx.solute_mass() - m0) / m0 - (st.solute_mass() - mo0) / mo0
    assert abs(dm) <= 1e-12
    ctx.close()
    st.close()
