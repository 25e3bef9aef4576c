"""Mixed-precision path (psi/phi and fine T in FP32; c, flux divergence,
ln terms and time accumulation in FP64) against the FP64 CPU oracle.

Stated tolerances (after 2000 steps of the 256^2 configuration):
  psi   L_inf <= 2e-6 absolute  (FP32 storage: ulp(1) = 1.2e-7 per step)
        on the interface (|psi| < 0.999) and, for a grain on the axes, on
        the whole field; for a rotated grain the near-bulk cells (psi
        within ~1e-5 of +-1, where the clamp phi = max(psi, -1) and an
        almost flat gradient make the normal direction ill-conditioned)
        are held to 3x the FP64 path's own sensitivity there: the FP64
        run with the grain angle gamma changed by 1e-5 degrees (3.3e-6 at
        step 2000, the same for a 1e-4 degree change)
  c     L_inf <= 1e-7 relative
  T     L_inf <= (L/cp) x the psi deviation, relative to max T (the
        latent-heat source telescopes; coarse T advanced in FP64)
  solute mass trajectory within 1e-10 of the oracle's (relative)
  tip x within 1e-4 cells
and exact interior conservation (<= 1e-13) with uniform T.
"""
import sys

import numpy as np
import pytest

from rsolid.abi import RS_FIELD_C, RS_FIELD_PSI, RS_FIELD_T, RS_PREC_FP64, RS_PREC_MIXED

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-mixed]", *a, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("tfc,rotated", [(False, False), (True, False), (True, True)])
def test_mixed_tracks_fp64_oracle(rs, oracle_mod, tfc, rotated):
    pm = rs.config_params(256, tfc, precision=RS_PREC_MIXED)
    pd = rs.config_params(256, tfc)
    if rotated:  # a grain turned away from the axes: the general-orientation kernels
        for p in (pm, pd):
            p.alpha0, p.beta0, p.gamma0 = 10.0, 15.0, 25.0
    say("mixed vs oracle, 256^2, tfc", tfc, "rotated", rotated)
    cm = rs.Context(pm)
    cm.init_config(7)
    info = cm.info()
    say("poly table", info)
    assert info["poly_err"] < 1e-3
    st = oracle_mod.State(pd)
    st.init_config(7)
    cs = None
    if rotated:  # the FP64 path with gamma off by 1e-5 degrees: the yardstick
        ps = rs.config_params(256, tfc)
        ps.alpha0, ps.beta0, ps.gamma0 = 10.0, 15.0, 25.0 + 1e-5
        cs = rs.Context(ps)
        cs.init_config(7)
    m0, mo0 = cm.solute_mass(), st.solute_mass()
    # The coupled coarse T differs between the paths only through its latent
    # heat source, (L/cp)/2 times the coarse-cell sum of psi_next - psi per
    # fine cell (fun.cu:1494-1535): the source telescopes over the steps, so
    # the T deviation is bounded by ~(L/cp) times the psi deviation seen so
    # far (a factor 2 over the telescoped sum for the diffusion between steps).
    Lcp = pd.latentHeat / pd.cp
    epsi_max = 0.0
    for n in (1, 199, 1800):
        cm.step(n)
        st.step(n)
        po = st.get(RS_FIELD_PSI)
        dpsi = np.abs(cm.download(RS_FIELD_PSI) - po)
        epsi = float(np.max(dpsi))
        eint = float(np.max(dpsi[np.abs(po) < 0.999]))
        epsi_max = max(epsi_max, epsi)
        ec = _rel(cm.download(RS_FIELD_C), st.get(RS_FIELD_C))
        dm = (cm.solute_mass() - m0) / m0 - (st.solute_mass() - mo0) / mo0
        say("step", cm.constants().step, "psi Linf", epsi, "interface", eint, "c rel", ec,
            "mass-drift diff", dm)
        assert eint <= 2e-6
        if cs is None:
            assert epsi <= 2e-6
        else:
            cs.step(n)
            esens = float(np.max(np.abs(cs.download(RS_FIELD_PSI) - po)))
            say("FP64 sensitivity (gamma + 1e-5 deg)", esens)
            assert epsi <= max(2e-6, 3.0 * esens)
        assert ec <= 1e-7
        assert abs(dm) <= 1e-10
        if tfc:
            To = st.get(RS_FIELD_T)
            eT = _rel(cm.download(RS_FIELD_T), To)
            bound = Lcp * epsi_max / float(np.max(np.abs(To)))
            say("T rel", eT, "latent-heat bound", bound)
            assert eT <= bound
    tm, to = cm.tip(), st.tip()
    assert tm["i"] == to["i"] and tm["j"] == to["j"]
    assert tm["x"] == pytest.approx(to["x"], abs=1e-4)
    cm.close()
    if cs is not None:
        cs.close()


def test_mixed_conserves_solute_exactly(rs):
    p = rs.config_params(256, False, precision=RS_PREC_MIXED)
    p.GG = 0.0
    p.q_solid = 0.0
    say("mixed conservation, uniform T")
    ctx = rs.Context(p)
    ctx.init_config(3)
    m0 = ctx.solute_mass()
    ctx.step(2000)
    c = ctx.download(RS_FIELD_C)
    assert c.min() > 1e-6
    drift = (ctx.solute_mass() - m0) / m0
    say("drift", drift)
    assert abs(drift) < 1e-13
    ctx.close()


def test_mixed_exact_fallback_outside_table(rs, oracle_mod):
    """A frozen-T run whose gradient spans more than the polynomial window:
    cells outside it take the exact FP64 routine; results still track the
    oracle."""
    pm = rs.config_params(256, False, precision=RS_PREC_MIXED)
    pm.GG = 2.0          # 256 cells * 0.2 nm * 2 K/nm = 102 K span
    pd = pm.copy()
    pd.precision = 0
    say("mixed exact fallback")
    cm = rs.Context(pm)
    cm.init_config(9)
    info = cm.info()
    st = oracle_mod.State(pd)
    st.init_config(9)
    cm.step(300)
    st.step(300)
    assert float(np.max(np.abs(cm.download(RS_FIELD_PSI) - st.get(RS_FIELD_PSI)))) <= 2e-6
    assert _rel(cm.download(RS_FIELD_C), st.get(RS_FIELD_C)) <= 1e-7
    say("window", info["T_lo"], info["T_hi"])
    cm.close()


def test_mixed_full_size_determinism(rs):
    p = rs.config_params(4096, True, precision=RS_PREC_MIXED)
    say("mixed 4096^2 determinism")
    outs = []
    for _ in range(2):
        ctx = rs.Context(p)
        ctx.init_config(5)
        ctx.step(100)
        outs.append((ctx.download(RS_FIELD_PSI), ctx.download(RS_FIELD_C)))
        ctx.close()
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
    assert np.isfinite(outs[0][0]).all() and np.isfinite(outs[0][1]).all()


@pytest.mark.parametrize("preset", ["config512", "reference768x272", "config512_fp64"])
def test_dispatch_order_and_band_height_change_no_bit(rs, monkeypatch, preset):
    """The FP32 exact-table kernel dispatches its (column tile, band group)
    items tile by tile from the side the front is held at (config preset:
    from the first tile up; reference preset, front at Nx - 268: from the
    last tile down) in 20-row bands; the FP64 kernel takes 20-row bands as
    well. Every item is computed once whatever the order and the band
    height, so the fields equal, bit for bit, those of 40-row bands in
    workgroups of one wavefront."""
    if preset.startswith("config512"):
        prec = RS_PREC_FP64 if preset.endswith("fp64") else RS_PREC_MIXED
        p = rs.config_params(512, True, precision=prec)
    else:
        p = rs.reference_params(768, 272)
        p.precision = RS_PREC_MIXED
    say(preset, "i_tip_target", p.i_tip_target, "Nx", p.Nx)
    outs = []
    for env in ({}, {"RS_BAND_H": "40", "RS_WPB": "1"}):
        for k in ("RS_BAND_H", "RS_WPB"):
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        ctx = rs.Context(p)
        say(" ", env, "band_h", ctx.info()["band_h"])
        if preset.startswith("config512"):
            ctx.init_config(11)
        else:
            ctx.init_reference()
        ctx.step(300)
        outs.append((ctx.download(RS_FIELD_PSI), ctx.download(RS_FIELD_C), ctx.download(RS_FIELD_T)))
        ctx.close()
    for k in ("RS_BAND_H", "RS_WPB"):
        monkeypatch.delenv(k, raising=False)
    for a, b in zip(outs[0], outs[1]):
        assert np.isfinite(a).all()
        assert np.array_equal(a, b)


def test_mixed_antitrapping_tracks_oracle(rs, oracle_mod):
    """The antitrapping / trapping extension on the mixed path (FP32 psi,
    FP64 correction of c) against the FP64 oracle, same tolerances as the
    plain mixed run."""
    pm = rs.config_params(256, True, precision=RS_PREC_MIXED)
    for p in (pm,):
        p.at_coef = 1.0 / (2.0 * np.sqrt(2.0))
        p.V_D = p.Vp / 0.5
    pd = pm.copy()
    pd.precision = 0
    say("mixed antitrapping vs oracle")
    cm = rs.Context(pm)
    cm.init_config(7)
    st = oracle_mod.State(pd)
    st.init_config(7)
    m0, mo0 = cm.solute_mass(), st.solute_mass()
    cm.step(2000)
    st.step(2000)
    epsi = float(np.max(np.abs(cm.download(RS_FIELD_PSI) - st.get(RS_FIELD_PSI))))
    ec = _rel(cm.download(RS_FIELD_C), st.get(RS_FIELD_C))
    dm = (cm.solute_mass() - m0) / m0 - (st.solute_mass() - mo0) / mo0
    say("psi Linf", epsi, "c rel", ec, "mass-drift diff", dm)
    assert epsi <= 2e-6
    assert ec <= 1e-7
    assert abs(dm) <= 1e-10
    cm.close()
