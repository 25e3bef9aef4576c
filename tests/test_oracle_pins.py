"""Pins of the CPU oracle against facts that do not come from the oracle:
values derived by hand from the reference's formulas, published SGTE data,
closed forms of the discrete operators, conservation laws and brute force.

Reference: amakthel/2d-rapid-solidification, fun.cu.
"""
import math

import numpy as np
import pytest

from rsolid.abi import RS_FIELD_C, RS_FIELD_PHI, RS_FIELD_PSI, RS_FIELD_T


# ---------------------------------------------------------------------------
# set_parameters (fun.cu:1984-2058): numbers derived independently (by hand
# from the formulas, with run.sh's flags)
# ---------------------------------------------------------------------------
def test_derived_reference_parameters(rs, oracle_mod):
    p = rs.reference_params(768, 832)
    k = oracle_mod.derive(p)
    assert oracle_mod.tau0(p) == pytest.approx(0.013549, rel=1e-4)
    assert k.dt == pytest.approx(0.054, rel=1e-12)          # 0.6 dx^2/4 cap active
    assert k.total_step == 164011377
    assert k.step_fields == 1366761
    assert k.step_pull_back == 1025
    assert k.ratio_dt_dtT == 1
    assert (k.NxT, k.NxT_all, k.NyT, k.NyT_all) == (6669, 6720, 85, 88)
    assert k.Thot == pytest.approx(3180.06, abs=0.01)
    assert k.i_left_end_initial == 8629


# ---------------------------------------------------------------------------
# CALPHAD (fun.cu:1628-1720) against SGTE facts (Dinsdale, CALPHAD 15 (1991))
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("el,Tm,dH", [("Ti", 1941.0, 14146.0), ("Nb", 2750.0, 30000.0)])
def test_sgte_melting_points_and_fusion_enthalpy(oracle_mod, el, Tm, dH):
    liq = lambda T: oracle_mod.G(f"G_{el}_Liq", T)
    bcc = lambda T: oracle_mod.G(f"G_{el}_BCC", T)
    # liquid and bcc Gibbs energies cross at the melting point
    assert abs(liq(Tm) - bcc(Tm)) < 0.05
    # enthalpy of fusion = -Tm d(G_liq - G_bcc)/dT at Tm (one-sided, low branch)
    h = 1e-3
    d = ((liq(Tm - h) - bcc(Tm - h)) - (liq(Tm - 2 * h) - bcc(Tm - 2 * h))) / h
    assert -Tm * d == pytest.approx(dH, rel=2e-3)


@pytest.mark.parametrize("name,Tb", [("G_Ti_BCC", 1941.0), ("G_Ti_Liq", 1941.0),
                                     ("G_Nb_BCC", 2750.0), ("G_Nb_Liq", 2750.0)])
def test_sgte_piecewise_continuity(oracle_mod, name, Tb):
    g = lambda T: oracle_mod.G(name, T)
    assert abs(g(Tb - 1e-9) - g(Tb)) < 0.2          # J/mol
    # entropy (first derivative) continuous too
    h = 1e-4
    s_lo = (g(Tb - h) - g(Tb - 2 * h)) / h
    s_hi = (g(Tb + 2 * h) - g(Tb + h)) / h
    assert abs(s_lo - s_hi) < 1e-2


def test_sgte_out_of_range_is_nan(oracle_mod):
    assert math.isnan(oracle_mod.G("G_Ti_BCC", 1000.0))
    assert math.isnan(oracle_mod.G("G_Ti_Liq", 4100.0))
    assert math.isnan(oracle_mod.G("G_Nb_BCC", 200.0))


@pytest.mark.parametrize("L1", [0.0, 1500.0])
@pytest.mark.parametrize("T,c", [(1900.0, 0.3), (2161.59, 0.6325), (2400.0, 0.9), (1960.0, 0.05)])
def test_chemical_potential_is_dG_dc(rs, oracle_mod, L1, T, c):
    p = rs.reference_params(768, 832)
    p.L1_liq = L1
    p.L1_bcc = -0.5 * L1
    h = 1e-6
    for G, mu in ((oracle_mod.Gl, oracle_mod.Glc), (oracle_mod.Gs, oracle_mod.Gsc)):
        fd = (G(p, T, c + h) - G(p, T, c - h)) / (2 * h)
        assert mu(p, T, c) == pytest.approx(fd, rel=2e-7, abs=1e-4)


def test_ideal_mixing_branches(rs, oracle_mod):
    p = rs.reference_params(768, 832)
    T = 2000.0
    # c -> 0: only the (1-c)ln(1-c) entropy term survives, no log(c/(1-c))
    c = 1e-9
    exact = p.RR * T * ((1 - c) * math.log(1 - c)) + (1 - c) * oracle_mod.G("G_Nb_Liq", T) + \
        c * oracle_mod.G("G_Ti_Liq", T) + c * (1 - c) * p.L0_liq
    assert oracle_mod.Gl(p, T, c) == pytest.approx(exact, rel=1e-14)
    mu = oracle_mod.G("G_Ti_Liq", T) - oracle_mod.G("G_Nb_Liq", T) + p.L0_liq * (1 - 2 * c)
    assert oracle_mod.Glc(p, T, c) == pytest.approx(mu, rel=1e-14)


def test_common_tangent_at_reference_liquidus(rs, oracle_mod):
    # hand-derived pin: tangent at 2161.59 K gives c_L = 0.6317, c_S = 0.5053; the
    # reference's own cinf/cs (fun.cu:62) are close to it
    p = rs.reference_params(768, 832)
    st, cl, cs = oracle_mod.common_tangent(p, 2161.59, 0.63, 0.5)
    assert st == 0
    assert cl == pytest.approx(0.6317, abs=2e-4)
    assert cs == pytest.approx(0.5053, abs=2e-4)
    assert cl == pytest.approx(0.632514105, abs=2e-3)
    # equal potentials and intercepts
    mu = oracle_mod.Glc(p, 2161.59, cl)
    assert oracle_mod.Gsc(p, 2161.59, cs) == pytest.approx(mu, abs=1e-6)
    assert oracle_mod.Gl(p, 2161.59, cl) - mu * cl == pytest.approx(
        oracle_mod.Gs(p, 2161.59, cs) - mu * cs, abs=1e-6)


# ---------------------------------------------------------------------------
# anisotropic operator (fun.cu:1030-1139)
# ---------------------------------------------------------------------------
def _params_eps(rs, eps, epk=0.0, w=(2 / 3, 1 / 3), dx=0.6):
    p = rs.reference_params(768, 832)
    p.eps1, p.epk1, p.dx = eps, epk, dx
    p.lap_w_axis, p.lap_w_diag = w
    return p


@pytest.mark.parametrize("theta", [0.0, 0.3, math.pi / 8, 1.0, 2.5])
def test_anisotropy_strength_closed_form(rs, oracle_mod, theta):
    # linear phi along angle theta: a_s = 1 + eps cos 4theta, a_k likewise,
    # and the operator vanishes (no curvature)
    p = _params_eps(rs, 0.05, 0.13)
    x = np.arange(3) - 1.0
    X, Y = np.meshgrid(x, x)
    ph = 0.1 * (math.cos(theta) * X + math.sin(theta) * Y) * p.dx
    v, a_s, a_k = oracle_mod.anisotropy(p, ph)
    assert a_s == pytest.approx(1 + 0.05 * math.cos(4 * theta), abs=1e-13)
    assert a_k == pytest.approx(1 + 0.13 * math.cos(4 * theta), abs=1e-13)
    assert abs(v) < 1e-12


PI_REF = 3.1415927  # fun.cu:59


def test_orientation_is_a_rotation(oracle_mod):
    assert np.array_equal(oracle_mod.orientation(0.0, 0.0, 0.0), np.eye(3))
    for ang in ((10.0, 15.0, 25.0), (45.0, 0.0, 30.0), (3.0, 44.0, 7.0)):
        R = oracle_mod.orientation(*ang)
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-15)
        assert np.linalg.det(R) == pytest.approx(1.0, abs=1e-15)
    # alpha = beta = 0: a rotation by gamma about z (fun.cu:750-771)
    g = 25.0 * PI_REF / 180.0
    Rz = np.array([[math.cos(g), -math.sin(g), 0.0], [math.sin(g), math.cos(g), 0.0],
                   [0.0, 0.0, 1.0]])
    assert np.allclose(oracle_mod.orientation(0.0, 0.0, 25.0), Rz, atol=1e-16)


@pytest.mark.parametrize("theta", [0.0, 0.3, 1.0, 2.5])
def test_anisotropy_of_a_rotated_grain(rs, oracle_mod, theta):
    """Linear phi along theta: the rotated normal n' = R (cos, sin, 0) enters
    a_s = 1 - 3 eps + 4 eps sum n'_k^4. About z (gamma): the 4-fold pattern
    turns, a_s = 1 + eps cos 4(theta + gamma); about y (beta): n' = (cos b
    cos t, sin t, -sin b cos t)."""
    eps, epk = 0.05, 0.13
    p = _params_eps(rs, eps, epk)
    x = np.arange(3) - 1.0
    X, Y = np.meshgrid(x, x)
    ph = 0.1 * (math.cos(theta) * X + math.sin(theta) * Y) * p.dx
    g = 25.0 * PI_REF / 180.0
    v, a_s, a_k = oracle_mod.anisotropy(p, ph, oracle_mod.orientation(0.0, 0.0, 25.0))
    assert a_s == pytest.approx(1 + eps * math.cos(4 * (theta + g)), abs=1e-13)
    assert a_k == pytest.approx(1 + epk * math.cos(4 * (theta + g)), abs=1e-13)
    assert abs(v) < 1e-12
    b = 15.0 * PI_REF / 180.0
    n = (math.cos(b) * math.cos(theta), math.sin(theta), -math.sin(b) * math.cos(theta))
    np4 = sum(q ** 4 for q in n)
    _, a_s, a_k = oracle_mod.anisotropy(p, ph, oracle_mod.orientation(0.0, 15.0, 0.0))
    assert a_s == pytest.approx(1 - 3 * eps + 4 * eps * np4, abs=1e-13)
    assert a_k == pytest.approx(1 - 3 * epk + 4 * epk * np4, abs=1e-13)


def test_anisotropy_isotropic_limits(rs, oracle_mod):
    rng = np.random.default_rng(1)
    ph = rng.uniform(-1, 1, (3, 3))
    dx = 0.6
    five = (ph[1, 2] + ph[1, 0] + ph[2, 1] + ph[0, 1] - 4 * ph[1, 1]) / dx ** 2
    diag = (ph[0, 0] + ph[0, 2] + ph[2, 0] + ph[2, 2] - 4 * ph[1, 1]) / (2 * dx ** 2)
    v, _, _ = oracle_mod.anisotropy(_params_eps(rs, 0.0, w=(1.0, 0.0)), ph)
    assert v == pytest.approx(five, rel=1e-13)
    v, _, _ = oracle_mod.anisotropy(_params_eps(rs, 0.0), ph)
    assert v == pytest.approx(2 / 3 * five + 1 / 3 * diag, rel=1e-13)


def test_anisotropy_matches_flux_form(rs, oracle_mod):
    """The discrete operator converges to div(a^2 grad phi) + d_x(-a a' phi_y)
    + d_y(a a' phi_x) with a(theta) = 1 + eps cos 4 theta, evaluated here
    independently in flux form on a fine grid (second order)."""
    eps = 0.05
    errs = []
    for h in (0.02, 0.01):
        p = _params_eps(rs, eps, dx=h)
        n = int(round(1.0 / h))
        x = (np.arange(-2 * n, 2 * n + 1)) * h
        X, Y = np.meshgrid(x, x)
        R0 = 1.3
        phi = -np.tanh((np.hypot(X - 0.17, Y + 0.11) - R0) / 0.9)

        def a_and_da(px, py):
            th = np.arctan2(py, px)
            return 1 + eps * np.cos(4 * th), -4 * eps * np.sin(4 * th)

        # x-faces (i+1/2)
        px = (phi[:, 1:] - phi[:, :-1]) / h
        py = ((phi[2:, 1:] + phi[2:, :-1]) - (phi[:-2, 1:] + phi[:-2, :-1])) / (4 * h)
        px = px[1:-1]
        a, da = a_and_da(px, py)
        Fx = a * a * px - a * da * py
        # y-faces (j+1/2)
        py2 = (phi[1:, :] - phi[:-1, :]) / h
        px2 = ((phi[1:, 2:] + phi[:-1, 2:]) - (phi[1:, :-2] + phi[:-1, :-2])) / (4 * h)
        py2 = py2[:, 1:-1]
        a2, da2 = a_and_da(px2, py2)
        Fy = a2 * a2 * py2 + a2 * da2 * px2
        div = (Fx[:, 1:] - Fx[:, :-1]) / h
        div = div + 0.0
        divy = (Fy[1:, :] - Fy[:-1, :]) / h
        ref = div + divy   # interior grid (rows 1..-2, cols 1..-2)
        # oracle at a set of probe points on the interface
        pts = [(2 * n + int(0.9 * n), 2 * n + int(0.4 * n)), (2 * n - int(0.6 * n), 2 * n + int(1.1 * n)),
               (2 * n + int(0.2 * n), 2 * n - int(1.2 * n))]
        e = 0.0
        for (jj, ii) in pts:
            v, _, _ = oracle_mod.anisotropy(p, phi[jj - 1:jj + 2, ii - 1:ii + 2])
            e = max(e, abs(v - ref[jj - 1, ii - 1]) / max(1.0, abs(ref[jj - 1, ii - 1])))
        errs.append(e)
    assert errs[1] < 5e-3
    assert errs[1] < errs[0] / 2.5   # second-order convergence


@pytest.mark.parametrize("preset", ["config", "reference"])
def test_phi_step_is_allen_cahn_and_keeps_the_tanh_profile(rs, oracle_mod, preset):
    """compute_phi + compute_psi (fun.cu:1143-1273) as a whole step. With no
    anisotropy (eps1 = epk1 = 0: a_s = a_k = 1) and no driving force (v0 so
    large that (Gl - Gs) g'(phi) / (2 hh v0) vanishes; v0 cancels from the
    solute equation, whose denominator h0 v0 = R Tmelt) the update is
    explicit Allen-Cahn, phi + dt (lap phi + phi - phi^3); on a field uniform
    in y the 5-point (configuration) and the 9-point (reference) Laplacians
    both reduce to the 1D second difference, evaluated here with numpy. The
    planar profile -tanh(x / sqrt 2) solves phi'' + phi - phi^3 = 0, so one
    step moves it by the second difference's truncation error,
    dt dx^2 / 12 times the largest fourth derivative to leading order."""
    if preset == "config":
        p = rs.config_params(128, tfc=False)
    else:
        p = rs.reference_params(512, 32)
    p.eps1 = 0.0
    p.epk1 = 0.0
    p.v0 = 1e200
    st = oracle_mod.State(p)
    if preset == "config":
        st.init_config(1)
    else:
        st.init_reference()
    k = st.constants()
    k.step = 1                        # no step-0 pull-back
    st.set_constants(k)
    dx, dt = p.dx, k.dt
    i0 = p.Nx // 2 + 0.37
    x = (np.arange(p.Nx) - i0) * dx
    phi = np.tile(-np.tanh(x / np.sqrt(2.0)), (p.Ny, 1))
    st.put(RS_FIELD_PSI, phi)
    st.put(RS_FIELD_PHI, phi)
    st.step(1)
    got = st.get(RS_FIELD_PSI)[1:-1, 1:-1]
    d2 = (phi[:, 2:] - 2.0 * phi[:, 1:-1] + phi[:, :-2]) / (dx * dx)
    p0 = phi[:, 1:-1]
    want = (p0 + dt * (d2 + p0 - p0 ** 3))[1:-1]
    assert np.max(np.abs(got - want)) <= 1e-13
    # fourth derivative of -tanh(x / sqrt 2): tanh''''(u) / 4 with
    # tanh''''(u) = 8 t (1 - t^2) (2 - 3 t^2), t = tanh u
    t = np.tanh(np.linspace(-4.0, 4.0, 200001))
    d4 = float(np.max(np.abs(8.0 * t * (1.0 - t * t) * (2.0 - 3.0 * t * t)))) / 4.0
    bound = dt * dx * dx / 12.0 * d4
    moved = float(np.max(np.abs(got - p0[1:-1])))
    assert 0.5 * bound <= moved <= 1.25 * bound, (moved, bound)


# ---------------------------------------------------------------------------
# whole-step invariants (oracle state)
# ---------------------------------------------------------------------------
def test_mass_conserved_without_gradient(rs, oracle_mod):
    # uniform T (G = 0, frozen), periodic y: the reference's face fluxes are
    # antisymmetric up to summation order and the x ghosts carry no flux
    p = rs.config_params(64, tfc=False)
    p.GG = 0.0
    p.q_solid = 0.0
    s = oracle_mod.State(p)
    s.init_config(3)
    m0 = s.solute_mass()
    s.step(200)
    c = s.get(RS_FIELD_C)
    assert c.min() > 1e-6                 # the clamp never acted
    assert abs(s.solute_mass() - m0) / m0 < 1e-13


def test_solute_sinusoid_decay_rate(rs, oracle_mod):
    """Bulk liquid (psi = -1), uniform T, small sinusoid in c along y: the
    explicit scheme multiplies the mode by 1 - dt*D_eff*(4/dx^2) sin^2(k dx/2)
    per step with D_eff = Dl/(Gamma muk0) (T/Tmelt - 2 L0 c(1-c)/(R Tmelt))."""
    p = rs.config_params(64, tfc=False)
    p.GG = 0.0
    s = oracle_mod.State(p)
    s.init_config(1)
    Ny, Nx = p.Ny, p.Nx
    psi = -np.ones((Ny, Nx))
    c0, A = 0.9, 1e-7
    j = np.arange(Ny)
    kk = 2 * np.pi / (Ny - 2)
    col = c0 + A * np.sin(kk * (j - 1))
    col[0], col[-1] = col[Ny - 2], col[1]
    c = np.repeat(col[:, None], Nx, axis=1)
    s.put(RS_FIELD_PSI, psi)
    s.put(1 + 1, psi)      # phi = psi = -1
    s.put(RS_FIELD_C, c)
    k = s.constants()
    k.step = 1             # avoid the step-0 pull-back
    s.set_constants(k)
    n = 50
    s.step(n)
    out = s.get(RS_FIELD_C)[1:-1, Nx // 2]
    T = p.Tmelt - p.deltaT0
    Deff = p.Dl / (p.Gamma * p.muk0) * (T / p.Tmelt - 2 * p.L0_liq * c0 * (1 - c0) / (p.RR * p.Tmelt))
    lam = 1 - k.dt * Deff * 4 / p.dx ** 2 * math.sin(kk * p.dx / p.dx / 2) ** 2 / 1.0
    # wavenumber per cell: kk (cells); sin^2(kk/2) with dx cancelling
    lam = 1 - k.dt * Deff * 4 / p.dx ** 2 * math.sin(kk / 2) ** 2
    amp = np.dot(out - c0, np.sin(kk * (np.arange(1, Ny - 1) - 1))) / ((Ny - 2) / 2)
    assert amp / A == pytest.approx(lam ** n, rel=2e-4)
    assert lam ** n < 0.999   # the test sees real decay


def test_latent_heat_conserves_energy(rs, oracle_mod):
    """Two runs differing only in cp: after one step the coarse T difference is
    the latent-heat source, whose sum over the coarse grid times the 100 fine
    cells per coarse cell equals Lcp * 1/2 * sum(delta psi) over the fine rows
    that the 10x10 blocks cover."""
    pA = rs.config_params(128, tfc=True)
    pB = pA.copy()
    pB.cp = 1e200
    sA, sB = oracle_mod.State(pA), oracle_mod.State(pB)
    for s in (sA, sB):
        s.init_config(11)
        k = s.constants()
        k.step = 1
        s.set_constants(k)
    psi0 = sA.get(RS_FIELD_PSI)
    sA.step(1)
    sB.step(1)
    psi1 = sA.get(RS_FIELD_PSI)
    dT = sA.get(RS_FIELD_T) - sB.get(RS_FIELD_T)
    kA = sA.constants()
    rows = 10 * (kA.NyT - 2)
    dpsi = (psi1 - psi0)[1:rows + 1, 1:pA.Nx - 1].sum()
    Lcp = pA.latentHeat / pA.cp
    got = dT[1:kA.NyT - 1, 1:kA.NxT - 1].sum() * 100 * kA.ratio_dt_dtT / Lcp
    assert abs(dpsi) > 1e-3
    assert got == pytest.approx(0.5 * dpsi, rel=1e-9)


def test_temperature_linear_profile_translates(rs, oracle_mod):
    """All-liquid state (no latent heat): a linear coarse T profile is
    advected by dtT * V tau0/(S W0) * dT/dx per step, diffusion adds nothing;
    fine T interpolated from it stays exactly linear."""
    p = rs.config_params(128, tfc=True)
    s = oracle_mod.State(p)
    s.init_config(5)
    psi = -np.ones((p.Ny, p.Nx))
    s.put(RS_FIELD_PSI, psi)
    s.put(2, psi)
    s.put(RS_FIELD_C, np.full((p.Ny, p.Nx), 0.9))
    k = s.constants()
    cellG = p.S * p.W0 * p.dx * 10 * p.GG
    Tf = s.get(4)
    i = np.arange(p.Nx)
    X = (i - 1.0 + k.i_left_end) / 10
    assert np.allclose(Tf, k.Tcold + (X - 1) * cellG, rtol=0, atol=1e-9)
    k.step = 1
    s.set_constants(k)
    T0 = s.get(RS_FIELD_T)
    s.step(1)
    T1 = s.get(RS_FIELD_T)
    tau0 = oracle_mod.tau0(p)
    vadv = p.Vp * tau0 / (p.S * p.W0)
    expect = k.dtT * vadv * cellG / (p.dx * 10)
    inner = (T1 - T0)[1:k.NyT - 1, 3:k.NxT - 3]
    assert np.allclose(inner, expect, rtol=1e-6, atol=1e-9)
    assert np.all(T1[:k.NyT, 0] == k.Tcold) and np.all(T1[:k.NyT, k.NxT - 1] == k.Thot)


# ---------------------------------------------------------------------------
# tip fit (fun.cu:2628-2954) on a synthetic parabolic dendrite
# ---------------------------------------------------------------------------
def test_tip_fit_recovers_parabola(rs, oracle_mod):
    p = rs.reference_params(512, 272)
    Ny, Nx = p.Ny, p.Nx
    x0, y0, R = 300.37, 131.6, 40.0
    j = np.arange(Ny)[:, None]
    i = np.arange(Nx)[None, :]
    X = x0 - (j - y0) ** 2 / (2 * R)
    # linear psi across the interface: x(psi) is linear, so the quartic fit is
    # exact and the parabola through exact points is recovered exactly
    psi = np.clip(-(i - X) * 0.2, -1.0, 1.0)
    st, t = oracle_mod.precise_tip(p, psi)
    assert st == 0
    assert t["j"] in (131, 132)
    assert t["x"] == pytest.approx(x0, abs=1e-8)
    assert t["y"] == pytest.approx(y0, abs=1e-7)
    # normal equations in raw row indices (y ~ 130) are ill-conditioned
    assert t["R"] == pytest.approx(R * p.dx * p.W0 * p.S, rel=1e-6)
    # tanh profile: quartic truncation error only
    psi = -np.tanh((i - X) * 0.6 / np.sqrt(2.0))
    st, t = oracle_mod.precise_tip(p, psi)
    assert st == 0
    assert t["x"] == pytest.approx(x0, abs=3e-2)
    assert t["R"] == pytest.approx(R * p.dx * p.W0 * p.S, rel=5e-2)


def test_tip_column_brute_force(rs, oracle_mod):
    p = rs.reference_params(512, 272)
    s = oracle_mod.State(p)
    s.init_reference()
    rng = np.random.default_rng(2)
    psi = -np.ones((p.Ny, p.Nx))
    L = rng.integers(5, 400, size=p.Ny)
    for jj in range(p.Ny):
        psi[jj, :L[jj]] = 0.5
    s.put(RS_FIELD_PSI, psi)
    assert s.tip_column() == L.max() - 1


# ---------------------------------------------------------------------------
# antitrapping + solute-trapping extension (absent from the reference):
# closed-form properties of the face flux and of the whole correction
# ---------------------------------------------------------------------------
AT = 1.0 / (2.0 * math.sqrt(2.0))


def _at_params(rs, k=0.8):
    p = rs.config_params(64, False)
    p.at_coef = AT
    p.k_part = k
    return p


def test_antitrap_flux_closed_forms(rs, oracle_mod):
    p = _at_params(rs)
    dt = 0.1
    # flat interface, solid (phi = 1) at P, liquid at Q: normal into the
    # liquid is +x, so a solidifying face (dphi/dt > 0) drives solute +x
    pP, pQ, rate = 0.3, -0.2, 0.05
    nP, nQ = pP + rate * dt, pQ + rate * dt
    c = 0.9
    J = oracle_mod.at_flux(p, 0.0, dt, pP, pQ, nP, nQ, c, c)
    hf = 0.5 * (1.0 + 0.5 * (pP + pQ))
    cl = c / (1.0 - hf * (1.0 - p.k_part))
    assert J == pytest.approx(AT * (1.0 - p.k_part) * cl * rate, rel=1e-14)
    assert J > 0.0
    # mirror image: the same face seen from the other side carries -J
    assert oracle_mod.at_flux(p, 0.0, dt, pQ, pP, nQ, nP, c, c) == pytest.approx(-J, rel=1e-14)
    # melting reverses it; no interface motion, no current
    assert oracle_mod.at_flux(p, 0.0, dt, pP, pQ, pP - rate * dt, pQ - rate * dt, c, c) == \
        pytest.approx(-J, rel=1e-12)
    assert oracle_mod.at_flux(p, 0.0, dt, pP, pQ, pP, pQ, c, c) == 0.0
    # no partitioning (k = 1): no current
    assert oracle_mod.at_flux(_at_params(rs, 1.0), 0.0, dt, pP, pQ, nP, nQ, c, c) == 0.0
    # c_L limits: all liquid (h = 0) gives c, all solid (h = 1) gives c / k
    Jl = oracle_mod.at_flux(p, 0.0, dt, -1.0, -1.0 - 1e-3, -1.0 + 1e-3, -1.0, c, c)
    Js = oracle_mod.at_flux(p, 0.0, dt, 1.0 + 1e-3, 1.0, 1.0 + 2e-3, 1.0 + 1e-3, c, c)
    assert Jl == pytest.approx(AT * (1 - p.k_part) * c * 1e-3 / dt, rel=2e-3)
    assert Js == pytest.approx(AT * (1 - p.k_part) * (c / p.k_part) * 1e-3 / dt, rel=2e-3)
    # a face tangential to the interface (g_n = 0) carries nothing
    assert oracle_mod.at_flux(p, 0.0, dt, 0.1, 0.1, 0.2, 0.2, c, c, 0.5, 0.5, 0.0, 0.0) == 0.0
    # an oblique interface: the normal component is g_n / |g|
    Jo = oracle_mod.at_flux(p, 0.0, dt, pP, pQ, nP, nQ, c, c, 0.4, 0.4, 0.0, 0.0)
    gn = (pQ - pP) / p.dx
    gt = 0.8 / (4.0 * p.dx)
    assert Jo == pytest.approx(J * abs(gn) / math.hypot(gn, gt), rel=1e-13)


def test_antitrap_trapping_factor(rs, oracle_mod):
    p = _at_params(rs)
    dt, pP, pQ, rate = 0.1, 0.3, -0.2, 0.05
    nP, nQ = pP + rate * dt, pQ + rate * dt
    J0 = oracle_mod.at_flux(p, 0.0, dt, pP, pQ, nP, nQ, 0.9, 0.9)
    V = rate / abs((pQ - pP) / p.dx)            # interface speed dphi/dt / |grad phi|
    # at V = V_D the current halves: the partition k(V) = (k + 1)/2
    assert oracle_mod.at_flux(p, V, dt, pP, pQ, nP, nQ, 0.9, 0.9) == pytest.approx(0.5 * J0, rel=1e-13)
    # V_D -> infinity: no trapping; V_D -> 0: complete trapping
    assert oracle_mod.at_flux(p, 1e12, dt, pP, pQ, nP, nQ, 0.9, 0.9) == pytest.approx(J0, rel=1e-10)
    assert abs(oracle_mod.at_flux(p, 1e-12, dt, pP, pQ, nP, nQ, 0.9, 0.9)) < 1e-9 * abs(J0)


def test_antitrap_run_conserves_and_moves_solute_to_liquid(rs, oracle_mod):
    """A growing front (uniform T, so no flux leaves the box): the corrected
    run conserves solute to rounding, and against the uncorrected run it has
    moved solute from the solid side to the liquid side."""
    p = rs.config_params(96, False)
    p.GG = 0.0
    p.q_solid = 0.0
    runs = {}
    for at in (0.0, AT):
        q = p.copy()
        q.at_coef = at
        st = oracle_mod.State(q)
        st.init_config(4)
        m0 = st.solute_mass()
        st.step(300)
        runs[at] = (st.get(RS_FIELD_C), st.get(RS_FIELD_PSI), (st.solute_mass() - m0) / m0)
        st.close()
    c0, psi0, d0 = runs[0.0]
    c1, psi1, d1 = runs[AT]
    assert abs(d1) < 1e-13 and abs(d0) < 1e-13
    dc = (c1 - c0)[1:-1, 1:-1]
    x = np.arange(1, p.Nx - 1)[None, :]
    # the solid is at small i: the correction's first moment points to +x
    assert float(np.sum(dc * x)) > 0.0
    assert abs(float(np.sum(dc))) < 1e-10 * float(np.sum(np.abs(dc)))
