"""Pins of the oracle's state functions against facts fixed outside the
oracle: closed forms of the initial profiles, exact reproduction of linear
fields by the bilinear interpolation, brute-force block assignment of the
latent heat, exact shift and solute bookkeeping of the pull-back, and the
tip-velocity and tip-temperature bookkeeping of set_tip_output from known
displacements and fields. Each check is chosen so that a plausible misreading
of the reference (an index off by one, a wrong sign of the dxT row offset,
`>` for `>=`, math.pi for the float-precision PI, the tip temperature sampled
at the tip row) fails it.

Reference lines: initial fun.cu:983-1021, initial_temperature :1023-1028,
pull_back :1469-1492 with set_pull_back :697-704, integrate_phi :1494-1535,
interpolate_temperature :1580-1608, set_tip_output :3024-3072.
"""
import math

import numpy as np
import pytest

from rsolid.abi import RS_FIELD_C, RS_FIELD_PHI, RS_FIELD_PSI, RS_FIELD_T

PI_REF = 3.1415927  # the reference's float-precision PI macro (fun.cu:59)


def _splitmix64(state):
    """splitmix64 (Steele, Lea, Flood 2014) on Python integers."""
    m = (1 << 64) - 1
    state = (state + 0x9E3779B97F4A7C15) & m
    z = state
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return state, z ^ (z >> 31)


def _front_from_psi(psi, dx, band=0.95):
    """Per row, the front position implied by psi = -tanh((i - x0) dx / sqrt 2)
    at every cell with |psi| < band (x0 = i + sqrt2 atanh(psi) / dx), and the
    spread of those estimates."""
    Ny, Nx = psi.shape
    ii = np.arange(Nx)
    est, spread = np.empty(Ny), np.empty(Ny)
    for j in range(Ny):
        m = np.abs(psi[j]) < band
        x0 = ii[m] + math.sqrt(2.0) * np.arctanh(psi[j, m]) / dx
        est[j] = x0.mean()
        spread[j] = x0.max() - x0.min()
    return est, spread


# ---------------------------------------------------------------- initial
def test_init_reference_front_and_profiles(rs, oracle_mod):
    """initial: psi = phi = -tanh((i - itip) dx / sqrt 2) with
    itip = i_tip_target + 0.5 sin(2 PI j / Ny), PI = 3.1415927 (not pi);
    c = cinf (0.3184 tanh(((i - itip) dx - 0.08) / 0.67) + 0.6816), i.e.
    c -> cinf in the far liquid and 0.3632 cinf in the far solid, with the
    c profile centred 0.08 / dx cells behind the psi front; T linear
    T = Tcold + (iT - 1) S W0 dx 10 GG."""
    p = rs.reference_params(512, 272)
    st = oracle_mod.State(p)
    st.init_reference()
    psi, phi, c = st.get(RS_FIELD_PSI), st.get(RS_FIELD_PHI), st.get(RS_FIELD_C)
    Ny, Nx = psi.shape
    assert np.array_equal(psi, phi)
    est, spread = _front_from_psi(psi, p.dx)
    jj = np.arange(Ny)
    want = p.i_tip_target + 0.5 * np.sin(2.0 * PI_REF * jj / Ny)
    assert np.max(spread) < 1e-9
    assert np.max(np.abs(est - want)) < 1e-10
    # the float-precision PI is visible: math.pi moves rows by > 1e-8
    alt = p.i_tip_target + 0.5 * np.sin(2.0 * math.pi * jj / Ny)
    assert np.max(np.abs(est - alt)) > 1e-8
    # c: far fields and the profile's own centre
    assert c[:, -1] == pytest.approx(np.full(Ny, p.cinf), rel=1e-12)
    assert c[:, 0] == pytest.approx(np.full(Ny, 0.3632 * p.cinf), rel=1e-12)
    u = (c / p.cinf - 0.6816) / 0.3184
    for j in (0, 17, Ny // 4, Ny - 1):
        m = np.abs(u[j]) < 0.9
        xc = np.arange(Nx)[m] - (np.arctanh(u[j, m]) * 0.67 + 0.08) / p.dx
        assert np.max(np.abs(xc - want[j])) < 1e-9
    # coarse T: linear in iT from Tcold at iT = 1, uniform in jT
    k = st.constants()
    T = st.get(RS_FIELD_T)
    slope = p.S * p.W0 * p.dx * p.ratio_dxT_dx * p.GG
    iT = np.arange(k.NxT_all)
    assert T[3] == pytest.approx(k.Tcold + (iT - 1) * slope, rel=1e-14)
    assert np.array_equal(T[0], T[k.NyT_all - 1])
    st.close()


def test_init_config_front_from_the_seed_stream(rs, oracle_mod):
    """Seeded configuration front: row j = 1..Ny-2 in order draws
    u = (splitmix64 >> 11) 2^-53 and sits at x0_j = i_tip_target + 2u - 1;
    ghost rows are the periodic images; c = c_sol0 in the far solid and
    c_liq0 in the far liquid."""
    p = rs.config_params(256, False)
    seed = 987654321
    st = oracle_mod.State(p)
    st.init_config(seed)
    psi, c = st.get(RS_FIELD_PSI), st.get(RS_FIELD_C)
    Ny = psi.shape[0]
    s, want = seed, np.empty(Ny)
    for j in range(1, Ny - 1):
        s, z = _splitmix64(s)
        want[j] = p.i_tip_target + 2.0 * ((z >> 11) * 2.0 ** -53) - 1.0
    want[0], want[Ny - 1] = want[Ny - 2], want[1]
    est, spread = _front_from_psi(psi, p.dx)
    assert np.max(spread) < 1e-9
    assert np.max(np.abs(est - want)) < 1e-10
    assert np.all(np.abs(want - p.i_tip_target) <= 1.0)
    assert c[:, 0] == pytest.approx(np.full(Ny, p.c_sol0), rel=1e-12)
    assert c[:, -1] == pytest.approx(np.full(Ny, p.c_liq0), rel=1e-12)
    st.close()


# -------------------------------------------------- interpolate_temperature
@pytest.mark.parametrize("ile", [1, 7, 123])
def test_interpolation_reproduces_a_plane_with_the_row_offset(rs, oracle_mod, ile):
    """A coarse T that is a plane T = A + B iT + C jT is reproduced exactly
    at every fine cell, at the coarse position the reference maps the cell
    to: x = (i - 1 + i_left_end) / 10 and y = (j - 1 + dxT) / 10 with
    dxT = 10 dx used as a row offset (fun.cu:1590-1594). A sign or unit slip
    in either offset moves the plane by C dxT / 10 or B / 10."""
    p = rs.config_params(256, True)
    st = oracle_mod.State(p)
    st.init_config(3)
    k = st.constants()
    A, B, C = 1900.0, 0.37, -1.9
    jT, iT = np.mgrid[0:k.NyT_all, 0:k.NxT_all]
    st.put(RS_FIELD_T, A + B * iT + C * jT)
    # the read of row jT + 1 at the top carries weight 0 only where y = 0;
    # compare the rows whose four coarse values are all inside the grid
    Tf = st.interpolate_at(ile)
    Ny, Nx = Tf.shape
    dxT = p.dx * p.ratio_dxT_dx
    j, i = np.mgrid[0:Ny, 0:Nx]
    want = A + B * (i - 1.0 + ile) / 10.0 + C * (j - 1.0 + dxT) / 10.0
    rows = ((j - 1 + dxT) // 10 + 1) < k.NyT_all
    assert np.max(np.abs(Tf - want)[rows]) < 1e-10
    # the offset matters: without it the plane is off by |C| dxT / 10
    wrong = A + B * (i - 1.0 + ile) / 10.0 + C * (j - 1.0) / 10.0
    assert np.min(np.abs(Tf - wrong)[rows]) > 0.5 * abs(C) * dxT / 10.0
    st.close()


# ------------------------------------------------------------ integrate_phi
def test_latent_heat_of_one_cell_lands_in_one_block(rs, oracle_mod):
    """A psi change of 1 in a single fine cell (i, j) gives 0.5 in exactly
    one coarse cell: iT = floor((i - 1 + i_left_end) / 10) and
    jT = floor((j - 1) / 10) + 1 (rows 1 + 10 (jT - 1) .. 10 jT, no dxT
    offset), and nothing for rows beyond 10 (NyT - 2) (the reference drops
    them when (Ny - 2) % 10 != 0)."""
    p = rs.config_params(256, True)
    st = oracle_mod.State(p)
    st.init_config(3)
    k = st.constants()
    Ny, Nx = p.Ny, p.Nx
    rng = np.random.default_rng(5)
    z = np.zeros((Ny, Nx))
    ile = 37
    cells = [(1, 1), (Nx - 2, 1), (1, 10 * (k.NyT - 2)), (Nx - 2, Ny - 2), (5, 11), (13, 10)]
    cells += [(int(a), int(b)) for a, b in zip(rng.integers(1, Nx - 1, 40), rng.integers(1, Ny - 1, 40))]
    dropped = 0
    for i, j in cells:
        nxt = z.copy()
        nxt[j, i] = 1.0
        out = st.integrate_phi_at(z, nxt, ile)
        nz = np.argwhere(out != 0.0)
        if j > 10 * (k.NyT - 2):
            assert nz.size == 0, (i, j)
            dropped += 1
            continue
        assert nz.shape == (1, 2), (i, j, nz)
        jT, iT = nz[0]
        assert (iT, jT) == ((i - 1 + ile) // 10, (j - 1) // 10 + 1), (i, j)
        assert out[jT, iT] == 0.5
    assert 10 * (k.NyT - 2) < Ny - 2   # 256 rows: 4 rows have no block
    st.close()


# ------------------------------------------------------------ pull_back
@pytest.mark.parametrize("off", [0, 1, 3, 17, -4])
def test_pull_back_shift_and_solute_bookkeeping(rs, oracle_mod, off):
    """pull_back by `off` columns: every row is read from column
    clamp(i + off, 1, Nx - 2); c takes cinf wherever that source column is
    Nx - 2 or beyond (i + off >= Nx - 2). For off >= 0 the interior solute
    then changes by exactly the injected (off + 1) cinf per row minus the
    shifted-out columns 1..off and the last interior column Nx - 2, which
    `>` in place of `>=` would keep. The run state follows set_pull_back:
    i_offset, i_offset_history += off, i_left_end = i_left_end_initial +
    i_offset_history - int(l_offset_Vp / dx)."""
    p = rs.config_params(256, True)
    st = oracle_mod.State(p)
    st.init_config(3)
    rng = np.random.default_rng(100 + off)
    Ny, Nx = p.Ny, p.Nx
    psi0 = rng.uniform(-1.5, 1.0, (Ny, Nx))
    phi0 = np.maximum(psi0, -1.0)
    c0 = rng.uniform(0.5, 0.9, (Ny, Nx))
    st.put(RS_FIELD_PSI, psi0)
    st.put(RS_FIELD_PHI, phi0)
    st.put(RS_FIELD_C, c0)
    k = st.constants()
    k.l_offset_Vp = 7.3
    k.i_offset_history = 11
    st.set_constants(k)
    m0 = st.solute_mass()
    st.pull_back(off)
    psi1, phi1, c1 = st.get(RS_FIELD_PSI), st.get(RS_FIELD_PHI), st.get(RS_FIELD_C)
    src = np.clip(np.arange(Nx) + off, 1, Nx - 2)
    for j in (0, 1, Ny // 2, Ny - 1):
        for i in range(Nx):
            assert psi1[j, i] == psi0[j, src[i]]
            assert phi1[j, i] == phi0[j, src[i]]
            want_c = p.cinf if i + off >= Nx - 2 else c0[j, src[i]]
            assert c1[j, i] == want_c, (j, i)
    if off >= 0:
        rows = slice(1, Ny - 1)
        out = c0[rows, 1:off + 1].sum() + c0[rows, Nx - 2].sum()
        dm = (off + 1) * (Ny - 2) * p.cinf - out
        assert st.solute_mass() - m0 == pytest.approx(dm, rel=1e-9, abs=1e-9)
    k1 = st.constants()
    assert k1.i_offset == off and k1.i_offset_history == 11 + off
    assert k1.i_left_end == k.i_left_end_initial + 11 + off - int(7.3 / p.dx)
    st.close()


# ------------------------------------------------------------ set_tip_output
def test_tip_velocity_from_a_known_displacement_and_its_clamps(rs, oracle_mod):
    """set_tip_output: V_tip = (x_now - x_tip_history) S W0 / (step_tip tau0
    dt) with x_now = (i_offset_history + x_tip) dx; the first call reports
    NaN and only records x_tip_history; V < 0 is reported as 0.01 m/s and
    V > 10 as 10 m/s. A field moved by d whole columns moves the fitted tip
    by exactly d, and so does the same d booked in i_offset_history."""
    p = rs.config_params(256, False)
    st = oracle_mod.State(p)
    st.init_reference()
    k = st.constants()
    tau0 = oracle_mod.tau0(p)
    # a tip cadence long enough that one column per cadence is below 1 m/s
    k.step_tip = int(math.ceil(p.dx * p.S * p.W0 / (tau0 * k.dt)))
    k.if_begin_tip = 1
    st.set_constants(k)
    t0 = st.tip()
    assert math.isnan(t0["V_tip"])
    unit = p.dx * p.S * p.W0 / (k.step_tip * tau0 * k.dt)   # m/s per column
    assert 0.5 < unit <= 1.0
    psi = st.get(RS_FIELD_PSI)
    d = 3
    moved = np.empty_like(psi)
    moved[:, d:] = psi[:, :-d]
    moved[:, :d] = psi[:, :1]
    st.put(RS_FIELD_PSI, moved)
    t1 = st.tip()
    assert t1["x"] - t0["x"] == pytest.approx(d, abs=1e-9)
    want = d * unit
    assert want < 10.0
    assert t1["V_tip"] == pytest.approx(want, rel=1e-9)
    # the same displacement booked in i_offset_history instead
    k = st.constants()
    k.i_offset_history += 2
    st.set_constants(k)
    assert st.tip()["V_tip"] == pytest.approx(2 * unit, rel=1e-9)
    # a receding tip is reported as 0.01 m/s
    st.put(RS_FIELD_PSI, psi)
    assert st.tip()["V_tip"] == 0.01
    # a jump beyond 10 m/s is clamped to 10
    k = st.constants()
    k.i_offset_history += int(math.ceil(11.0 / unit))
    st.set_constants(k)
    assert st.tip()["V_tip"] == 10.0
    st.close()


def test_tip_temperature_is_read_at_row_zero(rs, oracle_mod):
    """Coupled T: T_tip is the interpolated fine T at (tip.i, j = 0), not at
    the tip row, and delta = (T_tip - Tmelt) / deltaT0. With a coarse T that
    varies along y both the row and the dxT row offset show: the expected
    value is the plane at x = (i - 1 + i_left_end) / 10, y = (dxT - 1) / 10."""
    p = rs.reference_params(512, 272)
    p.ifTFC = 1
    st = oracle_mod.State(p)
    st.init_reference()
    k = st.constants()
    A, B, C = 2150.0, 0.05, 0.8
    jT, iT = np.mgrid[0:k.NyT_all, 0:k.NxT_all]
    st.put(RS_FIELD_T, A + B * iT + C * jT)
    t = st.tip()
    assert t["j"] != 0
    dxT = p.dx * p.ratio_dxT_dx
    want = A + B * (t["i"] - 1.0 + k.i_left_end) / 10.0 + C * (dxT - 1.0) / 10.0
    assert t["T_tip"] == pytest.approx(want, abs=1e-9)
    at_tip_row = want + C * t["j"] / 10.0
    assert abs(t["T_tip"] - at_tip_row) > 0.5 * C * t["j"] / 10.0
    assert t["delta"] == pytest.approx((want - p.Tmelt) / p.deltaT0, abs=1e-12)
    st.close()


def test_frozen_tip_temperature_is_the_imposed_gradient(rs, oracle_mod):
    """Frozen T: T_tip = Tmelt + deltaT0 delta with delta = delta_initial
    + GG (i - i_tip_target + shift) dx S W0 / deltaT0 - GG l_offset S W0 /
    deltaT0 (get_delta, fun.cu:1610-1619): the tip temperature moves by the
    gradient GG times the lab-frame distance the tip moved relative to the
    isotherm, 1 column = dx S W0 nm."""
    p = rs.config_params(256, False)
    st = oracle_mod.State(p)
    st.init_reference()
    a = st.tip()
    k = st.constants()
    k.i_offset_history += 4
    k.l_offset_Vp += 1.5
    st.set_constants(k)
    b = st.tip()
    assert b["i"] == a["i"]
    dT = p.GG * (4 * p.dx - 1.5) * p.S * p.W0
    assert b["T_tip"] - a["T_tip"] == pytest.approx(dT, rel=1e-9)
    assert a["T_tip"] == pytest.approx(p.Tmelt + p.deltaT0 * a["delta"], rel=1e-14)
    st.close()


# ------------------------------------------------------------ solute ledger
def test_solute_ledger_closes_on_a_coupled_run_with_pull_backs(rs, oracle_mod):
    """Every change of the interior solute is one of: a pull-back's shift
    (shifted-out columns, injected cinf), the flux through the x or y
    boundary faces (the ghosts copy c and psi from their images but see
    their own T), or the clamp; the remainder is the rounding of the updates
    (< 1e-13 of the mass). With coupled T the y faces leak (the coarse T
    varies along y); the x and y terms and the pull-backs are all non-zero."""
    p = rs.config_params(256, True)
    st = oracle_mod.State(p)
    st.init_config(7)
    k = st.constants()
    k.step_pull_back = 150
    st.set_constants(k)
    st.step(600)
    L = st.solute_ledger()
    assert L["clamped"] == 0
    assert abs(L["unaccounted"]) < 1e-13 * L["mass"]
    assert abs(L["pull_back"]) > 1.0 and abs(L["edge_x"]) > 1e-4 and abs(L["edge_y"]) > 1e-2
    L2 = st.solute_ledger(reset=True)
    assert L2 == L
    L3 = st.solute_ledger()
    assert L3["mass0"] == L["mass"] and L3["pull_back"] == 0.0 and L3["edge_y"] == 0.0


def test_solute_ledger_frozen_T_has_no_y_leak(rs, oracle_mod):
    """Frozen T depends on x only: the periodic y boundary faces carry equal
    and opposite fluxes (to rounding), the x faces a net flux."""
    p = rs.config_params(128, False)
    st = oracle_mod.State(p)
    st.init_config(3)
    st.step(300)
    L = st.solute_ledger()
    assert abs(L["unaccounted"]) < 1e-13 * L["mass"]
    assert abs(L["edge_y"]) < 1e-12 * abs(L["edge_x"]) + 1e-15
    assert abs(L["edge_x"]) > 1e-4


def test_solute_ledger_books_a_pull_back_exactly(rs, oracle_mod):
    """One pull-back by `off` on random fields changes the interior solute
    by (off + 1)(Ny - 2) cinf minus the shifted-out columns 1..off and
    Nx - 2; the ledger books exactly that."""
    p = rs.config_params(256, True)
    st = oracle_mod.State(p)
    st.init_config(3)
    rng = np.random.default_rng(8)
    Ny, Nx = p.Ny, p.Nx
    c0 = rng.uniform(0.5, 0.9, (Ny, Nx))
    st.put(RS_FIELD_C, c0)
    off = 6
    st.pull_back(off)
    rows = slice(1, Ny - 1)
    want = (off + 1) * (Ny - 2) * p.cinf - c0[rows, 1:off + 1].sum() - c0[rows, Nx - 2].sum()
    L = st.solute_ledger()
    assert L["pull_back"] == pytest.approx(want, rel=1e-11)
    assert abs(L["unaccounted"]) < 1e-13 * L["mass"]
