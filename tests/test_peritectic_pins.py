"""Pins of the CPU restatement of the 3D three-phase peritectic Ti-Ag model
(oracle/tp_oracle.c, reference multi/peritectic_multiGPU/TiAg.cu) against
facts fixed outside it, each chosen so that a plausible misreading fails:

  g_i (TiAg.cu:71)               sum_i g_i = 1 on the simplex; g(1,.,.) = 1,
                                 g(0,.,.) = 0
  gp1, gp2 (TiAg.cu:72-73)       central differences of g
  bulk forces (TiAg.cu:498-514)  central differences of an independently
                                 written free energy (double wells,
                                 multi-well terms, lambda sum_i g_i omega_i)
  compositions (TiAg.cu:58-62)   the equal-grand-potential points of phase
                                 pairs 1-2 and 1-3; the 1-2 driving force
                                 vanishes at that mu on the whole 1-2 edge
  solute update (TiAg.cu:525-533) in a pure phase, phi stays put and a c
                                 sinusoid decays by the explicit-Euler
                                 factor of the 7-point Laplacian times the
                                 phase's mobility (1, Da, Db); exact solute
                                 conservation in a three-phase run
  boundaries (TiAg.cu:458-483, 538-579) every ghost equals its image
  initial (TiAg.cu:411-444)      spherical caps: base disc of radius R on
                                 plane zcut, height R tan(theta/2), volume
                                 pi h^2 (3 r - h) / 3, compositions by phase
The file formats of data2File2D and save/load_break_point round-trip.
"""
import math
import os

import numpy as np
import pytest

from rsolid import peritectic as tp
from rsolid.abi import TP_FIELD_C, TP_FIELD_MU, TP_FIELD_PHI1, TP_FIELD_PHI2


def small_params(Nx=40, Ny=36, Nz=30, Da=0.0, Db=0.0):
    """The model constants on a small box with one cap inside it."""
    p = tp.reference_params(Nx, Ny, Nz)
    p.zcut_from_top = Nz - 10        # zcut = 10
    p.R = 7.0
    p.ncaps = 1
    p.cap_x[0], p.cap_y[0] = Nx / 2.0, Ny / 2.0
    p.Da, p.Db = Da, Db
    return p


def coeffs(p):
    d = tp.derived(p)
    r = np.array([p.A1, -p.A1 - p.A2, -p.A1 + p.A2])
    s = np.array([d["B1"], -d["B1"] - d["B2"], -d["B1"] + d["B2"]])
    return r, s, d


def simplex_points(n, seed):
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < n:
        a, b = rng.uniform(-0.05, 1.05, 2)
        if a + b <= 1.05:
            pts.append((a, b, 1.0 - a - b))
    return pts


# --------------------------------------------------------------- g and its derivatives
def test_g_is_a_partition_of_unity(oracle_mod):
    for p1, p2, p3 in simplex_points(500, 1):
        s = oracle_mod.tp_g(p1, p2, p3) + oracle_mod.tp_g(p2, p3, p1) + oracle_mod.tp_g(p3, p1, p2)
        assert s == pytest.approx(1.0, abs=2e-15)
    for pj in (0.0, 0.3, -0.2):
        assert oracle_mod.tp_g(1.0, pj, -pj) == pytest.approx(1.0, abs=1e-15)
        assert oracle_mod.tp_g(0.0, pj, 1.0 - pj) == 0.0


def test_g_derivatives_are_its_partials(oracle_mod):
    h = 1e-6
    g = oracle_mod.tp_g
    for pi, pj, pk in [(0.3, 0.5, 0.2), (0.8, 0.05, 0.15), (0.1, 0.45, 0.45), (0.6, -0.1, 0.5)]:
        d_i = (g(pi + h, pj, pk) - g(pi - h, pj, pk)) / (2 * h)
        d_j = (g(pi, pj + h, pk) - g(pi, pj - h, pk)) / (2 * h)
        assert oracle_mod.tp_gp1(pi, pj, pk) == pytest.approx(d_i, abs=1e-8)
        assert oracle_mod.tp_gp2(pi, pj, pk) == pytest.approx(d_j, abs=1e-8)


# --------------------------------------------------------------- bulk driving forces
def _free_energy(p, r, s, mu, q):
    """f(p1, p2, p3) at fixed mu, the phases independent: double wells,
    the triple-junction multi-well terms and lambda sum_i g_i (s_i - r_i mu)
    with g_1 = g(p1,p2,p3), g_2 = g(p2,p3,p1), g_3 = g(p3,p1,p2)."""
    p1, p2, p3 = q

    def g(a, bb, c):
        d = c - bb
        return a * a / 4.0 * (15.0 * (1.0 - a) * (1.0 + a - d * d) + a * (9.0 * a * a - 5.0))

    wells = sum(x * x * (1.0 - x) ** 2 for x in q)
    a1, a2, a3, b = p.a1, p.a2, p.a3, p.b
    multi = (a1 * (p1 * p2 ** 2 * p3 ** 2 * (3.0 + b * p1) + 2.0 * p2 ** 3 * p3 ** 3)
             + a2 * (p2 * p1 ** 2 * p3 ** 2 * (3.0 + b * p2) + 2.0 * p1 ** 3 * p3 ** 3)
             + a3 * (p3 * p1 ** 2 * p2 ** 2 * (3.0 + b * p3) + 2.0 * p1 ** 3 * p2 ** 3))
    gs = (g(p1, p2, p3), g(p2, p3, p1), g(p3, p1, p2))
    chem = p.lambda_ * sum(gi * (si - ri * mu) for gi, si, ri in zip(gs, s, r))
    return wells + multi + chem


@pytest.mark.parametrize("a1", [0.0, 1.7])
def test_bulk_forces_are_the_free_energy_gradient(oracle_mod, a1):
    """fphi_i = d f / d p_i (the phases taken independent) at points of the
    simplex, for the reference's a1 = 0 and a non-zero a1."""
    p = small_params()
    p.a1 = a1
    r, s, _ = coeffs(p)
    h = 1e-5
    for (p1, p2, p3), mu in zip(simplex_points(40, 7), np.linspace(-0.6, 0.7, 40)):
        got = oracle_mod.tp_bulk_force(p, p1, p2, mu)
        q = np.array([p1, p2, p3])
        for m in range(3):
            e = np.zeros(3)
            e[m] = h
            fd = (_free_energy(p, r, s, mu, q + e) - _free_energy(p, r, s, mu, q - e)) / (2 * h)
            assert got[m] == pytest.approx(fd, abs=1e-6 * (1 + abs(fd))), (m, p1, p2, mu)


# --------------------------------------------------------------- compositions
def test_compositions_are_the_phase_coexistence_points(oracle_mod):
    """mu = c - r_i in pure phase i, and phases i, j coexist where their
    grand potentials s - r mu are equal: ca, cla at the 1-2 point, cb, clb
    at the 1-3 point. On the 1-2 edge (p3 = 0) the chemical part of the
    force along the edge, fcphi_1 - fcphi_2, then vanishes for every p1, and
    only at that mu: what remains of fphi_1 - fphi_2 is the double wells'
    4 p1 (1 - p1) (1 - 2 p1)."""
    p = small_params()
    r, s, d = coeffs(p)
    mu12 = (s[0] - s[1]) / (r[0] - r[1])
    mu13 = (s[0] - s[2]) / (r[0] - r[2])
    assert oracle_mod.tp_mu(p, 1.0, 0.0, d["cla"]) == pytest.approx(mu12, abs=1e-15)
    assert oracle_mod.tp_mu(p, 0.0, 1.0, d["ca"]) == pytest.approx(mu12, abs=1e-15)
    assert oracle_mod.tp_mu(p, 1.0, 0.0, d["clb"]) == pytest.approx(mu13, abs=1e-15)
    assert oracle_mod.tp_mu(p, 0.0, 0.0, d["cb"]) == pytest.approx(mu13, abs=1e-15)
    q = small_params()
    q.a1 = q.a2 = q.a3 = 0.0     # keep only the wells and the chemistry
    for p1 in (0.1, 0.35, 0.5, 0.8):
        wells = 4.0 * p1 * (1.0 - p1) * (1.0 - 2.0 * p1)
        f = oracle_mod.tp_bulk_force(q, p1, 1.0 - p1, mu12)
        w = oracle_mod.tp_bulk_force(q, p1, 1.0 - p1, mu12 + 0.05)
        assert abs((f[0] - f[1]) - wells) < 1e-12
        assert abs((w[0] - w[1]) - wells) > 1e-3


# --------------------------------------------------------------- solute update
@pytest.mark.parametrize("phase,Da,Db", [(0, 0.0, 0.0), (1, 0.5, 0.0), (2, 0.0, 0.25)])
def test_pure_phase_solute_diffusion_decays_a_sinusoid(oracle_mod, phase, Da, Db):
    """In a pure phase phi stays exactly put, mu = c - r_i, and c follows
    the explicit-Euler heat equation with the phase's mobility M (phase 1:
    1, phase 2: Da, phase 3: Db): a sinusoid of wavenumber m along x
    decays by (1 - DT alpha M DXin^2 4 sin^2(pi m / (Nx - 2)))^n."""
    p = small_params(Da=Da, Db=Db)
    st = oracle_mod.TpState(p)
    Nz, Ny, Nx = st.shape
    phi1 = np.full(st.shape, 1.0 if phase == 0 else 0.0)
    phi2 = np.full(st.shape, 1.0 if phase == 1 else 0.0)
    m, eps, c0 = 3, 1e-3, 0.2
    i = np.arange(Nx)
    c = c0 + eps * np.sin(2 * np.pi * m * (i - 1) / (Nx - 2))
    st.put(TP_FIELD_PHI1, phi1)
    st.put(TP_FIELD_PHI2, phi2)
    st.put(TP_FIELD_C, np.broadcast_to(c, st.shape))
    n = 40
    st.step(n)
    M = (1.0, Da, Db)[phase]
    fac = (1.0 - p.DT * p.alpha * M * p.DXin ** 2 * 4 * math.sin(math.pi * m / (Nx - 2)) ** 2) ** n
    got = st.get(TP_FIELD_C)
    want = c0 + eps * fac * np.sin(2 * np.pi * m * (i - 1) / (Nx - 2))
    assert np.max(np.abs(got - want)) < 1e-14
    assert np.array_equal(st.get(TP_FIELD_PHI1), phi1)
    assert np.array_equal(st.get(TP_FIELD_PHI2), phi2)
    r, _, _ = coeffs(p)
    assert np.max(np.abs(st.get(TP_FIELD_MU) - (got - r[phase]))) < 1e-15
    st.close()


def test_three_phase_run_conserves_solute_and_keeps_ghost_images(oracle_mod):
    p = small_params()
    st = oracle_mod.TpState(p)
    st.init_reference()
    m0 = st.get(TP_FIELD_C)[1:-1, 1:-1, 1:-1].sum()
    st.step(60)
    for f in (TP_FIELD_PHI1, TP_FIELD_PHI2, TP_FIELD_C, TP_FIELD_MU):
        a = st.get(f)
        assert np.array_equal(a[:, :, 0], a[:, :, -2]) and np.array_equal(a[:, :, -1], a[:, :, 1])
        assert np.array_equal(a[:, 0, :], a[:, -2, :]) and np.array_equal(a[:, -1, :], a[:, 1, :])
        assert np.array_equal(a[0], a[1]) and np.array_equal(a[-1], a[-2])
    c = st.get(TP_FIELD_C)
    m1 = c[1:-1, 1:-1, 1:-1].sum()
    assert abs(m1 - m0) <= 1e-13 * abs(m0) + 1e-12
    # the phases moved: the cap changes shape
    assert not np.array_equal(st.get(TP_FIELD_PHI2) > 0.5, np.zeros(st.shape, bool))
    st.close()


# --------------------------------------------------------------- initial
def test_initial_caps_have_the_spherical_cap_geometry(oracle_mod):
    """Two caps of base radius R and contact angle theta on plane zcut of
    the reference's 512 x 512 slab: the base disc is the lattice disc of
    radius R, the cap is floor(R tan(theta/2)) + 1 planes high and its cell
    count (less half the base plane) is the cap volume pi h^2 (3 rho - h) / 3
    (rho = R / sin theta) to lattice accuracy; below zcut phase 3 at cb, above it phase 1 at clb,
    phase 2 in the caps at ca."""
    p = tp.reference_params(512, 512, 128)
    st = oracle_mod.TpState(p)
    st.init_reference()
    phi1, phi2, c = st.get(TP_FIELD_PHI1), st.get(TP_FIELD_PHI2), st.get(TP_FIELD_C)
    d = tp.derived(p)
    zcut = p.Nz - p.zcut_from_top
    assert np.all(phi1[:zcut] == 0) and np.all(phi2[:zcut] == 0) and np.all(c[:zcut] == d["cb"])
    cap = phi2[zcut:] == 1.0
    assert np.all(c[zcut:][cap] == d["ca"]) and np.all(phi1[zcut:][cap] == 0)
    assert np.all(c[zcut:][~cap] == d["clb"]) and np.all(phi1[zcut:][~cap] == 1)
    R, th = p.R, p.theta
    rho, h = R / math.sin(th), R * math.tan(th / 2)
    jj, ii = np.mgrid[0:p.Ny, 0:p.Nx]
    for q in range(p.ncaps):
        near = (ii - p.cap_x[q]) ** 2 + (jj - p.cap_y[q]) ** 2
        base = phi2[zcut] == 1.0
        assert np.all(base[near < R * R - 1e-6])
        assert not np.any(base[(near > R * R + 1e-6) & (near < 100 ** 2)])
        box = (np.abs(ii - p.cap_x[q]) < 40) & (np.abs(jj - p.cap_y[q]) < 40)
        planes = [k for k in range(zcut, p.Nz) if np.any(phi2[k][box] == 1.0)]
        assert planes[0] == zcut and planes[-1] == zcut + int(math.floor(h))
        count = sum(int(np.sum(phi2[k][box] == 1.0)) for k in planes)
        # lattice cells stand for unit cubes: the base plane's cells reach
        # half a plane below zcut, outside the continuum cap
        base_n = int(np.sum(phi2[zcut][box] == 1.0))
        vol = math.pi * h * h * (3 * rho - h) / 3
        assert abs(count - 0.5 * base_n - vol) / vol < 0.005, (count, base_n, vol)
    st.close()


# --------------------------------------------------------------- files
def test_cross_section_and_break_point_files(tmp_path):
    rng = np.random.default_rng(3)
    a = rng.uniform(-1, 1, (6, 5, 7))
    f = str(tmp_path / "c_2D_xz2_step0.txt")
    tp.write_cross_section(f, a, 0, 2)
    lines = open(f).read().split("\n")
    assert len(lines) == 7 and lines[-1] == ""
    assert lines[0] == "".join("%.4f\t" % v for v in a[0, 2, :])
    g = str(tmp_path / "yz.txt")
    tp.write_cross_section(g, a, 1, 4)
    assert open(g).read().split("\n")[5] == "".join("%.4f\t" % v for v in a[5, :, 4])
    b = str(tmp_path / "c_break_point.vtk")
    tp.write_break_point(b, a)
    raw = open(b, "rb").read()
    head = raw.split(b"\n")[:10]
    assert head[0] == b"# vtk DataFile Version 3.0" and head[4] == b"DIMENSIONS 7 5 6"
    assert head[7] == b"POINT_DATA 210" and head[9] == b"LOOKUP_TABLE default"
    assert len(raw) == sum(len(x) + 1 for x in head) + 8 * a.size
    assert np.array_equal(tp.read_break_point(b, a.shape), a)
    os.remove(b)
