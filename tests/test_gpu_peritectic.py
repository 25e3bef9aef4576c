"""The 3D peritectic Ti-Ag stepper on the GPU (rsolid.peritectic, HIP kernel
tp_peritectic.hip) against its CPU restatement (oracle/tp_oracle.c).

The update has no transcendental function and both sides round every
operation in the reference's order without contraction, so the bar is
bitwise equality of phi1, phi2, c and mu after every run, on the
reference's one-GPU slab (512 x 512 x 128, its two-nuclei start), on a
small box, on ragged shapes that are not multiples of the 64 x 8 tile or
of the z segment, with the phase-2/3 mobilities and the a1 term switched
on, and from uploaded fields whose ghosts are not their images."""
import os
import sys

import numpy as np
import pytest

from rsolid import peritectic as tp
from rsolid.abi import TP_FIELD_C, TP_FIELD_MU, TP_FIELD_PHI1, TP_FIELD_PHI2

pytestmark = pytest.mark.gpu
FIELDS = (TP_FIELD_PHI1, TP_FIELD_PHI2, TP_FIELD_C, TP_FIELD_MU)


def say(*a):
    print("[gpu-peritectic]", *a, flush=True)
    sys.stdout.flush()


def box_params(Nx, Ny, Nz, Da=0.0, Db=0.0, a1=0.0):
    p = tp.reference_params(Nx, Ny, Nz)
    p.zcut_from_top = Nz - max(2, Nz // 3)
    p.R = max(2.0, min(Nx, Ny) / 5.0)
    p.ncaps = 2
    p.cap_x[0], p.cap_y[0] = Nx / 3.0, Ny / 3.0
    p.cap_x[1], p.cap_y[1] = 2.0 * Nx / 3.0 + 0.5, 0.6 * Ny
    p.Da, p.Db, p.a1 = Da, Db, a1
    return p


def _same(name, g, o):
    for f in FIELDS:
        a, b = g.download(f), o.get(f)
        if not np.array_equal(a, b):
            d = float(np.max(np.abs(a - b)))
            say(name, "field", f, "max |diff|", d)
        assert np.array_equal(a, b), (name, f)


@pytest.mark.parametrize("shape,steps,extra", [
    ((40, 36, 30), 200, {}),
    ((70, 37, 19), 120, {"Da": 0.3, "Db": 0.2, "a1": 1.7}),
    ((5, 6, 4), 50, {}),
    ((130, 9, 70), 60, {"Da": 0.1}),
])
def test_matches_oracle_bitwise(rs, oracle_mod, shape, steps, extra):
    p = box_params(*shape, **extra)
    say("box", shape, steps, "steps", extra)
    g = tp.Slab(p)
    g.init_reference()
    o = oracle_mod.TpState(p)
    o.init_reference()
    g.step(steps)
    o.step(steps)
    assert g.get_step() == steps
    _same(f"box{shape}", g, o)
    rg, ro = g.report(1), o.report(1)
    # the GPU's fixed-tree FP64 sum against the oracle's compensated sum
    assert rg["max_abs_c_plane"] == ro["max_abs_c_plane"]
    assert rg["mean_c"] == pytest.approx(ro["mean_c"], rel=1e-12)
    g.close()
    o.close()


def test_reference_slab_matches_oracle_bitwise(rs, oracle_mod):
    """512 x 512 x 128 (one GPU's slab of the reference's 4-GPU run, which
    is the whole grid on one GPU), the reference's two caps, 10 steps."""
    p = tp.reference_params(512, 512, 128)
    say("reference slab 512x512x128, 10 steps")
    g = tp.Slab(p)
    g.init_reference()
    o = oracle_mod.TpState(p)
    o.init_reference()
    _same("slab@0", g, o)
    g.step(10)
    o.step(10)
    _same("slab@10", g, o)
    rg, ro = g.report(20), o.report(20)
    say("report", rg, ro)
    assert rg["max_abs_c_plane"] == ro["max_abs_c_plane"]
    assert rg["mean_c"] == pytest.approx(ro["mean_c"], rel=1e-12)
    g.close()
    o.close()


def test_uploaded_fields_with_foreign_ghosts(rs, oracle_mod):
    """Random fields whose ghost cells are not their images: the stencil
    reads phi at the ghosts as stored and mu of the images (computeMu then
    boundaryMu), and after the step every ghost is its image's new value."""
    p = box_params(66, 10, 12, a1=0.5)
    rng = np.random.default_rng(9)
    shape = (p.Nz, p.Ny, p.Nx)
    f1 = rng.uniform(0.0, 0.6, shape)
    f2 = rng.uniform(0.0, 0.4, shape)
    c = rng.uniform(-0.4, 0.6, shape)
    g = tp.Slab(p)
    o = oracle_mod.TpState(p)
    for f, a in ((TP_FIELD_PHI1, f1), (TP_FIELD_PHI2, f2), (TP_FIELD_C, c)):
        g.upload(f, a)
        o.put(f, a)
    say("foreign ghosts")
    g.step(3)
    o.step(3)
    _same("foreign", g, o)
    g.close()
    o.close()


@pytest.mark.parametrize("zseg", ["1", "7", "1000"])
def test_z_segmentation_does_not_change_results(rs, oracle_mod, monkeypatch, zseg):
    p = box_params(40, 36, 30)
    monkeypatch.setenv("TP_ZSEG", zseg)
    g = tp.Slab(p)
    g.init_reference()
    o = oracle_mod.TpState(p)
    o.init_reference()
    g.step(40)
    o.step(40)
    _same(f"zseg{zseg}", g, o)
    g.close()
    o.close()


def test_solute_is_conserved_over_a_long_run(rs):
    """Periodic x/y and no-flux z with antisymmetric face fluxes: the
    interior solute sum changes only by rounding (2000 steps)."""
    p = box_params(64, 64, 40)
    g = tp.Slab(p)
    g.init_reference()
    c0 = g.download(TP_FIELD_C)[1:-1, 1:-1, 1:-1].sum()
    g.step(2000)
    c = g.download(TP_FIELD_C)
    ph2 = g.download(TP_FIELD_PHI2)
    c1 = c[1:-1, 1:-1, 1:-1].sum()
    say("solute drift over 2000 steps", (c1 - c0) / abs(c0))
    assert np.isfinite(c).all() and np.isfinite(ph2).all()
    assert abs(c1 - c0) <= 1e-12 * abs(c0)
    g.close()
