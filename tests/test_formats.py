"""The reference's on-disk formats written and read by the library's host
code (no GPU): %g field text, parameters.txt, raw checkpoints, tip.txt
header, pull-back history records."""
import ctypes as C
import os

import numpy as np
import pytest

import rsolid
from rsolid.abi import RsTipRow


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def test_field_text_matches_save_field(tmp_path):
    rng = np.random.default_rng(1)
    a = rng.normal(size=(7, 11)) * 10.0 ** rng.integers(-8, 8, size=(7, 11))
    path = str(tmp_path / "c_3.txt")
    rsolid.write_field_txt(path, a)
    lines = open(path).read().split("\n")
    assert len(lines) == 8 and lines[-1] == ""
    assert all(ln.endswith("\t") and ln.count("\t") == 11 for ln in lines[:-1])
    back = np.array([[float(v) for v in ln.split("\t")[:-1]] for ln in lines[:-1]])
    assert np.allclose(back, a, rtol=5e-6, atol=0)          # %g: 6 significant digits
    assert lines[0].split("\t")[0] == "%g" % a[0, 0]
    # output stride di, dj (fun.cu:44-45)
    rsolid.write_field_txt(path, a, 2, 3)
    back = np.loadtxt(path)
    assert back.shape == (3, 6)
    assert np.allclose(back, a[::3, ::2], rtol=5e-6)


def test_parameters_round_trip(tmp_path):
    for p in (rsolid.reference_params(768, 832), rsolid.config_params(256, True)):
        k = rsolid.derive(p)
        k.step = 123456
        k.i_offset_history = 789
        k.l_offset_Vp = 4321.123456789
        k.delta = -0.4567891234
        k.i_left_end = k.i_left_end_initial + 17
        path = str(tmp_path / "parameters.txt")
        rsolid.write_parameters(path, p, k)
        text = open(path).read().split("\n")
        assert text[0].startswith("-----------------------------------------")
        assert text[1].split()[0] == "123456" and text[1].split()[1] == "step"
        k2 = rsolid.read_parameters(path, p)
        assert k2.step == k2.initial_step == 123456
        assert k2.i_offset_history == k2.i_offset_history_initial == 789
        assert k2.l_offset_Vp == pytest.approx(4321.12, rel=1e-6)    # %g, as the reference
        assert k2.delta == k2.delta_initial == pytest.approx(-0.456789, rel=1e-6)
        assert k2.dt == pytest.approx(k.dt, rel=1e-5)
        if p.ifTFC:
            assert k2.i_left_end == k.i_left_end
            assert k2.i_left_end_initial == k.i_left_end_initial
            assert (k2.NxT, k2.NyT, k2.NxT_all, k2.NyT_all) == (k.NxT, k.NyT, k.NxT_all, k.NyT_all)
            assert k2.Tcold == pytest.approx(k.Tcold, rel=5e-6)       # %g


def test_checkpoint_round_trip_is_bitwise(tmp_path):
    rng = np.random.default_rng(2)
    a = rng.normal(size=5000)
    path = str(tmp_path / "check_point_c.bin").encode()
    assert rsolid.lib.rs_write_checkpoint(path, _d(a), a.size) == 0
    assert os.path.getsize(path) == 8 * a.size
    b = np.empty_like(a)
    assert rsolid.lib.rs_read_checkpoint(path, _d(b), b.size) == 0
    assert np.array_equal(a, b)
    assert np.array_equal(np.fromfile(path.decode(), dtype=np.float64), a)


def test_tip_header_and_row(tmp_path):
    path = str(tmp_path / "tip.txt").encode()
    t = RsTipRow()
    t.step, t.t_us, t.delta, t.T_tip, t.V_tip, t.Vp, t.R = 16, 0.5, -0.3, 1931.5, 0.09, 0.0842, 12.5
    t.i, t.j, t.x, t.y = 27, 100, 27.25, 100.5
    assert rsolid.lib.rs_write_tip_row(path, C.byref(t), 1) == 0
    assert rsolid.lib.rs_write_tip_row(path, C.byref(t), 0) == 0
    lines = open(path, encoding="utf-8").read().split("\n")
    assert lines[0].split() == [f"({q})" for q in range(1, 12)]
    assert lines[1].split() == ["step", "t[us]", "Δtip", "Ttip[K]", "Vtip[m/s]", "Vp[m/s]",
                                "tip.R[nm]", "tip.i", "tip.j", "tip.x", "tip.y"]
    vals = lines[2].split()
    assert vals == ["16", "0.5", "-0.3", "1931.5", "0.09", "0.0842", "12.5", "27", "100", "27.25",
                    "100.5"]
    assert lines[2][:11] == "16         "      # %-10 columns


def test_history_records_skip_and_stride(tmp_path):
    path = str(tmp_path / "c_history_temp.txt").encode()
    Ny, n = 5, 4
    cols = np.arange(Ny * n, dtype=np.float64).reshape(Ny, n)
    ig = C.c_int32(0)
    assert rsolid.lib.rs_append_history(path, _d(cols), n, Ny, 100, 1, 1, C.byref(ig)) == 0
    rows = np.loadtxt(path.decode())
    assert rows.shape == (n, Ny)
    assert np.array_equal(rows, cols.T)             # one line per shifted-out column
    # i_ignore > 0 skips that many columns first; di = 2 keeps even lab columns
    os.remove(path)
    ig = C.c_int32(n + 2)
    assert rsolid.lib.rs_append_history(path, _d(cols), n, Ny, 0, 2, 1, C.byref(ig)) == 0
    assert not os.path.exists(path) or open(path).read() == ""
    assert ig.value == 0
    ig = C.c_int32(0)
    assert rsolid.lib.rs_append_history(path, _d(cols), n, Ny, 1, 2, 2, C.byref(ig)) == 0
    rows = np.loadtxt(path.decode(), ndmin=2)
    assert np.array_equal(rows, cols[::2, 1::2].T)  # (q + 1) % 2 == 0 -> q = 1, 3; rows 0, 2, 4
