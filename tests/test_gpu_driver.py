"""The command-line driver (2d-rapid-solidification_amd/bin/rsolid_run, C++
over the C ABI): reference output files, fields equal to the same run made
through the Python binding, exact restart from the state file, restart from
the reference's checkpoint files, and pull-back history records."""
import os
import subprocess
import sys

import numpy as np
import pytest

from rsolid.abi import RS_FIELD_C, RS_FIELD_PHI

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "2d-rapid-solidification_amd", "bin", "rsolid_run")


def say(*a):
    print("[gpu-driver]", *a, flush=True)
    sys.stdout.flush()


def _total_time(rs, steps, N=64):
    p = rs.config_params(N, True)
    k = rs.derive(p)
    tau0 = (p.S * p.W0) ** 2 / (p.Gamma * p.muk0)
    return (steps + 0.5) * tau0 * k.dt


def _run(args, cwd):
    say("rsolid_run", " ".join(args))
    r = subprocess.run([DRIVER] + args, cwd=cwd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_driver_outputs_match_python_run(rs, tmp_path):
    T = _total_time(rs, 300)
    d = str(tmp_path / "a")
    out = _run(["preset=config", "N=64", "tfc=1", "seed=9", f"total_time={T!r}", f"path_input={d}"],
               str(tmp_path))
    assert "step" in out
    assert os.path.exists(os.path.join(d, "parameters.txt"))
    tip = open(os.path.join(d, "tip.txt"), encoding="utf-8").read().split("\n")
    assert tip[1].split()[0] == "step" and len(tip) > 10
    # the field files of step 0 and the final checkpoint against the binding
    p = rs.config_params(64, True)
    p.total_time = T
    ctx = rs.Context(p)
    ctx.init_config(9)
    c0 = np.loadtxt(os.path.join(d, "data", "c_0.txt"))
    assert np.allclose(c0, ctx.download(RS_FIELD_C), rtol=5e-6)
    ctx.step(ctx.constants().total_step)
    cfin = np.fromfile(os.path.join(d, "data", "check_point_c.bin"))
    phifin = np.fromfile(os.path.join(d, "data", "check_point_phi.bin"))
    assert np.array_equal(cfin, ctx.download(RS_FIELD_C).ravel())
    assert np.array_equal(phifin, ctx.download(RS_FIELD_PHI).ravel())
    ctx.close()


def test_driver_exact_restart(rs, tmp_path):
    T300, T150 = _total_time(rs, 300), _total_time(rs, 150)
    common = ["preset=config", "N=64", "tfc=1", "seed=4", "state=1"]
    a, b = str(tmp_path / "a"), str(tmp_path / "b")
    _run(common + [f"total_time={T300!r}", f"path_input={a}"], str(tmp_path))
    _run(common + [f"total_time={T150!r}", f"path_input={b}"], str(tmp_path))
    state = os.path.join(b, "data", "state.bin")
    assert os.path.exists(state)
    _run(common + [f"total_time={T300!r}", f"path_input={b}", f"load_state={state}"], str(tmp_path))
    for f in ("check_point_c.bin", "check_point_phi.bin", "check_point_temperature.bin"):
        x = np.fromfile(os.path.join(a, "data", f))
        y = np.fromfile(os.path.join(b, "data", f))
        assert np.array_equal(x, y), f


def test_driver_reference_restart_and_history(rs, tmp_path):
    """Restart from the reference's files (parameters.txt in %g, psi := phi)
    stays close to the uninterrupted run; pull-backs every 40 steps write
    offset.txt and c_history_temp.txt records."""
    T300, T150 = _total_time(rs, 300), _total_time(rs, 150)
    common = ["preset=config", "N=64", "tfc=1", "seed=5", "pull_back=40"]
    a, b = str(tmp_path / "a"), str(tmp_path / "b")
    _run(common + [f"total_time={T300!r}", f"path_input={a}"], str(tmp_path))
    _run(common + [f"total_time={T150!r}", f"path_input={b}"], str(tmp_path))
    _run(common + [f"total_time={T300!r}", f"path_input={b}", "if_load=1"], str(tmp_path))
    x = np.fromfile(os.path.join(a, "data", "check_point_c.bin"))
    y = np.fromfile(os.path.join(b, "data", "check_point_c.bin"))
    say("reference-format restart max rel diff", float(np.max(np.abs(x - y)) / np.max(np.abs(x))))
    assert np.max(np.abs(x - y)) <= 1e-5 * np.max(np.abs(x))
    off = np.loadtxt(os.path.join(a, "data", "offset.txt"), ndmin=2)
    assert off.shape[1] == 4 and off[0, 0] == 0 and np.all(off[:, 3] == 40)
    assert np.array_equal(np.cumsum(off[:, 1]), off[:, 2])
    hist = os.path.join(a, "data", "c_history_temp.txt")
    nshift = int(np.sum(np.maximum(off[:, 1], 0)))
    if nshift:
        h = np.loadtxt(hist, ndmin=2)
        assert h.shape == (nshift, 64)


def test_local_launcher_runs_and_logs(rs, tmp_path):
    """driver/run_local.sh with the reference script's variables: the run
    directory gets the reference's files and the out.txt log."""
    # the reference preset (run.sh values) on a 320 x 322 grid, 200 steps
    p = rs.reference_params(320, 322)
    k = rs.derive(p)
    tau0 = (p.S * p.W0) ** 2 / (p.Gamma * p.muk0)
    T = 200.5 * tau0 * k.dt
    d = str(tmp_path / "run")
    script = os.path.join(ROOT, "2d-rapid-solidification_amd", "driver", "run_local.sh")
    env = dict(os.environ, total_time=repr(T), path_input=d, Nx_all="320", Ny_all="322")
    say("run_local.sh", d)
    r = subprocess.run(["bash", script], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    log = open(os.path.join(d, "out.txt"), encoding="utf-8").read()
    assert log.count("########################################") == 2 and "step" in log
    assert os.path.exists(os.path.join(d, "tip.txt"))
    assert os.path.exists(os.path.join(d, "data", "check_point_c.bin"))
