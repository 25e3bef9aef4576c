"""Throughput benchmark of the coupled phi / c / T time-stepper.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--grid N] [--path mixed|fp64]

One "step" is one pass of the reference loop body (fun.cu:448-509) over the
whole N x N grid: fused phase-field/solute update, latent-heat block sums,
coarse-grid heat equation and the pull-back check. Inputs are synthetic
(the seeded planar Ti-10at%Nb front) and resident in HBM before timing.

The metric compares the reduced-precision path with the plain FP64 path
on the same 4096^2 workload, so both are timed in the same run: `value` is
the headline path's MLUPS, `plain_fp64` holds the FP64 path's.
Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "2d-rapid-solidification_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "MLUPS @4096² Ti-Nb: surrogate+FP32 vs plain FP64; tip-vel err ≤1%"
HBM_PEAK_GBS = 8000.0
# algorithmic HBM bytes per cell update of the fused step kernel: read +
# write of the persistent state (phi = max(psi,-1) is not stored, fine T is
# rebuilt from the coarse grid)
BYTES_PER_CELL = {"fp64": 2 * 8 + 2 * 8,   # psi f64, c f64
                  "mixed": 2 * 4 + 2 * 8}  # psi f32, c f64
KERNEL = {"fp64": "step_fp64_kernel", "mixed": "step_mixed_kernel"}
PROFILE = {"fp64": "step_fp64_kernel_4096.md", "mixed": "step_mixed_kernel_4096.md"}


def _traffic_from_profile(path: str):
    """HBM bytes per launch of the fused kernel as measured by rocprofv3 PMC
    (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel calibration) and
    recorded in profiles/; None when not measured."""
    try:
        for line in open(os.path.join(ROOT, "profiles", PROFILE[path])):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def cpu_baseline(grid: int, seconds: float):
    """The CPU oracle (C, FP64, OpenMP over rows) on the host cores, timed on
    a bounded sample of the same workload: the first steps of the same grid."""
    import oracle
    import rsolid
    cores = min(16, os.cpu_count() or 1)
    oracle.set_threads(cores)
    p = rsolid.config_params(grid, tfc=True)
    st = oracle.State(p)
    st.init_config(5)
    st.step(1)                     # step 0 (pull-back) outside the timing
    n, t0 = 0, time.perf_counter()
    while True:
        st.step(1)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    st.close()
    return {"value": round(grid * grid * n / el / 1e6, 3), "unit": "MLUPS", "cores": cores,
            "kind": "port",
            "sample": f"CPU FP64 oracle, {grid}x{grid} coupled Ti-Nb, {n} steps ({el:.1f} s)"}


def run_path(path: str, grid: int, steps: int, warmup: int, calibrate: bool):
    import rsolid
    prec = rsolid.RS_PREC_MIXED if path == "mixed" else rsolid.RS_PREC_FP64
    p = rsolid.config_params(grid, tfc=True, precision=prec)
    ctx = rsolid.Context(p)
    ctx.init_config(5)
    ctx.step(warmup)
    ctx.sync()
    ctx.timing_enable(True)
    t0 = time.perf_counter()
    ctx.step(steps)
    ctx.sync()
    el = time.perf_counter() - t0
    k_ms, k_n = ctx.timing(0)
    t_ms, _ = ctx.timing(1)
    ctx.timing_enable(False)
    kc = ctx.constants()
    info = ctx.info()
    copy_ms = ctx.calibrate_copy(grid * grid, 5) if calibrate else None
    ctx.close()
    cells = grid * grid
    bpc = BYTES_PER_CELL[path]
    achieved = bpc * cells / (k_ms * 1e-3) / 1e9 if k_ms > 0 else 0.0
    return {
        "mlups": cells * steps / el / 1e6,
        "ms_per_step": el * 1e3 / steps,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": _traffic_from_profile(path), "kernel": KERNEL[path],
                     "kernel_avg_ms": round(k_ms, 5), "kernel_launches": k_n,
                     "bytes_per_cell": bpc},
        "coarse_T_kernel_avg_ms": round(t_ms, 5),
        "coarse_T": [kc.NxT, kc.NyT],
        "info": info,
        "copy_ms": copy_ms,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--grid", type=int, default=4096)
    ap.add_argument("--path", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64-reference", action="store_true",
                    help="skip timing the plain FP64 path beside the headline path")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the timed steps, run a plain FP64 copy of grid^2 doubles "
                         "(byte-counter calibration for rocprofv3 --pmc)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the y-band decomposition is a design note, not code).", file=sys.stderr)
        return 2

    head = run_path(a.path, a.grid, a.steps, a.warmup, a.calibrate_copy)
    fp64 = None
    if a.path != "fp64" and not a.no_fp64_reference:
        fp64 = run_path("fp64", a.grid, max(1, a.steps // 2), a.warmup, False)
    dtype = {"fp64": "f64", "mixed": "f32+f64"}[a.path]
    desc = {"fp64": "plain FP64 path",
            "mixed": "FP32 psi/phi and fine T, FP64 c / flux divergence / ln terms / "
                     "time accumulation, exact CALPHAD (no surrogate)"}[a.path]
    out = {
        "metric": METRIC,
        "value": round(head["mlups"], 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(head["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic: seeded planar Ti-10at%Nb front (configuration preset), no dataset",
        "config": {"workload": f"coupled phi/c/T step, {a.grid}x{a.grid} Ti-10at%Nb, coarse-grid "
                               f"heat equation (alpha = 100 D_l); {desc}",
                   "grid": a.grid, "path": a.path, "coarse_T": head["coarse_T"]},
        "roofline": head["roofline"],
        "coarse_T_kernel_avg_ms": head["coarse_T_kernel_avg_ms"],
    }
    if fp64 is not None:
        out["plain_fp64"] = {"value": round(fp64["mlups"], 2), "unit": "MLUPS",
                             "ms_per_step": round(fp64["ms_per_step"], 5),
                             "roofline": fp64["roofline"]}
        out["speedup_vs_plain_fp64"] = round(head["mlups"] / fp64["mlups"], 3)
    if head["copy_ms"] is not None:
        nb = 2 * 8 * a.grid * a.grid
        out["copy_calibration"] = {"bytes_per_launch": nb, "avg_ms": round(head["copy_ms"], 5),
                                   "GBps": round(nb / (head["copy_ms"] * 1e-3) / 1e9, 1)}
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(a.grid, a.cpu_seconds)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
