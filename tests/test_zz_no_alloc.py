"""rs_solute_mass_noalloc (no context, no allocation; its partial sums live
in a static device array) against rs_solute_mass on the same fields:
bitwise equal, on two calls in a row, the second on a smaller grid; sizes
beyond the array are rejected on the host before anything runs."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

from rsolid.abi import RS_ERR_ARG, RS_FIELD_C, RS_PREC_MIXED

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-noalloc]", *a, flush=True)
    sys.stdout.flush()


def _hip():
    for name in ("libamdhip64.so", os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib",
                                                "libamdhip64.so")):
        try:
            return C.CDLL(name)
        except OSError:
            continue
    pytest.skip("HIP runtime library not found")


class DeviceBuffer:
    """Device memory for the test's own inputs and outputs (the entry point
    under test allocates nothing; the test does)."""

    def __init__(self, hip, nbytes):
        self.hip = hip
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), C.c_size_t(nbytes)) == 0
        self.n = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.nbytes <= self.n
        assert self.hip.hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0

    def get(self, n):
        out = np.empty(n, dtype=np.float64)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _field(rs, N, seed, steps):
    p = rs.config_params(N, True, precision=RS_PREC_MIXED)
    ctx = rs.Context(p)
    ctx.init_config(seed)
    ctx.step(steps)
    c = ctx.download(RS_FIELD_C)
    m = ctx.solute_mass()
    ctx.close()
    return c, m


def test_noalloc_mass_matches_rs_solute_mass(rs):
    hip = _hip()
    big, small = _field(rs, 1024, 3, 50), _field(rs, 256, 4, 30)
    buf = DeviceBuffer(hip, big[0].nbytes)
    out = DeviceBuffer(hip, 8)
    try:
        for name, (c, want) in (("1024^2", big), ("256^2", small)):
            Ny, Nx = c.shape
            buf.put(c)
            say(name, "noalloc call")
            rc = rs.lib.rs_solute_mass_noalloc(buf.p, Nx, Ny, out.p, None)
            assert rc == 0, rs.lib.rs_last_error(None)
            assert hip.hipDeviceSynchronize() == 0
            got = float(out.get(1)[0])
            say(name, "noalloc", repr(got), "rs_solute_mass", repr(want))
            assert got == want
        rc = rs.lib.rs_solute_mass_noalloc(buf.p, 4, 65537, out.p, None)
        assert rc == RS_ERR_ARG
        msg = rs.lib.rs_last_error(None).decode()
        say("rejected:", msg)
        assert "65536" in msg
        assert rs.lib.rs_solute_mass_noalloc(None, 8, 8, out.p, None) == RS_ERR_ARG
    finally:
        buf.free()
        out.free()
