"""Pins of the surrogate network's CPU restatement (oracle/mlp.py).

The network has no counterpart in the reference, so it is pinned by facts
of the arithmetic itself: BF16 and FP16 rounding of known values, the
ordering of the activation formats' errors, the exactness of
the three-part input split, closed forms of the activation, and the
analytic gradient of the exact network against central finite differences.
"""
import numpy as np
import pytest

from oracle import mlp


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2.0**-8, 1.0 + 3 * 2.0**-8, 1.0 + 2.0**-7, -2.5, 3.0e-3],
                 dtype=np.float32)
    got = mlp.bf16(x)
    # ties go to the even mantissa: 1 + 2^-8 -> 1, 1 + 1.5 * 2^-7 -> 1 + 2^-6
    assert got[0] == 1.0 and got[1] == 1.0
    assert got[2] == 1.0 + 2.0**-6
    assert got[3] == 1.0 + 2.0**-7
    assert got[4] == -2.5
    # 3e-3 = 1.536 * 2^-9: 7 mantissa bits keep 1.5390625 * 2^-9
    assert got[5] == pytest.approx(1.5390625 * 2.0**-9, rel=0, abs=0)


def test_fp16_rounding_known_values():
    # IEEE binary16: 10 stored mantissa bits, ties to even, subnormals down
    # to 2^-24, largest finite 65504
    x = np.array([1.0 + 2.0**-11, 1.0 + 3 * 2.0**-11, 1.0 / 3.0, 65504.0, 2.0**-24, 2.0**-25,
                  3 * 2.0**-25, 2.0**-14 * (1 + 2.0**-10), -0.1], dtype=np.float32)
    got = mlp.fp16(x)
    assert got[0] == 1.0                         # tie -> even
    assert got[1] == 1.0 + 2.0**-9               # tie -> even (upward)
    assert got[2] == 0.333251953125              # 1365 * 2^-12
    assert got[3] == 65504.0
    assert got[4] == 2.0**-24                    # smallest subnormal
    assert got[5] == 0.0                         # half of it: tie -> even (0)
    assert got[6] == 2.0**-23                    # 1.5 subnormal steps: tie -> even (2)
    assert got[7] == 2.0**-14 * (1 + 2.0**-10)   # normal, exact
    assert got[8] == -0.0999755859375            # -1638 * 2^-14
    # BF16 values of normal FP16 magnitude are exact in FP16
    rng = np.random.default_rng(4)
    w = mlp.bf16(rng.uniform(-2, 2, 10000).astype(np.float32))
    w = w[np.abs(w) >= 2.0**-14]
    assert np.array_equal(mlp.fp16(w), w)


def test_packed_fp16_activation_known_values():
    # fp16(1/6) = 1365 * 2^-13; at z = 1: u = fp16(1365 * 2^-13 + 1/2) rounds
    # 0.66662598 to 1365 * 2^-11, and h = u exactly
    assert float(np.float16(1.0 / 6.0)) == 1365 * 2.0**-13
    got = mlp.act_fp16_packed(np.array([1.0, 3.0, -3.0, 0.0, 10.0, -10.0]))
    assert got[0] == 1365 * 2.0**-11
    assert got[1] == 3.0                       # u rounds up to 1
    # at z = -3 the FP16 constant leaves u = 2^-13 (1 - ...) > 0: h != 0,
    # unlike the exact activation
    assert got[2] == pytest.approx(-3.0 * 2.0**-13, rel=1e-3) and got[2] < 0
    assert got[3] == 0.0 and got[4] == 10.0 and got[5] == 0.0


def test_fp16_network_between_bf16_and_split():
    """FP16 activations carry 11 bits: their error against the FP32-activation
    network lies between the BF16 (8 bits) and the split BF16 (~16 bits)
    formats'."""
    rng = np.random.default_rng(6)
    P = mlp.init_params(2)
    X = rng.uniform(-1, 1, (4096, 2)).astype(np.float32)
    W1, b1, W2, b2, W3, b3 = mlp.unpack(P)
    # reference: the quantized weights and inputs with unrounded activations
    Pq = P.astype(np.float64).copy()
    for off, end in ((mlp.OFF_W1, mlp.OFF_B1), (mlp.OFF_W2, mlp.OFF_B2), (mlp.OFF_W3, mlp.OFF_B3)):
        Pq[off:end] = mlp.bf16(P[off:end])
    yr = mlp.forward(Pq, X, quantized=False)["y"]
    e = {}
    for name, kw in (("bf16", {}), ("split", {"split": True}), ("fp16", {"fp16_acts": True})):
        e[name] = np.max(np.abs(mlp.forward(P, X, True, **kw)["y"] - yr))
    assert e["split"] < e["fp16"] < e["bf16"]
    # each extra mantissa bit roughly halves the error: 3 bits -> ~8x
    assert 3.0 < e["bf16"] / e["fp16"] < 20.0


def test_split3_carries_fp32_exactly():
    rng = np.random.default_rng(1)
    x = rng.uniform(-1, 1, 100000).astype(np.float32)
    hi, mid, lo = mlp.split3(x)
    for part in (hi, mid, lo):
        assert np.array_equal(mlp.bf16(part), part)
    rec = (hi + mid) + lo
    assert np.max(np.abs(rec - x.astype(np.float64)) / np.maximum(np.abs(x), 1e-30)) <= 2.0**-24


def test_activation_closed_forms():
    x = np.array([-10.0, -3.0, -1.5, 0.0, 1.5, 3.0, 7.0])
    assert np.allclose(mlp.act(x), [0.0, 0.0, -1.5 * 0.25, 0.0, 1.5 * 0.75, 3.0, 7.0])
    # slope: 0 below -3, x/3 + 1/2 between, 1 above
    assert np.allclose(mlp.dact(np.array([-4.0, -1.5, 0.0, 1.5, 4.0])), [0.0, 0.0, 0.5, 1.0, 1.0])
    # the slope is the derivative of act away from the kinks
    h = 1e-6
    z = np.array([-2.0, -0.3, 0.7, 2.9, 5.0])
    fd = (mlp.act(z + h) - mlp.act(z - h)) / (2 * h)
    assert np.allclose(fd, mlp.dact(z), atol=1e-8)


def test_exact_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    P = mlp.init_params(11).astype(np.float64)
    X = rng.uniform(-1, 1, (64, 2))
    Y = rng.normal(size=(64, 4))
    _, g = mlp.loss_grad(P, X, Y, quantized=False)
    idx = np.concatenate([rng.choice(np.arange(mlp.OFF_W1, mlp.OFF_B1), 20, replace=False),
                          rng.choice(np.arange(mlp.OFF_B1, mlp.OFF_W2), 10, replace=False),
                          rng.choice(np.arange(mlp.OFF_W2, mlp.OFF_B2), 60, replace=False),
                          rng.choice(np.arange(mlp.OFF_B2, mlp.OFF_W3), 10, replace=False),
                          rng.choice(np.arange(mlp.OFF_W3, mlp.OFF_B3), 20, replace=False),
                          np.arange(mlp.OFF_B3, mlp.NPARAM)])
    h = 1e-6
    for q in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[q] += h
        Pm[q] -= h
        fd = (mlp.loss_grad(Pp, X, Y, False)[0] - mlp.loss_grad(Pm, X, Y, False)[0]) / (2 * h)
        assert fd == pytest.approx(g[q], rel=1e-5, abs=1e-9), q


def test_quantized_network_close_to_exact():
    rng = np.random.default_rng(5)
    P = mlp.init_params(2)
    X = rng.uniform(-1, 1, (512, 2)).astype(np.float32)
    yq = mlp.forward(P, X, True)["y"]
    ye = mlp.forward(P, X, False)["y"]
    assert np.max(np.abs(yq - ye)) < 0.05 * np.max(np.abs(ye))
    Y = rng.normal(size=(512, 4))
    _, gq = mlp.loss_grad(P, X, Y, True)
    _, ge = mlp.loss_grad(P, X, Y, False)
    assert np.linalg.norm(gq - ge) < 0.05 * np.linalg.norm(ge)
