"""CPU restatement of the CALPHAD surrogate network and its training
gradient (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the HIP
training and inference kernels, never by the product path.

The network (no counterpart in the reference, which evaluates Gl, Gs, Glc,
Gsc of fun.cu:1671-1720 exactly): 2 -> 64 -> 64 -> 4 with the hard SiLU
act(x) = x * clamp(x/6 + 1/2, 0, 1), parameters in one FP32 vector
W1[64][2] b1[64] W2[64][64] b2[64] W3[4][64] b3[4].

quantized=True states the arithmetic of the HIP kernels: weights rounded to
BF16 (round to nearest even), each FP32 input split into three BF16 parts
(hi, mid, lo: each the rounding of what the previous parts leave), the
activations h1, h2 rounded to BF16 before the next product, and in the
backward pass dy, dz2, dz1 rounded to BF16 where they enter a product while
the bias gradients take them unrounded (straight-through: the roundings are
treated as the identity when differentiating). Accumulation is float64 here
and FP32 on the GPU, so the two agree to accumulation rounding, plus the rare
element whose BF16 rounding falls the other way.

split=True (with quantized) states the split-activation kernels: each
activation enters the next product as hi + lo, hi = bf16(h), lo = bf16(h - hi)
(h in FP32), while the weight gradients of W2 and W3 take hi only.

fp16_acts=True (with quantized; inference only) states the FP16-activation
kernels: h1 and h2 rounded to FP16 (nearest even) instead of BF16, and W2,
W3 enter as their BF16 values held in FP16 (exact unless below 2^-14, where
FP16 subnormals round them); layer 1 is as above. fp16_packed=True (with
fp16_acts) computes the activation itself in FP16: p = fp16(z),
u = clamp(fp16(p/6 + 1/2), 0, 1) with the 1/6 constant in FP16 and one
rounding of the fused multiply-add, h = fp16(p u).

quantized=False is the exact float64 network; its gradient is pinned against
central finite differences in the tests.
"""
from __future__ import annotations

import numpy as np

H, NIN, NOUT = 64, 2, 4
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * NIN
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + NOUT * H
NPARAM = OFF_B3 + NOUT


def bf16(x) -> np.ndarray:
    """Round FP32 values to BF16 (nearest, ties to even); returns float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def fp16(x) -> np.ndarray:
    """Round FP32 values to FP16 (nearest, ties to even, with subnormals);
    returns float64."""
    return np.asarray(x, dtype=np.float32).astype(np.float16).astype(np.float64)


def split3(x) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Three BF16 parts of FP32 x, as the kernels form them."""
    x32 = np.asarray(x, dtype=np.float32)
    hi = bf16(x32)
    r = (x32 - hi.astype(np.float32)).astype(np.float32)
    mid = bf16(r)
    lo = bf16((r - mid.astype(np.float32)).astype(np.float32))
    return hi, mid, lo


def act(x: np.ndarray) -> np.ndarray:
    return x * np.clip(x / 6.0 + 0.5, 0.0, 1.0)


def dact(x: np.ndarray) -> np.ndarray:
    return np.where(x <= -3.0, 0.0, np.where(x >= 3.0, 1.0, x / 3.0 + 0.5))


def unpack(P: np.ndarray):
    P = np.asarray(P, dtype=np.float64)
    W1 = P[OFF_W1:OFF_B1].reshape(H, NIN)
    b1 = P[OFF_B1:OFF_W2]
    W2 = P[OFF_W2:OFF_B2].reshape(H, H)
    b2 = P[OFF_B2:OFF_W3]
    W3 = P[OFF_W3:OFF_B3].reshape(NOUT, H)
    b3 = P[OFF_B3:NPARAM]
    return W1, b1, W2, b2, W3, b3


def _f32(x):
    return np.asarray(x, dtype=np.float32).astype(np.float64)


def act_fp16_packed(z: np.ndarray) -> np.ndarray:
    """Hard SiLU in FP16 arithmetic: each product of two FP16 values and the
    sum with 1/2 are exact in float64, so one FP16 rounding per operation."""
    p = fp16(np.asarray(z, dtype=np.float32))
    u = np.clip(fp16(p * float(np.float16(1.0 / 6.0)) + 0.5), 0.0, 1.0)
    return fp16(p * u)


def _acts(z: np.ndarray, quantized: bool, split: bool, half: bool = False,
          packed: bool = False):
    """(value entering the next product, value the weight gradient takes)."""
    if quantized and half and packed:
        h = act_fp16_packed(z)
        return h, h
    h = act(z)
    if not quantized:
        return h, h
    h32 = np.asarray(h, dtype=np.float32)
    if half:
        return fp16(h32), fp16(h32)
    hi = bf16(h32)
    if not split:
        return hi, hi
    lo = bf16((h32 - hi.astype(np.float32)).astype(np.float32))
    return hi + lo, hi


def forward(P: np.ndarray, X: np.ndarray, quantized: bool = True, split: bool = False,
            fp16_acts: bool = False, fp16_packed: bool = False) -> dict:
    """X: n x 2 inputs (a, b). Returns the intermediates and y (n x 4)."""
    W1, b1, W2, b2, W3, b3 = unpack(P)
    if quantized:
        W1q, W2q, W3q = bf16(W1), bf16(W2), bf16(W3)
        if fp16_acts:
            W2q, W3q = fp16(W2q), fp16(W3q)
        b1, b2, b3 = _f32(b1), _f32(b2), _f32(b3)
        pa = split3(X[:, 0])
        pb = split3(X[:, 1])
        xa = (pa[0] + pa[1]) + pa[2]
        xb = (pb[0] + pb[1]) + pb[2]
    else:
        W1q, W2q, W3q = W1, W2, W3
        xa = np.asarray(X[:, 0], dtype=np.float64)
        xb = np.asarray(X[:, 1], dtype=np.float64)
        pa = pb = None
    z1 = np.outer(xa, W1q[:, 0]) + np.outer(xb, W1q[:, 1]) + b1
    h1, g1 = _acts(z1, quantized, split, fp16_acts, fp16_packed)
    z2 = h1 @ W2q.T + b2
    h2, g2 = _acts(z2, quantized, split, fp16_acts, fp16_packed)
    y = h2 @ W3q.T + b3
    return {"z1": z1, "h1": h1, "z2": z2, "h2": h2, "g1": g1, "g2": g2, "y": y, "pa": pa, "pb": pb,
            "xa": xa, "xb": xb, "W1q": W1q, "W2q": W2q, "W3q": W3q}


def loss_grad(P: np.ndarray, X: np.ndarray, Yn: np.ndarray, quantized: bool = True,
              split: bool = False):
    """Mean squared error over n x 4 outputs and its gradient (parameter
    layout), with the kernels' rounding points when quantized."""
    f = forward(P, X, quantized, split)
    n = X.shape[0]
    e = f["y"] - np.asarray(Yn, dtype=np.float64)
    loss = float(np.mean(e * e))
    gscale = 2.0 / (n * NOUT)
    dy = gscale * e
    if quantized:
        dy = _f32(dy)
    dyb = bf16(dy) if quantized else dy
    dh2 = dyb @ f["W3q"]
    dz2 = dh2 * dact(f["z2"])
    dz2b = bf16(dz2) if quantized else dz2
    dh1 = dz2b @ f["W2q"]
    dz1 = dh1 * dact(f["z1"])
    dz1b = bf16(dz1) if quantized else dz1
    g = np.zeros(NPARAM)
    g[OFF_W2:OFF_B2] = (dz2b.T @ f["g1"]).ravel()
    g[OFF_B2:OFF_W3] = dz2.sum(axis=0)
    g[OFF_W3:OFF_B3] = (dyb.T @ f["g2"]).ravel()
    g[OFF_B3:NPARAM] = dy.sum(axis=0)
    if quantized:
        # W1 multiplies each of the three BF16 parts of its input
        ga = sum(dz1b.T @ part for part in f["pa"])
        gb = sum(dz1b.T @ part for part in f["pb"])
    else:
        ga = dz1b.T @ f["xa"]
        gb = dz1b.T @ f["xb"]
    W1g = np.stack([ga, gb], axis=1)
    g[OFF_W1:OFF_B1] = W1g.ravel()
    g[OFF_B1:OFF_W2] = dz1.sum(axis=0)
    return loss, g


def init_params(seed: int) -> np.ndarray:
    """Seeded Glorot-uniform weights with small random biases (tests)."""
    rng = np.random.default_rng(seed)
    P = np.zeros(NPARAM)

    def fill(off, fan_in, fan_out, count):
        r = np.sqrt(6.0 / (fan_in + fan_out))
        P[off:off + count] = rng.uniform(-r, r, count)

    fill(OFF_W1, NIN, H, H * NIN)
    fill(OFF_W2, H, H, H * H)
    fill(OFF_W3, H, NOUT, NOUT * H)
    P[OFF_B1:OFF_W2] = rng.uniform(-0.5, 0.5, H)
    P[OFF_B2:OFF_W3] = rng.uniform(-0.5, 0.5, H)
    P[OFF_B3:NPARAM] = rng.uniform(-0.5, 0.5, NOUT)
    return P.astype(np.float32)
