"""NumPy reference of the learned-preconditioner MLP (16 -> 64 -> 64 -> 1, GELU) and
its loss gradient.

TEST INFRASTRUCTURE ONLY: the checker for the lf_mlp_* kernels.  With ``bf16=True`` it
rounds (to nearest even) at exactly the points where the kernels use BF16 operands:
inputs, W1, W2, the layer-1 activations, the layer-2 pre-activation gradient and the
layer-1 pre-activation gradient; everything else is FP32.  With ``bf16=False`` it is an
FP64 reference whose gradient is pinned by central finite differences
(tests/test_oracle_mlp.py).  Parameter layout: W1[16][64], b1[64], W2[64][64], b2[64],
w3[64], b3.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

IN, HID = 16, 64
OFF_W1 = 0
OFF_B1 = IN * HID
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + HID * HID
OFF_W3 = OFF_B2 + HID
OFF_B3 = OFF_W3 + HID
NPARAM = OFF_B3 + 1


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def unpack(p):
    p = np.asarray(p)
    W1 = p[OFF_W1:OFF_B1].reshape(IN, HID)
    b1 = p[OFF_B1:OFF_W2]
    W2 = p[OFF_W2:OFF_B2].reshape(HID, HID)
    b2 = p[OFF_B2:OFF_W3]
    w3 = p[OFF_W3:OFF_B3]
    b3 = p[OFF_B3]
    return W1, b1, W2, b2, w3, b3


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2 * np.pi)


def forward(p, X, bf=False):
    dt = np.float32 if bf else np.float64
    q = (lambda a: bf16(a)) if bf else (lambda a: np.asarray(a, dt))
    W1, b1, W2, b2, w3, b3 = unpack(np.asarray(p, dt))
    Xb = q(X)
    u1 = (Xb @ q(W1)).astype(dt) + b1
    h1 = gelu(u1).astype(dt)
    H1b = q(h1)
    u2 = (H1b @ q(W2)).astype(dt) + b2
    h2 = gelu(u2).astype(dt)
    y = h2 @ w3 + b3
    return y.astype(dt), dict(Xb=Xb, u1=u1, H1b=H1b, u2=u2, h2=h2, W2b=q(W2))


def loss_grad(p, X, t, bf=False):
    dt = np.float32 if bf else np.float64
    q = (lambda a: bf16(a)) if bf else (lambda a: np.asarray(a, dt))
    n = len(X)
    W1, b1, W2, b2, w3, b3 = unpack(np.asarray(p, dt))
    y, c = forward(p, X, bf)
    e = y - np.asarray(t, dt)
    loss = float(np.sum(e * e) / n)
    dy = 2.0 * e / n
    gw3 = c["h2"].T @ dy
    gb3 = dy.sum()
    dz2 = (dy[:, None] * w3[None, :] * gelu_d(c["u2"])).astype(dt)
    gb2 = dz2.sum(0)
    dZ2b = q(dz2)
    dH1 = dZ2b @ c["W2b"].T
    dz1 = (dH1 * gelu_d(c["u1"])).astype(dt)
    gb1 = dz1.sum(0)
    gW2 = c["H1b"].T @ dZ2b
    gW1 = c["Xb"].T @ q(dz1)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gw3, [gb3]]).astype(dt)
    return loss, g


def init_params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM)
    p[OFF_W1:OFF_B1] = rng.normal(0, np.sqrt(2 / IN), IN * HID)
    p[OFF_W2:OFF_B2] = rng.normal(0, np.sqrt(2 / HID), HID * HID)
    p[OFF_W3:OFF_B3] = rng.normal(0, np.sqrt(1 / HID), HID)
    return p
