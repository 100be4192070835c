"""The MLP reference (oracle/mlp_oracle.py): its FP64 gradient against central finite
differences on every parameter block, and the BF16 rounding helper against the exact
round-to-nearest-even rule."""
import numpy as np
import pytest

import mlp_oracle as M


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.14159, 0.0], np.float32)
    r = M.bf16(x)
    # 8 significant bits: 1 + 2^-8 is a tie -> even (1.0); 1 + 3*2^-9 rounds up to 1 + 2^-7
    np.testing.assert_array_equal(r, np.array([1.0, 1.0, 1.0 + 2 ** -7, -2.5, 3.140625, 0.0],
                                              np.float32))


def test_fp64_gradient_vs_central_fd():
    rng = np.random.default_rng(0)
    p = M.init_params(1)
    X = rng.normal(size=(37, 16))
    t = rng.normal(size=37)
    _, g = M.loss_grad(p, X, t)
    h = 1e-6
    idx = np.concatenate([rng.choice(np.arange(M.OFF_W1, M.OFF_B1), 6),
                          rng.choice(np.arange(M.OFF_B1, M.OFF_W2), 4),
                          rng.choice(np.arange(M.OFF_W2, M.OFF_B2), 8),
                          rng.choice(np.arange(M.OFF_B2, M.OFF_W3), 4),
                          rng.choice(np.arange(M.OFF_W3, M.OFF_B3), 4), [M.OFF_B3]])
    for k in idx:
        e = np.zeros_like(p)
        e[k] = h
        fd = (M.loss_grad(p + e, X, t)[0] - M.loss_grad(p - e, X, t)[0]) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=1e-6, abs=1e-9)
