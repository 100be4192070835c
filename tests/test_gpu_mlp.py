"""The MLP kernels (BF16 MFMA, FP32 accumulation) against the NumPy reference with the
same BF16 rounding points (oracle/mlp_oracle.py), and a training run.

Tolerances: forward 2e-3 of the output scale; gradient 3e-3 of each parameter block's
largest entry (FP32 summation order and rare one-ulp BF16 rounding differences of
intermediate activations)."""
import numpy as np
import pytest

import mlp_oracle as M
from lithoinv3d_amd import mlp as L

pytestmark = pytest.mark.gpu

BLOCKS = [("W1", M.OFF_W1, M.OFF_B1), ("b1", M.OFF_B1, M.OFF_W2), ("W2", M.OFF_W2, M.OFF_B2),
          ("b2", M.OFF_B2, M.OFF_W3), ("w3", M.OFF_W3, M.OFF_B3), ("b3", M.OFF_B3, M.NPARAM)]


def say(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def net():
    say("mlp create")
    n = L.MLP(seed=3)
    yield n
    n.close()


def test_param_roundtrip(net):
    p = M.init_params(7).astype(np.float32)
    net.params = p
    np.testing.assert_array_equal(net.params, p)


@pytest.mark.parametrize("n", [1, 16, 37, 1000])
def test_forward_matches_reference(net, n):
    rng = np.random.default_rng(n)
    p = M.init_params(11).astype(np.float32)
    net.params = p
    X = rng.normal(size=(n, 16)).astype(np.float32)
    say(f"mlp forward n={n}")
    y = net.forward(X)
    yr, _ = M.forward(p, X, bf=True)
    scale = max(1.0, np.abs(yr).max())
    assert np.abs(y - yr).max() <= 2e-3 * scale


@pytest.mark.parametrize("n", [16, 53, 4096])
def test_loss_grad_matches_reference(net, n):
    rng = np.random.default_rng(100 + n)
    p = M.init_params(5).astype(np.float32)
    net.params = p
    X = rng.normal(size=(n, 16)).astype(np.float32)
    t = rng.normal(size=n).astype(np.float32)
    say(f"mlp loss_grad n={n}")
    loss, g = net.loss_grad(X, t)
    lr, gr = M.loss_grad(p, X, t, bf=True)
    assert loss == pytest.approx(lr, rel=2e-3)
    for name, a, b in BLOCKS:
        err = np.abs(g[a:b] - gr[a:b]).max() / max(np.abs(gr[a:b]).max(), 1e-12)
        print(f"n={n} {name}: rel err {err:.2e}")
        assert err <= 3e-3, name


def test_training_reduces_loss(net):
    rng = np.random.default_rng(9)
    n = 20000
    X = rng.normal(size=(n, 16)).astype(np.float32)
    a = rng.normal(size=16)
    t = (np.sin(X @ a * 0.5) + 0.3 * X[:, 0] * X[:, 1]).astype(np.float32)
    net.params = M.init_params(2).astype(np.float32)
    say("mlp training")
    losses = net.train(X, t, steps=400, lr=3e-3)
    print(f"loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    assert losses[-1] < 0.15 * losses[0]
    # the trained network's own loss agrees with the reference at its parameters
    lr, _ = M.loss_grad(net.params, X, t, bf=True)
    l2, _ = net.loss_grad(X, t)
    assert l2 == pytest.approx(lr, rel=5e-3)


def test_learned_h0_gain_over_plain_solve():
    """What the trained network gains: trained (on the GPU) on 48 seeded small contact
    models (32x32x16 cells, 64x64 stations) to correct the 1/1024-station-subset estimate
    of the log Gauss-Newton diagonal, and used as H0 on 4 held-out ones.  Passes when
    (1) on the held-out models the corrected estimate is closer to the exact log diagonal
    than the subset estimate alone (RMS after removing the constant offset, which the
    solver's gamma absorbs), and (2) the learned-H0 FP32C runs reach the plain FP32C run's
    30-iteration misfit in fewer iterations than the plain runs needed (median over the 4
    models).  Every number is printed."""
    from lithoinv3d_amd import experiments as E

    shape, n_st = (32, 32, 16), 64
    say("learned H0: training set")
    X, t = E.training_set(range(5000, 5048), shape, n_st)
    net, norm, losses = E.train_mlp(X, t, steps=1500)
    print(f"trained on {len(t)} parameters: loss {losses[0]:.3f} -> {losses[-1]:.3f}")
    plain_it, learned_it, err_l, err_b = [], [], [], []
    for s in range(6000, 6004):
        say(f"learned H0: held-out seed {s}")
        prob = E.contact_problem(s, shape, n_st)
        E.synth_data(prob)
        rp, _ = E.run_solver(prob, "fp32c", 30)
        _, exact, _ = E.features_targets(prob)
        pred, base = E.predict_log_diag(prob, net, norm)
        err_l.append(np.std(pred - exact))
        err_b.append(np.std(base - exact))
        rl, _ = E.run_solver(prob, "fp32c", 90, target=rp.f, h0=np.exp(-E.H0_POWER * pred))
        hit = E.time_to(rl, rp.f)
        plain_it.append(rp.iterations)
        learned_it.append(hit[0] if hit else 90)
        print(f"seed {s}: log-diagonal RMS error learned {err_l[-1]:.3f}, subset only "
              f"{err_b[-1]:.3f}; plain {rp.iterations} iterations to f={rp.f:.4e}; learned H0 "
              f"{learned_it[-1]} ({rl.message})")
    net_gain = 1.0 - np.median(learned_it) / np.median(plain_it)
    print(f"median iterations: plain {np.median(plain_it)}, learned {np.median(learned_it)} "
          f"({100 * net_gain:.0f} % fewer)")
    net.close()
    assert np.median(err_l) < np.median(err_b)
    assert np.median(learned_it) < np.median(plain_it)
