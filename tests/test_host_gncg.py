"""The host Gauss-Newton CG optimizer (GNCGOptimizer, gncg_optimizer.cpp:13-194) on the
CPU.  The oracle's restatement is pinned first by cases worked out by hand -- the Newton
step of a quadratic, the CG branch for non-positive curvature, the rejected line search,
the stalled step -- each of which the product must reproduce too; then the product and
the oracle are compared on bounded problems whose iterations exercise every branch."""
import numpy as np
import pytest

from lithoinv3d_amd import inversion


@pytest.fixture(scope="module")
def go():
    import gncg_oracle
    return gncg_oracle


def counted(fun, grad):
    calls = {"f": 0, "g": 0}

    def f(x):
        calls["f"] += 1
        return fun(x)

    def g(x):
        calls["g"] += 1
        return grad(x)

    return f, g, calls


def both(go, fun, grad, x0, lb, ub, **kw):
    """(oracle result, product result, product call counts)"""
    ref = go.gncg_minimize(fun, grad, x0, lb, ub, **kw)
    f, g, calls = counted(fun, grad)
    r = inversion.gncg_minimize(f, g, x0, lb, ub, **kw)
    return ref, r, calls


def test_quadratic_newton_step(go):
    """CG on a 3x3 SPD system ends after 3 iterations at A^-1 b (central differences of a
    linear gradient are exact), so one outer iteration lands on the minimiser and the
    second stops on the projected-gradient test: gradients 1 + 2*3 + 1, objectives
    f(x0) + one trial + the final value."""
    A = np.array([[4.0, 1.0, 0.0], [1.0, 3.0, 0.5], [0.0, 0.5, 2.0]])
    b = np.array([1.0, -2.0, 0.5])
    ref, r, calls = both(go, lambda x: 0.5 * x @ A @ x - b @ x, lambda x: A @ x - b,
                         np.zeros(3), np.full(3, -10.0), np.full(3, 10.0), max_iter=5)
    xs = np.linalg.solve(A, b)
    for x, it, msg, cg in ((ref["x"], ref["iterations"], ref["message"], ref["cg_iterations"]),
                           (r.x, r.iterations, r.message, list(r.cg_iterations))):
        np.testing.assert_allclose(x, xs, rtol=0, atol=1e-12)
        assert it == 2 and cg == [3]
        assert msg == "Converged: projected gradient below tolerance"
    assert (ref["gradient_calls"], ref["objective_calls"]) == (8, 3)
    assert (calls["g"], calls["f"]) == (8, 3)
    # history: the start and the minimiser, with f and |pg| at each
    np.testing.assert_allclose(r.history[:, 1], [0.0, -0.5 * b @ xs], rtol=1e-12, atol=1e-15)
    assert r.history[0, 2] == pytest.approx(np.linalg.norm(b)) and r.history[1, 2] < 1e-12


def test_nonpositive_curvature_branch(go):
    """f = -x^2 from x = 1, two CG iterations.  Every Hessian-vector product is exactly
    -2 v (|h v| = fd_step = 2 keeps both points inside the box).  j = 0: p = 2, Hp = -4,
    p.Hp < 0 so the step is |r|^2/|p|^2 = 1: x = 2, r = 2 + 4 = 6, beta = 9, p = 24;
    j = 1: Hp = -48, step 36/576: x = 3.5, r = 9.  d = 3.5 descends (g = -2), the full
    step is taken: x = 4.5 after the one allowed iteration."""
    kw = dict(max_iter=1, cg_max_iter=2)
    ref, r, calls = both(go, lambda x: -float(x @ x), lambda x: -2.0 * x, np.array([1.0]),
                         np.array([-100.0]), np.array([100.0]), **kw)
    for x, msg, cg in ((ref["x"], ref["message"], ref["cg_iterations"]),
                       (r.x, r.message, list(r.cg_iterations))):
        assert x[0] == 4.5 and cg == [2] and msg == "Maximum iterations reached"
    assert (calls["g"], calls["f"]) == (ref["gradient_calls"], ref["objective_calls"]) == (5, 3)
    # the same with the upper bound at 3: the step is projected onto it, and there the
    # outward gradient does not count, so the next iteration stops
    ref, r, _ = both(go, lambda x: -float(x @ x), lambda x: -2.0 * x, np.array([1.0]),
                     np.array([-10.0]), np.array([3.0]), max_iter=5, cg_max_iter=2)
    for x, it, msg in ((ref["x"], ref["iterations"], ref["message"]),
                       (r.x, r.iterations, r.message)):
        assert x[0] == 3.0 and it == 2
        assert msg == "Converged: projected gradient below tolerance"


def test_rejected_line_search(go):
    """A gradient that points the wrong way (f = x^2, g = +1 at 0): H = 0, CG returns
    d = -1, and every one of the 30 halved steps raises f, so the optimizer stops where it
    started after f(x0) + 30 trials + the final value."""
    ref, r, calls = both(go, lambda x: float(x @ x), lambda x: np.ones_like(x), np.zeros(1),
                         np.full(1, -5.0), np.full(1, 5.0), max_iter=10, cg_max_iter=1)
    for x, it, msg in ((ref["x"], ref["iterations"], ref["message"]),
                       (r.x, r.iterations, r.message)):
        assert x[0] == 0.0 and it == 1 and msg == "Converged: no step taken"
    assert (calls["f"], calls["g"]) == (ref["objective_calls"], ref["gradient_calls"]) == (32, 3)


def test_stalled_steps(go):
    """f = x with a gradient of 1e-9: each iteration steps by 1e-9 (< 1e-8); the third
    such step is not taken (gncg_optimizer.cpp:169-180), so x = -2e-9 after 3
    iterations."""
    kw = dict(max_iter=10, cg_max_iter=1, ftol=1e-12)
    ref, r, calls = both(go, lambda x: float(x[0]), lambda x: np.full_like(x, 1e-9),
                         np.zeros(1), np.full(1, -1.0), np.full(1, 1.0), **kw)
    for x, it, msg, cg in ((ref["x"], ref["iterations"], ref["message"], ref["cg_iterations"]),
                           (r.x, r.iterations, r.message, list(r.cg_iterations))):
        assert x[0] == pytest.approx(-2e-9, rel=1e-12) and it == 3
        assert msg == "Converged: step stalled" and cg == [1, 1, 1]
    assert (calls["f"], calls["g"]) == (ref["objective_calls"], ref["gradient_calls"]) == (7, 9)


def rosenbrock(x):
    return float(np.sum(100.0 * (x[1:] - x[:-1] ** 2) ** 2 + (1 - x[:-1]) ** 2))


def rosenbrock_grad(x):
    gr = np.zeros_like(x)
    gr[:-1] = -400 * x[:-1] * (x[1:] - x[:-1] ** 2) - 2 * (1 - x[:-1])
    gr[1:] += 200 * (x[1:] - x[:-1] ** 2)
    return gr


@pytest.mark.parametrize("case", ["rosenbrock", "rosenbrock_bounded", "bounded_lsq"])
def test_product_matches_oracle(go, case):
    """Iterates, stop reason, CG iterations per outer iteration and call counts equal to
    the restatement's on problems with indefinite Hessians (Rosenbrock's start), active
    bounds and a projected Hessian-vector difference."""
    if case.startswith("rosenbrock"):
        n = 6
        fun, grad = rosenbrock, rosenbrock_grad
        x0 = np.full(n, -1.2)
        lb, ub = (np.full(n, -5.0), np.full(n, 5.0)) if case == "rosenbrock" else (
            np.full(n, -0.5), np.full(n, 0.8))
        kw = dict(max_iter=40, cg_max_iter=20, fd_step=1e-4, ftol=1e-10)
    else:
        rng = np.random.default_rng(2)
        m, n = 30, 12
        J = rng.normal(size=(m, n))
        y = rng.normal(size=m)
        fun = lambda x: 0.5 * float(np.sum((np.tanh(J @ x) - y) ** 2))
        grad = lambda x: J.T @ ((np.tanh(J @ x) - y) * (1 - np.tanh(J @ x) ** 2))
        x0 = np.zeros(n)
        lb, ub = np.full(n, -0.3), np.full(n, 0.3)
        kw = dict(max_iter=25, cg_max_iter=30, fd_step=0.05, ftol=1e-9)
    ref, r, calls = both(go, fun, grad, x0, lb, ub, **kw)
    print(case, r.iterations, r.message, list(r.cg_iterations), fun(r.x))
    assert r.message == ref["message"] and r.iterations == ref["iterations"]
    assert list(r.cg_iterations) == ref["cg_iterations"]
    assert (calls["f"], calls["g"]) == (ref["objective_calls"], ref["gradient_calls"])
    np.testing.assert_allclose(r.x, ref["x"], rtol=1e-9, atol=1e-12)
    assert (r.x >= lb).all() and (r.x <= ub).all()
    assert fun(r.x) < fun(x0)
    # history rows: f falls monotonically (the line search never accepts a rise)
    assert np.all(np.diff(r.history[:, 1]) <= 0.0)


def test_options_rejected():
    f = lambda x: float(x @ x)
    g = lambda x: 2 * x
    with pytest.raises(Exception, match="fd_step"):
        inversion.gncg_minimize(f, g, np.ones(2), -np.ones(2), np.ones(2), fd_step=0.0)
    with pytest.raises(Exception, match="iteration limits"):
        inversion.gncg_minimize(f, g, np.ones(2), -np.ones(2), np.ones(2), cg_max_iter=-1)
