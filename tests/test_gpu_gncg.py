"""Gauss-Newton CG (gncg_optimizer.cpp:105-194, the runner's solver = "gncg",
runner.cpp:192-202) over the GPU objective and gradient:

  * Problem.gncg_minimize against the oracle's restatement of the same algorithm driving
    the same problem's objective and gradient: the same iterates (rel. 1e-8), inner CG
    iterations, stop reason and call counts;
  * the inversion loop with solver="gncg": history entry 0 is the start's objective, the
    returned objective is the problem's at the returned parameters, the log names the
    solver, and the misfit falls.
"""
import numpy as np
import pytest

from lithoinv3d_amd import inversion, synth

pytestmark = pytest.mark.gpu

AXIS_SCALE = 10.0


def say(msg):
    print(f"[case] {msg}", flush=True)


def setup(seed):
    rng = np.random.default_rng(seed)
    labels = rng.integers(0, 3, size=(6, 5, 4)).astype(np.int32)
    m = synth.label_grid_model(labels, (-75.0, -62.5, -31.0), (25.0, 25.0, 25.0),
                               [0.25, -0.15, 0.4], [0.02, 0.003, 0.04])
    st = synth.topo_stations(7, 180.0)
    M = inversion.Model(m)
    M.fix_exterior(inversion.Z_ONLY)
    M.set_axis_scale(AXIS_SCALE)
    n = M.set_cp_stride(2, 25.0)
    p0 = M.params()
    ptrue = p0 + np.random.default_rng(seed + 1).normal(0, 0.6, n)
    P = inversion.Problem(M, st, np.zeros(len(st)), tmi_obs=np.zeros(len(st)))
    gz, tm = P.forward(ptrue)
    P.close()
    M.apply(p0)
    k = len(st)
    P = inversion.Problem(M, st, gz, np.full(k, 0.05), tm, np.full(k, 1.0))
    return M, P, p0


def test_problem_gncg_matches_reference_algorithm():
    import gncg_oracle

    M, P, p0 = setup(21)
    lb, ub = p0 - 20.0, p0 + 20.0
    kw = dict(max_iter=3, cg_max_iter=8, fd_step=1.0)
    say("GN-CG on the GPU problem: oracle restatement")
    calls = {"f": 0, "g": 0}

    def f(x):
        calls["f"] += 1
        return float(P.evaluate(x)[3])

    def g(x):
        calls["g"] += 1
        return P.gradient(x)

    ref = gncg_oracle.gncg_minimize(f, g, p0, lb, ub, **kw)
    say("GN-CG on the GPU problem: product")
    n_f0, n_a0 = P.pass_counts()
    r = P.gncg_minimize(p0, lb, ub, **kw)
    n_f1, n_a1 = P.pass_counts()
    print(r.iterations, r.message, list(r.cg_iterations), ref["cg_iterations"])
    print("objective", f(p0), "->", ref["f"], r.history[:, 1])
    assert r.message == ref["message"] and r.iterations == ref["iterations"]
    assert list(r.cg_iterations) == ref["cg_iterations"]
    np.testing.assert_allclose(r.x, ref["x"], rtol=1e-8, atol=1e-9)
    # one adjoint pass per gradient: 1 per outer iteration + 2 per CG iteration
    assert n_a1 - n_a0 == ref["gradient_calls"]
    assert r.history[-1, 1] == pytest.approx(ref["f"], rel=1e-10)
    assert ref["f"] < 0.1 * r.history[0, 1]
    P.close()


def test_runner_with_gncg_solver():
    M, P, p0 = setup(22)
    c0 = P.evaluate(p0)
    say("runner, solver gncg")
    r = P.run(max_iterations=4, property_interval=4, validate=True, eigenvalue_scaling=False,
              bounds_margin=200.0, solver="gncg", gncg_cg_max_iter=10, fd_step=1.0)
    print(r["log"])
    h = r["history"]
    np.testing.assert_allclose(h[0, 1:5], c0, rtol=1e-12)
    assert "geometry phase (gncg)" in r["log"]
    cf = P.evaluate(r["params"])
    assert r["final_misfit"] == pytest.approx(cf[3], rel=1e-12)
    assert cf[3] < 0.1 * c0[3]
    with pytest.raises(ValueError):
        P.run(max_iterations=1, solver="newton")
    P.close()
