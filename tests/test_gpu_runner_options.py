"""InversionConfig options of the loop's set-up and solver (runner.h:27-170) on the GPU
objective:

  * gravityUncertainty / magneticUncertainty (runner.cpp:88-101): data whose std is 0
    take the default before the objective is built -- the same objective and gradient as
    with those stds given explicitly;
  * disableLineSearch (lbfgsb_optimizer.cpp:90-96): one geometry iteration from p0 is the
    full projected step along -g, clip(p0 - g(p0), lower, upper);
  * iterationExportDir / groupExportNames (runner.cpp:319-348, 571-575): one directory
    per accepted geometry phase holding every unit's closed surface at that phase's
    geometry;
  * the reference's discrete constraint penalty (ConstraintHandler, weight omega) in the
    objective: value and gradient add exactly.
"""
import os

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
    return M, st, gz, tm, p0, m


def test_default_uncertainties():
    M, st, gz, tm, p0, _ = setup(31)
    k = len(st)
    say("default data uncertainties")
    gstd = np.full(k, 0.05)
    gstd[::3] = 0.0  # some given, some not
    A = inversion.Problem(M, st, gz, gstd, tm, None, gravity_uncertainty=0.02,
                          magnetic_uncertainty=2.0)
    B = inversion.Problem(M, st, gz, np.where(gstd > 0, gstd, 0.02), tm, np.full(k, 2.0))
    C = inversion.Problem(M, st, gz, gstd, tm, None)  # no defaults: weight 1 where std <= 0
    p = p0 + 0.3
    np.testing.assert_allclose(A.evaluate(p), B.evaluate(p), rtol=1e-14)
    np.testing.assert_allclose(A.gradient(p), B.gradient(p), rtol=1e-12, atol=0)
    assert abs(C.evaluate(p)[1] - A.evaluate(p)[1]) > 0.1 * A.evaluate(p)[1]
    for P in (A, B, C):
        P.close()


def test_runner_without_line_search_takes_the_full_step():
    M, st, gz, tm, p0, _ = setup(32)
    k = len(st)
    P = inversion.Problem(M, st, gz, np.full(k, 0.05), tm, np.full(k, 1.0))
    g0 = P.gradient(p0)
    lo, hi = M.bounds(p0, 200.0)
    say("runner, disable_line_search, one iteration")
    r = P.run(max_iterations=1, property_interval=1, validate=False, eigenvalue_scaling=False,
              bounds_margin=200.0, disable_line_search=True)
    print(r["log"])
    np.testing.assert_allclose(r["params"], np.clip(p0 - g0, lo, hi), rtol=0, atol=1e-12)
    assert r["final_misfit"] == pytest.approx(P.evaluate(r["params"])[3], rel=1e-12)
    P.close()


def test_runner_iteration_export(tmp_path):
    from lithoinv3d_amd import litho_io

    M, st, gz, tm, p0, m = setup(33)
    k = len(st)
    P = inversion.Problem(M, st, gz, np.full(k, 0.05), tm, np.full(k, 1.0))
    say("runner, iteration export")
    out = tmp_path / "iters"
    r = P.run(max_iterations=4, property_interval=2, validate=False, eigenvalue_scaling=False,
              bounds_margin=200.0, iteration_export_dir=out, group_export_names=["a", "b"])
    print(r["log"])
    phases = r["log"].count("geometry phase")
    dirs = sorted(os.listdir(out))
    assert dirs == [f"iter_{i:03d}" for i in range(phases)] and phases >= 1
    last = out / dirs[-1]
    assert sorted(os.listdir(last)) == ["invert_a.ts", "invert_b.ts", "invert_group_2.ts"]
    V = M.apply(r["params"])
    for g, f in enumerate(("invert_a.ts", "invert_b.ts", "invert_group_2.ts")):
        mesh = litho_io.read_surface(last / f)
        tri = m.group_triangles(g)
        used = np.unique(tri)
        np.testing.assert_allclose(mesh.vertices, V[used], rtol=1e-8, atol=1e-6)
        np.testing.assert_array_equal(used[mesh.triangles], tri)
    P.close()


def test_problem_with_the_reference_constraint_penalty():
    """Problem.add_constraints (the reference's ConstraintHandler with the runner's omega):
    the objective's regularization part and its gradient grow by exactly the term's own
    value and gradient (Model.constraint_penalty)."""
    from lithoinv3d_amd import litho_io

    M, st, gz, tm, p0, m = setup(34)
    k = len(st)
    lo, hi = m.vertices.min(0), m.vertices.max(0)
    xy = np.array([[0.5 * (lo[0] + hi[0]) + 3.3, 0.5 * (lo[1] + hi[1]) - 2.1]] * 3)
    depth = -np.array([hi[2] - 10.0, hi[2] - 40.0, hi[2] - 70.0])
    c = litho_io.Constraints(xy, depth - 5.0, depth + 5.0, np.array([0, 1, 2]))
    A = inversion.Problem(M, st, gz, np.full(k, 0.05), tm, np.full(k, 1.0))
    B = inversion.Problem(M, st, gz, np.full(k, 0.05), tm, np.full(k, 1.0))
    say("problem with the reference constraint penalty")
    B.add_constraints(c, weight=2.0)
    p = p0 + 0.2
    val, viol, g = M.constraint_penalty(p, c, weight=2.0, gradient=True)
    ca, cb = A.evaluate(p), B.evaluate(p)
    print("violations", viol, "penalty", val)
    np.testing.assert_allclose(cb[:2], ca[:2], rtol=1e-14)
    assert cb[2] == pytest.approx(ca[2] + val, rel=1e-14, abs=1e-12)
    np.testing.assert_allclose(B.gradient(p), A.gradient(p) + g, rtol=1e-12, atol=1e-12)
    A.close()
    B.close()
