"""The C++ host layer (ForwardModel surface, objectives, L-BFGS-B) on the GPU.

  * GravityForward/MagneticForward::computeJacobian (through the adjoint) against the
    reference's per-face FD Jacobian (oracle) mapped to Z-only DOFs x axisScale
    (gravity_forward.cpp:269-288): rel. tolerance 2e-4 of the largest entry
  * computeGroupUnitResponse against the oracle with unit properties
  * the joint objective's gradient against central differences of the objective with
    control points active (exact chain rule through the interpolation): rel. 1e-5
  * a small L-BFGS-B inversion recovers a perturbed interface (FP64 and FP32C)
"""
import numpy as np
import pytest

from lithoinv3d_amd import inversion, synth

pytestmark = pytest.mark.gpu

AXIS_SCALE = 10.0


def say(msg):
    print(f"[case] {msg}", flush=True)


def model_and_stations(seed=3):
    rng = np.random.default_rng(seed)
    labels = rng.integers(0, 3, size=(6, 5, 4)).astype(np.int32)
    m = synth.label_grid_model(labels, (-75.0, -62.5, -31.0), (25.0, 25.0, 25.0),
                               [0.25, -0.15, 0.4], [0.02, 0.003, 0.04])
    st = synth.topo_stations(7, 180.0)
    return m, st


def test_jacobian_vs_reference_fd(oracle):
    m, st = model_and_stations()
    fh = synth.field_unit_vector(-60.0, 0.0)
    M = inversion.Model(m)
    M.fix_exterior(inversion.Z_ONLY)
    M.set_axis_scale(AXIS_SCALE)
    v, a, _ = M.dofs()
    p0 = M.params()
    om = oracle.OracleModel.from_litho(m, 50000.0, fh)
    Jg_ref = om.gravity_vertex_jacobian(st)[:, v, 2] * AXIS_SCALE
    Jm_ref = om.magnetic_vertex_jacobian(fh, st)[:, v, 2] * AXIS_SCALE
    for prec in ("fp64", "fp32c"):
        say(f"host jacobian {prec}")
        P = inversion.Problem(M, st, np.zeros(len(st)), tmi_obs=np.zeros(len(st)), inc=-60.0,
                              dec=0.0, precision=prec)
        Jg = P.jacobian(p0, "gz")
        Jm = P.jacobian(p0, "tmi")
        eg = np.abs(Jg - Jg_ref).max() / np.abs(Jg_ref).max()
        em = np.abs(Jm - Jm_ref).max() / np.abs(Jm_ref).max()
        print(f"{prec}: J_gz rel {eg:.2e}, J_tmi rel {em:.2e}")
        assert eg < 2e-4 and em < 2e-3
        P.close()


def test_unit_response(oracle):
    m, st = model_and_stations(4)
    fh = synth.field_unit_vector(-60.0, 0.0)
    M = inversion.Model(m)
    P = inversion.Problem(M, st, np.zeros(len(st)), tmi_obs=np.zeros(len(st)))
    say("unit responses")
    for g in range(m.n_units):
        d = np.zeros(m.n_units)
        d[g] = 1.0
        s = np.zeros(m.n_units)
        s[g] = 1.0
        mm = synth.LithoModel(m.vertices, m.faces, m.front, m.back, d, s)
        om = oracle.OracleModel.from_litho(mm, 50000.0, fh)
        np.testing.assert_allclose(P.unit_response(g, "gz"), om.gravity(st), atol=1e-10, rtol=1e-9)
        np.testing.assert_allclose(P.unit_response(g, "tmi"), om.magnetic(fh, st), atol=1e-7,
                                   rtol=1e-9)


def test_objective_gradient_with_control_points():
    m, st = model_and_stations(5)
    M = inversion.Model(m)
    M.fix_exterior(inversion.Z_ONLY)
    M.set_axis_scale(AXIS_SCALE)
    n = M.set_cp_stride(2, 25.0)
    p0 = M.params()
    rng = np.random.default_rng(1)
    ptrue = p0 + rng.normal(0, 0.8, n)
    say("objective gradient with control points fp64")
    P = inversion.Problem(M, st, np.zeros(len(st)), tmi_obs=np.zeros(len(st)), alpha_mag=0.5,
                          ref_weight=0.01)
    gz_t, tm_t = P.forward(ptrue)
    P.close()
    P = inversion.Problem(M, st, gz_t, np.full(len(st), 0.05), tm_t, np.full(len(st), 1.0),
                          alpha_mag=0.5, ref_weight=0.01)
    g = P.gradient(p0)
    h = 1e-3  # parameter units (0.01 m)
    worst = 0.0
    for j in rng.choice(n, min(10, n), replace=False):
        e = np.zeros(n)
        e[j] = h
        fd = (P.evaluate(p0 + e)[3] - P.evaluate(p0 - e)[3]) / (2 * h)
        worst = max(worst, abs(fd - g[j]) / np.abs(g).max())
    print(f"gradient vs central FD: {worst:.2e}")
    assert worst < 1e-5
    P.close()


@pytest.mark.parametrize("cp_mode", ["position", "displacement"])
def test_objective_gradient_along_vector_and_cp_modes(cp_mode):
    """The joint objective's gradient (adjoint pulled back through the parameter map)
    against central differences of the objective with ALONG_VECTOR vertices (parameter =
    projection on a move vector) among Z_ONLY ones, control points in the reference's
    position form and in the displacement form: rel. 1e-5 of the largest component."""
    m, st = model_and_stations(7)
    M = inversion.Model(m)
    M.fix_exterior(inversion.Z_ONLY)
    M.set_axis_scale(AXIS_SCALE)
    fr = M.freedoms()
    rng = np.random.default_rng(3)
    inner = np.nonzero(fr == inversion.Z_ONLY)[0]
    along = rng.choice(inner, len(inner) // 3, replace=False)
    mv = np.zeros((m.n_vertices, 3))
    mv[along] = rng.normal(size=(len(along), 3)) + np.array([0.0, 0.0, 2.0])
    M.set_move_vectors(mv)
    fr[along] = inversion.ALONG_VECTOR
    M.set_freedom(fr)
    n = M.set_cp_stride(2, 25.0)
    M.set_cp_interpolation(cp_mode)
    p0 = M.params()
    ptrue = p0 + rng.normal(0, 0.5, n)
    say(f"objective gradient, ALONG_VECTOR vertices, {cp_mode} control points")
    P = inversion.Problem(M, st, np.zeros(len(st)), tmi_obs=np.zeros(len(st)), alpha_mag=0.5)
    gz_t, tm_t = P.forward(ptrue)
    P.close()
    P = inversion.Problem(M, st, gz_t, np.full(len(st), 0.05), tm_t, np.full(len(st), 1.0),
                          alpha_mag=0.5, smooth_weight=0.1)
    g = P.gradient(p0)
    h = 1e-3
    worst = 0.0
    for j in rng.choice(n, min(12, n), replace=False):
        e = np.zeros(n)
        e[j] = h
        fd = (P.evaluate(p0 + e)[3] - P.evaluate(p0 - e)[3]) / (2 * h)
        worst = max(worst, abs(fd - g[j]) / np.abs(g).max())
    print(f"{cp_mode}: gradient vs central FD {worst:.2e}")
    assert worst < 1e-5
    P.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
def test_small_inversion_recovers(prec):
    m, st = model_and_stations(6)
    M = inversion.Model(m)
    M.fix_exterior(inversion.Z_ONLY)
    M.set_axis_scale(AXIS_SCALE)
    n = M.set_cp_stride(2, 25.0)
    p0 = M.params()
    rng = np.random.default_rng(2)
    ptrue = p0 + rng.normal(0, 0.6, n)
    P = inversion.Problem(M, st, np.zeros(len(st)), tmi_obs=np.zeros(len(st)), precision=prec)
    gz_t, tm_t = P.forward(ptrue)
    P.close()
    say(f"small inversion {prec}")
    P = inversion.Problem(M, st, gz_t, np.full(len(st), 0.05), tm_t, np.full(len(st), 1.0),
                          alpha_mag=1.0, precision=prec)
    lo, hi = M.bounds(p0, 200.0)
    f0 = P.evaluate(p0)[3]
    rows = []
    P.set_iteration_callback(lambda kind, row: rows.append((kind, row)))
    r = P.minimize(p0, lo, hi, max_iter=40)
    print(f"{prec}: f0 {f0:.3e} -> {r.f:.3e} in {r.iterations} it ({r.message})")
    assert r.f < 1e-2 * f0
    # the iteration callback saw every history row as it was recorded
    assert [k for k, _ in rows] == ["minimize"] * len(r.history)
    np.testing.assert_array_equal(np.array([row for _, row in rows]), r.history)
    P.set_iteration_callback(None)
    P.minimize(p0, lo, hi, max_iter=2)
    assert len(rows) == len(r.history)
    P.close()


def test_fused_objective_matches_reference_shaped():
    """One GPU context for gz and TMI (fused) against the reference-shaped JointObjective
    over separate forwards: same value and gradient."""
    m, st = model_and_stations(8)
    M = inversion.Model(m)
    M.fix_exterior(inversion.Z_ONLY)
    M.set_axis_scale(AXIS_SCALE)
    n = M.set_cp_stride(2, 25.0)
    p0 = M.params()
    rng = np.random.default_rng(4)
    gz_obs = rng.normal(0, 0.5, len(st))
    tm_obs = rng.normal(0, 20.0, len(st))
    say("fused vs separate objective fp64")
    P = inversion.Problem(M, st, gz_obs, np.full(len(st), 0.05), tm_obs, np.full(len(st), 1.0),
                          alpha_mag=0.7, ref_weight=0.01)
    p = p0 + rng.normal(0, 0.3, n)
    P.set_fused(False)
    c_sep, g_sep = P.evaluate(p), P.gradient(p)
    P.set_fused(True)
    c_fus, g_fus = P.evaluate(p), P.gradient(p)
    np.testing.assert_allclose(c_fus, c_sep, rtol=1e-11)
    np.testing.assert_allclose(g_fus, g_sep, rtol=1e-9, atol=1e-9 * np.abs(g_sep).max())
    f, a = P.pass_counts()
    assert (f, a) == (1, 1)  # the gradient reused the forward of the evaluation
    P.close()


def test_property_inversion_recovers_units():
    """The property phase (runner.cpp:708-935): unit densities from gravity and unit
    susceptibilities from TMI, started from wrong values, recover the values that made the
    noise-free data (relative 1e-5), and the problem's forward takes them."""
    m, st = model_and_stations(11)
    M0 = inversion.Model(m)
    P = inversion.Problem(M0, st, np.zeros(len(st)), tmi_obs=np.zeros(len(st)))
    gz, tm = P.forward(M0.params())
    P.close()
    wrong = synth.LithoModel(m.vertices, m.faces, m.front, m.back, 0.5 * m.density + 0.05,
                             2.0 * m.susceptibility)
    M = inversion.Model(wrong)
    P = inversion.Problem(M, st, gz, np.full(len(st), 0.05), tm, np.full(len(st), 1.0))
    say("property inversion: densities")
    d = P.invert_properties("density", lower=-1.0, upper=1.0, max_iter=200, tol=1e-14)
    print("densities", d, "true", m.density)
    np.testing.assert_allclose(d, m.density, rtol=1e-5)
    say("property inversion: susceptibilities")
    s = P.invert_properties("susceptibility", lower=0.0, upper=1.0, max_iter=200, tol=1e-14)
    print("susceptibilities", s, "true", m.susceptibility)
    np.testing.assert_allclose(s, m.susceptibility, rtol=1e-5)
    g2, t2 = P.forward(M.params())
    np.testing.assert_allclose(g2, gz, rtol=0, atol=1e-6 * np.abs(gz).max())
    np.testing.assert_allclose(t2, tm, rtol=0, atol=1e-5 * np.abs(tm).max())
    comp = P.evaluate(M.params())
    print("objective components after the property phase", comp)
    P.close()


def test_calibrate_magnetic_weight_vs_dense_eigenvalues(oracle):
    """JointObjective::calibrateDataWeights (joint_objective.cpp:281-393) done matrix-free
    (power iteration, J v by central differences of the FP64 forward, J^T u by the
    adjoint) against numpy's largest eigenvalue of J^T diag(1/std^2) J with J the
    reference's per-face FD Jacobian (oracle) on Z-only DOFs x axisScale: rel. 1e-3.  The
    new alpha_mag is alpha_mag x the clipped ratio and the objective takes it; a gravity
    std that makes the ratio exceed 1e6 is clipped to 1e6."""
    m, st = model_and_stations(12)
    fh = synth.field_unit_vector(-60.0, 0.0)
    M = inversion.Model(m)
    M.fix_exterior(inversion.Z_ONLY)
    M.set_axis_scale(AXIS_SCALE)
    v, _, _ = M.dofs()
    p0 = M.params()
    om = oracle.OracleModel.from_litho(m, 50000.0, fh)
    Jg = om.gravity_vertex_jacobian(st)[:, v, 2] * AXIS_SCALE
    Jm = om.magnetic_vertex_jacobian(fh, st)[:, v, 2] * AXIS_SCALE
    rng = np.random.default_rng(9)
    gstd = rng.uniform(0.02, 0.08, len(st))
    tstd = rng.uniform(0.5, 2.0, len(st))
    eg = np.linalg.eigvalsh(Jg.T @ (Jg / gstd[:, None] ** 2))
    em = np.linalg.eigvalsh(Jm.T @ (Jm / tstd[:, None] ** 2))
    print(f"dense: lambda_g {eg[-1]:.6e} (next/largest {eg[-2] / eg[-1]:.3f}), "
          f"lambda_m {em[-1]:.6e} (next/largest {em[-2] / em[-1]:.3f})")
    alpha0 = 0.3
    say("calibrate alpha_mag")
    P = inversion.Problem(M, st, rng.normal(0, 0.5, len(st)), gstd, rng.normal(0, 30, len(st)),
                          tstd, alpha_mag=alpha0)
    c0 = P.evaluate(p0)
    r = P.calibrate(p0, seed=3)
    print(r)
    assert not r["fallback"]
    assert abs(r["eig_gravity"] / eg[-1] - 1) < 1e-3
    assert abs(r["eig_magnetic"] / em[-1] - 1) < 1e-3
    ratio = r["eig_gravity"] / r["eig_magnetic"]
    assert abs(r["ratio"] / ratio - 1) < 1e-12
    assert abs(r["alpha_mag"] / (alpha0 * np.clip(ratio, 1e-6, 1e6)) - 1) < 1e-12
    np.testing.assert_allclose(M.params(), p0, rtol=0, atol=0)  # the model is left at p0
    c1 = P.evaluate(p0)
    mag0, mag1 = c0[3] - c0[0] - c0[2], c1[3] - c1[0] - c1[2]
    np.testing.assert_allclose(mag1, mag0 * r["alpha_mag"] / alpha0, rtol=1e-10)
    P.close()
    say("calibrate alpha_mag, clipped ratio")
    P = inversion.Problem(M, st, np.zeros(len(st)), np.full(len(st), 1e-6), np.zeros(len(st)),
                          np.full(len(st), 1e3), alpha_mag=alpha0)
    r = P.calibrate(p0, seed=4)
    print(r)
    assert r["ratio"] > 1e6 and abs(r["alpha_mag"] / (alpha0 * 1e6) - 1) < 1e-12
    P.close()
