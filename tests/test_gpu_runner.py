"""The inversion loop (runner.cpp:217-641) over the GPU objective.

  * geometry only, with calibration and the validation gate: history entry 0 is the
    objective at the start (same components as Problem.evaluate), the calibrated
    alpha_mag equals Problem.calibrate's at the start, the final objective and gravity RMS
    equal Problem.evaluate / the forward at the returned parameters, the last entry's RMS
    and Durbin-Watson pairs match the residuals there when it is the final point, and
    the final objective is below 1e-2 of the start's (noise-free data of a perturbed
    interface);
  * geometry with property phases: unit densities started 20 % off come back within 2 %
    while the misfit falls.
"""
import numpy as np
import pytest

from lithoinv3d_amd import inversion, synth

pytestmark = pytest.mark.gpu

AXIS_SCALE = 10.0


def say(msg):
    print(f"[case] {msg}", flush=True)


def setup(seed, dens_factor=1.0):
    rng = np.random.default_rng(seed)
    labels = rng.integers(0, 3, size=(6, 5, 4)).astype(np.int32)
    dens = np.array([0.25, -0.15, 0.4])
    sus = np.array([0.02, 0.003, 0.04])
    m = synth.label_grid_model(labels, (-75.0, -62.5, -31.0), (25.0, 25.0, 25.0), dens, sus)
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
    if dens_factor != 1.0:
        m2 = synth.LithoModel(m.vertices, m.faces, m.front, m.back, dens * dens_factor, sus)
        M = inversion.Model(m2)
        M.fix_exterior(inversion.Z_ONLY)
        M.set_axis_scale(AXIS_SCALE)
        M.set_cp_stride(2, 25.0)
    return M, st, gz, tm, p0, dens


def test_runner_geometry_with_calibration_and_gate():
    M, st, gz, tm, p0, _ = setup(21)
    n = len(st)
    P = inversion.Problem(M, st, gz, np.full(n, 0.05), tm, np.full(n, 1.0), alpha_mag=1.0)
    c0 = P.evaluate(p0)
    alpha = P.calibrate(p0, seed=1)["alpha_mag"]
    P.close()
    say("runner: geometry, calibration, validation gate")
    P = inversion.Problem(M, st, gz, np.full(n, 0.05), tm, np.full(n, 1.0), alpha_mag=1.0)
    rows = []
    P.set_iteration_callback(lambda kind, row: rows.append((kind, row)))
    r = P.run(max_iterations=30, property_interval=10, validate=True, eigenvalue_scaling=True,
              bounds_margin=200.0)
    print(r["log"])
    # the iteration callback (runner.h:200-201) saw every history entry as it was recorded
    assert len(rows) == r["total_entries"] and all(k == "run" for k, _ in rows)
    np.testing.assert_array_equal(np.array([row for _, row in rows])[:len(r["history"])],
                                  r["history"])
    print({k: v for k, v in r.items() if k not in ("history", "params", "log")})
    h = r["history"]
    np.testing.assert_allclose(h[0, 1:5], [c0[0], c0[1], c0[2], c0[3]], rtol=1e-12)
    assert r["alpha_mag"] == pytest.approx(alpha, rel=1e-12)
    cf = P.evaluate(r["params"])
    assert r["final_misfit"] == pytest.approx(cf[3], rel=1e-12)
    g_pred, t_pred = P.forward(r["params"])
    rms = np.sqrt(np.mean((gz - g_pred) ** 2))
    assert r["final_rms"] == pytest.approx(rms, rel=1e-12)
    # the last entry is the last evaluation at new parameters; recompute its statistics
    # when it is the final point (the run reaches residuals of ~1e-9 mGal, where the
    # fused forward and the separate one differ in the sixth digit)
    if np.isclose(h[-1, 4], cf[3], rtol=1e-13):
        dwg = inversion.durbin_watson(st, gz - g_pred)
        dwm = inversion.durbin_watson(st, tm - t_pred)
        np.testing.assert_allclose(h[-1, 5:10], [rms, *dwg, *dwm], rtol=1e-6)
    assert r["total_entries"] == len(h)
    # the objective at the start with the calibrated weight (entry 0 precedes calibration)
    start = c0[0] + r["alpha_mag"] * c0[1] + c0[2]
    assert cf[3] < 1e-2 * start
    P.close()


def test_runner_property_phases():
    M, st, gz, tm, p0, dens = setup(22, dens_factor=0.8)
    n = len(st)
    say("runner: geometry and property phases")
    P = inversion.Problem(M, st, gz, np.full(n, 0.05), tm, np.full(n, 1.0), alpha_mag=1.0)
    c0 = P.evaluate(p0)
    r = P.run(max_iterations=40, property_interval=10, properties=True, validate=True,
              eigenvalue_scaling=False, density_bounds=(-1.0, 1.0),
              susceptibility_bounds=(0.0, 1.0), property_max_iter=100)
    print(r["log"])
    print("densities", r["densities"], "true", dens)
    assert r["final_misfit"] < 1e-2 * c0[3]
    np.testing.assert_allclose(r["densities"], dens, rtol=2e-2)
    P.close()
