"""A run driven by the reference's own configuration file: apps/forrestania/test_config.ini
(tests/golden/forrestania) read by InversionConfig.from_ini, its settings handed to the
model (Z_ONLY with the exterior classification, axis scale 10, control points every 5th
cell), the problem (inducing field -66.2 / -0.1 / 58 874 nT, magnetic weight 0.5,
smoothness weight 0.01) and the inversion loop (L-BFGS-B history 20, tolerance 1e-5,
Armijo 1e-3, 30 line-search trials, calibration, property phases with densities bounded
to 1.5..5.0 g/cc), on the reference's Forrestania units 1-4 with synthetic data of the
true model and a start whose z parameters are raised by 10 m.  The run is shortened to 4 geometry iterations in
phases of 2, so that property phases run between them.

Checks: history entry 0 is the start's objective with the calibrated magnetic weight as
Problem.calibrate finds it, the objective falls, the unit densities stay inside the
file's bounds, and the log shows the property phases.
"""
import os

import numpy as np
import pytest

from lithoinv3d_amd import inversion
from lithoinv3d_amd import litho_io as IO

pytestmark = pytest.mark.gpu

FOR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "forrestania")


def say(msg):
    print(f"[case] {msg}", flush=True)


def test_run_from_the_reference_configuration_file():
    cfg = inversion.InversionConfig.from_ini(os.path.join(FOR, "test_config.ini"))
    meshes = [IO.read_surface(os.path.join(FOR, f"volume_group_{k}.ts")) for k in range(1, 5)]
    m = IO.litho_model(meshes, [2.67, 2.80, 3.00, 3.50], [0.0, 0.005, 0.02, 0.08])
    lo, hi = m.vertices.min(0), m.vertices.max(0)
    X, Y = np.meshgrid(np.linspace(lo[0] + 1, hi[0] - 1, 8), np.linspace(lo[1] + 1, hi[1] - 1, 8))
    st = np.column_stack([X.ravel(), Y.ravel(), np.full(64, 0.5)])
    pk = cfg.problem_kwargs()
    M0 = inversion.Model(m)
    P0 = inversion.Problem(M0, st, np.zeros(64), tmi_obs=np.zeros(64), inc=pk["inc"],
                           dec=pk["dec"], field=pk["field"])
    gz, tm = P0.forward(M0.params())
    P0.close()
    M = inversion.Model(m)
    n = cfg.apply_to_model(M)
    v, ax, _ = M.dofs()
    # the exterior classification gives some boundary vertices x or y instead of z
    assert n == 64 and (ax == 2).sum() == 24
    p0 = M.params()
    p0[ax == 2] += 1.0  # +10 m on every z parameter (axis scale 10)
    M.apply(p0)
    say("config run: calibration at the start")
    P = inversion.Problem(M, st, gz, np.full(64, 0.05), tm, np.full(64, 1.0), **pk)
    cal = P.calibrate(p0, seed=1)
    P.close()
    M.apply(p0)
    P = inversion.Problem(M, st, gz, np.full(64, 0.05), tm, np.full(64, 1.0), **pk)
    c0 = P.evaluate(p0)
    rk = cfg.run_kwargs()
    rk.update(max_iterations=4, property_interval=2)
    say("config run: 4 geometry iterations with property phases")
    r = P.run(**rk)
    print(r["log"])
    h = r["history"]
    np.testing.assert_allclose(h[0, 1:5], c0, rtol=1e-12)
    assert r["alpha_mag"] == pytest.approx(cal["alpha_mag"], rel=1e-12)
    assert r["final_misfit"] < h[0, 4]
    assert r["log"].count("property phase") >= 1
    dens = r["densities"]
    assert (dens >= cfg.density_min - 1e-12).all() and (dens <= cfg.density_max + 1e-12).all()
    P.close()
