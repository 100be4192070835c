"""The reference's Forrestania integration test (apps/forrestania/tests/
test_forrestania_integration.cpp:588-729) restated over the GPU inversion loop.

Setup as there (:53-82, :264-291, :588-627, :691-722): units 0-3 are the reference's
volume_group_1..4 (tests/golden/forrestania) with densities 2.67, 2.80, 3.00, 3.50 g/cc
and susceptibilities 0, 0.005, 0.02, 0.08 SI; 8 x 8 stations over the model's (x, y)
extent at z = 0.5 m with 0.05 mGal standard deviation, inset by 1 m from the extent (the
reference's grid at z = 0 puts stations exactly on the model's side faces and on its
horizontal contact facets at z = 0, inside the model, where the solid angle is the
sign of a rounded zero); gz and TMI (inclination -66.2, declination
-0.1, 58 874 nT) of the true model as data; the starting model with every vertex that
can move in z raised by 10 m after finalizeModelSetup (XYZ_FREE, then the exterior
classes of fixExteriorFaces) and bounds from the model box grown by 1 %
(applyModelBounds); axis scale 10, smoothness weight 1 (the runner's lambda), magnetic
weight 1, eigenvalue calibration, L-BFGS history 20, tolerance 1e-6, one geometry
iteration per run.

Checks after every run, as runAllChecks (:480-577): largest displacement from the
start below 500 m; the top and bottom (vertices within 1 m of the extreme z) flat to
0.1 m; more than half of the non-FIXED vertices moved; non-adjacent overlaps at most
baseline + 120 and adjacent ones at most baseline + 10; every vertex inside the start's
bounding box grown by 2 % of its largest side; the validation gate passes at the start's
baselines.  Shared-vertex gaps and edge tearing cannot occur with one shared vertex
array.  Each run's objective (with its calibrated magnetic weight and the smoothness
term, which dominates at the start: the raised interior meets the fixed exterior in a
10 m step) falls.
"""
import os

import numpy as np
import pytest

from lithoinv3d_amd import inversion
from lithoinv3d_amd import litho_io as IO

pytestmark = pytest.mark.gpu

FOR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "forrestania")
INC, DEC, FIELD = -66.2, -0.1, 58874.0


def say(msg):
    print(f"[case] {msg}", flush=True)


def cluster_model():
    meshes = [IO.read_surface(os.path.join(FOR, f"volume_group_{k}.ts")) for k in range(1, 5)]
    return IO.litho_model(meshes, [2.67, 2.80, 3.00, 3.50], [0.0, 0.005, 0.02, 0.08])


def stations(m, n=8, inset=1.0, z=0.5):
    lo, hi = m.vertices.min(0), m.vertices.max(0)
    x = np.linspace(lo[0] + inset, hi[0] - inset, n)
    y = np.linspace(lo[1] + inset, hi[1] - inset, n)
    X, Y = np.meshgrid(x, y)
    return np.column_stack([X.ravel(), Y.ravel(), np.full(n * n, z)])


def setup():
    m = cluster_model()
    st = stations(m)
    M0 = inversion.Model(m)
    P = inversion.Problem(M0, st, np.zeros(len(st)), tmi_obs=np.zeros(len(st)), inc=INC,
                          dec=DEC, field=FIELD)
    gz, tm = P.forward(M0.params())
    P.close()
    M = inversion.Model(m)
    M.fix_exterior_faces(inversion.XYZ_FREE)  # finalizeModelSetup (model_setup.h:9-41)
    M.set_axis_scale(10.0)
    v, ax, _ = M.dofs()
    p = M.params()
    p[ax == 2] += 10.0 / 10.0  # +10 m on every vertex free in z
    V0 = M.apply(p)
    return m, st, gz, tm, M, p, V0


def checks(m, M, V0, V, base, label):
    disp = np.linalg.norm(V - V0, axis=1)
    assert disp.max() < 500.0, label
    zmax, zmin = V0[:, 2].max(), V0[:, 2].min()
    top = np.abs(V0[:, 2] - zmax) <= 1.0
    bot = np.abs(V0[:, 2] - zmin) <= 1.0
    assert np.abs(V[top, 2] - zmax).max() <= 0.1 and np.abs(V[bot, 2] - zmin).max() <= 0.1, label
    fixed = M.freedoms() == inversion.FIXED
    moved = disp > 1e-6
    assert (~fixed).sum() > 0 and moved.sum() > 0.5 * (~fixed).sum(), (label, moved.sum())
    lo, hi = V0.min(0), V0.max(0)
    margin = 0.02 * max((hi - lo).max(), 1.0)
    assert np.all(V >= lo - margin) and np.all(V <= hi + margin), label
    non, adj = base
    r = M.validate(baselines=(non, adj))
    print(f"{label}: max displacement {disp.max():.2f} m, moved {moved.sum()} of "
          f"{(~fixed).sum()} non-FIXED, overlaps adjacent {r['counts'][0]} (baseline {adj}), "
          f"non-adjacent {r['counts'][1]} (baseline {non}), validation {r['passed']}")
    assert r["counts"][1] <= non + 120 and r["counts"][0] <= adj + 10, label
    assert r["passed"], (label, r["reason"])


def problem(M, st, gz, tm):
    n = len(st)
    return inversion.Problem(M, st, gz, np.full(n, 0.05), tm, None, inc=INC, dec=DEC,
                             field=FIELD, alpha_mag=1.0, smooth_weight=1.0)


@pytest.mark.parametrize("n_runs", [1, 5])
def test_forrestania_runner_comprehensive(n_runs):
    """One runner per geometry iteration, carrying the geometry forward, as the
    reference's multi-iteration case does (a fresh runner calibrates alpha_mag from the
    configured weight each time)."""
    m, st, gz, tm, M, p, V0 = setup()
    adj, non = M.validate()["counts"]
    for k in range(n_runs):
        say(f"forrestania runner: run {k + 1} of {n_runs}")
        P = problem(M, st, gz, tm)
        c = P.evaluate(M.params())  # gravity, magnetic (weight 1), regularization, total
        r = P.run(max_iterations=1, lbfgs_history=20, tolerance=1e-6, eigenvalue_scaling=True,
                  validate=True, model_bounds=1.0)
        P.close()
        print(r["log"].strip())
        V = M.apply(r["params"])
        checks(m, M, V0, V, (non, adj), f"run {k + 1}")
        # the run's objective (calibrated magnetic weight, smoothness) falls
        start = c[0] + r["alpha_mag"] * c[1] + c[2]
        print(f"run {k + 1}: objective {start:.6e} -> {r['final_misfit']:.6e}")
        assert r["final_misfit"] < start
