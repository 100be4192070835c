"""The FP32C accuracy contract on the 192x192x64-cell contact model (2.24 M facets,
65 536 stations): every station's FP32C gz and TMI against the FP64 GPU path (itself held
to the FP64 oracle by test_gpu_forward.py and test_gpu_forrestania.py), in both TMI
coefficient modes, and the FP32C adjoint against the FP64 adjoint with all 65 536
stations weighted.

Bounds (absolute, over all stations): gz 1e-3 mGal; TMI 1e-2 nT, on the reference
formula's field of about -1.2e6 nT (8e-9 relative) as well as on the textbook field of a
few hundred nT.  Adjoint: 1e-5 of the largest gradient component."""
import numpy as np
import pytest

from lithoinv3d_amd import synth, workload
from lithoinv3d_amd.lithofwd import ForwardContext

pytestmark = pytest.mark.gpu

GZ_BOUND = 1e-3    # mGal
TMI_BOUND = 1e-2   # nT


def say(msg):
    print(f"[case] {msg}", flush=True)


@pytest.fixture(scope="module")
def scale_model():
    c = workload.build("contact192x192x64_st256")
    c["fh"] = synth.field_unit_vector(c["inc"], c["dec"])
    return c


@pytest.mark.parametrize("tmi_mode", ["reference", "textbook"])
def test_fp32c_forward_all_stations(scale_model, tmi_mode):
    c = scale_model
    m, st, fh = c["model"], c["stations"], c["fh"]
    out = {}
    for prec in ("fp64", "fp32c"):
        say(f"192x192x64 forward, {len(st)} stations, {prec}, TMI {tmi_mode}")
        with ForwardContext(m, st, fh, c["field"], precision=prec, tmi_mode=tmi_mode) as ctx:
            out[prec] = ctx.forward()
    dg = np.abs(out["fp32c"][0] - out["fp64"][0]).max()
    dt = np.abs(out["fp32c"][1] - out["fp64"][1]).max()
    t64 = out["fp64"][1]
    print(f"TMI {tmi_mode}: max |dgz| {dg:.2e} mGal, max |dTMI| {dt:.2e} nT "
          f"(TMI {t64.min():.1f} .. {t64.max():.1f} nT)")
    assert dg <= GZ_BOUND
    assert dt <= TMI_BOUND


def test_fp32c_adjoint_all_stations(scale_model):
    c = scale_model
    m, st, fh = c["model"], c["stations"], c["fh"]
    rng = np.random.default_rng(7)
    wg = rng.normal(size=len(st))
    wt = rng.normal(size=len(st)) * 1e-3
    g = {}
    for prec in ("fp64", "fp32c"):
        say(f"192x192x64 adjoint, {len(st)} stations, {prec}")
        with ForwardContext(m, st, fh, c["field"], precision=prec) as ctx:
            g[prec] = ctx.adjoint(wg, wt)
    scale = np.abs(g["fp64"]).max()
    e = np.abs(g["fp32c"] - g["fp64"]).max() / scale
    print(f"FP32C vs FP64 adjoint over {m.n_vertices} vertices x 3: {e:.2e} of the largest "
          f"component ({scale:.3e})")
    assert e <= 1e-5
