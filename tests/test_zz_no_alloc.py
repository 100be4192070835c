"""lf_forward_noalloc (the forward with no device allocation, scratch in a static device
buffer) against lf_create + lf_forward_device on the same model and stations: identical
kernels and chunking, so the results must be identical bit for bit.  Two calls in a row,
the second on a smaller model with fewer stations, in both precisions and both TMI
modes; a model larger than the scratch is refused with LF_ENOMEM."""
import numpy as np
import pytest

from lithoinv3d_amd import lithofwd, synth
from lithoinv3d_amd.lithofwd import ForwardContext

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)


def device_model(m, st, fh, field):
    import torch

    dev = torch.device("cuda:0")
    amp = m.susceptibility * field / 100.0
    mag = amp[:, None] * np.asarray(fh)[None, :]
    t = dict(
        v=torch.tensor(m.vertices.reshape(-1), dtype=torch.float64, device=dev),
        f=torch.tensor(m.faces.astype(np.int32).reshape(-1), dtype=torch.int32, device=dev),
        fr=torch.tensor(m.front, dtype=torch.int32, device=dev),
        bk=torch.tensor(m.back, dtype=torch.int32, device=dev),
        st=torch.tensor(np.ascontiguousarray(st).reshape(-1), dtype=torch.float64, device=dev),
        den=torch.tensor(m.density, dtype=torch.float64, device=dev),
        mag=torch.tensor(mag.reshape(-1), dtype=torch.float64, device=dev),
        gz=torch.full((len(st),), np.nan, dtype=torch.float64, device=dev),
        tmi=torch.full((len(st),), np.nan, dtype=torch.float64, device=dev),
    )
    torch.cuda.synchronize()
    return t


def noalloc(m, st, fh, t, precision, tmi_mode):
    import torch

    lithofwd.forward_noalloc(
        m.n_vertices, m.n_faces, len(st), m.n_units, t["v"].data_ptr(), t["f"].data_ptr(),
        t["fr"].data_ptr(), t["bk"].data_ptr(), t["st"].data_ptr(), t["den"].data_ptr(),
        t["mag"].data_ptr(), fh, gz=t["gz"].data_ptr(), tmi=t["tmi"].data_ptr(),
        precision=precision, tmi_mode=tmi_mode)
    torch.cuda.synchronize()
    return t["gz"].cpu().numpy(), t["tmi"].cpu().numpy()


def with_context(m, st, fh, field, precision, tmi_mode):
    with ForwardContext(m, st, fh, field, precision=precision, tmi_mode=tmi_mode) as ctx:
        return ctx.forward()


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
@pytest.mark.parametrize("tmi_mode", ["reference", "textbook"])
def test_noalloc_matches_context_forward(precision, tmi_mode):
    big = synth.config_contact((24, 20, 12), n_st=40, seed=3, prop_seed=4)
    c0 = synth.config0()
    cases = [(big["model"], big["stations"], "24x20x12 contact model, 1600 stations"),
             (c0["model"], c0["stations"][::5], "icosphere, 205 stations")]
    fh = synth.field_unit_vector(-60.0, 10.0)
    for m, st, name in cases:
        say(f"noalloc vs context forward: {name}, {precision}, TMI {tmi_mode}")
        t = device_model(m, st, fh, 50000.0)
        g1, t1 = noalloc(m, st, fh, t, precision, tmi_mode)
        g0, t0 = with_context(m, st, fh, 50000.0, precision, tmi_mode)
        print(f"{name}: max |dgz| {np.abs(g1 - g0).max():.3e}, max |dTMI| {np.abs(t1 - t0).max():.3e}")
        assert np.array_equal(g1, g0)
        assert np.array_equal(t1, t0)


def test_noalloc_refuses_a_model_larger_than_its_scratch():
    import torch

    say("noalloc: model larger than the scratch")
    dev = torch.device("cuda:0")
    n_f = (lithofwd_scratch_bytes() // 128) + 1
    z = torch.zeros(16, dtype=torch.float64, device=dev)
    zi = torch.zeros(16, dtype=torch.int32, device=dev)
    with pytest.raises(lithofwd.LithoFwdError, match="scratch"):
        lithofwd.forward_noalloc(3, n_f, 4, 1, z.data_ptr(), zi.data_ptr(), zi.data_ptr(),
                                 zi.data_ptr(), z.data_ptr(), z.data_ptr(), z.data_ptr(),
                                 (0.0, 0.0, 1.0), gz=z.data_ptr(), precision="fp32c")


def lithofwd_scratch_bytes():
    import os
    import re

    hdr = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                            "include", "lithofwd.h")).read()
    m = re.search(r"#define LF_NOALLOC_SCRATCH_BYTES \((\d+)ull << (\d+)\)", hdr)
    return int(m.group(1)) << int(m.group(2))
