"""Host-side checks (no GPU): the C-ABI library loads and exports every entry point the
header declares, the generators build closed consistently oriented meshes, and the
Python layer rejects malformed input before anything reaches a kernel."""
import os
import re

import numpy as np
import pytest

from lithoinv3d_amd import synth
from lithoinv3d_amd import lithofwd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(prefix="lf_"):
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\w[\w\s\*]*?\b(" + prefix + r"\w+)\s*\(", txt, re.M))
    return names


def test_library_exports_declared_symbols():
    L = lithofwd.load_library()
    names = declared_symbols()
    assert {"lf_create", "lf_forward", "lf_adjoint", "lf_destroy"} <= names
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    assert set(lithofwd.EXPORTED) | set(lithofwd.EXPORTED_MLP) == names


def test_host_library_exports_declared_symbols():
    """liblithohost.so exports every li_ entry point include/lithohost.h declares, and
    declares every li_ function its C source defines."""
    from lithoinv3d_amd import inversion
    L = inversion.load_host()
    names = declared_symbols("li_")
    assert {"li_model_create", "li_problem_create", "li_problem_minimize",
            "li_problem_calibrate", "li_assemble"} <= names
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    src = open(os.path.join(ROOT, "lithoinv3d_amd", "host", "litho_capi.cpp")).read()
    defined = set(re.findall(r"^(?:int|void|int64_t|const char\*) (li_\w+)\(", src, re.M))
    assert defined == names, sorted(defined ^ names)


def directed_edges(F):
    return np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]], 0)


def assert_closed_oriented(tris):
    """Every directed edge a->b is matched by as many b->a (a closed, consistently
    oriented surface; voxel meshes may be non-manifold where a unit touches itself
    along an edge, so counts can exceed one)."""
    e = directed_edges(tris.astype(np.int64))
    key = e[:, 0] * (1 << 32) + e[:, 1]
    rkey = e[:, 1] * (1 << 32) + e[:, 0]
    k, kc = np.unique(key, return_counts=True)
    r, rc = np.unique(rkey, return_counts=True)
    assert np.array_equal(k, r) and np.array_equal(kc, rc), "open or inconsistently oriented"


def signed_volume(V, F):
    return np.einsum("ij,ij->i", V[F[:, 0]], np.cross(V[F[:, 1]], V[F[:, 2]])).sum() / 6.0


def test_icosphere():
    V, F = synth.icosphere(3)
    assert V.shape == (642, 3) and F.shape == (1280, 3)
    assert_closed_oriented(F)
    assert signed_volume(V, F) > 0
    np.testing.assert_allclose(np.linalg.norm(V, axis=1), 1.0, atol=1e-14)


@pytest.mark.parametrize("seed", [0, 1])
def test_label_grid_units_closed(seed):
    rng = np.random.default_rng(seed)
    shape = (7, 6, 4)
    labels = rng.integers(0, 3, size=shape).astype(np.int32)
    labels[rng.random(shape) < 0.15] = -1
    cell = (25.0, 20.0, 10.0)
    m = synth.label_grid_model(labels, (100.0, -50.0, -5.0), cell, [0.1, 0.2, 0.3], [0, 0, 0])
    assert m.front.min() >= 0 and (m.back >= -1).all()
    assert (m.front != m.back).all()
    for g in range(3):
        T = m.group_triangles(g)
        if (labels == g).any():
            assert_closed_oriented(T)
            vol = signed_volume(m.vertices, T)
            assert vol == pytest.approx((labels == g).sum() * np.prod(cell), rel=1e-12)
    # contact faces are stored once; the reference layout counts them twice
    assert m.reference_face_count() == m.n_faces + int((m.back >= 0).sum())
    # the reference emits, per quad, (v0,v2,v1),(v0,v3,v2) for the unit on the
    # lower-index side: outward normal of a front facet points from front to back
    V = m.vertices
    n = np.cross(V[m.faces[:, 1]] - V[m.faces[:, 0]], V[m.faces[:, 2]] - V[m.faces[:, 0]])
    assert (np.abs(n).max(axis=1) > 0).all()


def test_config_contact_small():
    c = synth.config_contact(shape=(16, 16, 8), n_st=16)
    m = c["model"]
    assert m.n_units == 4
    assert len(c["stations"]) == 256
    assert (m.density >= -0.3).all() and (m.density <= 0.5).all()
    assert (m.susceptibility >= 1e-4).all() and (m.susceptibility <= 5e-2).all()
    # every station above the model top
    assert c["stations"][:, 2].min() > m.vertices[:, 2].max()
    for g in range(4):
        assert_closed_oriented(m.group_triangles(g))


def test_forward_context_rejects_bad_input():
    pytest.importorskip("ctypes")
    m = synth.box_model((0, 0, -10), (10, 10, 0))
    bad = synth.LithoModel(m.vertices, m.faces.copy(), m.front, m.back, m.density,
                           m.susceptibility)
    bad.faces[0, 0] = 999  # out-of-range vertex id: rejected on the host
    with pytest.raises(lithofwd.LithoFwdError, match="references vertex"):
        lithofwd.ForwardContext(bad, np.zeros((1, 3)))
    bad2 = synth.LithoModel(m.vertices, m.faces, m.front.copy(), m.back, m.density,
                            m.susceptibility)
    bad2.front[0] = 5
    with pytest.raises(lithofwd.LithoFwdError, match="invalid unit id"):
        lithofwd.ForwardContext(bad2, np.zeros((1, 3)))


def test_station_subset():
    """One station per cell of a grid `stride` spacings wide, the one nearest the cell
    centre: a regular 64 x 64 grid at stride 32 keeps 4 stations, one per quadrant, each
    within one spacing of its quadrant's centre; stride 1 keeps every station of the
    grid; an irregular set keeps one station per occupied cell."""
    from lithoinv3d_amd import mlp
    st = synth.topo_stations(64, 1600.0)
    idx, cell = mlp.station_subset(st, 32)
    assert len(idx) == 4 and len(set(idx)) == 4
    xy = st[idx, :2]
    lo = st[:, :2].min(0)
    d = (st[1, 0] - st[0, 0]) if st[1, 0] != st[0, 0] else (st[1, 1] - st[0, 1])
    q = np.floor((xy - lo + abs(d) / 2) / cell).astype(int)
    assert sorted(map(tuple, q)) == [(0, 0), (0, 1), (1, 0), (1, 1)]
    np.testing.assert_array_less(np.abs(xy - (lo - abs(d) / 2 + (q + 0.5) * cell)), abs(d) + 1e-9)
    idx1, _ = mlp.station_subset(st, 1)
    assert len(idx1) == len(st)
    rng = np.random.default_rng(1)
    pts = rng.uniform(0, 1000, (500, 3))
    idx, cell = mlp.station_subset(pts, 4)
    org = pts[:, :2].min(0) - cell / 8
    occupied = {tuple(c) for c in np.floor((pts[:, :2] - org) / cell).astype(int)}
    assert len(idx) == len(occupied)
