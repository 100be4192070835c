"""File formats (host C++ through ctypes, no GPU): GOCAD TSurf / OBJ surfaces, CSV
stations, the binary label grid, and the assembly of per-unit meshes into the shared
facet set, against hand-written files with known content, round trips, the Forrestania
data files of the reference (tests/golden/forrestania, copied by copy_fixtures.sh), and
the oracle forward of the per-unit (reference) layout."""
import os

import numpy as np
import pytest

from lithoinv3d_amd import litho_io as IO
from lithoinv3d_amd import synth

HERE = os.path.dirname(os.path.abspath(__file__))
FOR = os.path.join(HERE, "golden", "forrestania")


def write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return p


def test_tsurf_parse(tmp_path):
    p = write(tmp_path, "a.ts", """GOCAD TSurf 1
HEADER {
name:demo
*solid*color: 1 0 0 1
}
# a comment
TFACE
VRTX 1 0.0 0.0 0.0
PVRTX 2 1.0 0.0 0.0 7.5
  VRTX 3 0.0 1.0 0.0
VRTX 4 0.0 0.0 1.0
VRTX 5 not a vertex
TRGL 1 3 2
TRGL 1 2 4
TRGL 2 3 4
TRGL 1 4 3
END
VRTX 9 5 5 5
TRGL 1 2 9
""")
    m = IO.read_surface(p)
    np.testing.assert_array_equal(m.vertices, [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]])
    np.testing.assert_array_equal(m.triangles, [[0, 2, 1], [0, 1, 3], [1, 2, 3], [0, 3, 2]])
    assert "invalid vertex" in m.warnings  # reported, skipped


def test_tsurf_rejects_bad_index(tmp_path):
    p = write(tmp_path, "b.ts", "GOCAD TSurf 1\nTFACE\nVRTX 1 0 0 0\nVRTX 2 1 0 0\nTRGL 1 2 3\nEND\n")
    with pytest.raises(RuntimeError, match="references vertex 3"):
        IO.read_surface(p)
    with pytest.raises(RuntimeError, match="no vertices"):
        IO.read_surface(write(tmp_path, "c.ts", "GOCAD TSurf 1\nTFACE\nEND\n"))
    with pytest.raises(RuntimeError, match="unrecognized extension"):
        IO.read_surface(write(tmp_path, "d.stl", "solid x\n"))


def test_tsurf_write_read_roundtrip(tmp_path):
    V, F = synth.icosphere(2)
    V = V * 200.0 + np.array([744400.123, 6414800.456, -500.0])
    p = tmp_path / "sphere.ts"
    IO.write_ts(p, V, F, name="sphere")
    text = p.read_text().splitlines()
    assert text[:5] == ["GOCAD TSurf 1", "HEADER {", "  name: sphere", "}", "TFACE"]
    assert text[5].startswith("VRTX 1 ") and "e+05" in text[5] and text[-1] == "END"
    m = IO.read_surface(p)
    np.testing.assert_array_equal(m.triangles, F)
    # 8 digits after the point in scientific notation: 1e-8 relative
    np.testing.assert_allclose(m.vertices, V, rtol=6e-9, atol=0)


def test_obj_parse(tmp_path):
    p = write(tmp_path, "q.obj", """# quad and triangle
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
vn 0 0 1
f 1/1/1 2/2/1 3/3/1 4/4/1
fo 1 3 2
f 1 2
""")
    m = IO.read_surface(p)
    np.testing.assert_array_equal(m.triangles, [[0, 1, 2], [0, 2, 3], [0, 2, 1]])
    assert "invalid face" in m.warnings


def test_station_csv(tmp_path):
    p = write(tmp_path, "g.csv", """X,Y,Z,FGrav_mgal
# comment line
100.5, 200.25 ,300,0.125
101,201,301,0.5,0.05

102,202,302
,103,203,303,0.75
104,205,305,12abc
nan,1,2,3
105,206,306,bad,1
""")
    s = IO.read_stations(p)
    np.testing.assert_array_equal(s.xyz, [[100.5, 200.25, 300], [101, 201, 301], [103, 203, 303],
                                          [104, 205, 305]])
    # the fifth field is the standard deviation (0 = unweighted); a leading empty field is
    # ignored; a numeric prefix ("12abc") is read as the number, as std::stod does
    np.testing.assert_array_equal(s.obs, [0.125, 0.5, 0.75, 12.0])
    np.testing.assert_array_equal(s.stdev, [0.0, 0.05, 0.0, 0.0])
    assert "expected at least 4 values" in s.warnings and "'bad' is not a number" in s.warnings


def test_label_grid_roundtrip(tmp_path):
    rng = np.random.default_rng(3)
    lab = rng.integers(-1, 5, size=(7, 5, 3)).astype(np.int32)
    p = tmp_path / "lg.bin"
    IO.write_label_grid(p, lab, [744200.0, 6414800.0, 800.0], [200.0, 200.0, 200.0])
    header = p.read_bytes().split(b"\n", 1)[0].split()
    assert [int(x) for x in header[:3]] == [7, 5, 3]
    g = IO.read_label_grid(p)
    np.testing.assert_array_equal(g.labels, lab)
    np.testing.assert_array_equal(g.origin, [744200.0, 6414800.0, 800.0])
    # flat order: i fastest (label (i, j, k) at i + nx*(j + ny*k))
    flat = np.frombuffer(p.read_bytes().split(b"\n", 1)[1], np.int32)
    assert flat[1 + 7 * (2 + 5 * 1)] == lab[1, 2, 1]


def cube(x0):
    """Closed outward unit cube [x0, x0+1] x [0, 1] x [0, 1] as 12 triangles."""
    V = np.array([[x, y, z] for z in (0, 1) for y in (0, 1) for x in (x0, x0 + 1)], float)
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    F = []
    for a, b, c, d in quads:
        F += [(a, b, c), (a, c, d)]
    return V, np.array(F, np.uint32)


def test_assemble_two_cubes(oracle):
    A, B = cube(0.0), cube(1.0)
    a = IO.assemble([A, B])
    assert a.vertices.shape == (12, 3) and a.merged_vertices == 4
    assert len(a.faces) == 22 and a.contacts == 2
    contact = a.back >= 0
    np.testing.assert_array_equal(a.front[contact], [0, 0])
    np.testing.assert_array_equal(a.back[contact], [1, 1])
    # the contact facets keep unit 0's winding: normal +x
    V, F = a.vertices, a.faces[contact].astype(np.int64)
    n = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    assert np.all(n[:, 0] > 0)
    # forward of the assembled model == forward of the two separate closed meshes
    dens = np.array([0.3, -0.2])
    m = synth.LithoModel(a.vertices, a.faces, a.front, a.back, dens, np.array([0.01, 0.02]))
    st = np.array([[0.3, 0.4, 2.5], [1.7, -0.6, 1.9], [-1.0, 2.0, 3.0]])
    ref = oracle.OracleModel(np.concatenate([A[0], B[0]]),
                             [A[1].astype(np.int64), B[1].astype(np.int64) + 8], dens)
    np.testing.assert_allclose(oracle.OracleModel.from_litho(m).gravity(st), ref.gravity(st),
                               rtol=1e-12, atol=1e-15)


def forrestania_meshes():
    return [IO.read_surface(os.path.join(FOR, f"volume_group_{k}.ts")) for k in range(1, 9)]


def test_forrestania_files(oracle):
    meshes = forrestania_meshes()
    # vertex and triangle counts of the eight files (their VRTX and TRGL lines)
    assert [len(m.vertices) for m in meshes] == [1733, 3540, 1087, 199, 253, 1663, 210, 88]
    assert [len(m.triangles) for m in meshes] == [3568, 7236, 2232, 392, 508, 3368, 412, 172]
    assert all(m.warnings == "" for m in meshes)
    a = IO.assemble(meshes)
    assert a.contacts == 6696 and len(a.faces) == 17888 - 6696
    g = IO.read_label_grid(os.path.join(FOR, "label_grid.bin"))
    assert g.labels.shape == (28, 25, 10)
    np.testing.assert_array_equal(g.origin, [744200.0, 6414800.0, 800.0])
    np.testing.assert_array_equal(g.cell, [200.0, 200.0, 200.0])
    assert set(np.unique(g.labels)) == set(range(-1, 8))
    s = IO.read_stations(os.path.join(FOR, "Forrestania_Gravity_Station_trim_.csv"))
    assert len(s.obs) == 992 and s.warnings == ""
    np.testing.assert_array_equal(s.xyz[0], [748151.81, 6416320.50, 380.31])
    assert s.obs[0] == 0.147
    # the assembled facet set reproduces the per-unit layout's forward (oracle, FP64)
    dens = np.linspace(-0.1, 0.2, 8)
    sus = np.linspace(1e-3, 3e-2, 8)
    fh = synth.field_unit_vector(-65.0, 0.0)
    V = np.concatenate([m.vertices for m in meshes])
    off = np.cumsum([0] + [len(m.vertices) for m in meshes])
    ref = oracle.OracleModel(V, [m.triangles.astype(np.int64) + off[k]
                                 for k, m in enumerate(meshes)],
                             dens, sus * 500.0, np.tile(fh, (8, 1)))
    mm = IO.litho_model(meshes, dens, sus)
    om = oracle.OracleModel.from_litho(mm, 50000.0, fh)
    st = s.xyz[::40]
    np.testing.assert_allclose(om.gravity(st), ref.gravity(st), rtol=0, atol=1e-12)
    np.testing.assert_allclose(om.magnetic(fh, st), ref.magnetic(fh, st), rtol=1e-12, atol=1e-8)


def test_tsurf_writer_reference_cases(tmp_path):
    """modules/litho-io/tests/test_ts_content.cpp restated for the TSurf writer: 8
    digits after the point in scientific notation (123.456789012345 -> 1.23456789e+02),
    1-based triangle indices, every vertex written once, an empty mesh still has the
    header, TFACE and END with no VRTX or TRGL line, printable ASCII only, and the same
    mesh written twice gives identical files."""
    V = np.array([[123.456789012345, 0.0, -50.0], [0.0, 0.0, 0.0], [1.0, 1.0, 1.0]])
    T = np.array([[0, 1, 2]], np.uint32)
    p = tmp_path / "precision.ts"
    IO.write_ts(p, V, T, name="precision")
    lines = p.read_text().splitlines()
    v1 = next(l for l in lines if l.startswith("VRTX 1 "))
    assert v1.split()[2] == "1.23456789e+02"
    assert [l for l in lines if l.startswith("TRGL")] == ["TRGL 1 2 3"]
    assert [l.split()[1] for l in lines if l.startswith("VRTX")] == ["1", "2", "3"]
    raw = p.read_bytes()
    assert all(b in (10, 13) or 0x20 <= b <= 0x7E for b in raw)
    q = tmp_path / "again.ts"
    IO.write_ts(q, V, T, name="precision")
    assert q.read_bytes() == raw
    e = tmp_path / "empty.ts"
    IO.write_ts(e, np.zeros((0, 3)), np.zeros((0, 3), np.uint32), name="empty")
    text = e.read_text()
    assert "GOCAD TSurf 1" in text and "TFACE" in text and text.rstrip().endswith("END")
    assert not any(l.startswith(("VRTX", "TRGL")) for l in text.splitlines())


def test_unit_volume_export(tmp_path, oracle):
    """li_model_export_units (InversionExporter::exportClosedVolume over the units,
    exporters.cpp:769-813): one closed GOCAD TSurf per unit at the current geometry, front
    facets as stored and back facets reversed.  Reading the files back and assembling them
    gives the model again (same facet set, same units on each side), and the oracle's
    gravity of the per-unit files equals the model's."""
    from lithoinv3d_amd import inversion

    rng = np.random.default_rng(5)
    labels = rng.integers(0, 3, size=(4, 3, 3)).astype(np.int32)
    dens = np.array([0.25, -0.15, 0.4])
    m = synth.label_grid_model(labels, (-50.0, -37.5, -30.0), (25.0, 25.0, 20.0), dens,
                               [0.02, 0.003, 0.04])
    M = inversion.Model(m)
    M.fix_exterior(inversion.Z_ONLY)
    p = M.params() + rng.normal(0, 0.3, M.n_dof)
    moved = M.apply(p)
    out = tmp_path / "units"
    assert M.export_units(out, names=["granite", "sulfide"]) == 3
    files = sorted(os.listdir(out))
    assert files == ["invert_granite.ts", "invert_group_2.ts", "invert_sulfide.ts"]
    text = (out / "invert_granite.ts").read_text().splitlines()
    assert text[:3] == ["GOCAD TSurf 1", "HEADER {", "  name: invert_granite"]
    assert text[4] == "TFACE" and text[-1] == "END"
    meshes = [IO.read_surface(out / f) for f in
              ("invert_granite.ts", "invert_sulfide.ts", "invert_group_2.ts")]
    for g, mesh in enumerate(meshes):
        tri = m.group_triangles(g)
        used = np.unique(tri)
        np.testing.assert_allclose(mesh.vertices, moved[used], rtol=1e-8, atol=1e-6)
        np.testing.assert_array_equal(used[mesh.triangles], tri)
    st = np.array([[0.0, 0.0, 15.0], [30.0, -20.0, 40.0], [-45.0, 35.0, 5.0]])
    per_unit = oracle.OracleModel(np.concatenate([mm.vertices for mm in meshes]),
                                  [mm.triangles.astype(np.int64) + off for mm, off in
                                   zip(meshes, np.cumsum([0] + [len(mm.vertices)
                                                               for mm in meshes[:-1]]))],
                                  dens)
    moved_model = synth.LithoModel(moved, m.faces, m.front, m.back, dens, m.susceptibility)
    np.testing.assert_allclose(per_unit.gravity(st),
                               oracle.OracleModel.from_litho(moved_model).gravity(st),
                               rtol=1e-7, atol=1e-9)
