"""File formats of LithoInv3D around the forward path, through the C++ host layer
(lithoinv3d_amd/host/litho_io.cpp): GOCAD TSurf / OBJ unit surfaces
(modules/litho-io/src/io/surface_loader.cpp, exporters.cpp:116-150), CSV gravity and
magnetic stations (gravity_loader.cpp:8-69, magnetic_loader.cpp:8-69), the binary label
grid (python/lithoseed/src/lithoseed/_surfaces.py:659-671), and the assembly of per-unit
closed meshes into the shared-vertex facet set the GPU forward consumes."""
from __future__ import annotations

import ctypes
import dataclasses
import os

import numpy as np

from . import inversion, synth

_P = ctypes.c_void_p
_i64 = ctypes.c_int64
_declared = False


def _lib():
    global _declared
    L = inversion.load_host()
    if not _declared:
        L.li_mesh_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(_P)]
        L.li_mesh_counts.argtypes = [_P, ctypes.POINTER(_i64), ctypes.POINTER(_i64)]
        L.li_mesh_get.argtypes = [_P, _P, _P]
        L.li_mesh_warnings.argtypes = [_P]
        L.li_mesh_warnings.restype = ctypes.c_char_p
        L.li_mesh_destroy.argtypes = [_P]
        L.li_mesh_destroy.restype = None
        L.li_mesh_write_ts.argtypes = [ctypes.c_char_p, ctypes.c_char_p, _i64, _P, _i64, _P]
        L.li_stations_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(_P)]
        L.li_stations_count.argtypes = [_P]
        L.li_stations_count.restype = _i64
        L.li_stations_get.argtypes = [_P, _P, _P, _P]
        L.li_stations_warnings.argtypes = [_P]
        L.li_stations_warnings.restype = ctypes.c_char_p
        L.li_stations_destroy.argtypes = [_P]
        L.li_stations_destroy.restype = None
        L.li_constraints_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(_P)]
        L.li_constraints_count.argtypes = [_P]
        L.li_constraints_count.restype = _i64
        L.li_constraints_get.argtypes = [_P, _P, _P, _P, _P]
        L.li_constraints_warnings.argtypes = [_P]
        L.li_constraints_warnings.restype = ctypes.c_char_p
        L.li_constraints_destroy.argtypes = [_P]
        L.li_constraints_destroy.restype = None
        L.li_label_grid_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(_P)]
        L.li_label_grid_info.argtypes = [_P, _P, _P, _P]
        L.li_label_grid_get.argtypes = [_P, _P]
        L.li_label_grid_destroy.argtypes = [_P]
        L.li_label_grid_destroy.restype = None
        L.li_label_grid_write.argtypes = [ctypes.c_char_p, _P, _P, _P, _P]
        L.li_assemble.argtypes = [ctypes.c_int32, _P, _P, _P, _P, ctypes.POINTER(_P)]
        L.li_assembled_counts.argtypes = [_P] + [ctypes.POINTER(_i64)] * 4
        L.li_assembled_get.argtypes = [_P, _P, _P, _P, _P]
        L.li_assembled_destroy.argtypes = [_P]
        L.li_assembled_destroy.restype = None
        cs = ctypes.c_char_p
        L.li_ini_load.argtypes = [cs, ctypes.POINTER(_P)]
        L.li_ini_parse.argtypes = [cs, ctypes.POINTER(_P)]
        L.li_ini_destroy.argtypes = [_P]
        L.li_ini_destroy.restype = None
        L.li_ini_get.argtypes = [_P, cs, cs, cs, _i64, ctypes.POINTER(ctypes.c_int32)]
        L.li_ini_get_double.argtypes = [_P, cs, cs, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
        L.li_ini_get_int.argtypes = [_P, cs, cs, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
        L.li_ini_get_bool.argtypes = [_P, cs, cs, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
        L.li_ini_set.argtypes = [_P, cs, cs, cs]
        L.li_ini_save.argtypes = [_P, cs]
        L.li_ini_dump.argtypes = [_P, cs, _i64, ctypes.POINTER(_i64)]
        _declared = True
    return L


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what}: {inversion.load_host().li_last_error().decode()}")


def _p(a):
    return a.ctypes.data_as(_P)


def _path(p):
    return os.fsencode(os.fspath(p))


@dataclasses.dataclass
class Mesh:
    vertices: np.ndarray   # (n, 3) float64
    triangles: np.ndarray  # (m, 3) uint32, 0-based
    warnings: str = ""


def read_surface(path) -> Mesh:
    """A .ts (GOCAD TSurf) or .obj unit surface, by extension."""
    L = _lib()
    h = _P()
    _chk(L.li_mesh_load(_path(path), ctypes.byref(h)), f"read_surface({path})")
    try:
        nv, nt = _i64(), _i64()
        L.li_mesh_counts(h, ctypes.byref(nv), ctypes.byref(nt))
        v = np.zeros((nv.value, 3))
        t = np.zeros((nt.value, 3), np.uint32)
        L.li_mesh_get(h, _p(v), _p(t))
        w = L.li_mesh_warnings(h).decode()
    finally:
        L.li_mesh_destroy(h)
    return Mesh(v, t, w)


def write_ts(path, vertices, triangles, name="surface"):
    """GOCAD TSurf in the reference exporter's layout (8-digit scientific coordinates)."""
    v = np.ascontiguousarray(vertices, np.float64).reshape(-1, 3)
    t = np.ascontiguousarray(triangles, np.uint32).reshape(-1, 3)
    _chk(_lib().li_mesh_write_ts(_path(path), name.encode(), len(v), _p(v), len(t), _p(t)),
         f"write_ts({path})")


@dataclasses.dataclass
class Stations:
    xyz: np.ndarray    # (n, 3)
    obs: np.ndarray    # mGal or nT
    stdev: np.ndarray  # 0 = unweighted
    warnings: str = ""


def read_stations(path) -> Stations:
    """x, y, z, observation[, standard deviation] per line; '#' comments and header
    lines (first field starting with a letter) skipped, short lines skipped."""
    L = _lib()
    h = _P()
    _chk(L.li_stations_load(_path(path), ctypes.byref(h)), f"read_stations({path})")
    try:
        n = L.li_stations_count(h)
        xyz, obs, sd = np.zeros((n, 3)), np.zeros(n), np.zeros(n)
        L.li_stations_get(h, _p(xyz), _p(obs), _p(sd))
        w = L.li_stations_warnings(h).decode()
    finally:
        L.li_stations_destroy(h)
    return Stations(xyz, obs, sd, w)


@dataclasses.dataclass
class Constraints:
    xy: np.ndarray        # (n, 2) borehole collars
    z_top: np.ndarray     # interval top depth (m, positive down)
    z_bottom: np.ndarray  # interval bottom depth
    unit: np.ndarray      # unit id of the interval
    warnings: str = ""


def read_constraints(path) -> Constraints:
    """Borehole intervals: x, y, z_top, z_bottom, unit per line (depths positive down)."""
    L = _lib()
    h = _P()
    _chk(L.li_constraints_load(_path(path), ctypes.byref(h)), f"read_constraints({path})")
    try:
        n = L.li_constraints_count(h)
        xy, zt, zb = np.zeros((n, 2)), np.zeros(n), np.zeros(n)
        u = np.zeros(n, np.int32)
        L.li_constraints_get(h, _p(xy), _p(zt), _p(zb), _p(u))
        w = L.li_constraints_warnings(h).decode()
    finally:
        L.li_constraints_destroy(h)
    return Constraints(xy, zt, zb, u, w)


@dataclasses.dataclass
class ContactPicks:
    xyz: np.ndarray    # (n, 3): collar x, y and the contact's elevation (z up)
    above: np.ndarray  # unit above the contact
    below: np.ndarray  # unit below it
    sigma: np.ndarray  # standard deviation of the elevation (m)


def contact_picks(c: Constraints, sigma=5.0, gap=1e-6) -> ContactPicks:
    """Depth-to-contact picks from borehole intervals: wherever two intervals of one
    borehole (same collar) touch (bottom of one = top of the next, within gap) and hold
    different units, a contact at elevation -depth between the upper and the lower unit."""
    xyz, above, below = [], [], []
    keys = {}
    for i, (x, y) in enumerate(c.xy):
        keys.setdefault((float(x), float(y)), []).append(i)
    for (x, y), idx in keys.items():
        idx = sorted(idx, key=lambda i: c.z_top[i])
        for i, j in zip(idx[:-1], idx[1:]):
            if abs(c.z_bottom[i] - c.z_top[j]) <= gap and c.unit[i] != c.unit[j]:
                xyz.append((x, y, -c.z_bottom[i]))
                above.append(c.unit[i])
                below.append(c.unit[j])
    n = len(xyz)
    return ContactPicks(np.array(xyz, np.float64).reshape(n, 3), np.array(above, np.int32),
                        np.array(below, np.int32), np.full(n, float(sigma)))


@dataclasses.dataclass
class LabelGrid:
    labels: np.ndarray  # (nx, ny, nz) int32
    origin: np.ndarray  # x0, y0, z0
    cell: np.ndarray    # dx, dy, dz


def read_label_grid(path) -> LabelGrid:
    L = _lib()
    h = _P()
    _chk(L.li_label_grid_load(_path(path), ctypes.byref(h)), f"read_label_grid({path})")
    try:
        dims = np.zeros(3, np.int32)
        origin, cell = np.zeros(3), np.zeros(3)
        L.li_label_grid_info(h, _p(dims), _p(origin), _p(cell))
        flat = np.zeros(int(np.prod(dims)), np.int32)
        L.li_label_grid_get(h, _p(flat))
    finally:
        L.li_label_grid_destroy(h)
    return LabelGrid(flat.reshape(tuple(dims), order="F"), origin, cell)


def write_label_grid(path, labels, origin, cell):
    lab = np.asarray(labels, np.int32)
    dims = np.array(lab.shape, np.int32)
    flat = np.ascontiguousarray(lab.ravel(order="F"))
    o = np.ascontiguousarray(origin, np.float64)
    c = np.ascontiguousarray(cell, np.float64)
    _chk(_lib().li_label_grid_write(_path(path), _p(dims), _p(o), _p(c), _p(flat)),
         f"write_label_grid({path})")


@dataclasses.dataclass
class Assembly:
    vertices: np.ndarray
    faces: np.ndarray
    front: np.ndarray
    back: np.ndarray
    contacts: int
    merged_vertices: int


def assemble(meshes) -> Assembly:
    """Per-unit meshes (each a Mesh or a (vertices, triangles) pair, unit u = position in
    the list) -> one shared-vertex facet set; a triangle whose reversed copy is in another
    unit becomes one facet with (front, back)."""
    vs, ts = [], []
    for m in meshes:
        v, t = (m.vertices, m.triangles) if isinstance(m, Mesh) else m
        vs.append(np.asarray(v, np.float64).reshape(-1, 3))
        ts.append(np.asarray(t, np.uint32).reshape(-1, 3))
    v_off = np.zeros(len(vs) + 1, np.int64)
    t_off = np.zeros(len(ts) + 1, np.int64)
    v_off[1:] = np.cumsum([len(v) for v in vs])
    t_off[1:] = np.cumsum([len(t) for t in ts])
    V = np.ascontiguousarray(np.concatenate(vs) if vs else np.zeros((0, 3)))
    T = np.ascontiguousarray(np.concatenate(ts) if ts else np.zeros((0, 3), np.uint32))
    L = _lib()
    h = _P()
    _chk(L.li_assemble(len(vs), _p(v_off), _p(V), _p(t_off), _p(T), ctypes.byref(h)), "assemble")
    try:
        nv, nf, nc, nm = _i64(), _i64(), _i64(), _i64()
        L.li_assembled_counts(h, ctypes.byref(nv), ctypes.byref(nf), ctypes.byref(nc),
                              ctypes.byref(nm))
        verts = np.zeros((nv.value, 3))
        faces = np.zeros((nf.value, 3), np.uint32)
        front = np.zeros(nf.value, np.int32)
        back = np.zeros(nf.value, np.int32)
        L.li_assembled_get(h, _p(verts), _p(faces), _p(front), _p(back))
    finally:
        L.li_assembled_destroy(h)
    return Assembly(verts, faces, front, back, int(nc.value), int(nm.value))


def litho_model(meshes, density, susceptibility) -> synth.LithoModel:
    """The forward's model from per-unit meshes and per-unit properties."""
    a = assemble(meshes)
    return synth.LithoModel(a.vertices, a.faces, a.front, a.back,
                            np.asarray(density, np.float64), np.asarray(susceptibility, np.float64))


class IniConfig:
    """A run configuration file (IniConfig, modules/litho-io/src/io/ini_config.cpp) held
    by the host library: typed getters with the reference's defaults and parsing rules."""

    def __init__(self, path=None, text=None):
        L = _lib()
        self._h = _P()
        if path is not None:
            _chk(L.li_ini_load(_path(path), ctypes.byref(self._h)), f"read_ini({path})")
        else:
            _chk(L.li_ini_parse(None if text is None else text.encode(), ctypes.byref(self._h)),
                 "parse_ini")

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib().li_ini_destroy(self._h)
            self._h = _P()

    def get(self, section, key, default=None):
        buf = ctypes.create_string_buffer(1 << 16)
        found = ctypes.c_int32()
        _chk(_lib().li_ini_get(self._h, section.encode(), key.encode(), buf, 1 << 16,
                               ctypes.byref(found)), "ini get")
        return buf.value.decode() if found.value else default

    def get_double(self, section, key, default):
        out = ctypes.c_double()
        _chk(_lib().li_ini_get_double(self._h, section.encode(), key.encode(), float(default),
                                      ctypes.byref(out)), "ini get_double")
        return out.value

    def get_int(self, section, key, default):
        out = ctypes.c_int32()
        _chk(_lib().li_ini_get_int(self._h, section.encode(), key.encode(), int(default),
                                   ctypes.byref(out)), "ini get_int")
        return out.value

    def get_bool(self, section, key, default):
        out = ctypes.c_int32()
        _chk(_lib().li_ini_get_bool(self._h, section.encode(), key.encode(), int(bool(default)),
                                    ctypes.byref(out)), "ini get_bool")
        return bool(out.value)

    def set(self, section, key, value):
        _chk(_lib().li_ini_set(self._h, section.encode(), key.encode(), str(value).encode()),
             "ini set")

    def save(self, path):
        _chk(_lib().li_ini_save(self._h, _path(path)), f"write_ini({path})")

    def entries(self):
        """{section: {key: value}} of every entry."""
        L = _lib()
        need = _i64()
        _chk(L.li_ini_dump(self._h, None, 0, ctypes.byref(need)), "ini dump")
        buf = ctypes.create_string_buffer(max(need.value, 1))
        _chk(L.li_ini_dump(self._h, buf, need.value, ctypes.byref(need)), "ini dump")
        out = {}
        for line in buf.value.decode().splitlines():
            sec, key, val = line.split("\t", 2)
            out.setdefault(sec, {})[key] = val
        return out
