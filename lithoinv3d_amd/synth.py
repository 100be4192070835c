"""Seeded synthetic lithology models and station layouts.

The model layout is the one the HIP path consumes: one shared vertex array, and one
face list in which a contact triangle between two units appears ONCE, oriented outward
for its ``front`` unit, with ``back`` naming the unit on the other side (-1 for the
outside).  ``to_group_meshes`` expands it back into the reference's layout -- one closed
mesh per lithology group, contact triangles present in both with opposite winding
(python/lithoseed/src/lithoseed/_surfaces.py:394-614, ``extract_group_volumes``).

Generators:
  * ``icosphere_model``      one closed icosphere (3 subdivisions -> 642 vertices /
                             1280 triangles), outward winding.
  * ``label_grid_model``     voxel-face boundary extraction of a 3D label grid, the same
                             algorithm as ``extract_group_volumes`` (shared quads emitted
                             to both units with opposite winding, quad split along the
                             v0-v2 diagonal, _surfaces.py:428-435), vectorised, except
                             that y-faces are wound outward (the reference winds them
                             inward, see label_grid_model).
  * ``gaussian_label_grid``  seeded white noise -> Gaussian filter -> equal-volume
                             quantile thresholds.
  * ``grid_stations`` / ``topo_stations``  regular station grids.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

__all__ = [
    "LithoModel",
    "icosphere",
    "icosphere_model",
    "box_model",
    "label_grid_model",
    "gaussian_label_grid",
    "grid_stations",
    "topo_stations",
    "field_unit_vector",
    "config0",
    "config_contact",
]


def field_unit_vector(inc_deg: float, dec_deg: float) -> np.ndarray:
    """Inducing-field unit vector, z up (modules/litho-forward/src/forward/magnetic_forward.cpp:61-66)."""
    i = math.radians(inc_deg)
    d = math.radians(dec_deg)
    return np.array([math.cos(i) * math.cos(d), math.cos(i) * math.sin(d), -math.sin(i)])


@dataclasses.dataclass
class LithoModel:
    """Shared-vertex lithology model with contact faces stored once.

    vertices    (n_v, 3) float64, metres, z up
    faces       (n_f, 3) uint32, outward winding for ``front``
    front/back  (n_f,) int32 unit index on the outward-normal's negative / positive side
                (back = -1: the face bounds the model)
    density     (n_u,) g/cc contrast per unit (signed)
    susceptibility (n_u,) SI per unit
    """

    vertices: np.ndarray
    faces: np.ndarray
    front: np.ndarray
    back: np.ndarray
    density: np.ndarray
    susceptibility: np.ndarray

    @property
    def n_units(self) -> int:
        return int(len(self.density))

    @property
    def n_faces(self) -> int:
        return int(len(self.faces))

    @property
    def n_vertices(self) -> int:
        return int(len(self.vertices))

    def group_triangles(self, g: int) -> np.ndarray:
        """Unit ``g``'s closed mesh as the reference holds it: its front faces as stored and
        its back faces with reversed winding (t0, t2, t1)."""
        f = self.faces[self.front == g]
        b = self.faces[self.back == g][:, [0, 2, 1]]
        return np.ascontiguousarray(np.concatenate([f, b], axis=0), dtype=np.uint32)

    def to_group_meshes(self):
        """(triangles concatenated over groups, group offsets) in the reference layout."""
        tris = [self.group_triangles(g) for g in range(self.n_units)]
        offs = np.zeros(self.n_units + 1, dtype=np.int64)
        offs[1:] = np.cumsum([len(t) for t in tris])
        all_t = np.concatenate(tris, axis=0) if tris else np.zeros((0, 3), np.uint32)
        return np.ascontiguousarray(all_t, dtype=np.uint32), offs

    def reference_face_count(self) -> int:
        """Triangles summed over the per-group meshes (contacts counted twice)."""
        return int((self.front >= 0).sum() + (self.back >= 0).sum())

    def magnetization(self, field_nT: float, fhat: np.ndarray):
        """Induced-only magnetization per unit: amplitude chi*F/100 along F-hat
        (magnetic_forward.cpp:105-106, 128-133, EffectiveSusceptibility mode)."""
        amp = self.susceptibility * field_nT / 100.0
        mdir = np.tile(np.asarray(fhat, dtype=np.float64), (self.n_units, 1))
        return amp.astype(np.float64), mdir


# ----------------------------------------------------------------------------- icosphere

def icosphere(subdivisions: int = 3):
    """Unit icosphere: icosahedron with vertices (0, +-1, +-phi) and cyclic permutations,
    each triangle split into 4 by edge midpoints ``subdivisions`` times, midpoints
    projected to the unit sphere.  Returns (vertices, outward triangles)."""
    phi = (1.0 + math.sqrt(5.0)) / 2.0
    v = [
        (-1, phi, 0), (1, phi, 0), (-1, -phi, 0), (1, -phi, 0),
        (0, -1, phi), (0, 1, phi), (0, -1, -phi), (0, 1, -phi),
        (phi, 0, -1), (phi, 0, 1), (-phi, 0, -1), (-phi, 0, 1),
    ]
    f = [
        (0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
        (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
        (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
        (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1),
    ]
    verts = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    faces = list(f)
    for _ in range(subdivisions):
        cache = {}

        def mid(i, j):
            key = (min(i, j), max(i, j))
            if key not in cache:
                m = verts[i] + verts[j]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for (a, b, c) in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(verts)
    F = np.array(faces, dtype=np.uint32)
    # make the winding outward
    cen = V[F].mean(axis=1)
    n = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    flip = (n * cen).sum(axis=1) < 0
    F[flip] = F[flip][:, [0, 2, 1]]
    return V, F


def icosphere_model(radius=200.0, center=(0.0, 0.0, -500.0), subdivisions=3,
                    density=0.3, susceptibility=0.02) -> LithoModel:
    V, F = icosphere(subdivisions)
    V = V * radius + np.asarray(center, dtype=np.float64)
    n = len(F)
    return LithoModel(V, F, np.zeros(n, np.int32), -np.ones(n, np.int32),
                      np.array([density], np.float64), np.array([susceptibility], np.float64))


def box_model(lo, hi, density=1.0, susceptibility=0.0) -> LithoModel:
    """Axis-aligned box as 12 outward triangles (a 1x1x1 label grid)."""
    lo = np.asarray(lo, np.float64)
    hi = np.asarray(hi, np.float64)
    labels = np.zeros((1, 1, 1), np.int32)
    m = label_grid_model(labels, origin=(lo[0], lo[1], hi[2]), cell=(hi - lo),
                         density=[density], susceptibility=[susceptibility])
    return m


# ----------------------------------------------------------------------- label grids

def gaussian_label_grid(shape, cell, sigma_m, n_units, seed):
    """Seeded Gaussian random field thresholded into ``n_units`` equal-volume units.

    White noise N(0,1) from numpy's PCG64 generator with ``seed``, smoothed by an
    isotropic Gaussian of standard deviation ``sigma_m`` metres (per-axis sigma in cells
    = sigma_m / cell), boundary mode 'reflect', cut at the (k/n_units) quantiles."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    w = rng.standard_normal(shape)
    sig = [sigma_m / c for c in cell]
    f = gaussian_filter(w, sigma=sig, mode="reflect")
    qs = np.quantile(f, [k / n_units for k in range(1, n_units)])
    return np.searchsorted(qs, f, side="right").astype(np.int32)


def label_grid_model(labels, origin, cell, density, susceptibility) -> LithoModel:
    """Voxel-face boundary extraction (python/lithoseed/src/lithoseed/_surfaces.py:394-614).

    ``labels[i, j, k]`` is the unit of voxel (i, j, k) (-1 = outside the model);
    voxel (i, j, k) spans x0+i*dx..x0+(i+1)*dx, y0+j*dy.., and z0-k*dz down to
    z0-(k+1)*dz (k counts downward, _surfaces.py:418-423).  Every face between two
    different labels becomes two triangles stored once, outward for the lower-index
    voxel's unit (``front`` = la, ``back`` = lb) exactly as ``_emit_quad(v0, v1, v2, v3,
    la, lb)`` orders them; a face against -1 is kept for the unit inside it.
    """
    labels = np.asarray(labels, dtype=np.int32)
    nx, ny, nz = labels.shape
    x0, y0, z0 = (float(t) for t in origin)
    dx, dy, dz = (float(t) for t in cell)
    pad = np.full((nx + 2, ny + 2, nz + 2), -1, np.int32)
    pad[1:-1, 1:-1, 1:-1] = labels

    def vid(i, j, k):
        return (i * (ny + 1) + j) * (nz + 1) + k

    quads = []  # (v0, v1, v2, v3) lattice ids CCW seen from `front`, front, back
    # x-faces at lattice i (between padded voxels i and i+1 -> real i-1 and i)
    la = pad[:-1, 1:-1, 1:-1]
    lb = pad[1:, 1:-1, 1:-1]
    ii, jj, kk = np.nonzero(la != lb)
    if len(ii):
        a, b = la[ii, jj, kk], lb[ii, jj, kk]
        v0, v1, v2, v3 = vid(ii, jj, kk), vid(ii, jj + 1, kk), vid(ii, jj + 1, kk + 1), vid(ii, jj, kk + 1)
        quads.append((v0, v1, v2, v3, a, b))
    la = pad[1:-1, :-1, 1:-1]
    lb = pad[1:-1, 1:, 1:-1]
    ii, jj, kk = np.nonzero(la != lb)
    if len(ii):
        a, b = la[ii, jj, kk], lb[ii, jj, kk]
        v0, v1, v2, v3 = vid(ii, jj, kk), vid(ii + 1, jj, kk), vid(ii + 1, jj, kk + 1), vid(ii, jj, kk + 1)
        # _surfaces.py:458-466 passes (v0, v1, v2, v3) here, which winds the y-faces
        # inward (k runs downward, a reflection of index space); (v0, v3, v2, v1) keeps
        # the same v0-v2 diagonal and makes them outward like the x- and z-faces
        quads.append((v0, v3, v2, v1, a, b))
    la = pad[1:-1, 1:-1, :-1]
    lb = pad[1:-1, 1:-1, 1:]
    ii, jj, kk = np.nonzero(la != lb)
    if len(ii):
        a, b = la[ii, jj, kk], lb[ii, jj, kk]
        v0, v1, v2, v3 = vid(ii, jj, kk), vid(ii + 1, jj, kk), vid(ii + 1, jj + 1, kk), vid(ii, jj + 1, kk)
        quads.append((v0, v1, v2, v3, a, b))

    v0 = np.concatenate([q[0] for q in quads])
    v1 = np.concatenate([q[1] for q in quads])
    v2 = np.concatenate([q[2] for q in quads])
    v3 = np.concatenate([q[3] for q in quads])
    fa = np.concatenate([q[4] for q in quads])
    fb = np.concatenate([q[5] for q in quads])
    # A quad is CCW seen from la toward lb; _emit_quad stores (v0,v2,v1), (v0,v3,v2)
    # for la.  When la is outside (-1) the face belongs to lb only, which the
    # reference emits with the reversed quad (v0,v3,v2,v1) -> (v0,v2,v3), (v0,v1,v2).
    own_a = fa >= 0
    t_front = np.where(own_a, fa, fb).astype(np.int32)
    t_back = np.where(own_a, fb, -1).astype(np.int32)
    tri1 = np.where(own_a[:, None], np.stack([v0, v2, v1], 1), np.stack([v0, v2, v3], 1))
    tri2 = np.where(own_a[:, None], np.stack([v0, v3, v2], 1), np.stack([v0, v1, v2], 1))
    tris = np.empty((2 * len(v0), 3), np.int64)
    tris[0::2] = tri1
    tris[1::2] = tri2
    front = np.repeat(t_front, 2)
    back = np.repeat(t_back, 2)

    used, inv = np.unique(tris.ravel(), return_inverse=True)
    faces = inv.reshape(-1, 3).astype(np.uint32)
    k = used % (nz + 1)
    j = (used // (nz + 1)) % (ny + 1)
    i = used // ((nz + 1) * (ny + 1))
    V = np.stack([x0 + i * dx, y0 + j * dy, z0 - k * dz], axis=1).astype(np.float64)
    return LithoModel(V, faces, front, back,
                      np.asarray(density, np.float64), np.asarray(susceptibility, np.float64))


# ------------------------------------------------------------------------- stations

def grid_stations(n, spacing, z, center=(0.0, 0.0)) -> np.ndarray:
    """n x n stations, ``spacing`` apart, centred on ``center``, at height ``z``."""
    c = (np.arange(n) - (n - 1) / 2.0) * spacing
    X, Y = np.meshgrid(c + center[0], c + center[1], indexing="ij")
    Z = np.full_like(X, z, dtype=np.float64)
    return np.ascontiguousarray(np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1))


def topo_stations(n, extent, amplitude=30.0, wavelength=1000.0, clearance=1.0,
                  center=(0.0, 0.0)) -> np.ndarray:
    """n x n cell-centred stations over a square footprint of side ``extent`` on the
    topography z = amplitude*sin(2 pi x/wavelength)*sin(2 pi y/wavelength), ``clearance``
    above it."""
    sp = extent / n
    c = -extent / 2.0 + sp * (np.arange(n) + 0.5)
    X, Y = np.meshgrid(c + center[0], c + center[1], indexing="ij")
    Z = amplitude * np.sin(2 * np.pi * X / wavelength) * np.sin(2 * np.pi * Y / wavelength) + clearance
    return np.ascontiguousarray(np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1))


# ----------------------------------------------------------------------- named configs

def config0():
    """One closed icosphere (R = 200 m, 1280 triangles) centred at 500 m depth,
    0.3 g/cc, 0.02 SI; field 50000 nT, I = -60 deg, D = 0; 32x32 stations at 50 m,
    z = +1 m."""
    m = icosphere_model(200.0, (0.0, 0.0, -500.0), 3, 0.3, 0.02)
    st = grid_stations(32, 50.0, 1.0)
    return dict(model=m, stations=st, inc=-60.0, dec=0.0, field=50000.0)


def config_contact(shape=(64, 64, 32), cell=25.0, n_units=4, seed=1, prop_seed=2,
                   n_st=128, model_top=-31.0, sigma_m=100.0):
    """Multi-lithology contact model: Gaussian-field label grid (``shape`` cells of
    ``cell`` m, 100 m smoothing, ``n_units`` equal-volume units, seed ``seed``), unit
    density contrasts U[-0.3, 0.5] g/cc and susceptibilities log-U[1e-4, 5e-2] SI
    (seed ``prop_seed``), ``n_st`` x ``n_st`` stations 1 m above sinusoidal topography
    (30 m amplitude, 1 km wavelength) over the footprint.  The model top sits at
    ``model_top`` (-31 m) so every station stays at least 1 m above it.
    (64,64,32)/128 is the 4-unit contact model; (192,192,64)/256 is the ~1M-triangle
    adjoint-at-scale model."""
    labels = gaussian_label_grid(shape, (cell, cell, cell), sigma_m, n_units, seed)
    rng = np.random.default_rng(prop_seed)
    dens = rng.uniform(-0.3, 0.5, n_units)
    sus = np.exp(rng.uniform(np.log(1e-4), np.log(5e-2), n_units))
    ext = (shape[0] * cell, shape[1] * cell)
    origin = (-ext[0] / 2.0, -ext[1] / 2.0, model_top)
    m = label_grid_model(labels, origin, (cell, cell, cell), dens, sus)
    st = topo_stations(n_st, ext[0])
    return dict(model=m, stations=st, inc=-60.0, dec=0.0, field=50000.0, labels=labels)
