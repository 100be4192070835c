"""Learned diagonal initial inverse Hessian for the projected L-BFGS.

The optimizer starts from H0 = gamma * diag(exp(-L/2)), L an estimate of the log of the
Gauss-Newton diagonal of the joint misfit per control-point parameter,

    d_j = sum_i (dgz_i/dp_j / sigma_gz)^2 + alpha * sum_i (dTMI_i/dp_j / sigma_tmi)^2

(one over the column norm; full Jacobi scaling, exp(-L), is slower on these problems).
L = B + r: B is the same sum over a sparse subset of the stations, one per cell of a
grid `stride` station spacings wide (1/1024 of a regular grid at stride 32), scaled by
n_stations / n_subset and evaluated by the GPU column-norm kernel (lf_colnorm2) at a
small fraction of the full cost; r is a correction predicted by a 16 -> 64 -> 64 -> 1
GELU MLP (trained and evaluated by the lf_mlp_* kernels of liblithofwd.so: BF16 MFMA,
FP32 accumulation) from local geometry and B itself.  The subset undersamples the
response of shallow parameters, whose footprint is narrower than the subset's spacing;
the network learns that, halving the error of B alone.
The objective stays exact, so the plain solve (H0 = gamma * I) remains both the fallback
and the yardstick of the gain.  Training targets are the exact column norms over all
stations (lf_colnorm2).
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np

from . import lithofwd

N_FEATURES = 16
FEATURE_NAMES = (
    "depth_km", "log_area", "log_chi_sum_area", "log_chi_sum_area_fpar", "n_axis",
    "log_stations_within_2_depths", "abs_drho", "log_abs_dchi", "log_subset_colnorm_gz",
    "log_subset_colnorm_tmi", "log_subset_spacing_over_depth", "log_area_over_depth2",
    "log_chi_sum_area_naxis", "abs_drho_naxis", "abs_dchi_naxis_scaled",
    "edge_distance_over_depth",
)
SUBSET_STRIDE = 32

_lib_ready = False


def _lib():
    global _lib_ready
    L = lithofwd.load_library()
    if not _lib_ready:
        P = ctypes.c_void_p
        i64 = ctypes.c_int64
        L.lf_mlp_param_count.restype = i64
        L.lf_mlp_param_count.argtypes = []
        L.lf_mlp_create.argtypes = [ctypes.POINTER(P), ctypes.c_int, ctypes.c_uint64]
        L.lf_mlp_create.restype = ctypes.c_int
        L.lf_mlp_destroy.argtypes = [P]
        L.lf_mlp_destroy.restype = None
        L.lf_mlp_last_error.argtypes = [P]
        L.lf_mlp_last_error.restype = ctypes.c_char_p
        for name, args in (("lf_mlp_set_params", [P, P]), ("lf_mlp_get_params", [P, P]),
                           ("lf_mlp_forward", [P, i64, P, P]),
                           ("lf_mlp_loss_grad", [P, i64, P, P, P, P]),
                           ("lf_mlp_train", [P, i64, P, P, ctypes.c_int, ctypes.c_float, P])):
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        _lib_ready = True
    return L


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


class MLP:
    """The GPU MLP (parameters resident on the device, FP32 master copy)."""

    def __init__(self, seed=0, device=0):
        self.L = _lib()
        self._h = ctypes.c_void_p()
        rc = self.L.lf_mlp_create(ctypes.byref(self._h), device, seed)
        if rc:
            raise lithofwd.LithoFwdError(self.L.lf_mlp_last_error(None).decode())
        self.n_params = int(self.L.lf_mlp_param_count())

    def _chk(self, rc, what):
        if rc:
            raise lithofwd.LithoFwdError(f"{what}: {self.L.lf_mlp_last_error(self._h).decode()}")

    def close(self):
        if self._h:
            self.L.lf_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self):
        p = np.zeros(self.n_params, np.float32)
        self._chk(self.L.lf_mlp_get_params(self._h, _p(p)), "get_params")
        return p

    @params.setter
    def params(self, p):
        p = _f32(p)
        assert p.size == self.n_params
        self._chk(self.L.lf_mlp_set_params(self._h, _p(p)), "set_params")

    def forward(self, x):
        x = _f32(x).reshape(-1, N_FEATURES)
        y = np.zeros(len(x), np.float32)
        self._chk(self.L.lf_mlp_forward(self._h, len(x), _p(x), _p(y)), "forward")
        return y

    def loss_grad(self, x, t):
        x = _f32(x).reshape(-1, N_FEATURES)
        t = _f32(t)
        loss = np.zeros(1, np.float32)
        g = np.zeros(self.n_params, np.float32)
        self._chk(self.L.lf_mlp_loss_grad(self._h, len(x), _p(x), _p(t), _p(loss), _p(g)),
                  "loss_grad")
        return float(loss[0]), g

    def train(self, x, t, steps, lr=1e-3):
        x = _f32(x).reshape(-1, N_FEATURES)
        t = _f32(t)
        losses = np.zeros(steps, np.float32)
        self._chk(self.L.lf_mlp_train(self._h, len(x), _p(x), _p(t), steps, lr, _p(losses)), "train")
        return losses


# --------------------------------------------------------------------- features

def vertex_areas_normals(m, field_dir=(0.0, 0.0, 1.0)):
    """Per-vertex sums over the incident facets of one third of each facet's area A_f, of
    A_f n_f, A_f |drho_f| and A_f |dchi_f| (contrasts across the facet), and of
    A_f (chi_front + chi_back) weighted by 1, by 1 - (F.n_f)^2 and by |n_f,x|, |n_f,y|,
    |n_f,z|.  The sum
    of susceptibilities follows the reference's TMI edge coefficient, which adds the
    magnetizations on both sides of a contact instead of differencing them."""
    V, F = m.vertices, m.faces.astype(np.int64)
    cr = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    area = 0.5 * np.linalg.norm(cr, axis=1)
    nrm = cr / np.maximum(2 * area, 1e-30)[:, None]
    has_back = m.back >= 0
    bk = np.maximum(m.back, 0)
    rb = np.where(has_back, m.density[bk], 0.0)
    cb = np.where(has_back, m.susceptibility[bk], 0.0)
    drho = np.abs(m.density[m.front] - rb)
    dchi = np.abs(m.susceptibility[m.front] - cb)
    csum = m.susceptibility[m.front] + cb
    fpar = 1.0 - (nrm @ np.asarray(field_dir, np.float64)) ** 2
    a3 = area / 3
    cols = np.stack([a3, nrm[:, 0] * a3, nrm[:, 1] * a3, nrm[:, 2] * a3, drho * a3,
                     dchi * a3, csum * a3, csum * fpar * a3, csum * np.abs(nrm[:, 0]) * a3,
                     csum * np.abs(nrm[:, 1]) * a3, csum * np.abs(nrm[:, 2]) * a3], 1)
    from scipy.sparse import csr_matrix

    nv, nf = m.n_vertices, len(F)
    # vertex x facet incidence (three corners per facet) times the facet columns
    inc = csr_matrix((np.ones(3 * nf), (F.T.ravel(), np.tile(np.arange(nf), 3))), shape=(nv, nf))
    acc = np.asarray(inc @ cols)
    return dict(A=acc[:, 0], N=acc[:, 1:4], R=acc[:, 4], C=acc[:, 5], S=acc[:, 6],
                S_par=acc[:, 7], S_n=acc[:, 8:11])


def station_subset(stations, stride=SUBSET_STRIDE):
    """Indices of a sparse station subset: the station nearest the centre of each occupied
    cell of a horizontal grid `stride` mean station spacings wide, and that cell width.
    The spacing s solves (W + s)(H + s) = n s^2 for the bounding box W x H (s is the
    pitch of a regular grid); the cells start half a spacing before the box, so a regular
    grid of k*stride stations per side gives k cells per side."""
    xy = np.asarray(stations, np.float64)[:, :2]
    n = len(xy)
    if n < 2:
        return np.arange(n), 0.0
    lo, hi = xy.min(0), xy.max(0)
    W, H = np.maximum(hi - lo, 0.0)
    sp = ((W + H) + np.sqrt((W + H) ** 2 + 4.0 * (n - 1) * W * H)) / (2.0 * (n - 1))
    if sp <= 0.0:
        return np.arange(n), 0.0
    cell = stride * sp
    org = lo - 0.5 * sp
    ij = np.floor((xy - org) / cell).astype(np.int64)
    centre = org + (ij + 0.5) * cell
    d2 = ((xy - centre) ** 2).sum(1)
    key = ij[:, 0] * (1 << 31) + ij[:, 1]
    order = np.lexsort((d2, key))
    first = np.ones(len(order), bool)
    first[1:] = key[order][1:] != key[order][:-1]
    return np.sort(order[first]), float(cell)


@dataclasses.dataclass
class SubsetDiagonal:
    gz: np.ndarray    # sum over the subset of (dgz/dp_j / sigma_gz)^2, scaled to all stations
    tmi: np.ndarray   # alpha * the same for TMI
    spacing: float    # subset spacing (m)

    def log_diag(self):
        return np.log(np.maximum(self.gz + self.tmi, 1e-300))


def subset_gn_diagonal(m, stations, dof_map, field_dir, field_nT=50000.0, sigma_gz=0.05,
                       sigma_tmi=1.0, alpha=1.0, axis=2, stride=SUBSET_STRIDE,
                       precision="fp32c", device=0):
    """The Gauss-Newton diagonal over station_subset(stations, stride), each part scaled
    by n_stations / n_subset (GPU column-norm kernel)."""
    idx, cell = station_subset(stations, stride)
    with lithofwd.ForwardContext(m, np.asarray(stations)[idx], field_dir, field_nT,
                                 precision=precision, device=device) as ctx:
        cg, ct = ctx.colnorm2(axis, *dof_map)
    scale = len(stations) / len(idx)
    return SubsetDiagonal(cg * scale / sigma_gz ** 2, alpha * ct * scale / sigma_tmi ** 2, cell)


def dof_features(m, dof_map, dof_vertex, stations, axis_scale, sub: SubsetDiagonal,
                 field_dir=(0.0, 0.0, 1.0), axis=2, va=None):
    """(n_dof, 16) features per control-point parameter (FEATURE_NAMES) moving its
    vertices along `axis`, from the parameter's vertices and their interpolation weights
    (dof_map), the stations, the inducing-field direction and the subset Gauss-Newton
    diagonal.  The normal-component features take the component along the parameter's
    axis (n_z for the depth parameters the network is trained on).  va: the model's
    vertex_areas_normals, when already computed."""
    from scipy.spatial import cKDTree

    off, vert, w = dof_map
    if va is None:
        va = vertex_areas_normals(m, field_dir)
    n = len(off) - 1
    ww = w / axis_scale
    seg = np.repeat(np.arange(n), np.diff(off))

    def dofsum(vals):
        if vals.ndim == 1:
            return np.bincount(seg, weights=ww * vals[vert], minlength=n)
        return np.stack([np.bincount(seg, weights=ww * vals[vert, j], minlength=n)
                         for j in range(vals.shape[1])], 1)

    Aeff = dofsum(va["A"])
    Neff = dofsum(va["N"])
    nn = Neff / np.maximum(np.linalg.norm(Neff, axis=1, keepdims=True), 1e-30)
    Reff = dofsum(va["R"]) / np.maximum(Aeff, 1e-30)
    Ceff = dofsum(va["C"]) / np.maximum(Aeff, 1e-30)
    P = m.vertices[dof_vertex]
    ztop = float(np.mean(stations[:, 2]))
    depth = np.maximum(ztop - P[:, 2], 1.0)
    # stations within a horizontal radius of two depths
    sxy = stations[:, :2]
    cnt = cKDTree(sxy).query_ball_point(P[:, :2], r=2 * depth, return_length=True)
    cnt = np.asarray(cnt, dtype=np.float64)
    xmin, ymin = sxy.min(0)
    xmax, ymax = sxy.max(0)
    edge = np.minimum.reduce([P[:, 0] - xmin, xmax - P[:, 0], P[:, 1] - ymin, ymax - P[:, 1]])
    tiny = 1e-12
    X = np.stack([
        depth / 1000.0,
        np.log1p(Aeff),
        np.log(np.abs(dofsum(va["S"])) + tiny),
        np.log(np.abs(dofsum(va["S_par"])) + tiny),
        nn[:, axis],
        np.log1p(cnt),
        Reff,
        np.log10(np.maximum(Ceff, 1e-6)),
        np.log(np.maximum(sub.gz, 1e-300)),
        np.log(np.maximum(sub.tmi, 1e-300)),
        np.full(n, np.log(sub.spacing)) - np.log(depth),
        np.log(np.maximum(Aeff, 1e-6) / depth ** 2),
        np.log(np.abs(dofsum(va["S_n"][:, axis])) + tiny),
        Reff * np.abs(nn[:, axis]),
        Ceff * np.abs(nn[:, axis]) * 100.0,
        edge / depth,
    ], axis=1)
    return X


@dataclasses.dataclass
class Normalizer:
    mean: np.ndarray
    std: np.ndarray
    t_mean: float
    t_std: float

    @classmethod
    def fit(cls, X, t):
        s = X.std(0)
        s[s < 1e-12] = 1.0
        return cls(X.mean(0), s, float(t.mean()), float(t.std() or 1.0))

    def x(self, X):
        return (X - self.mean) / self.std

    def t(self, t):
        return (t - self.t_mean) / self.t_std

    def t_inv(self, y):
        return y * self.t_std + self.t_mean


def gn_diagonal(ctx, dof_map, axis=2, sigma_gz=0.05, sigma_tmi=1.0, alpha=1.0):
    """Exact Gauss-Newton diagonal of the joint misfit for the parameters of dof_map."""
    off, vert, w = dof_map
    cg, ct = ctx.colnorm2(axis, off, vert, w)
    return cg / sigma_gz ** 2 + alpha * ct / sigma_tmi ** 2
