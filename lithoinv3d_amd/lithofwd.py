"""ctypes binding of liblithofwd.so (include/lithofwd.h).

The library is the product: this module only marshals arrays.  It raises
``LithoFwdUnavailable`` when the shared library is missing or cannot be loaded -- there
is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblithofwd.so")

LF_FP64, LF_FP32C = 0, 1
LF_TMI_REFERENCE, LF_TMI_TEXTBOOK = 0, 1
LF_WANT_GZ, LF_WANT_TMI = 1, 2
N_TIMING_CLASSES = 7
TIMING_CLASSES = ("face_setup", "forward", "forward_reduce", "adjoint_pairs",
                  "adjoint_finalize", "vertex_gather", "jvp")

EXPORTED = (
    "lf_create", "lf_destroy", "lf_last_error", "lf_set_vertices", "lf_set_vertices_device",
    "lf_set_properties", "lf_forward", "lf_forward_device", "lf_adjoint", "lf_adjoint_device",
    "lf_jvp", "lf_jvp_device", "lf_sync", "lf_set_timing", "lf_timing", "lf_stats", "lf_colnorm2", "lf_forward_noalloc",
)
# the learned-preconditioner MLP (declared in mlp.py)
EXPORTED_MLP = (
    "lf_mlp_param_count", "lf_mlp_create", "lf_mlp_destroy", "lf_mlp_last_error",
    "lf_mlp_set_params", "lf_mlp_get_params", "lf_mlp_forward", "lf_mlp_loss_grad",
    "lf_mlp_train",
)


class LithoFwdUnavailable(RuntimeError):
    pass


class LithoFwdError(RuntimeError):
    pass


_D = ctypes.POINTER(ctypes.c_double)


class _Desc(ctypes.Structure):
    _fields_ = [
        ("n_vertices", ctypes.c_int64),
        ("n_faces", ctypes.c_int64),
        ("n_stations", ctypes.c_int64),
        ("n_units", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("tmi_mode", ctypes.c_int32),
        ("want", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("vertices", _D),
        ("faces", ctypes.POINTER(ctypes.c_uint32)),
        ("front", ctypes.POINTER(ctypes.c_int32)),
        ("back", ctypes.POINTER(ctypes.c_int32)),
        ("stations", _D),
        ("density", _D),
        ("magnetization", _D),
        ("field_dir", ctypes.c_double * 3),
    ]


_lib = None


def load_library(path: str = LIB_PATH):
    """Loads liblithofwd.so and declares its signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise LithoFwdUnavailable(
            f"{path} is missing: build it with `python -m lithoinv3d_amd.build` "
            "(there is no CPU fallback)")
    # PyTorch ships its own HIP runtime (same soname, libamdhip64.so.7).  Loading it
    # first makes this library bind to that one instance; loading ours first would let
    # torch bring a second runtime into the process, which then finds no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    try:
        L = ctypes.CDLL(path)
    except OSError as e:  # pragma: no cover - depends on the runtime
        raise LithoFwdUnavailable(f"cannot load {path}: {e}") from e
    P = ctypes.c_void_p
    i64 = ctypes.c_int64
    L.lf_create.argtypes = [ctypes.POINTER(P), ctypes.POINTER(_Desc)]
    L.lf_destroy.argtypes = [P]
    L.lf_destroy.restype = None
    L.lf_last_error.argtypes = [P]
    L.lf_last_error.restype = ctypes.c_char_p
    L.lf_set_vertices.argtypes = [P, P]
    L.lf_set_vertices_device.argtypes = [P, P]
    L.lf_set_properties.argtypes = [P, P, P]
    L.lf_forward.argtypes = [P, P, P]
    L.lf_forward_device.argtypes = [P, P, P]
    L.lf_adjoint.argtypes = [P, P, P, P]
    L.lf_adjoint_device.argtypes = [P, P, P, P]
    L.lf_jvp.argtypes = [P, P, P, P]
    L.lf_jvp_device.argtypes = [P, P, P, P]
    L.lf_sync.argtypes = [P]
    L.lf_set_timing.argtypes = [P, ctypes.c_int]
    L.lf_timing.argtypes = [P, _D, ctypes.POINTER(i64), ctypes.c_int]
    L.lf_stats.argtypes = [P, ctypes.POINTER(i64)]
    L.lf_colnorm2.argtypes = [P, ctypes.c_int, i64, P, P, P, P, P]
    L.lf_forward_noalloc.argtypes = [ctypes.POINTER(_Desc), P, P, P]
    for name in EXPORTED:
        if name not in ("lf_destroy", "lf_last_error"):
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


class ForwardContext:
    """One lithology model + station set resident on a GPU.

    model: synth.LithoModel (shared vertices, contact faces stored once);
    stations (n_s, 3); field_dir: inducing-field unit vector; field_nT: its strength.
    ``magnetization`` overrides the induced-only (n_u, 3) magnetization vectors
    (reference units: chi*F/100 along the direction)."""

    def __init__(self, model, stations, field_dir=(0.0, 0.0, 1.0), field_nT=50000.0,
                 precision="fp64", tmi_mode="reference", want=("gz", "tmi"), device=0,
                 magnetization=None):
        self._lib = load_library()
        self._ctx = ctypes.c_void_p()
        self.n_v = model.n_vertices
        self.n_f = model.n_faces
        self.n_u = model.n_units
        st = _c(stations, np.float64).reshape(-1, 3)
        self.n_s = len(st)
        fd = np.asarray(field_dir, dtype=np.float64)
        if magnetization is None:
            amp = model.susceptibility * field_nT / 100.0
            magnetization = amp[:, None] * fd[None, :]
        self._keep = dict(
            v=_c(model.vertices, np.float64), f=_c(model.faces, np.uint32),
            fr=_c(model.front, np.int32), bk=_c(model.back, np.int32), st=st,
            den=_c(model.density, np.float64), mag=_c(magnetization, np.float64).reshape(-1, 3))
        k = self._keep
        d = _Desc()
        d.n_vertices, d.n_faces, d.n_stations, d.n_units = self.n_v, self.n_f, self.n_s, self.n_u
        d.precision = {"fp64": LF_FP64, "fp32c": LF_FP32C}[precision]
        d.tmi_mode = {"reference": LF_TMI_REFERENCE, "textbook": LF_TMI_TEXTBOOK}[tmi_mode]
        d.want = (LF_WANT_GZ if "gz" in want else 0) | (LF_WANT_TMI if "tmi" in want else 0)
        d.device = device
        d.vertices = k["v"].ctypes.data_as(_D)
        d.faces = k["f"].ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        d.front = k["fr"].ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        d.back = k["bk"].ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        d.stations = k["st"].ctypes.data_as(_D)
        d.density = k["den"].ctypes.data_as(_D)
        d.magnetization = k["mag"].ctypes.data_as(_D)
        d.field_dir[:] = list(fd)
        rc = self._lib.lf_create(ctypes.byref(self._ctx), ctypes.byref(d))
        if rc != 0:
            raise LithoFwdError(f"lf_create failed ({rc}): {self._lib.lf_last_error(None).decode()}")
        self.precision = precision
        self.tmi_mode = tmi_mode
        self.want = want

    # -------------------------------------------------------------- lifecycle
    def close(self):
        if getattr(self, "_ctx", None) and self._ctx.value:
            self._lib.lf_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, what):
        if rc != 0:
            raise LithoFwdError(f"{what} failed ({rc}): {self._lib.lf_last_error(self._ctx).decode()}")

    # ---------------------------------------------------------------- updates
    def set_vertices(self, xyz):
        v = _c(xyz, np.float64).reshape(-1, 3)
        if len(v) != self.n_v:
            raise ValueError(f"expected {self.n_v} vertices, got {len(v)}")
        self._check(self._lib.lf_set_vertices(self._ctx, _ptr(v)), "lf_set_vertices")

    def set_vertices_device(self, ptr: int):
        self._check(self._lib.lf_set_vertices_device(self._ctx, ctypes.c_void_p(ptr)),
                    "lf_set_vertices_device")

    def set_properties(self, density=None, magnetization=None):
        den = None if density is None else _c(density, np.float64)
        mag = None if magnetization is None else _c(magnetization, np.float64).reshape(-1, 3)
        if den is not None and len(den) != self.n_u:
            raise ValueError("density length != n_units")
        if mag is not None and len(mag) != self.n_u:
            raise ValueError("magnetization rows != n_units")
        self._check(self._lib.lf_set_properties(self._ctx, _ptr(den), _ptr(mag)),
                    "lf_set_properties")

    # ---------------------------------------------------------------- compute
    def forward(self):
        """(gz mGal, tmi nT) as numpy arrays (None for an output not wanted)."""
        gz = np.zeros(self.n_s) if "gz" in self.want else None
        tm = np.zeros(self.n_s) if "tmi" in self.want else None
        self._check(self._lib.lf_forward(self._ctx, _ptr(gz), _ptr(tm)), "lf_forward")
        return gz, tm

    def forward_device(self, gz_ptr: int = 0, tmi_ptr: int = 0):
        self._check(self._lib.lf_forward_device(self._ctx, ctypes.c_void_p(gz_ptr or None),
                                                ctypes.c_void_p(tmi_ptr or None)),
                    "lf_forward_device")

    def adjoint(self, w_gz=None, w_tmi=None):
        """sum_i w_gz[i] dgz_i/dx + w_tmi[i] dTMI_i/dx for every vertex, (n_v, 3)."""
        wg = None if w_gz is None else _c(w_gz, np.float64)
        wt = None if w_tmi is None else _c(w_tmi, np.float64)
        for w in (wg, wt):
            if w is not None and len(w) != self.n_s:
                raise ValueError("weight length != n_stations")
        g = np.zeros((self.n_v, 3))
        self._check(self._lib.lf_adjoint(self._ctx, _ptr(wg), _ptr(wt), _ptr(g)), "lf_adjoint")
        return g

    def adjoint_device(self, wg_ptr: int, wt_ptr: int, grad_ptr: int):
        self._check(self._lib.lf_adjoint_device(self._ctx, ctypes.c_void_p(wg_ptr or None),
                                                ctypes.c_void_p(wt_ptr or None),
                                                ctypes.c_void_p(grad_ptr)),
                    "lf_adjoint_device")

    def jvp(self, vtan):
        """(dgz, dTMI) along the vertex tangent vtan (n_v, 3): J v in one pass."""
        v = _c(vtan, np.float64).reshape(-1)
        if len(v) != 3 * self.n_v:
            raise ValueError("tangent length != 3 * n_vertices")
        gz = np.zeros(self.n_s)
        tmi = np.zeros(self.n_s)
        self._check(self._lib.lf_jvp(self._ctx, _ptr(v), _ptr(gz), _ptr(tmi)), "lf_jvp")
        return gz, tmi

    def jvp_device(self, vtan_ptr: int, gz_ptr: int = 0, tmi_ptr: int = 0):
        self._check(self._lib.lf_jvp_device(self._ctx, ctypes.c_void_p(vtan_ptr),
                                            ctypes.c_void_p(gz_ptr or None),
                                            ctypes.c_void_p(tmi_ptr or None)), "lf_jvp_device")

    def colnorm2(self, axis, dof_off, dof_vertex, dof_weight):
        """(sum_i (dgz_i/dp_j)^2, sum_i (dTMI_i/dp_j)^2) for parameters p_j that move
        coordinate `axis` of vertices dof_vertex[dof_off[j]:dof_off[j+1]] by the weights."""
        off = _c(dof_off, np.int64)
        vert = _c(dof_vertex, np.int32)
        w = _c(dof_weight, np.float64)
        n = len(off) - 1
        if n < 0 or len(vert) != off[-1] or len(w) != off[-1]:
            raise ValueError("inconsistent DOF map")
        og = np.zeros(n)
        ot = np.zeros(n)
        self._check(self._lib.lf_colnorm2(self._ctx, int(axis), n, _ptr(off), _ptr(vert), _ptr(w),
                                          _ptr(og), _ptr(ot)), "lf_colnorm2")
        return og, ot

    def sync(self):
        self._check(self._lib.lf_sync(self._ctx), "lf_sync")

    def set_timing(self, on: bool):
        self._check(self._lib.lf_set_timing(self._ctx, int(bool(on))), "lf_set_timing")

    def timing(self):
        """{class: (ms_total, launches)} since the previous call (resets)."""
        ms = np.zeros(N_TIMING_CLASSES)
        n = np.zeros(N_TIMING_CLASSES, dtype=np.int64)
        self._check(self._lib.lf_timing(self._ctx, ms.ctypes.data_as(_D),
                                        n.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                        N_TIMING_CLASSES), "lf_timing")
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(TIMING_CLASSES)}

    def fp64_edges(self) -> int:
        v = ctypes.c_int64()
        self._check(self._lib.lf_stats(self._ctx, ctypes.byref(v)), "lf_stats")
        return int(v.value)


def forward_noalloc(n_vertices, n_faces, n_stations, n_units, vertices, faces, front, back,
                    stations, density, magnetization, field_dir, gz=0, tmi=0,
                    precision="fp64", tmi_mode="reference", want=("gz", "tmi"), device=0,
                    stream=0):
    """lf_forward_noalloc: the forward of a model held in device memory, with no device
    allocation and no context.  Every array argument is a DEVICE pointer (int) on
    `device`, which must be the current device: vertices (n_vertices*3 FP64), faces
    (n_faces*3 uint32), front/back (n_faces int32), stations (n_stations*3 FP64), density
    (n_units FP64), magnetization (n_units*3 FP64); gz / tmi are device outputs of
    n_stations FP64 (0: not wanted).  Enqueued on `stream` (a hipStream_t as int, 0: the
    null stream); the caller synchronizes."""
    L = load_library()
    d = _Desc()
    d.n_vertices, d.n_faces, d.n_stations, d.n_units = n_vertices, n_faces, n_stations, n_units
    d.precision = {"fp64": LF_FP64, "fp32c": LF_FP32C}[precision]
    d.tmi_mode = {"reference": LF_TMI_REFERENCE, "textbook": LF_TMI_TEXTBOOK}[tmi_mode]
    d.want = (LF_WANT_GZ if "gz" in want else 0) | (LF_WANT_TMI if "tmi" in want else 0)
    d.device = device
    d.vertices = ctypes.cast(ctypes.c_void_p(vertices), _D)
    d.faces = ctypes.cast(ctypes.c_void_p(faces), ctypes.POINTER(ctypes.c_uint32))
    d.front = ctypes.cast(ctypes.c_void_p(front), ctypes.POINTER(ctypes.c_int32))
    d.back = ctypes.cast(ctypes.c_void_p(back), ctypes.POINTER(ctypes.c_int32))
    d.stations = ctypes.cast(ctypes.c_void_p(stations), _D)
    d.density = ctypes.cast(ctypes.c_void_p(density), _D)
    d.magnetization = ctypes.cast(ctypes.c_void_p(magnetization), _D)
    d.field_dir[:] = [float(x) for x in field_dir]
    rc = L.lf_forward_noalloc(ctypes.byref(d), ctypes.c_void_p(gz or None),
                              ctypes.c_void_p(tmi or None), ctypes.c_void_p(stream or None))
    if rc != 0:
        raise LithoFwdError(f"lf_forward_noalloc failed ({rc}): {L.lf_last_error(None).decode()}")

