"""The real Forrestania model of the reference (eight unit surfaces read from their GOCAD
TSurf files, tests/golden/forrestania) at its 992 gravity stations, in UTM coordinates
(x ~ 7.4e5 m, y ~ 6.4e6 m): the GPU forward and adjoint against the oracle evaluated on
the per-unit (reference) layout, contacts carried twice.

Stations lying exactly on a facet (one of the 992, inside a vertical contact facet) are
left out: the reference's TMI there depends on the sign of a rounded zero.

Unit properties: the cluster table's median density minus 2.67 g/cc and its median
susceptibility, unit k taking row k-1 (the files do not say which cluster each volume
holds; any values exercise the same arithmetic).  Tolerances: FP64 1e-4 mGal and
1e-3 nT absolute; FP32C 1e-3 mGal and 1e-2 nT in both TMI formulas; adjoint against the oracle's one-sided-FD J^T w 2e-4 (FP64) and 2e-3 (FP32C) of
the largest component.
"""
import os

import numpy as np
import pytest

from lithoinv3d_amd import litho_io as IO
from lithoinv3d_amd import synth
from lithoinv3d_amd.lithofwd import ForwardContext

pytestmark = pytest.mark.gpu

FOR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "forrestania")
FIELD = 50000.0


def say(msg):
    print(f"[case] {msg}", flush=True)


def on_facet(model, st, tol=1e-6):
    """Stations lying in the plane of a facet and inside it: the solid angle there is
    +-2 pi by the sign of a rounded zero in each of the reference's two copies of a contact
    facet (their vertex orders differ), so the reference's own TMI is arbitrary."""
    V, F = model.vertices, model.faces.astype(np.int64)
    a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    n = np.cross(b - a, c - a)
    n /= np.linalg.norm(n, axis=1)[:, None]
    hit = np.zeros(len(st), bool)
    for i, o in enumerate(st):
        h = np.einsum("ij,ij->i", a - o, n)
        k = np.nonzero(np.abs(h) < tol)[0]
        for f in k:
            # barycentric test in the facet plane
            ra, rb, rc = a[f] - o, b[f] - o, c[f] - o
            s1 = np.dot(np.cross(ra, rb), n[f])
            s2 = np.dot(np.cross(rb, rc), n[f])
            s3 = np.dot(np.cross(rc, ra), n[f])
            if (s1 >= 0 and s2 >= 0 and s3 >= 0) or (s1 <= 0 and s2 <= 0 and s3 <= 0):
                hit[i] = True
                break
    return hit


@pytest.fixture(scope="module")
def forrestania(oracle):
    meshes = [IO.read_surface(os.path.join(FOR, f"volume_group_{k}.ts")) for k in range(1, 9)]
    props = np.genfromtxt(os.path.join(FOR, "cluster_properties.csv"), delimiter=",",
                          names=True, dtype=None, encoding=None)
    dens = np.array([r["density_median_gcc"] for r in props])[:8] - 2.67
    sus = np.array([r["susceptibility_median_SI"] for r in props])[:8]
    st = IO.read_stations(os.path.join(FOR, "Forrestania_Gravity_Station_trim_.csv")).xyz
    fh = synth.field_unit_vector(-65.0, 1.0)
    V = np.concatenate([m.vertices for m in meshes])
    off = np.cumsum([0] + [len(m.vertices) for m in meshes])
    tris = [m.triangles.astype(np.int64) + off[k] for k, m in enumerate(meshes)]
    ref = oracle.OracleModel(V, tris, dens, sus * FIELD / 100.0, np.tile(fh, (8, 1)))
    model = IO.litho_model(meshes, dens, sus)
    skip = on_facet(model, st)
    print(f"{skip.sum()} of {len(st)} stations lie on a facet: {np.nonzero(skip)[0]}")
    st = st[~skip]
    return dict(meshes=meshes, model=model, st=st, fh=fh, ref=ref, dens=dens, sus=sus,
                gz=ref.gravity(st), tmi=ref.magnetic(fh, st),
                tmi_tb=ref.magnetic(fh, st, textbook=True))


def test_forward_vs_oracle(forrestania):
    d = forrestania
    print(f"{d['model'].n_faces} facets ({sum(len(m.triangles) for m in d['meshes'])} in the "
          f"per-unit files), {len(d['st'])} stations; gz {d['gz'].min():.4f} .. "
          f"{d['gz'].max():.4f} mGal, TMI {d['tmi'].min():.1f} .. {d['tmi'].max():.1f} nT")
    for prec, tg, tt in (("fp64", 1e-4, 1e-3), ("fp32c", 1e-3, 1e-2)):
        say(f"forrestania forward {prec}")
        with ForwardContext(d["model"], d["st"], d["fh"], FIELD, precision=prec) as ctx:
            gz, tm = ctx.forward()
        eg, et = np.abs(gz - d["gz"]).max(), np.abs(tm - d["tmi"]).max()
        print(f"{prec}: max |dgz| {eg:.2e} mGal, max |dTMI| {et:.2e} nT")
        assert eg < tg and et < tt
        say(f"forrestania textbook TMI {prec}")
        with ForwardContext(d["model"], d["st"], d["fh"], FIELD, precision=prec,
                            tmi_mode="textbook", want=("tmi",)) as ctx:
            _, tb = ctx.forward()
        e = np.abs(tb - d["tmi_tb"]).max()
        print(f"{prec}: textbook TMI {d['tmi_tb'].min():.2f} .. {d['tmi_tb'].max():.2f} nT, "
              f"max |dTMI| {e:.2e} nT")
        assert e < (1e-3 if prec == "fp64" else 1e-2)


def test_adjoint_vs_oracle_fd(forrestania):
    d = forrestania
    st = d["st"][::31]
    rng = np.random.default_rng(17)
    wg = rng.normal(size=len(st))
    wt = rng.normal(size=len(st)) * 1e-3
    say("forrestania oracle J^T w")
    # the oracle differentiates the per-unit vertex arrays; sum each merged vertex's copies
    gu = d["ref"].gravity_vertex_jtw(st, wg) + d["ref"].magnetic_vertex_jtw(d["fh"], st, wt)
    m = d["model"]
    Vall = np.concatenate([mm.vertices for mm in d["meshes"]])
    q = {tuple(np.round(v * 1e6).astype(np.int64)): i for i, v in enumerate(m.vertices)}
    idx = np.array([q[tuple(np.round(v * 1e6).astype(np.int64))] for v in Vall])
    gref = np.zeros_like(m.vertices)
    np.add.at(gref, idx, gu)
    for prec, tol in (("fp64", 2e-4), ("fp32c", 2e-3)):
        say(f"forrestania adjoint {prec}")
        with ForwardContext(m, st, d["fh"], FIELD, precision=prec) as ctx:
            g = ctx.adjoint(wg, wt)
        e = np.abs(g - gref).max() / np.abs(gref).max()
        print(f"adjoint {prec}: rel err vs oracle one-sided-FD J^T w {e:.2e}")
        assert e < tol
