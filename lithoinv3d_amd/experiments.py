"""Inversion experiments on the seeded contact models.

  * ``contact_problem``   a four-unit contact model, its control-point parameters
                          (Z only, stride 4, axis scale 10 m, exterior vertices fixed,
                          bounds +-200 m) and synthetic data: the forward of a smoothly
                          perturbed "true" interface geometry plus Gaussian noise
                          (0.05 mGal, 1 nT).
  * ``training_set``      features and targets over many seeded models: the exact log
                          Gauss-Newton diagonal (GPU column-norm kernel over all stations)
                          minus its estimate from a 1/1024 station subset.
  * ``compare_solvers``   plain FP64, plain FP32C and learned-H0 FP32C L-BFGS-B runs on one
                          held-out model (and H0 from the subset estimate alone and from
                          the exact diagonal); every FP32C run stops at the plain FP64
                          run's final misfit, and reports iterations and seconds to it.

  * ``large_inversion``   a few FP32C iterations of the joint XYZ_FREE inversion at the
                          large workload with 50 boreholes' depth-to-contact picks and the
                          learned H0, resident on one GPU.

  * ``compare_gncg``      the runner's two geometry solvers, L-BFGS-B and Gauss-Newton CG,
                          from the same start on one model: objective against seconds and
                          gradient evaluations.
  * ``jvp_timing``        one Jacobian-vector product (FP32C, FP64) on a named workload
                          against two FP64 forwards; ``calibration_timing``: the
                          magnetic-weight calibration with each J v.

CLI:  python -m lithoinv3d_amd.experiments learned-h0 --train 200 --test 20 --out FILE
      python -m lithoinv3d_amd.experiments solvers --seed 2000 --out FILE
      python -m lithoinv3d_amd.experiments jvp --out FILE
      python -m lithoinv3d_amd.experiments large --max-iter 3
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import inversion, mlp, synth, workload
from .lithofwd import ForwardContext

AXIS_SCALE = 10.0
SIGMA_GZ, SIGMA_TMI = 0.05, 1.0
# H0 = gamma * diag(GN)^-H0_POWER.  Full Jacobi scaling (power 1) slows the projected
# L-BFGS-B on these problems (97-114 iterations to the plain FP64 40-iteration misfit
# against 42-43 unscaled); the square root reaches it in 24-31.
H0_POWER = 0.5


def log(msg):
    print(f"[exp] {msg}", file=sys.stderr, flush=True)


def true_displacement(xy, extent, seed, amp=40.0, n_bumps=6, width=300.0):
    """Smooth vertical displacement field (m): Gaussian bumps of +-amp, tapered to zero
    within one bump width of the model sides."""
    rng = np.random.default_rng(seed)
    half = extent / 2.0
    dz = np.zeros(len(xy))
    for _ in range(n_bumps):
        c = rng.uniform(-half + width, half - width, 2)
        a = rng.uniform(-amp, amp)
        dz += a * np.exp(-((xy - c) ** 2).sum(1) / (2 * width ** 2))
    taper = np.clip((half - np.abs(xy)).min(1) / width, 0.0, 1.0)
    return dz * taper


def contact_problem(seed, shape=(64, 64, 32), n_st=128, stride=4, noise_seed=None,
                    perturb=True):
    cfg = synth.config_contact(shape=shape, n_st=n_st, seed=seed, prop_seed=seed + 100000)
    m, st = cfg["model"], cfg["stations"]
    M = inversion.Model(m)
    M.fix_exterior(inversion.Z_ONLY)
    M.set_axis_scale(AXIS_SCALE)
    n = M.set_cp_stride(stride, 25.0)
    v, _, _ = M.dofs()
    p0 = M.params()
    extent = shape[0] * 25.0
    ptrue = p0 + (true_displacement(m.vertices[v, :2], extent, seed + 7) / AXIS_SCALE
                  if perturb else 0.0)
    lo, hi = M.bounds(p0, 200.0)
    ptrue = np.clip(ptrue, lo, hi)
    return dict(model=m, M=M, stations=st, p0=p0, ptrue=ptrue, lo=lo, hi=hi, n_dof=n,
                dof_vertex=v, cfg=cfg, seed=seed,
                noise_seed=seed + 3 if noise_seed is None else noise_seed)


def synth_data(prob, precision="fp64"):
    st = prob["stations"]
    P = inversion.Problem(prob["M"], st, np.zeros(len(st)), tmi_obs=np.zeros(len(st)),
                          precision=precision)
    gz, tm = P.forward(prob["ptrue"])
    P.close()
    rng = np.random.default_rng(prob["noise_seed"])
    gz = gz + rng.normal(0, SIGMA_GZ, len(gz))
    tm = tm + rng.normal(0, SIGMA_TMI, len(tm))
    prob["gz_obs"], prob["tmi_obs"] = gz, tm
    return gz, tm


def make_problem(prob, precision, smooth_weight=0.0):
    st = prob["stations"]
    n = len(st)
    return inversion.Problem(prob["M"], st, prob["gz_obs"], np.full(n, SIGMA_GZ),
                             prob["tmi_obs"], np.full(n, SIGMA_TMI), alpha_mag=1.0,
                             precision=precision, smooth_weight=smooth_weight)


def features_targets(prob, with_targets=True):
    """Features (n_dof, 16), the exact log Gauss-Newton diagonal (or None) and its
    station-subset estimate B, at the geometry the solver starts from."""
    m, M = prob["model"], prob["M"]
    M.apply(prob["p0"])
    dmap = M.dof_map()
    fh = synth.field_unit_vector(-60.0, 0.0)
    sub = mlp.subset_gn_diagonal(m, prob["stations"], dmap, fh, 50000.0, SIGMA_GZ, SIGMA_TMI)
    X = mlp.dof_features(m, dmap, prob["dof_vertex"], prob["stations"], AXIS_SCALE, sub, fh)
    t = None
    if with_targets:
        with ForwardContext(m, prob["stations"], fh, 50000.0, precision="fp32c") as ctx:
            d = mlp.gn_diagonal(ctx, dmap, 2, SIGMA_GZ, SIGMA_TMI, 1.0)
        t = np.log(np.maximum(d, 1e-30))
    return X, t, sub.log_diag()


def training_set(seeds, shape=(64, 64, 32), n_st=128):
    """Features and the network's targets: the exact log diagonal minus its subset
    estimate."""
    Xs, ts = [], []
    for k, s in enumerate(seeds):
        t0 = time.time()
        prob = contact_problem(s, shape, n_st, perturb=False)
        X, t, base = features_targets(prob)
        Xs.append(X)
        ts.append(t - base)
        log(f"training model {k + 1}/{len(seeds)} seed {s}: {len(t)} parameters "
            f"({time.time() - t0:.1f}s)")
    return np.concatenate(Xs), np.concatenate(ts)


def train_mlp(X, t, steps=3000, lr=2e-3, seed=0):
    norm = mlp.Normalizer.fit(X, t)
    net = mlp.MLP(seed=seed)
    losses = net.train(norm.x(X), norm.t(t), steps=steps, lr=lr)
    return net, norm, losses


def predict_log_diag(prob, net, norm):
    """Learned log Gauss-Newton diagonal: subset estimate + the network's correction;
    also returns the subset estimate alone."""
    X, _, base = features_targets(prob, with_targets=False)
    return base + norm.t_inv(net.forward(norm.x(X))).astype(np.float64), base


def run_solver(prob, precision, max_iter, target=None, h0=None):
    P = make_problem(prob, precision)
    t0 = time.perf_counter()
    r = P.minimize(prob["p0"], prob["lo"], prob["hi"], max_iter=max_iter, target=target, h0=h0)
    wall = time.perf_counter() - t0
    P.close()
    return r, wall


def time_to(r, target):
    """(iterations, seconds) at the first iterate with f <= target, or None."""
    ok = np.nonzero(r.history[:, 1] <= target)[0]
    if len(ok) == 0:
        return None
    k = ok[0]
    return int(r.history[k, 0]), float(r.history[k, 4])


def compare_solvers(prob, net=None, norm=None, max_iter=40, exact_diag=False,
                    h0_power=H0_POWER, extra_powers=()):
    synth_data(prob)
    out = {"seed": prob["seed"], "n_dof": prob["n_dof"]}
    r64, w64 = run_solver(prob, "fp64", max_iter)
    f_star = r64.f
    out["fp64"] = dict(iterations=r64.iterations, seconds=r64.seconds, f_final=f_star,
                       f0=float(r64.history[0, 1]), message=r64.message)
    r32, _ = run_solver(prob, "fp32c", 3 * max_iter, target=f_star)
    hit = time_to(r32, f_star)
    out["fp32c"] = dict(iterations=r32.iterations, seconds=r32.seconds, f_final=r32.f,
                        reached=hit is not None, to_target=hit)
    if net is not None:
        t0 = time.perf_counter()
        pred, base = predict_log_diag(prob, net, norm)
        h0 = np.exp(-h0_power * pred)
        t_h0 = time.perf_counter() - t0
        rl, _ = run_solver(prob, "fp32c", 3 * max_iter, target=f_star, h0=h0)
        hit = time_to(rl, f_star)
        out["learned_fp32c"] = dict(iterations=rl.iterations, seconds=rl.seconds + t_h0,
                                    h0_seconds=t_h0, f_final=rl.f, reached=hit is not None,
                                    to_target=None if hit is None else (hit[0], hit[1] + t_h0))
        rs, _ = run_solver(prob, "fp32c", 3 * max_iter, target=f_star,
                           h0=np.exp(-h0_power * base))
        hit = time_to(rs, f_star)
        out["subset_only_fp32c"] = dict(iterations=rs.iterations, reached=hit is not None,
                                        to_target=None if hit is None else
                                        (hit[0], hit[1] + t_h0))
        for pw in extra_powers:
            rp, _ = run_solver(prob, "fp32c", 3 * max_iter, target=f_star,
                               h0=np.exp(-pw * pred))
            hit = time_to(rp, f_star)
            out[f"learned_fp32c_power{pw:g}"] = dict(
                iterations=rp.iterations, reached=hit is not None,
                to_target=None if hit is None else (hit[0], hit[1] + t_h0))
        rl64, _ = run_solver(prob, "fp64", 3 * max_iter, target=f_star, h0=h0)
        hit = time_to(rl64, f_star)
        out["learned_fp64"] = dict(iterations=rl64.iterations, seconds=rl64.seconds + t_h0,
                                   f_final=rl64.f, reached=hit is not None,
                                   to_target=None if hit is None else (hit[0], hit[1] + t_h0))
    if exact_diag:
        t0 = time.perf_counter()
        M = prob["M"]
        M.apply(prob["p0"])
        with ForwardContext(prob["model"], prob["stations"], synth.field_unit_vector(-60.0, 0.0),
                            50000.0, precision="fp32c") as ctx:
            t = np.log(np.maximum(mlp.gn_diagonal(ctx, M.dof_map(), 2, SIGMA_GZ, SIGMA_TMI,
                                                  1.0), 1e-30))
        t_diag = time.perf_counter() - t0
        re, _ = run_solver(prob, "fp32c", 3 * max_iter, target=f_star,
                           h0=np.exp(-h0_power * t))
        hit = time_to(re, f_star)
        out["exact_diag_fp32c"] = dict(iterations=re.iterations, seconds=re.seconds + t_diag,
                                       diag_seconds=t_diag, reached=hit is not None,
                                       to_target=None if hit is None else
                                       (hit[0], hit[1] + t_diag))
    return out


def compare_gncg(prob, lbfgs_iter=100, gncg_iter=8, cg_max_iter=50, fd_step=1.0,
                 precision="fp32c", smooth_weight=0.0):
    """The runner's two geometry solvers from the same start on one problem: L-BFGS-B
    (history 10) and Gauss-Newton CG with the runner's settings (50 CG iterations,
    tolerance 1e-6, fd_step 1 parameter unit), each Hessian-vector product two gradients.
    Reports each GN-CG iterate's objective, gradient count and seconds beside the first
    L-BFGS-B iterate at or below that objective.  smooth_weight: the Laplacian smoothness
    term of the objective (the runner's lambda, 1 by default there)."""
    synth_data(prob)
    out = {"seed": prob["seed"], "n_dof": prob["n_dof"], "precision": precision,
           "cg_max_iter": cg_max_iter, "fd_step": fd_step, "smooth_weight": smooth_weight}
    runs = {}
    for name in ("lbfgsb", "gncg"):
        P = make_problem(prob, precision, smooth_weight)
        P.gradient(prob["p0"])  # context set-up outside the timed solve
        log(f"{name}: solving")
        if name == "lbfgsb":
            r = P.minimize(prob["p0"], prob["lo"], prob["hi"], max_iter=lbfgs_iter)
        else:
            r = P.gncg_minimize(prob["p0"], prob["lo"], prob["hi"], max_iter=gncg_iter,
                                cg_max_iter=cg_max_iter, fd_step=fd_step)
        P.close()
        runs[name] = r
        out[name] = dict(iterations=r.iterations, message=r.message, f0=float(r.history[0, 1]),
                         f_final=r.f, seconds=r.seconds,
                         gradients=int(r.history[-1, 6]),
                         history=r.history[:, [0, 1, 3, 4, 5, 6]].tolist())
        if name == "gncg":
            out[name]["cg_iterations"] = [int(c) for c in r.cg_iterations]
        log(f"{name}: {r.iterations} iterations, f {r.history[0, 1]:.4e} -> {r.f:.4e} in "
            f"{r.seconds:.2f} s ({r.message})")
    rows = []
    lb = runs["lbfgsb"]
    for it, f, sec, _, ng in runs["gncg"].history[:, [0, 1, 4, 5, 6]]:
        hit = time_to(lb, f)
        k = None if hit is None else int(np.nonzero(lb.history[:, 1] <= f)[0][0])
        rows.append(dict(gncg_iteration=int(it), f=float(f), gncg_seconds=float(sec),
                         gncg_gradients=int(ng),
                         lbfgsb_iteration=None if hit is None else hit[0],
                         lbfgsb_seconds=None if hit is None else hit[1],
                         lbfgsb_gradients=None if k is None else int(lb.history[k, 6])))
    out["at_equal_objective"] = rows
    return out


def jvp_timing(name="contact192x192x64_st256", reps=2, seed=1):
    """One Jacobian-vector product along a random vertex tangent on a named workload, FP32C
    and FP64, against the two FP64 forwards a central difference needs: seconds per call,
    the per-kernel milliseconds (HIP events) and the FP32C product's deviation from the
    FP64 one."""
    cfg = workload.build(name)
    m, st = cfg["model"], cfg["stations"]
    fh = synth.field_unit_vector(cfg["inc"], cfg["dec"])
    vt = np.random.default_rng(seed).normal(size=(m.n_vertices, 3))
    out = {"workload": name, "n_vertices": m.n_vertices, "n_facets": m.n_faces,
           "n_stations": len(st), "pairs": m.n_faces * len(st)}
    res = {}
    for prec in ("fp32c", "fp64"):
        with ForwardContext(m, st, fh, cfg["field"], precision=prec) as ctx:
            ctx.forward()
            ctx.sync()
            ctx.timing()
            ctx.set_timing(True)
            n = reps if prec == "fp32c" else 1
            t0 = time.perf_counter()
            for _ in range(n):
                res[prec] = ctx.jvp(vt)
            t1 = time.perf_counter()
            for _ in range(2):
                ctx.forward()
            ctx.sync()
            t2 = time.perf_counter()
            ctx.set_timing(False)
            tim = {k: v for k, v in ctx.timing().items() if v[1]}
        out[prec] = dict(jvp_seconds=(t1 - t0) / n, two_forwards_seconds=t2 - t1,
                         kernel_ms=tim)
        log(f"{prec}: J v {(t1 - t0) / n:.3f} s, two forwards {t2 - t1:.3f} s, kernels {tim}")
    for i, k in enumerate(("gz", "tmi")):
        a, b = res["fp32c"][i], res["fp64"][i]
        out[f"fp32c_vs_fp64_{k}"] = float(np.abs(a - b).max() / np.abs(b).max())
    return out


def calibration_timing(prob):
    """Problem.calibrate at the start of one contact problem: the exact Jacobian-vector
    product in FP32C and FP64 against the central differences of FP64 forwards."""
    synth_data(prob)
    out = {"seed": prob["seed"], "n_dof": prob["n_dof"]}
    for prec, eps in (("fp32c", 0.0), ("fp64", 0.0), ("fp64", 1e-3)):
        P = make_problem(prob, prec)
        P.gradient(prob["p0"])
        t0 = time.perf_counter()
        r = P.calibrate(prob["p0"], seed=1, eps=eps)
        sec = time.perf_counter() - t0
        P.close()
        key = f"{prec}_{'exact' if eps <= 0 else 'central_differences'}"
        out[key] = dict(seconds=sec, **{k: (float(v) if not isinstance(v, bool) else v)
                                        for k, v in r.items()})
        log(f"calibration {key}: {sec:.2f} s, {r}")
    return out


def borehole_picks(m, M, ptrue, n=50, seed=5, sigma=5.0, step=5.0, margin=100.0):
    """Depth-to-contact picks of n seeded vertical boreholes at the true geometry: along
    each vertical line the unit is sampled every `step` metres (point classification);
    wherever it changes between two units, the elevation of their contact on that line
    (the depth-to-contact term's own intersection at ptrue) becomes a pick with the given
    standard deviation.  Returns litho_io.ContactPicks."""
    from .litho_io import ContactPicks

    rng = np.random.default_rng(seed)
    V = m.vertices
    lo, hi = V.min(0), V.max(0)
    xy = rng.uniform(lo[:2] + margin, hi[:2] - margin, (n, 2))
    z = np.arange(hi[2] - step / 2, lo[2], -step)
    M.apply(ptrue)
    pts = np.concatenate([np.column_stack([np.full(len(z), x), np.full(len(z), y), z])
                          for x, y in xy])
    lab = M.classify(pts).reshape(n, len(z))
    cand, above, below = [], [], []
    for b in range(n):
        for k in range(len(z) - 1):
            u, w = lab[b, k], lab[b, k + 1]
            if u >= 0 and w >= 0 and u != w:
                cand.append((xy[b, 0], xy[b, 1], 0.5 * (z[k] + z[k + 1])))
                above.append(u)
                below.append(w)
    picks = ContactPicks(np.array(cand, np.float64).reshape(-1, 3), np.array(above, np.int32),
                         np.array(below, np.int32), np.full(len(cand), float(sigma)))
    # the contact's elevation on the line at ptrue, from the term itself
    _, _, zc = M.boreholes(picks, ptrue, weight=1.0, radius=2 * step)
    ok = np.isfinite(zc)
    picks = ContactPicks(np.column_stack([picks.xyz[ok, :2], zc[ok]]), picks.above[ok],
                         picks.below[ok], picks.sigma[ok])
    return picks


def learned_log_diag_xyz(m, M, stations, net, norm, field_dir, field_nT):
    """Learned log Gauss-Newton diagonal for every parameter of M (any axes): per axis,
    the subset estimate for the parameters moving that axis plus the network's
    correction from features along that axis."""
    v, ax, _ = M.dofs()
    off, vert, w = M.dof_map()
    out = np.zeros(len(v))
    va = mlp.vertex_areas_normals(m, field_dir)
    for a in range(3):
        sel = np.nonzero(ax == a)[0]
        if len(sel) == 0:
            continue
        cnt = off[sel + 1] - off[sel]
        so = np.zeros(len(sel) + 1, np.int64)
        so[1:] = np.cumsum(cnt)
        idx = np.concatenate([np.arange(off[j], off[j + 1]) for j in sel])
        dm = (so, vert[idx], w[idx])
        sub = mlp.subset_gn_diagonal(m, stations, dm, field_dir, field_nT, SIGMA_GZ,
                                     SIGMA_TMI, axis=a)
        X = mlp.dof_features(m, dm, v[sel], stations, AXIS_SCALE, sub, field_dir, axis=a, va=va)
        out[sel] = sub.log_diag() + norm.t_inv(net.forward(norm.x(X))).astype(np.float64)
    return out


def large_inversion(name="contact256x256x96_st512", stride=8, max_iter=3, ref_weight=1e-2,
                    seed=5, n_boreholes=50, h0="learned", train_models=40, target=None):
    """Joint gz + TMI inversion at the large workload, everything resident on one GPU:
    XYZ_FREE control points (every ``stride``-th cell), FP32C forward and adjoint, the
    reference-model term, n_boreholes seeded vertical boreholes with depth-to-contact
    picks (5 m), synthetic data from a perturbed interface.  h0 = "learned": the
    diagonal H0 from the network trained on `train_models` small contact models (the
    training time is reported); "scaled": the identity.  The first step is scaled so the
    largest parameter moves one unit (an unscaled step would cost a long series of
    halvings, each a full forward); the diagonal's shape is kept.  Reports setup,
    per-iteration times, misfit and device memory in use.  target: stop once the
    objective is at or below it (seconds to another run's final misfit)."""
    import torch

    out = {"workload": name, "h0": h0}
    net = norm = None
    if h0 == "learned":
        t0 = time.perf_counter()
        X, t = training_set(range(1000, 1000 + train_models), (32, 32, 16), 64)
        net, norm, losses = train_mlp(X, t, steps=2000)
        out["h0_training_s"] = time.perf_counter() - t0
        out["h0_training_samples"] = int(len(t))
        log(f"large: H0 network trained on {len(t)} parameters in {out['h0_training_s']:.1f}s")
    t0 = time.perf_counter()
    cfg = workload.build(name)
    m, st = cfg["model"], cfg["stations"]
    out["n_vertices"], out["n_facets"], out["n_stations"] = m.n_vertices, m.n_faces, len(st)
    out["pairs_per_forward"] = float(m.n_faces) * len(st)
    cell = float(workload.WORKLOADS[name].get("cell", 25.0))
    M = inversion.Model(m)
    M.fix_exterior(inversion.XYZ_FREE)
    M.set_axis_scale(AXIS_SCALE)
    n = M.set_cp_stride(stride, cell)
    v, ax, _ = M.dofs()
    p0 = M.params()
    extent = workload.WORKLOADS[name]["shape"][0] * cell
    ptrue = p0.copy()
    zsel = ax == 2
    ptrue[zsel] += true_displacement(m.vertices[v[zsel], :2], extent, seed, amp=30.0,
                                     width=600.0) / AXIS_SCALE
    lo, hi = M.bounds(p0, 200.0)
    ptrue = np.clip(ptrue, lo, hi)
    out["n_dof"] = int(n)
    out["setup_s"] = time.perf_counter() - t0
    log(f"large: {m.n_vertices} vertices, {m.n_faces} facets, {len(st)} stations, {n} "
        f"parameters; setup {out['setup_s']:.1f}s")
    picks = None
    if n_boreholes > 0:
        t0 = time.perf_counter()
        picks = borehole_picks(m, M, ptrue, n_boreholes, seed)
        M.apply(p0)
        out["boreholes"] = n_boreholes
        out["borehole_picks"] = int(len(picks.xyz))
        out["borehole_setup_s"] = time.perf_counter() - t0
        log(f"large: {len(picks.xyz)} contact picks from {n_boreholes} boreholes "
            f"({out['borehole_setup_s']:.1f}s)")
    inc, dec, fld = cfg["inc"], cfg["dec"], cfg["field"]
    P = inversion.Problem(M, st, np.zeros(len(st)), tmi_obs=np.zeros(len(st)), inc=inc,
                          dec=dec, field=fld, precision="fp32c")
    t1 = time.perf_counter()
    gz, tm = P.forward(ptrue)
    out["forward_s"] = time.perf_counter() - t1
    P.close()
    M.apply(p0)
    rng = np.random.default_rng(seed + 1)
    gz = gz + rng.normal(0, SIGMA_GZ, len(gz))
    tm = tm + rng.normal(0, SIGMA_TMI, len(tm))
    log(f"large: synthetic data in {out['forward_s']:.1f}s")
    P = inversion.Problem(M, st, gz, np.full(len(st), SIGMA_GZ), tm, np.full(len(st), SIGMA_TMI),
                          inc=inc, dec=dec, field=fld, alpha_mag=1.0, precision="fp32c",
                          ref_weight=ref_weight, boreholes=picks, borehole_weight=1.0)
    t2 = time.perf_counter()
    g0 = P.gradient(p0)
    out["first_gradient_s"] = time.perf_counter() - t2
    free, total = torch.cuda.mem_get_info(0)
    out["device_memory_in_use_gb"] = (total - free) / 1e9
    D = np.ones(n)
    if net is not None:
        t3 = time.perf_counter()
        fh = synth.field_unit_vector(inc, dec)
        D = np.exp(-H0_POWER * learned_log_diag_xyz(m, M, st, net, norm, fh, fld))
        out["h0_seconds"] = time.perf_counter() - t3
        net.close()
        log(f"large: learned H0 in {out['h0_seconds']:.1f}s")
    h0 = D / max(float(np.abs(D * g0).max()), 1e-300)
    log(f"large: first gradient {out['first_gradient_s']:.1f}s; "
        f"{out['device_memory_in_use_gb']:.1f} GB in use")
    P.set_iteration_callback(lambda kind, row: log(
        f"large: iteration {int(row[0])}: f {row[1]:.5e}, |pg| {row[2]:.3e}, {row[4]:.1f}s, "
        f"{int(row[5])} objective / {int(row[6])} gradient calls"))
    t4 = time.perf_counter()
    r = P.minimize(p0, lo, hi, max_iter=max_iter, h0=h0, target=target)
    out["solve_s"] = time.perf_counter() - t4
    out["target"] = target
    comp = P.evaluate(r.x)
    P.close()
    h = r.history
    out["history"] = [dict(iteration=int(row[0]), f=float(row[1]), seconds=float(row[4]),
                           n_obj=int(row[5]), n_grad=int(row[6])) for row in h]
    out["final_components"] = dict(gravity=float(comp[0]), magnetic=float(comp[1]),
                                   regularization=float(comp[2]), total=float(comp[3]))
    out["message"] = r.message
    return out


def summarize_h0(results, meta, extra_powers=()):
    """Medians over the held-out models of compare_solvers' results."""
    def med(key, field):
        vals = [r[key][field] for r in results if key in r]
        return float(np.median(vals)) if vals else math.nan

    def med_to(key, idx):
        vals = [r[key]["to_target"][idx] for r in results if r.get(key, {}).get("to_target")]
        return float(np.median(vals)) if vals else math.nan

    summary = {
        **meta,
        "median_fp64_seconds": med("fp64", "seconds"),
        "median_fp64_iterations": med("fp64", "iterations"),
        "median_fp32c_seconds_to_target": med_to("fp32c", 1),
        "median_fp32c_iterations_to_target": med_to("fp32c", 0),
        "median_learned_fp32c_seconds_to_target": med_to("learned_fp32c", 1),
        "median_learned_fp32c_iterations_to_target": med_to("learned_fp32c", 0),
        "median_learned_fp64_iterations_to_target": med_to("learned_fp64", 0),
        "median_exact_diag_fp32c_seconds_to_target": med_to("exact_diag_fp32c", 1),
        "median_exact_diag_fp32c_iterations_to_target": med_to("exact_diag_fp32c", 0),
        "median_exact_diag_seconds": med("exact_diag_fp32c", "diag_seconds"),
        "median_h0_seconds": med("learned_fp32c", "h0_seconds"),
        "learned_fp32c_reached": sum(1 for r in results if r["learned_fp32c"]["reached"]),
        "fp32c_reached": sum(1 for r in results if r["fp32c"]["reached"]),
        "median_time_ratio_learned_fp32c_vs_fp64": float(np.median([
            r["learned_fp32c"]["to_target"][1] / r["fp64"]["seconds"]
            for r in results if r["learned_fp32c"]["to_target"]])) if results else math.nan,
        "median_time_ratio_learned_vs_plain_fp32c": float(np.median([
            r["learned_fp32c"]["to_target"][1] / r["fp32c"]["to_target"][1]
            for r in results if r["learned_fp32c"]["to_target"] and r["fp32c"]["to_target"]]))
        if results else math.nan,
        **{f"median_learned_fp32c_power{pw:g}_iterations_to_target":
           med_to(f"learned_fp32c_power{pw:g}", 0) for pw in extra_powers},
        "median_subset_only_fp32c_iterations_to_target": med_to("subset_only_fp32c", 0),
        "median_log_diag_rmse": float(np.median([r["log_diag_rmse"] for r in results])),
        "median_log_diag_rmse_subset_only": float(np.median(
            [r["log_diag_rmse_subset_only"] for r in results])),
        "median_log_diag_std": float(np.median([r["log_diag_std"] for r in results])),
    }
    return summary


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("learned-h0")
    a.add_argument("--train", type=int, default=200)
    a.add_argument("--test", type=int, default=20)
    a.add_argument("--shape", type=int, nargs=3, default=(64, 64, 32))
    a.add_argument("--n-st", type=int, default=128)
    a.add_argument("--steps", type=int, default=3000)
    a.add_argument("--max-iter", type=int, default=300,
                   help="iterations of the plain FP64 run whose final misfit is the target "
                        "(300: within about 0.5 %% of its value after 1000)")
    a.add_argument("--h0-power", type=float, default=H0_POWER)
    a.add_argument("--extra-powers", type=float, nargs="*", default=())
    a.add_argument("--out", default=None)
    a.add_argument("--test-first", type=int, default=2000, help="first held-out seed")
    a.add_argument("--save-net", default=None, help="write the trained network (.npz)")
    a.add_argument("--load-net", default=None, help="use this network instead of training")
    a.add_argument("--jsonl", default=None, help="append each held-out model's result")
    b = sub.add_parser("invert")
    b.add_argument("--seed", type=int, default=1)
    b.add_argument("--shape", type=int, nargs=3, default=(64, 64, 32))
    b.add_argument("--n-st", type=int, default=128)
    b.add_argument("--max-iter", type=int, default=40)
    b.add_argument("--out", default=None)
    d = sub.add_parser("summarize-h0", help="medians over the results of learned-h0 runs")
    d.add_argument("--jsonl", required=True)
    d.add_argument("--net", required=True, help="the network file of the runs (.npz)")
    d.add_argument("--out", default=None)
    jv = sub.add_parser("jvp", help="Jacobian-vector product timing and calibration cost")
    jv.add_argument("--workload", default="contact192x192x64_st256")
    jv.add_argument("--calibration-seed", type=int, default=2000,
                    help="the 64x64x32 contact problem calibrated (negative: skip)")
    jv.add_argument("--out", default=None)
    sv = sub.add_parser("solvers", help="L-BFGS-B against Gauss-Newton CG on one model")
    sv.add_argument("--seed", type=int, default=2000)
    sv.add_argument("--shape", type=int, nargs=3, default=(64, 64, 32))
    sv.add_argument("--n-st", type=int, default=128)
    sv.add_argument("--lbfgs-iter", type=int, default=100)
    sv.add_argument("--gncg-iter", type=int, default=8)
    sv.add_argument("--cg-max-iter", type=int, default=50)
    sv.add_argument("--fd-step", type=float, default=1.0)
    sv.add_argument("--precision", choices=("fp64", "fp32c"), default="fp32c")
    sv.add_argument("--smooth", type=float, default=0.0, help="Laplacian smoothness weight")
    sv.add_argument("--out", default=None)
    c = sub.add_parser("large")
    c.add_argument("--workload", default="contact256x256x96_st512")
    c.add_argument("--stride", type=int, default=8)
    c.add_argument("--max-iter", type=int, default=3)
    c.add_argument("--boreholes", type=int, default=50)
    c.add_argument("--h0", choices=("learned", "scaled"), default="learned")
    c.add_argument("--train-models", type=int, default=40)
    c.add_argument("--target", type=float, default=None,
                   help="stop at this objective value (another run's final misfit)")
    c.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    if args.cmd == "large":
        res = large_inversion(args.workload, args.stride, args.max_iter,
                              n_boreholes=args.boreholes, h0=args.h0,
                              train_models=args.train_models, target=args.target)
        if args.out:
            json.dump(res, open(args.out, "w"), indent=1)
        print(json.dumps(res))
        return
    if args.cmd == "jvp":
        res = {"jvp": jvp_timing(args.workload)}
        if args.calibration_seed >= 0:
            res["calibration"] = calibration_timing(contact_problem(args.calibration_seed))
        if args.out:
            json.dump(res, open(args.out, "w"), indent=1)
        print(json.dumps(res))
        return
    if args.cmd == "solvers":
        prob = contact_problem(args.seed, tuple(args.shape), args.n_st)
        res = compare_gncg(prob, args.lbfgs_iter, args.gncg_iter, args.cg_max_iter,
                           args.fd_step, args.precision, args.smooth)
        if args.out:
            json.dump(res, open(args.out, "w"), indent=1)
        print(json.dumps({k: v for k, v in res.items() if k != "lbfgsb"}))
        return
    if args.cmd == "summarize-h0":
        results = [json.loads(line) for line in open(args.jsonl) if line.strip()]
        z = np.load(args.net)
        summary = summarize_h0(results, dict(
            train_samples=int(z["n_samples"]), test_models=len(results),
            train_loss_first_last=[float(z["losses"][0]), float(z["losses"][-1])],
            train_seconds=float(z["train_seconds"]), data_seconds=float(z["data_seconds"])))
        out = {"summary": summary, "models": results}
        if args.out:
            json.dump(out, open(args.out, "w"), indent=1)
        print(json.dumps(summary, indent=1))
        return
    shape = tuple(args.shape)
    if args.cmd == "invert":
        prob = contact_problem(args.seed, shape, args.n_st)
        res = compare_solvers(prob, max_iter=args.max_iter)
        log(json.dumps(res))
        if args.out:
            json.dump(res, open(args.out, "w"), indent=1)
        print(json.dumps(res))
        return
    if args.load_net:
        z = np.load(args.load_net)
        net = mlp.MLP(seed=0)
        net.params = z["params"]
        norm = mlp.Normalizer(z["mean"], z["std"], float(z["t_mean"]), float(z["t_std"]))
        losses = z["losses"]
        t = np.zeros(int(z["n_samples"]))
        t_data, t_train = float(z["data_seconds"]), float(z["train_seconds"])
        log(f"network from {args.load_net}")
    else:
        t0 = time.time()
        X, t = training_set(range(1000, 1000 + args.train), shape, args.n_st)
        t_data = time.time() - t0
        t1 = time.time()
        net, norm, losses = train_mlp(X, t, steps=args.steps)
        t_train = time.time() - t1
        log(f"trained on {len(t)} parameters from {args.train} models: loss {losses[0]:.4f} "
            f"-> {losses[-1]:.4f} ({t_train:.1f}s; data {t_data:.1f}s)")
        if args.save_net:
            np.savez(args.save_net, params=net.params, mean=norm.mean, std=norm.std,
                     t_mean=norm.t_mean, t_std=norm.t_std, losses=np.asarray(losses),
                     n_samples=len(t), data_seconds=t_data, train_seconds=t_train)
    results = []
    for k, s in enumerate(range(args.test_first, args.test_first + args.test)):
        prob = contact_problem(s, shape, args.n_st)
        # prediction quality on the held-out model (a constant offset does not change
        # the solver's path, gamma absorbs it: it is removed before the RMS)
        _, th, base = features_targets(prob)
        pred, _ = predict_log_diag(prob, net, norm)
        r = compare_solvers(prob, net, norm, max_iter=args.max_iter, exact_diag=True,
                            h0_power=args.h0_power, extra_powers=args.extra_powers)
        e, eb = pred - th, base - th
        r["log_diag_rmse"] = float(np.std(e))
        r["log_diag_rmse_subset_only"] = float(np.std(eb))
        r["log_diag_std"] = float(np.std(th))
        results.append(r)
        log(f"held-out {k + 1}/{args.test}: {json.dumps(r)}")
        if args.jsonl:
            with open(args.jsonl, "a") as fh:
                fh.write(json.dumps(r) + "\n")

    summary = summarize_h0(results, dict(
        train_models=args.train, test_models=len(results), train_samples=int(len(t)),
        train_loss_first_last=[float(losses[0]), float(losses[-1])],
        train_seconds=t_train, data_seconds=t_data), args.extra_powers)
    out = {"summary": summary, "models": results}
    if args.out:
        json.dump(out, open(args.out, "w"), indent=1)
    print(json.dumps(summary))


if __name__ == "__main__":
    main()
