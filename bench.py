#!/usr/bin/env python3
"""Throughput of the FP32C polyhedral forward (gz + TMI) plus the matrix-free adjoint
gradient (d/dx, d/dy, d/dz of every vertex) on one MI355X.

One step = new vertex positions (facet setup) -> forward over every station x facet
pair -> adjoint over every pair, on the contact model with 192x192x64 cells (about 2.2M
facets, contact facets stored once) and 256x256 stations; inputs are resident in HBM
before the timed region.  Prints one JSON line; with --gpus N > 1 it exits with a
message (this build runs on one GPU).

value = station x facet pairs per second for forward + adjoint
      = n_stations * n_facets / (seconds per step)

After the timed region (not part of value): the accuracy of the FP32C forward against
the FP64 path, the CPU baseline, and the second half of the headline metric (seconds of
the learned-H0 FP32C L-BFGS-B to the plain FP64 run's final misfit, "inversion").
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from lithoinv3d_amd import synth, workload  # noqa: E402
from lithoinv3d_amd.lithofwd import ForwardContext  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline(model, st, fh, field, seconds):
    """The oracle (FP64 restatement of the reference forward and its per-face FD
    Jacobian contracted with weights), single-threaded like the reference, on a sample
    of stations x every facet of the reference layout (contact facets twice)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # the checker; timed here as the CPU baseline only

    om = O.OracleModel.from_litho(model, field, fh)
    ref_faces = model.reference_face_count()
    rng = np.random.default_rng(11)

    def run(idx, threads=1):
        s = st[idx]
        w = rng.normal(size=len(idx))
        t0 = time.perf_counter()
        om.gravity(s, threads=threads)
        om.magnetic(fh, s, threads=threads)
        t1 = time.perf_counter()
        om.gravity_vertex_jtw(s, w, threads=threads)
        om.magnetic_vertex_jtw(fh, s, w, threads=threads)
        t2 = time.perf_counter()
        return t1 - t0, t2 - t1

    probe = run(np.array([len(st) // 2]))
    per_station = sum(probe)
    n = int(max(1, min(64, seconds // max(per_station, 1e-9))))
    idx = rng.choice(len(st), n, replace=False)
    tf, tg = run(idx)
    pairs = n * ref_faces
    # the same code on the host's cores (one station per thread at a time), for scale;
    # 16 threads at most, the CPU share of one GPU on the GPU hosts
    nt = max(1, min(16, os.cpu_count() or 1))
    idx_t = rng.choice(len(st), n * nt, replace=False)
    tft, tgt = run(idx_t, threads=nt)
    pairs_t = n * nt * ref_faces
    return {
        "value": pairs / (tf + tg),
        "unit": "station x facet pairs/s (fwd+grad, reference facet layout)",
        "cores": 1,
        "kind": "port",
        "sample": f"{n} stations x {ref_faces} facets (all units' meshes, contacts twice): "
                  f"forward gz+TMI {tf:.2f} s, per-face FD Jacobian J^T w gz+TMI {tg:.2f} s; "
                  f"host {cpu_model()}",
        "forward_pairs_per_s": pairs / tf,
        "grad_pairs_per_s": pairs / tg,
        "threaded": {"threads": nt, "nproc": os.cpu_count(), "value": pairs_t / (tft + tgt),
                     "sample": f"{n * nt} stations, forward {tft:.2f} s, J^T w {tgt:.2f} s"},
    }


def inversion_check(n_models, n_train=24):
    """The second half of the headline metric on the config-1-style contact models
    (64x64x32 cells, 128x128 stations, about 1 660 control-point parameters): the H0
    network trained on n_train seeded models, then per held-out model the plain FP64
    L-BFGS-B run of 40 iterations and the FP32C runs to its final misfit with and without
    the learned H0 (seconds include the learned H0's subset column norms, features and
    network).  Medians over the held-out models."""
    from lithoinv3d_amd import experiments as E

    t0 = time.perf_counter()
    X, t = E.training_set(range(1000, 1000 + n_train), (64, 64, 32), 128)
    net, norm, _ = E.train_mlp(X, t, steps=3000)
    t_train = time.perf_counter() - t0
    rows = []
    for s in range(2000, 2000 + n_models):
        r = E.compare_solvers(E.contact_problem(s, (64, 64, 32), 128), net, norm, max_iter=40)
        rows.append(r)
        log(f"inversion seed {s}: {json.dumps(r)}")
    net.close()

    def med(f):
        v = [f(r) for r in rows]
        v = [x for x in v if x is not None]
        return float(np.median(v)) if v else None

    hit = lambda r, k, i: r[k]["to_target"][i] if r[k]["to_target"] else None
    return {
        "models": "contact64x64x32_st128, seeds 2000..%d held out, %d trained" % (
            1999 + n_models, n_train),
        "train_seconds": t_train,
        "fp64_seconds": med(lambda r: r["fp64"]["seconds"]),
        "fp64_iterations": med(lambda r: r["fp64"]["iterations"]),
        "fp32c_seconds_to_fp64_misfit": med(lambda r: hit(r, "fp32c", 1)),
        "learned_fp32c_seconds_to_fp64_misfit": med(lambda r: hit(r, "learned_fp32c", 1)),
        "learned_fp32c_iterations_to_fp64_misfit": med(lambda r: hit(r, "learned_fp32c", 0)),
        "learned_time_ratio_vs_fp64": med(
            lambda r: hit(r, "learned_fp32c", 1) / r["fp64"]["seconds"]
            if r["learned_fp32c"]["to_target"] else None),
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="contact192x192x64_st256",
                    choices=sorted(workload.WORKLOADS))
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--inversion", type=int, default=2,
                    help="held-out models for the L-BFGS-B seconds-to-target check (0: skip)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first run one plain device copy of this many MiB (a kernel of known "
                         "traffic for scaling the HBM byte counters of a profiler run)")
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "traffic.json"),
                    help="measured HBM bytes per launch of the dominant kernels, if present")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU design is facet-sharded with one all-reduce per forward, "
              "not built)", file=sys.stderr)
        sys.exit(2)

    import torch

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)
    dev = torch.device("cuda:0")
    if args.copy_calibration_mib > 0:
        n = args.copy_calibration_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy calibration: {args.copy_calibration_mib} MiB read + written once")
        del src, dst

    t0 = time.time()
    cfg = workload.build(args.workload)
    m, st = cfg["model"], cfg["stations"]
    fh = synth.field_unit_vector(cfg["inc"], cfg["dec"])
    field = cfg["field"]
    n_s, n_f = len(st), m.n_faces
    pairs = n_s * n_f
    log(f"workload {args.workload}: {m.n_vertices} vertices, {n_f} facets "
        f"({m.reference_face_count()} in the reference layout), {n_s} stations, "
        f"{pairs:.3e} pairs; generated in {time.time() - t0:.1f}s")

    ctx = ForwardContext(m, st, fh, field, precision="fp32c")
    rng = np.random.default_rng(7)
    wg = torch.tensor(rng.normal(size=n_s), dtype=torch.float64, device=dev)
    wt = torch.tensor(rng.normal(size=n_s), dtype=torch.float64, device=dev)
    gz = torch.zeros(n_s, dtype=torch.float64, device=dev)
    tm = torch.zeros(n_s, dtype=torch.float64, device=dev)
    grad = torch.zeros(m.n_vertices * 3, dtype=torch.float64, device=dev)
    verts = torch.tensor(m.vertices.reshape(-1), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    def step():
        ctx.set_vertices_device(verts.data_ptr())
        ctx.forward_device(gz.data_ptr(), tm.data_ptr())
        ctx.adjoint_device(wg.data_ptr(), wt.data_ptr(), grad.data_ptr())

    for i in range(args.warmup):
        step()
        ctx.sync()
        log(f"warmup step {i + 1}/{args.warmup} done")
    ctx.timing()  # reset
    ctx.set_timing(True)
    torch.cuda.synchronize()
    ctx.sync()
    t1 = time.perf_counter()
    for i in range(args.steps):
        step()
    ctx.sync()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    ctx.set_timing(False)
    tim = ctx.timing()
    sec = (t2 - t1) / args.steps
    log(f"{args.steps} steps in {t2 - t1:.3f}s; kernels: " +
        ", ".join(f"{k} {v[0] / max(v[1], 1):.2f} ms x{v[1]}" for k, v in tim.items() if v[1]))
    fb = ctx.fp64_edges()
    # the host-buffer entry points (vertices in, data and gradient out through host memory,
    # as the reference's call sites would use them), after the timed region: PCIe included
    wg_h, wt_h = wg.cpu().numpy(), wt.cpu().numpy()
    ctx.sync()
    t3 = time.perf_counter()
    for _ in range(2):
        ctx.set_vertices(m.vertices)
        ctx.forward()
        ctx.adjoint(wg_h, wt_h)
    host_sec = (time.perf_counter() - t3) / 2
    log(f"host-buffer step {host_sec * 1e3:.1f} ms")

    fwd_ms = tim["forward"][0] / max(tim["forward"][1], 1)
    adj_ms = tim["adjoint_pairs"][0] / max(tim["adjoint_pairs"][1], 1)
    kern = {
        "forward_f32c": fwd_ms,
        "adjoint_pairs_f32": adj_ms,
    }
    dom = max(kern, key=kern.get)
    flops = workload.flop_eq(dom) * pairs
    achieved = flops / (kern[dom] * 1e-3) / 1e12
    traffic = None
    if os.path.exists(args.traffic_file):
        tr = json.load(open(args.traffic_file))
        rec = tr.get("kernels", {}).get(dom)
        if rec and tr.get("workload") == args.workload:
            traffic = rec.get("hbm_bytes_per_launch")
    per_kernel = {
        k: {
            "avg_ms": v,
            "tflops": workload.flop_eq(k) * pairs / (v * 1e-3) / 1e12,
            "arith_tflops": workload.arith_flops(k) * pairs / (v * 1e-3) / 1e12,
            "pairs_per_s": pairs / (v * 1e-3),
        }
        for k, v in kern.items()
    }

    acc = None
    if not args.no_accuracy:
        # FP32C against the FP64 GPU path (itself held to the oracle at 1e-4 mGal /
        # 1e-3 nT by the parity tests) over every station
        g32 = gz.cpu().numpy().copy()
        t32 = tm.cpu().numpy().copy()
        with ForwardContext(m, st, fh, field, precision="fp64") as c64:
            ta = time.perf_counter()
            g64, t64 = c64.forward()
            tb = time.perf_counter()
        acc = {
            "vs": "FP64 GPU path, all stations",
            "max_abs_dgz_mGal": float(np.abs(g32 - g64).max()),
            "max_abs_dtmi_nT": float(np.abs(t32 - t64).max()),
            "gz_range_mGal": [float(g64.min()), float(g64.max())],
            "tmi_range_nT": [float(t64.min()), float(t64.max())],
            "fp64_forward_s": tb - ta,
            "fp32c_edges_on_fp64_branch": fb,
        }
        # the reference's TMI edge coefficient leaves a field of order 1e6 nT on a closed
        # body; the textbook coefficient gives the physical field (a few hundred nT): the
        # precision mode is held to 1e-2 nT in both
        with ForwardContext(m, st, fh, field, precision="fp32c", tmi_mode="textbook",
                            want=("tmi",)) as c:
            _, tt32 = c.forward()
        with ForwardContext(m, st, fh, field, precision="fp64", tmi_mode="textbook",
                            want=("tmi",)) as c:
            _, tt64 = c.forward()
        acc["max_rel_dtmi"] = float(np.abs(t32 - t64).max() / np.abs(t64).max())
        acc["textbook_tmi"] = {
            "max_abs_dtmi_nT": float(np.abs(tt32 - tt64).max()),
            "tmi_range_nT": [float(tt64.min()), float(tt64.max())],
        }
        log(f"accuracy: {acc}")

    cpu = None
    if not args.no_cpu_baseline:
        log("timing the CPU baseline")
        cpu = cpu_baseline(m, st, fh, field, args.cpu_seconds)
        log(f"cpu baseline: {cpu}")

    inv = None
    if args.inversion > 0:
        log("L-BFGS-B seconds to the plain FP64 misfit, learned H0")
        try:
            inv = inversion_check(args.inversion)
        except Exception as e:  # the throughput line is still reported
            inv = {"error": repr(e)}
        log(f"inversion: {inv}")

    out = {
        "metric": "station x facet evals/s (fwd+grad)",
        "value": pairs / sec,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded Gaussian-field label grid, voxel-face facets)",
        "config": {
            "workload": args.workload,
            "n_stations": n_s,
            "n_facets": n_f,
            "n_facets_reference_layout": m.reference_face_count(),
            "n_vertices": m.n_vertices,
            "precision": "fp32c (FP32 facet frames, Kahan facet sums, FP64 near-edge branch)",
            "step": "facet setup + forward gz,TMI + adjoint dPhi/dxyz",
            "reference_layout_pairs_per_s": n_s * m.reference_face_count() / sec,
        },
        "host_buffers": {"ms_per_step": host_sec * 1e3, "pairs_per_s": pairs / host_sec,
                         "note": "lf_set_vertices + lf_forward + lf_adjoint with host arrays "
                                 "(PCIe copies included), 2 steps after the timed region"},
        "roofline": {
            "bound": "valu",
            "kernel": dom,
            "achieved": achieved,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": traffic,
            "flop_eq_per_pair": workload.flop_eq(dom),
            # the same kernel on arithmetic alone (FMA 2, add/mul 1, transcendentals 0)
            "arith_flops_per_pair": workload.arith_flops(dom),
            "arith_achieved": workload.arith_flops(dom) * pairs / (kern[dom] * 1e-3) / 1e12,
            "arith_frac": (workload.arith_flops(dom) * pairs / (kern[dom] * 1e-3) / 1e12
                           / FP32_PEAK_TFLOPS),
        },
        "kernels": per_kernel,
        "accuracy": acc,
        "cpu_baseline": cpu,
        "inversion": inv,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
