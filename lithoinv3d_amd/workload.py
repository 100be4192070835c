"""Named workloads and the per-pair operation counts used for roofline figures.

Operation counts are algorithmic: counted from the per-pair source of each kernel
(csrc/lithofwd.hip k_forward_f32c, csrc/adjoint.h k_adjoint_pairs) with a fused
multiply-add as 2 flops and each transcendental (sqrt, reciprocal, log2) as 8
flop-equivalents -- on gfx950 these issue at a quarter of the FMA rate, so one occupies
the issue slots of four FMAs.  A pair is one (station, facet) combination with the
facet stored once (contact facets are not duplicated).
"""
from __future__ import annotations

from . import synth

TRANSCENDENTAL_FLOP_EQ = 8

# per (station, facet) pair
OPS = {
    # forward, gz + TMI fused, FP32C: station offset 3, frame projection 15, squared
    # distances 14, solid-angle D and N 19, atan2 20, edge (s, dperp, rho^2) 18,
    # three stable edge logs 12, FP64-branch tests 6, gz combine + Kahan 12,
    # TMI combine + Kahan 11  -> 130 arithmetic; 3 sqrt + 4 rcp + 3 log2 = 10 special
    "forward_f32c": dict(arith=130, special=10),
    # adjoint, gz + TMI seeds: forward part 112, coefficient seeds 13, reverse sweep 141,
    # accumulation of the 22 facet seeds 25 -> 291 arithmetic;
    # 3 sqrt + 1 rcp (atan2) + 15 (three edge logs with their reciprocals) + 4 reverse rcp
    "adjoint_pairs_f32": dict(arith=291, special=23),
}


def flop_eq(kernel: str) -> int:
    o = OPS[kernel]
    return o["arith"] + TRANSCENDENTAL_FLOP_EQ * o["special"]


WORKLOADS = {
    # four-unit contact model, 64x64x32 cells of 25 m, 128x128 stations
    "contact64x64x32_st128": dict(shape=(64, 64, 32), n_st=128),
    # the same generator at 192x192x64 cells (about 2.2M facets), 256x256 stations:
    # FP32C forward + adjoint gradient with respect to x, y, z of every vertex
    "contact192x192x64_st256": dict(shape=(192, 192, 64), n_st=256),
    # eight units, 256x256x96 cells of 20 m, 512x512 stations
    "contact256x256x96_st512": dict(shape=(256, 256, 96), n_st=512, cell=20.0, n_units=8),
}


def build(name: str):
    kw = dict(WORKLOADS[name])
    return synth.config_contact(**kw)
