"""Named workloads and the per-pair operation counts used for roofline figures.

Operation counts are algorithmic, per (station, facet) pair, counted from the per-pair
source of each kernel (csrc/lithofwd.hip k_forward_f32c, csrc/adjoint.h k_adjoint_pairs;
the compiled loops hold the same mix): a fused multiply-add is 2 flops, an add, subtract
or multiply 1, and each transcendental (sqrt, reciprocal, log2) 8 flop-equivalents: on
gfx950 these issue at 32 lanes/clock/CU, a quarter of the 128 FMA lanes/clock/CU that the
157.3 TFLOP/s FP32 peak assumes, so one occupies the issue slots of four FMAs.
Compares, selects, sign and min/max operations and conversions are not counted.
A pair is one (station, facet) combination with the facet stored once (contact facets are
not duplicated).
"""
from __future__ import annotations

from . import synth

TRANSCENDENTAL_FLOP_EQ = 8

# per (station, facet) pair, counted over the common path of the compiled loops: a wave
# whose stations are all more than about five edge lengths from its facet (forward) or
# whose facets are all that far from its station (adjoint), with the solid angle's
# tan(Omega/2) = N/D below 1/8 -- nearly every wave of the named workloads.  The branch
# for nearer pairs (octant-reduced atan2, case-form edge logs, FP64 edges) is not counted;
# the three FP64 subtractions of the station offset count as add/mul
OPS = {
    # forward, gz + TMI fused, FP32C: station offset from the record's origin and its frame
    # projection (nested fmas), squared distances, three Newton-refined square roots (rsq +
    # 2 fmas), solid-angle N and D, small-angle atan series of N/D, z = ell / (d1 + d2) for
    # the three edges from one Newton-refined reciprocal of the product of the three sums,
    # the 2 atanh(z) series per edge, gz and TMI combined into the station's FP32 sums of
    # the 16-facet group (the group sums' Kahan update, 8 operations per 16 pairs, is not
    # counted); 3 rsq + 2 rcp.  The compiled loop runs two stations per lane and computes
    # the facet's 2A once for both (108 FMA + 93 add/mul + 10 transcendentals per two pairs)
    "forward_f32c": dict(fma=54, addmul=47, special=5),
    # adjoint, gz + TMI seeds: the forward part (plain rsq distances and rcp edge ratios,
    # nested-fma projection), the small-angle atan series and its derivative, the series of
    # 2 atanh(z) and its derivative, coefficient seeds, reverse sweep through the edge
    # ratios, solid angle and distances (the near-edge seeds left to their own branch,
    # doubled seeds for the squared distances; the far-form edge seeds taken for every
    # edge and removed again in the near branch), accumulation of the 19 facet seeds (their
    # FP64 flush every 16 stations is not counted); 3 rsq + 4 rcp (three edge ratios and
    # the solid angle's N/D)
    "adjoint_pairs_f32": dict(fma=105, addmul=79, special=7),
}


def flop_eq(kernel: str) -> int:
    o = OPS[kernel]
    return 2 * o["fma"] + o["addmul"] + TRANSCENDENTAL_FLOP_EQ * o["special"]


def arith_flops(kernel: str) -> int:
    """Flops of the pair without the transcendental weighting (FMA = 2)."""
    o = OPS[kernel]
    return 2 * o["fma"] + o["addmul"]


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
