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

# per (station, facet) pair
OPS = {
    # forward, gz + TMI fused, FP32C: FP64 station offset (3 adds), frame projection,
    # squared distances, solid-angle N and D, atan2 polynomial, edge (s, dperp, rho^2),
    # three cancellation-free edge logs (ln(1+y) with the rounding correction of 1+y,
    # weighted by sat(2 - u)), gz and TMI combine with Kahan sums; 3 sqrt + 4 rcp + 3 log2
    "forward_f32c": dict(fma=50, addmul=64, special=10),
    # adjoint, gz + TMI seeds: the forward part, coefficient seeds, reverse sweep through
    # the edge logs, solid angle and distances, accumulation of the 20 facet seeds (one
    # angle seed per in-plane edge direction, torques by two FMAs each);
    # 3 rsqrt (distances and their reciprocals) + 8 rcp (atan2, two per edge log with its
    # derivative, the atan2 derivative) + 3 log2
    "adjoint_pairs_f32": dict(fma=106, addmul=142, special=14),
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
