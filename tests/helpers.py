"""Shared input builders for the tests (seeded CPS-layout fields)."""
import numpy as np

from cps_public_amd import fields as F
from cps_public_amd import layout as LY


def wilson_gauge(L, eps, seed, bc=(0, 0, 0, 1)):
    """WILSON-order gauge with the boundary signs applied (as the DiracOp
    constructor leaves it)."""
    u = F.gauge_field(L, eps, seed)
    return LY.gauge_canon_to_wilson(LY.bond_cond(u, L, bc), L)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
