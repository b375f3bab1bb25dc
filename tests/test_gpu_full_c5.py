"""GPU, full size, workload C5: the solve service on 8^4, 16^3x32 and 24^3x48 in one mix.
Tolerances and shared inputs: tests/fullsize_common.py."""
import numpy as np
import pytest

import cps_public_amd as C
from cps_public_amd import fields as F
from tests.fullsize_common import ORA_RES_TOL
from tests.helpers import rel, wilson_gauge

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [C.PREC_DOUBLE, C.PREC_MIXED_HALF])
def test_c5_service_three_lattices(oracle, prec):
    """Requests on 8^4 (Wilson), 16^3x32 (DWF Ls 8) and 24^3x48 (DWF Ls 16),
    masses 0.005-0.05, 1-3 rhs, FP64 or mixed precision: each result meets
    the FP64 true residual 1e-10, the oracle's M^dag M of its operator
    reproduces the source, and it is the same alone and in the mix -- bit
    for bit in FP64 (per-slot arithmetic), within 1e-8 relative in mixed
    precision (an engine's reliable updates fall on the same iterations for
    all its slots, so the mix moves a slot's updates)."""
    lats = {"w8": ([8, 8, 8, 8], 1, 0.2, 501), "d16": ([16, 16, 16, 32], 8, 0.2, 502),
            "d24": ([24, 24, 24, 48], 16, 0.2, 503)}
    svc = C.Service(200 << 30, max_slots=4)
    gid, us = {}, {}
    for k, (L, ls, eps, seed) in lats.items():
        us[k] = wilson_gauge(L, eps, seed)
        gid[k] = svc.add_gauge(L, ls, us[k])
    plan = [("w8", C.OP_WILSON, 0.05, 2), ("d16", C.OP_DWF, 0.02, 1), ("d24", C.OP_DWF, 0.05, 1),
            ("w8", C.OP_WILSON, 0.01, 3), ("d16", C.OP_DWF, 0.005, 2)]
    reqs = []
    for i, (k, kind, mass, nrhs) in enumerate(plan):
        L, ls = lats[k][0], lats[k][1]
        vcb = int(np.prod(L)) // 2
        op = C.make_op(kind, mass=mass, m5=1.8)
        src = np.concatenate([F.gaussian_fermion(vcb * ls, 510 + 10 * i + r) for r in range(nrhs)])
        reqs.append((k, op, src, nrhs, mass))
    alone = [svc.wait(svc.submit(gid[k], op, src, nrhs, precision=prec)) for k, op, src, nrhs, _ in reqs]
    ids = [svc.submit(gid[k], op, src, nrhs, precision=prec) for k, op, src, nrhs, _ in reqs]
    mix = [svc.wait(rid) for rid in ids]
    st = svc.stats()
    print("c5 service stats", st, flush=True)
    for (k, op, src, nrhs, mass), (sa, ra, la), (sm, rm, lm) in zip(reqs, alone, mix):
        L, ls = lats[k][0], lats[k][1]
        n = len(src) // nrhs
        print("c5 request", prec, k, mass, nrhs, [r.iterations for r in rm], max(r.true_res for r in rm),
              "latency alone/mix", la, lm, flush=True)
        if prec == C.PREC_DOUBLE:
            assert np.array_equal(sa, sm)
        else:
            assert rel(sm, sa) < 1e-8
        assert max(r.true_res for r in rm) <= 1e-10
        r0 = sm[:n]
        if k == "w8":
            ax = oracle.wilson_mdagm(us[k], r0, 1.0 / (2 * (mass + 4.0)), L)[0]
        else:
            ax = oracle.dwf_mdagm(us[k], r0, mass, oracle.dwf_struct(L, ls, 1.8))[0]
        assert rel(ax, src[:n]) <= ORA_RES_TOL
    assert st["last_error"] == 0
    svc.close()
