"""GPU, full size, workload C2: Shamir DWF 16^3x32, Ls 16, M5 1.8, mf 0.01, links exp(i 0.2 H), seeded, t antiperiodic.
Tolerances and shared inputs: tests/fullsize_common.py."""
import pytest

import cps_public_amd as C
from cps_public_amd import fields as F
from tests.fullsize_common import DOT_TOL, ORA_RES_TOL, SOL_TOL
from tests.helpers import rel, wilson_gauge

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c2():
    L, ls = [16, 16, 16, 32], 16
    u = wilson_gauge(L, 0.2, 202)
    lat = C.Lattice(L, ls, (0, 0, 0, 1))
    lat.import_gauge(u)
    yield lat, u, L, ls
    lat.close()


def test_c2_mdagm_full_size(oracle, c2):
    lat, u, L, ls = c2
    d = oracle.dwf_struct(L, ls, 1.8)
    inp = F.gaussian_fermion(lat.vcb * ls, 203)
    ref, dref = oracle.dwf_mdagm(u, inp, 0.01, d)
    got, dp = lat.dwf_mdagm(inp, 0.01)
    print("c2 mdagm", rel(got, ref), abs(dp - dref) / dref, flush=True)
    assert rel(got, ref) < 1e-13
    assert abs(dp - dref) / dref < DOT_TOL


def test_c2_fp64_cg_fixed_iterations(oracle, c2):
    """Ten FP64 CG iterations (max_num_iter 11: InvCg runs max_num_iter - 1,
    inv_cg.C:275,293) against the oracle's InvCg stopped at the same count."""
    lat, u, L, ls = c2
    d = oracle.dwf_struct(L, ls, 1.8)
    src = F.gaussian_fermion(lat.vcb * ls, 204)
    ref, it_ref, _ = oracle.dwf_invcg(u, src, 0.01, 1e-10, 11, d)
    got, res = lat.invcg(C.make_op(C.OP_DWF, mass=0.01, m5=1.8), src, C.make_cg_arg(1e-10, 11))
    print("c2 cg 10 iterations", it_ref, res[0].iterations, rel(got, ref), flush=True)
    assert res[0].iterations == it_ref
    assert rel(got, ref) < 1e-11


def test_c2_mixed_solve(oracle, c2):
    """The C2 workload: one Gaussian source, mixed (FP16 spinors and links,
    FP64 reliable updates every 100 iterations) and FP64, both to 1e-10."""
    lat, u, L, ls = c2
    d = oracle.dwf_struct(L, ls, 1.8)
    src = F.gaussian_fermion(lat.vcb * ls, 203)
    op = C.make_op(C.OP_DWF, mass=0.01, m5=1.8)
    x64, r64 = lat.invcg(op, src, C.make_cg_arg(1e-10, 5000))
    xm, rm = lat.invcg(op, src, C.make_cg_arg(1e-10, 5000, precision=C.PREC_MIXED_HALF, reliable_every=100))
    e = rel(xm, x64)
    res_ora = rel(oracle.dwf_mdagm(u, xm, 0.01, d)[0], src)
    print("c2 solves: fp64", r64[0].iterations, r64[0].true_res, "mixed", rm[0].iterations, rm[0].true_res,
          "err vs fp64", e, "oracle residual", res_ora, flush=True)
    assert r64[0].true_res <= 1e-10 and rm[0].true_res <= 1e-10
    assert res_ora <= ORA_RES_TOL
    assert e < SOL_TOL
