"""GPU, full size, workload C4: clover 32^3x64, c_sw 1.0, links exp(i 0.15 H), 12 point sources x 4 kappas
(0.1360-0.1375) as 48 rhs.
Tolerances and shared inputs: tests/fullsize_common.py."""
import numpy as np
import pytest

import cps_public_amd as C
from cps_public_amd import fields as F
from tests.fullsize_common import DOT_TOL, ORA_RES_TOL, SOL_TOL, point_sources
from tests.helpers import rel, wilson_gauge

pytestmark = pytest.mark.gpu


C4_KAPPAS = [0.1360, 0.1365, 0.1370, 0.1375]


@pytest.fixture(scope="module")
def c4():
    L = [32, 32, 32, 64]
    u = wilson_gauge(L, 0.15, 404)
    lat = C.Lattice(L, 1, (0, 0, 0, 1))
    lat.import_gauge(u)
    lat.clover_build(C4_KAPPAS, 1.0)
    yield lat, u, L
    lat.close()


def _clover_op(kappa):
    return C.make_op(C.OP_CLOVER, mass=1.0 / (2 * kappa) - 4.0, csw=1.0)


@pytest.fixture(scope="module")
def c4_oracle_clover(oracle, c4):
    lat, u, L = c4
    kappa = C4_KAPPAS[1]
    return kappa, oracle.clover_mat_chkb(u, 1, 0, kappa, 1.0, L), oracle.clover_mat_chkb(u, 0, 1, kappa, 1.0, L)


def test_c4_clover_and_mdagm_full_size(oracle, c4, c4_oracle_clover):
    """The device-built A_oo and A_ee^-1 of kappa 0.1365 and M^dag M of that
    kappa (the second of the four sets) against the oracle."""
    lat, u, L = c4
    kappa, A_oo, A_ee_inv = c4_oracle_clover
    assert rel(lat.clover_export(1, 1), A_oo) < 1e-13
    assert rel(lat.clover_export(1, 0), A_ee_inv) < 1e-13
    inp = F.gaussian_fermion(lat.vcb, 405)
    ref, dref = oracle.clover_mdagm(u, A_oo, A_ee_inv, inp, kappa, L)
    # rhs r of a batch uses kappa set r // (nrhs / 4): rhs 1 of 4 is set 1
    others = [F.gaussian_fermion(lat.vcb, 406 + k) for k in range(3)]
    batch = np.concatenate([others[0], inp, others[1], others[2]])
    got, dp = lat.mdagm(_clover_op(C4_KAPPAS[0]), batch, nrhs=4)
    n = len(inp)
    print("c4 mdagm", rel(got[n:2 * n], ref), abs(dp[1] - dref) / dref, flush=True)
    assert rel(got[n:2 * n], ref) < 1e-13
    assert abs(dp[1] - dref) / dref < DOT_TOL


def test_c4_fp64_cg_fixed_iterations(oracle, c4, c4_oracle_clover):
    lat, u, L = c4
    kappa, A_oo, A_ee_inv = c4_oracle_clover
    src = point_sources(lat.vcb, 1, 1)
    ref, it_ref, _ = oracle.clover_invcg(u, A_oo, A_ee_inv, src, kappa, 1e-10, 6, L)
    others = [F.gaussian_fermion(lat.vcb, 407 + k) for k in range(3)]
    batch = np.concatenate([others[0], src, others[1], others[2]])
    got, res = lat.invcg(_clover_op(C4_KAPPAS[0]), batch, C.make_cg_arg(1e-10, 6), nrhs=4)
    n = len(src)
    print("c4 cg 5 iterations", it_ref, res[1].iterations, rel(got[n:2 * n], ref), flush=True)
    assert res[1].iterations == it_ref
    assert rel(got[n:2 * n], ref) < 1e-11


def test_c4_mixed_solve_48rhs(oracle, c4, c4_oracle_clover):
    """The C4 workload: 48 rhs (12 point sources x 4 kappas) in one batched
    mixed solve (FP16 spinors and links, hipGraph-replayed iterations), every
    rhs to FP64 true residual 1e-10; one rhs per kappa against the FP64 GPU
    CG, and rhs 12 through the oracle's M^dag M of its kappa."""
    lat, u, L = c4
    kappa, A_oo, A_ee_inv = c4_oracle_clover
    n = 24 * lat.vcb
    src = point_sources(lat.vcb, 1, 48)
    arg = C.make_cg_arg(1e-10, 5000, precision=C.PREC_MIXED_HALF, reliable_every=100, reliable_delta=0.01)
    x, res = lat.invcg(_clover_op(C4_KAPPAS[0]), src, arg, nrhs=48)
    print("c4 mixed", [r.iterations for r in res], max(r.true_res for r in res), flush=True)
    assert all(r.status == 1 for r in res)
    assert max(r.true_res for r in res) <= 1e-10
    # rhs 0, 12, 24, 36 (kappa sets 0..3) in FP64: a batch of 4 uses set r
    pick = np.concatenate([src[k * 12 * n:(k * 12 + 1) * n] for k in range(4)])
    x64, r64 = lat.invcg(_clover_op(C4_KAPPAS[0]), pick, C.make_cg_arg(1e-10, 5000), nrhs=4)
    errs = [rel(x[k * 12 * n:(k * 12 + 1) * n], x64[k * n:(k + 1) * n]) for k in range(4)]
    res_ora = rel(oracle.clover_mdagm(u, A_oo, A_ee_inv, x[12 * n:13 * n], kappa, L)[0], src[12 * n:13 * n])
    print("c4 fp64", [r.iterations for r in r64], "err vs fp64", errs, "oracle residual", res_ora, flush=True)
    assert max(r.true_res for r in r64) <= 1e-10
    assert res_ora <= ORA_RES_TOL
    assert max(errs) < SOL_TOL
