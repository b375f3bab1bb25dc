"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on the
same seeded inputs.  Tolerances: FP64 operators 1e-13 relative (rounding of a
different summation order only); FP64 CG solutions 1e-10 relative (the
bound the solver is specified to); mixed-precision CG true residual <= stop_rsd, solution
within 1e-8 of the FP64 oracle."""
import numpy as np
import pytest

import cps_public_amd as C
from cps_public_amd import fields as F
from tests.helpers import rel, wilson_gauge

pytestmark = pytest.mark.gpu

L4 = [4, 4, 4, 4]
L8 = [8, 8, 8, 8]


@pytest.fixture(scope="module")
def lat4():
    lat = C.Lattice(L4, ls=1, bc=(0, 0, 0, 1))
    u = wilson_gauge(L4, 0.3, 11)
    lat.import_gauge(u)
    yield lat, u
    lat.close()


@pytest.mark.parametrize("cb", [0, 1])
@pytest.mark.parametrize("dag", [0, 1])
def test_wilson_dslash(oracle, lat4, cb, dag):
    lat, u = lat4
    psi = F.gaussian_fermion(lat.vcb, 100 + cb)
    ref = oracle.wilson_dslash(u, psi, cb, dag, L4)
    got = lat.wilson_dslash(psi, cb, dag)
    print("wilson_dslash", cb, dag, rel(got, ref), flush=True)
    assert rel(got, ref) < 1e-13


def test_wilson_dslash_passes_gauge(oracle):
    # the gauge passed as an argument replaces the resident one (wilson.h semantics)
    lat = C.Lattice(L4, 1, (0, 0, 0, 1))
    u = wilson_gauge(L4, 0.3, 12)
    psi = F.gaussian_fermion(lat.vcb, 5)
    got = lat.wilson_dslash(psi, 1, 0, u=u)
    assert rel(got, oracle.wilson_dslash(u, psi, 1, 0, L4)) < 1e-13
    lat.close()


def test_wilson_mdagm(oracle, lat4):
    lat, u = lat4
    psi = F.gaussian_fermion(lat.vcb, 7)
    kappa = 0.12
    ref, mp_ref = oracle.wilson_mdagm(u, psi, kappa, L4)
    got, mp = lat.wilson_mdagm(psi, kappa)
    assert rel(got, ref) < 1e-13
    assert abs(mp - mp_ref) / mp_ref < 1e-13
    assert rel(lat.wilson_m(psi, kappa), oracle.wilson_m(u, psi, kappa, L4)) < 1e-13
    assert rel(lat.wilson_m(psi, kappa, dag=1), oracle.wilson_mdag(u, psi, kappa, L4)) < 1e-13


def test_wilson_cg_c1(oracle):
    """Wilson 8^4, links exp(i 0.3 H), kappa 0.12, CG to 1e-10."""
    lat = C.Lattice(L8, 1, (0, 0, 0, 1))
    u = wilson_gauge(L8, 0.3, 2024)
    lat.import_gauge(u)
    src = F.gaussian_fermion(lat.vcb, 2025)
    mass = 1.0 / (2 * 0.12) - 4.0
    ref, it_ref, tr_ref = oracle.wilson_invcg(u, src, mass, 1e-10, 5000, L8)
    op = C.make_op(C.OP_WILSON, mass=mass)
    got, res = lat.invcg(op, src, C.make_cg_arg(1e-10, 5000))
    print("wilson cg", it_ref, res[0].iterations, tr_ref, res[0].true_res, rel(got, ref), flush=True)
    assert abs(res[0].iterations - it_ref) <= 1
    assert res[0].true_res < 1e-10
    assert rel(got, ref) < 1e-10
    lat.close()


@pytest.fixture(scope="module")
def dwf4():
    L = [4, 4, 4, 8]
    ls = 8
    lat = C.Lattice(L, ls=ls, bc=(0, 0, 0, 1))
    u = wilson_gauge(L, 0.2, 31)
    lat.import_gauge(u)
    yield lat, u, L, ls
    lat.close()


@pytest.mark.parametrize("cb", [0, 1])
@pytest.mark.parametrize("dag", [0, 1])
def test_dwf_dslash(oracle, dwf4, cb, dag):
    lat, u, L, ls = dwf4
    d = oracle.dwf_struct(L, ls, 1.8)
    inp = F.gaussian_fermion(lat.vcb * ls, 40 + cb)
    ref = oracle.dwf_dslash(u, inp, 0.01, cb, dag, d)
    got = lat.dwf_dslash(inp, 0.01, cb, dag)
    print("dwf_dslash", cb, dag, rel(got, ref), flush=True)
    assert rel(got, ref) < 1e-13


def test_dwf_mdagm(oracle, dwf4):
    lat, u, L, ls = dwf4
    d = oracle.dwf_struct(L, ls, 1.8)
    inp = F.gaussian_fermion(lat.vcb * ls, 43)
    ref, dp_ref = oracle.dwf_mdagm(u, inp, 0.01, d)
    got, dp = lat.dwf_mdagm(inp, 0.01)
    assert rel(got, ref) < 1e-13
    assert abs(dp - dp_ref) / dp_ref < 1e-13


def test_dwf_cg(oracle, dwf4):
    lat, u, L, ls = dwf4
    d = oracle.dwf_struct(L, ls, 1.8)
    src = F.gaussian_fermion(lat.vcb * ls, 44)
    ref, it_ref, tr_ref = oracle.dwf_invcg(u, src, 0.01, 1e-10, 5000, d)
    op = C.make_op(C.OP_DWF, mass=0.01, m5=1.8)
    got, res = lat.invcg(op, src, C.make_cg_arg(1e-10, 5000))
    print("dwf cg", it_ref, res[0].iterations, tr_ref, res[0].true_res, rel(got, ref), flush=True)
    assert abs(res[0].iterations - it_ref) <= 1
    assert res[0].true_res < 1e-10
    assert rel(got, ref) < 1e-10


@pytest.mark.parametrize("prec", [C.PREC_MIXED_HALF, C.PREC_MIXED_BF16, C.PREC_MIXED_SINGLE])
def test_dwf_mixed_cg(oracle, dwf4, prec):
    lat, u, L, ls = dwf4
    d = oracle.dwf_struct(L, ls, 1.8)
    src = F.gaussian_fermion(lat.vcb * ls, 45)
    ref, it_ref, _ = oracle.dwf_invcg(u, src, 0.01, 1e-12, 5000, d)
    op = C.make_op(C.OP_DWF, mass=0.01, m5=1.8)
    got, res = lat.invcg(op, src, C.make_cg_arg(1e-10, 5000, precision=prec, reliable_every=50))
    print("dwf mixed", prec, it_ref, res[0].iterations, res[0].true_res, res[0].reliable_updates,
          rel(got, ref), flush=True)
    assert res[0].status == 1
    assert res[0].true_res <= 1e-10
    assert rel(got, ref) < 1e-8


@pytest.mark.parametrize("vectors", ["xhalf", "half"])
def test_dwf_mixed_cg_half_sloppy_vectors(oracle, monkeypatch, dwf4, vectors):
    """The sloppy solution (and with "half" also the sloppy residual) stored
    in FP16 with per-site norms: the FP64 residual still meets the stopping
    rule and the solution matches the FP64 oracle."""
    monkeypatch.setenv("CPS_HIP_SLOPPY_VECTORS", vectors)
    lat, u, L, ls = dwf4
    d = oracle.dwf_struct(L, ls, 1.8)
    src = F.gaussian_fermion(lat.vcb * ls, 46)
    ref, it_ref, _ = oracle.dwf_invcg(u, src, 0.01, 1e-12, 5000, d)
    op = C.make_op(C.OP_DWF, mass=0.01, m5=1.8)
    got, res = lat.invcg(op, src, C.make_cg_arg(1e-10, 5000, precision=C.PREC_MIXED_HALF))
    print("dwf mixed sloppy vectors", vectors, it_ref, res[0].iterations, res[0].true_res, flush=True)
    assert res[0].status == 1
    assert res[0].true_res <= 1e-10
    assert rel(got, ref) < 1e-8


def test_dwf_multi_rhs_matches_single(dwf4):
    lat, u, L, ls = dwf4
    op = C.make_op(C.OP_DWF, mass=0.01, m5=1.8)
    n = 24 * lat.vcb * ls
    srcs = [F.gaussian_fermion(lat.vcb * ls, 60 + k) for k in range(3)]
    arg = C.make_cg_arg(1e-10, 5000)
    batched, res = lat.invcg(op, np.concatenate(srcs), arg, nrhs=3)
    for k in range(3):
        single, r1 = lat.invcg(op, srcs[k], arg)
        assert res[k].iterations == r1[0].iterations
        assert rel(batched[k * n:(k + 1) * n], single) < 1e-13


@pytest.mark.parametrize("prec", [C.PREC_DOUBLE, C.PREC_MIXED_HALF])
def test_active_rhs_window_bit_identical(monkeypatch, dwf4, prec):
    """rhs 0 and 3 start from solutions good to 1e-5 and converge first, so
    the batch shrinks to rhs 1..2 for the rest of the solve; the solutions and
    iteration counts are bit-identical to the solve that keeps all 4 rhs."""
    lat, u, L, ls = dwf4
    op = C.make_op(C.OP_DWF, mass=0.01, m5=1.8)
    n = 24 * lat.vcb * ls
    src = np.concatenate([F.gaussian_fermion(lat.vcb * ls, 70 + k) for k in range(4)])
    rough, _ = lat.invcg(op, src, C.make_cg_arg(1e-5, 5000), nrhs=4)
    guess = np.zeros_like(src)
    for k in (0, 3):
        guess[k * n:(k + 1) * n] = rough[k * n:(k + 1) * n]
    arg = C.make_cg_arg(1e-10, 5000, precision=prec)
    out = {}
    for w in ("1", "0"):
        monkeypatch.setenv("CPS_HIP_RHS_WINDOW", w)
        out[w] = lat.invcg(op, src, arg, guess=guess, nrhs=4)
    it = [r.iterations for r in out["1"][1]]
    print("rhs window", prec, it, flush=True)
    assert it == [r.iterations for r in out["0"][1]]
    assert max(it[0], it[3]) < min(it[1], it[2])
    assert np.array_equal(out["1"][0], out["0"][0])
    assert max(r.true_res for r in out["1"][1]) <= 1e-10


def test_active_rhs_window_clover_sets(monkeypatch, clov4):
    """Two kappas, rhs r on kappa r // 2: with rhs 0 and 3 started near their
    solutions the window holds rhs 1 (kappa 0) and rhs 2 (kappa 1), so each
    rhs must keep its own clover set inside the window."""
    lat, u, L = clov4
    lat.clover_build(CLOVER_KAPPAS, CLOVER_CSW)
    op = clover_op(CLOVER_KAPPAS[0])
    n = 24 * lat.vcb
    src = np.concatenate([F.gaussian_fermion(lat.vcb, 80 + k) for k in range(4)])
    rough, _ = lat.invcg(op, src, C.make_cg_arg(1e-5, 5000), nrhs=4)
    guess = np.zeros_like(src)
    for k in (0, 3):
        guess[k * n:(k + 1) * n] = rough[k * n:(k + 1) * n]
    out = {}
    for w in ("1", "0"):
        monkeypatch.setenv("CPS_HIP_RHS_WINDOW", w)
        out[w] = lat.invcg(op, src, C.make_cg_arg(1e-10, 5000), guess=guess, nrhs=4)
    it = [r.iterations for r in out["1"][1]]
    assert it == [r.iterations for r in out["0"][1]]
    assert max(it[0], it[3]) < min(it[1], it[2])
    assert np.array_equal(out["1"][0], out["0"][0])


@pytest.mark.parametrize("kind", ["dwf", "mobius16"])
def test_mixed_lagged_solution_step_bit_identical(monkeypatch, kind):
    """The mixed CG applying x += alpha p one iteration late (in the direction
    update; pending steps flushed before each reliable update) gives the same
    bits and iteration counts as applying it with r, for rhs that converge
    at different iterations."""
    if kind == "dwf":
        L, ls = [4, 4, 4, 8], 8
        op = C.make_op(C.OP_DWF, mass=0.01, m5=1.8)
    else:
        L, ls = [4, 4, 4, 8], 16
        op = C.make_op(C.OP_MOBIUS, mass=0.02, m5=1.8, b=1.5, c=0.5)
    u = wilson_gauge(L, 0.2, 280)
    vcb = int(np.prod(L)) // 2
    src = np.concatenate([F.gaussian_fermion(vcb * ls, 281 + k) * (1.0 + 9.0 * k) for k in range(3)])
    arg = C.make_cg_arg(1e-10, 5000, precision=C.PREC_MIXED_HALF, reliable_every=30)
    monkeypatch.setenv("CPS_HIP_R_IN_HOP", "0")   # (it rounds A p differently)
    out = {}
    for lag in ("1", "0"):
        monkeypatch.setenv("CPS_HIP_LAGGED_X", lag)
        lat = C.Lattice(L, ls, (0, 0, 0, 1))
        lat.import_gauge(u)
        out[lag] = lat.invcg(op, src, arg, nrhs=3)
        lat.close()
    assert [r.iterations for r in out["1"][1]] == [r.iterations for r in out["0"][1]]
    assert [r.reliable_updates for r in out["1"][1]] == [r.reliable_updates for r in out["0"][1]]
    assert np.array_equal(out["1"][0], out["0"][0])
    assert max(r.true_res for r in out["1"][1]) <= 1e-10


@pytest.mark.parametrize("kind,check_every", [("dwf", 0), ("mobius16", 0), ("dwf", 7)])
def test_mixed_paired_solution_steps_bit_identical(monkeypatch, kind, check_every):
    """Iterations in pairs with the direction in two buffers and the solution
    updated every second iteration (both pending steps, in order) give the
    same bits and iteration counts as one step per iteration; an odd check
    interval is rounded up to whole pairs."""
    if kind == "dwf":
        L, ls = [4, 4, 4, 8], 8
        op = C.make_op(C.OP_DWF, mass=0.01, m5=1.8)
    else:
        L, ls = [4, 4, 4, 8], 16
        op = C.make_op(C.OP_MOBIUS, mass=0.02, m5=1.8, b=1.5, c=0.5)
    u = wilson_gauge(L, 0.2, 310)
    vcb = int(np.prod(L)) // 2
    src = np.concatenate([F.gaussian_fermion(vcb * ls, 311 + k) * (1.0 + 5.0 * k) for k in range(3)])
    arg = C.make_cg_arg(1e-10, 5000, precision=C.PREC_MIXED_HALF, reliable_every=40)
    arg.check_every = check_every
    out = {}
    for paired in ("1", "0"):
        monkeypatch.setenv("CPS_HIP_PAIRED_X", paired)
        lat = C.Lattice(L, ls, (0, 0, 0, 1))
        lat.import_gauge(u)
        out[paired] = lat.invcg(op, src, arg, nrhs=3)
        lat.close()
    it = [r.iterations for r in out["1"][1]]
    print("paired", kind, check_every, it, [r.iterations for r in out["0"][1]], flush=True)
    if check_every == 0:
        assert it == [r.iterations for r in out["0"][1]]
        assert np.array_equal(out["1"][0], out["0"][0])
    assert max(r.true_res for r in out["1"][1]) <= 1e-10
    assert all(r.status == 1 for r in out["1"][1])


def test_mixed_s5_in_direction_update_bit_identical(monkeypatch):
    """Moebius, Ls 16, FP16 directions: the direction update writing the
    first step of M (b p + c P p, slots s-1, s+1 from the neighbouring lanes)
    gives the same solve, bit for bit, as the operator's own k_s5 pass."""
    L, ls = [4, 4, 4, 8], 16
    op = C.make_op(C.OP_MOBIUS, mass=0.02, m5=1.8, b=1.5, c=0.5)
    u = wilson_gauge(L, 0.2, 290)
    vcb = int(np.prod(L)) // 2
    src = np.concatenate([F.gaussian_fermion(vcb * ls, 291 + k) for k in range(2)])
    arg = C.make_cg_arg(1e-10, 5000, precision=C.PREC_MIXED_HALF)
    monkeypatch.setenv("CPS_HIP_R_IN_HOP", "0")   # (it rounds A p differently)
    out = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("CPS_HIP_S5_XPAY", fused)
        lat = C.Lattice(L, ls, (0, 0, 0, 1))
        lat.import_gauge(u)
        out[fused] = lat.invcg(op, src, arg, nrhs=2)
        lat.close()
    print("s5 in xpay", [r.iterations for r in out["1"][1]], [r.iterations for r in out["0"][1]], flush=True)
    assert [r.iterations for r in out["1"][1]] == [r.iterations for r in out["0"][1]]
    assert np.array_equal(out["1"][0], out["0"][0])
    assert max(r.true_res for r in out["1"][1]) <= 1e-10


@pytest.mark.parametrize("nrhs", [1, 3])
def test_mixed_residual_step_in_hop(oracle, monkeypatch, nrhs):
    """Moebius, Ls 16, half precision: r -= alpha M^dag M p and |r|^2 in the
    epilogue of M^dag's last hop (alpha from the first half), A p never
    stored.  It rounds A p differently (FP32 instead of FP16), so the check
    is the FP64 stopping rule, the oracle's solution, and iteration counts
    within 5% of the separate residual pass."""
    L, ls = [4, 4, 4, 8], 16
    op = C.make_op(C.OP_MOBIUS, mass=0.02, m5=1.8, b=1.5, c=0.5)
    u = wilson_gauge(L, 0.2, 300)
    vcb = int(np.prod(L)) // 2
    n = 24 * vcb * ls
    src = np.concatenate([F.gaussian_fermion(vcb * ls, 301 + k) * (1.0 + 4.0 * k) for k in range(nrhs)])
    arg = C.make_cg_arg(1e-10, 5000, precision=C.PREC_MIXED_HALF, reliable_every=40)
    out = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("CPS_HIP_R_IN_HOP", fused)
        lat = C.Lattice(L, ls, (0, 0, 0, 1))
        lat.import_gauge(u)
        out[fused] = lat.invcg(op, src, arg, nrhs=nrhs)
        lat.close()
    it1 = [r.iterations for r in out["1"][1]]
    it0 = [r.iterations for r in out["0"][1]]
    print("r in hop", it1, it0, flush=True)
    assert max(r.true_res for r in out["1"][1]) <= 1e-10
    assert all(r.status == 1 for r in out["1"][1])
    for a, b_ in zip(it1, it0):
        assert abs(a - b_) <= 0.05 * b_ + 2
    m = oracle.mobius_struct(L, ls, 1.8, 1.5, 0.5)
    for k in range(nrhs):
        ref, _, _ = oracle.mobius_invcg(u, src[k * n:(k + 1) * n], 0.02, 1e-12, 5000, m)
        assert rel(out["1"][0][k * n:(k + 1) * n], ref) < 1e-8


def test_invcg_device_matches_host(dwf4):
    """cps_hip_invcg_device (vectors resident in HBM) gives the host entry
    point's result bit for bit, for FP64 and mixed precision, two rhs."""
    lat, u, L, ls = dwf4
    op = C.make_op(C.OP_DWF, mass=0.01, m5=1.8)
    src = np.concatenate([F.gaussian_fermion(lat.vcb * ls, 90 + k) for k in range(2)])
    for prec in (C.PREC_DOUBLE, C.PREC_MIXED_HALF):
        arg = C.make_cg_arg(1e-10, 5000, precision=prec)
        host, rh = lat.invcg(op, src, arg, nrhs=2)
        s_dev = C.DeviceBuffer.from_array(src)
        x_dev = C.DeviceBuffer.from_array(np.zeros_like(src))
        rd = lat.invcg_device(op, x_dev, s_dev, arg, nrhs=2)
        got = x_dev.download()
        s_dev.free()
        x_dev.free()
        assert [r.iterations for r in rd] == [r.iterations for r in rh]
        assert np.array_equal(got, host)


@pytest.fixture(scope="module")
def mob4():
    L = [4, 4, 4, 4]
    ls = 8
    lat = C.Lattice(L, ls=ls, bc=(0, 0, 0, 1))
    u = wilson_gauge(L, 0.2, 77)
    lat.import_gauge(u)
    yield lat, u, L, ls
    lat.close()


def test_mobius_m_and_mdagm(oracle, mob4):
    lat, u, L, ls = mob4
    m = oracle.mobius_struct(L, ls, 1.8, 1.5, 0.5)
    op = C.make_op(C.OP_MOBIUS, mass=0.005, m5=1.8, b=1.5, c=0.5)
    inp = F.gaussian_fermion(lat.vcb * ls, 78)
    assert rel(lat.matpc(op, inp, 0), oracle.mobius_m(u, inp, 0.005, m)) < 1e-13
    assert rel(lat.matpc(op, inp, 1), oracle.mobius_mdag(u, inp, 0.005, m)) < 1e-13
    ref, dp_ref = oracle.mobius_mdagm(u, inp, 0.005, m)
    got, dp = lat.mdagm(op, inp)
    assert rel(got, ref) < 1e-13
    assert abs(dp[0] - dp_ref) / dp_ref < 1e-13


@pytest.mark.parametrize("ls,fused", [(4, "1"), (16, "1"), (16, "0"), (12, "1")])
def test_mobius_mdag_s5_fusion(oracle, monkeypatch, ls, fused):
    """M^dag with its last fifth-dimension step fused into the hop epilogue
    (slots s-1, s+1 from neighbouring lanes; Ls 4 and 16), kept as its own
    pass (CPS_HIP_S5_FUSED=0, and Ls 12 which does not divide a wavefront),
    against the oracle; the mixed-precision M^dag M agrees with FP64 to the
    FP16 storage precision either way."""
    monkeypatch.setenv("CPS_HIP_S5_FUSED", fused)
    L = [4, 4, 4, 4]
    u = wilson_gauge(L, 0.2, 240)
    lat = C.Lattice(L, ls, (0, 0, 0, 1))
    lat.import_gauge(u)
    m = oracle.mobius_struct(L, ls, 1.8, 1.5, 0.5)
    op = C.make_op(C.OP_MOBIUS, mass=0.02, m5=1.8, b=1.5, c=0.5)
    inp = np.concatenate([F.gaussian_fermion(lat.vcb * ls, 241 + k) for k in range(2)])
    n = 24 * lat.vcb * ls
    got = lat.matpc(op, inp, 1, nrhs=2)
    for k in range(2):
        assert rel(got[k * n:(k + 1) * n], oracle.mobius_mdag(u, inp[k * n:(k + 1) * n], 0.02, m)) < 1e-13
    d64, _ = lat.mdagm(op, inp, nrhs=2)
    dh, _ = lat.mdagm_prec(op, inp, C.PREC_MIXED_HALF, nrhs=2)
    assert rel(dh, d64) < 5e-3
    lat.close()


@pytest.mark.parametrize("nrhs", [1, 3])
def test_mobius_m5_fused_in_hop(oracle, monkeypatch, nrhs):
    """Half precision, Ls = 16: M5^-1 and its s-term run across the 16 lanes
    of a site in the first hop's epilogue.  The fused M^dag M agrees with
    FP64 at least as well as the separate M5^-1 pass does (it skips one
    FP16 rounding of the intermediate), and a mixed solve with it reaches the
    FP64 stopping rule and the oracle's solution."""
    L, ls = [4, 4, 4, 8], 16
    u = wilson_gauge(L, 0.2, 260)
    op = C.make_op(C.OP_MOBIUS, mass=0.05, m5=1.8, b=1.5, c=0.5)
    vcb = int(np.prod(L)) // 2
    src = np.concatenate([F.gaussian_fermion(vcb * ls, 261 + k) for k in range(nrhs)])
    out = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("CPS_HIP_M5_FUSED", fused)
        lat = C.Lattice(L, ls, (0, 0, 0, 1))
        lat.import_gauge(u)
        out[fused] = lat.mdagm_prec(op, src, C.PREC_MIXED_HALF, nrhs=nrhs)[0]
        d64 = lat.mdagm(op, src, nrhs=nrhs)[0]
        if fused == "1":
            x, res = lat.invcg(op, src, C.make_cg_arg(1e-10, 5000, precision=C.PREC_MIXED_HALF), nrhs=nrhs)
        lat.close()
    e1, e0 = rel(out["1"], d64), rel(out["0"], d64)
    print("m5 fused", nrhs, e1, e0, [r.iterations for r in res], flush=True)
    assert e1 < 5e-3 and e0 < 5e-3
    assert e1 <= 1.5 * e0
    assert max(r.true_res for r in res) <= 1e-10
    m = oracle.mobius_struct(L, ls, 1.8, 1.5, 0.5)
    n = 24 * vcb * ls
    ref, _, _ = oracle.mobius_invcg(u, src[:n], 0.05, 1e-12, 5000, m)
    assert rel(x[:n], ref) < 1e-8


def test_mobius_s5_row_exchange_bit_identical(monkeypatch):
    """Half precision, Ls = 16: the s-term epilogue taking the slots s-1, s+1
    of X and of the hop from the neighbouring lanes (DPP row rotations) gives
    the same bits as loading them and exchanging through lane permutes."""
    L, ls, nrhs = [4, 4, 4, 8], 16, 2
    u = wilson_gauge(L, 0.2, 270)
    op = C.make_op(C.OP_MOBIUS, mass=0.05, m5=1.8, b=1.5, c=0.5)
    vcb = int(np.prod(L)) // 2
    src = np.concatenate([F.gaussian_fermion(vcb * ls, 271 + k) for k in range(nrhs)])
    out = {}
    for row in ("1", "0"):
        monkeypatch.setenv("CPS_HIP_S5X_ROW", row)
        lat = C.Lattice(L, ls, (0, 0, 0, 1))
        lat.import_gauge(u)
        out[row] = lat.mdagm_prec(op, src, C.PREC_MIXED_HALF, nrhs=nrhs)
        lat.close()
    assert np.array_equal(out["1"][0], out["0"][0])
    assert np.array_equal(out["1"][1], out["0"][1])


def test_mobius_cg(oracle, mob4):
    lat, u, L, ls = mob4
    m = oracle.mobius_struct(L, ls, 1.8, 1.5, 0.5)
    op = C.make_op(C.OP_MOBIUS, mass=0.005, m5=1.8, b=1.5, c=0.5)
    src = F.gaussian_fermion(lat.vcb * ls, 79)
    ref, it_ref, _ = oracle.mobius_invcg(u, src, 0.005, 1e-10, 5000, m)
    got, res = lat.invcg(op, src, C.make_cg_arg(1e-10, 5000))
    print("mobius cg", it_ref, res[0].iterations, res[0].true_res, rel(got, ref), flush=True)
    assert abs(res[0].iterations - it_ref) <= 1
    assert rel(got, ref) < 1e-10
    got2, res2 = lat.invcg(op, src, C.make_cg_arg(1e-10, 5000, precision=C.PREC_MIXED_HALF))
    assert res2[0].true_res <= 1e-10


def test_wilson_matinv(oracle):
    lat = C.Lattice(L4, 1, (0, 0, 0, 1))
    u = wilson_gauge(L4, 0.3, 13)
    lat.import_gauge(u)
    b = F.gaussian_fermion(2 * lat.vcb, 14)
    mass = 0.1
    ref, it_ref, _ = oracle.wilson_matinv(u, b, mass, 1e-12, 500, L4)
    op = C.make_op(C.OP_WILSON, mass=mass)
    got, res = lat.matinv(op, b, C.make_cg_arg(1e-12, 500))
    assert abs(res[0].iterations - it_ref) <= 1
    assert rel(got, ref) < 1e-10
    lat.close()


def test_dwf_matinv(oracle, dwf4):
    lat, u, L, ls = dwf4
    d = oracle.dwf_struct(L, ls, 1.8)
    b = F.gaussian_fermion(2 * lat.vcb * ls, 15)
    ref, it_ref, _ = oracle.dwf_matinv(u, b, 0.05, 1e-12, 5000, d)
    op = C.make_op(C.OP_DWF, mass=0.05, m5=1.8)
    got, res = lat.matinv(op, b, C.make_cg_arg(1e-12, 5000))
    assert abs(res[0].iterations - it_ref) <= 1
    assert rel(got, ref) < 1e-10


SLOPPY_TOL = {C.PREC_MIXED_SINGLE: 1e-5, C.PREC_MIXED_HALF: 5e-3, C.PREC_MIXED_BF16: 3e-2}


@pytest.mark.parametrize("prec", [C.PREC_MIXED_SINGLE, C.PREC_MIXED_HALF, C.PREC_MIXED_BF16])
@pytest.mark.parametrize("kind", ["wilson", "dwf", "mobius", "clover"])
def test_sloppy_operator(oracle, prec, kind):
    """The reduced-precision MdagM the mixed CG iterates with, against the FP64
    oracle; the tolerance is the storage rounding (FP32 ~1e-7, FP16 spinors
    2^-11 of the site maximum, BF16 links 2^-8) compounded over four hops."""
    L = [4, 4, 4, 8]
    ls = 1 if kind == "wilson" else 8
    lat = C.Lattice(L, ls, (0, 0, 0, 1))
    u = wilson_gauge(L, 0.2, 90)
    lat.import_gauge(u)
    inp = F.gaussian_fermion(lat.vcb * ls, 91)
    if kind == "wilson":
        op = C.make_op(C.OP_WILSON, mass=0.1)
        ref, dref = oracle.wilson_mdagm(u, inp, 1.0 / (2 * 4.1), L)
    elif kind == "dwf":
        op = C.make_op(C.OP_DWF, mass=0.01, m5=1.8)
        ref, dref = oracle.dwf_mdagm(u, inp, 0.01, oracle.dwf_struct(L, ls, 1.8))
    elif kind == "mobius":
        op = C.make_op(C.OP_MOBIUS, mass=0.01, m5=1.8, b=1.5, c=0.5)
        ref, dref = oracle.mobius_mdagm(u, inp, 0.01, oracle.mobius_struct(L, ls, 1.8, 1.5, 0.5))
    else:
        kappa, csw = 0.13, 1.0
        op = C.make_op(C.OP_CLOVER, mass=1.0 / (2 * kappa) - 4.0, csw=csw)
        lat.clover_build([kappa], csw)
        A_oo = oracle.clover_mat_chkb(u, 1, 0, kappa, csw, L)
        A_ee_inv = oracle.clover_mat_chkb(u, 0, 1, kappa, csw, L)
        ref, dref = oracle.clover_mdagm(u, A_oo, A_ee_inv, inp, kappa, L)
    got, d = lat.mdagm_prec(op, inp, prec)
    print("sloppy", kind, prec, rel(got, ref), abs(d[0] - dref) / dref, flush=True)
    assert rel(got, ref) < SLOPPY_TOL[prec]
    assert abs(d[0] - dref) / dref < SLOPPY_TOL[prec]
    lat.close()


# ---------------------------------------------------------------------------
# clover (DiracOpClover, d_op_clover.C:107-352)

CLOVER_KAPPAS = [0.1360, 0.1375]
CLOVER_CSW = 1.0


@pytest.fixture(scope="module")
def clov4():
    L = [4, 4, 4, 8]
    lat = C.Lattice(L, 1, (0, 0, 0, 1))
    u = wilson_gauge(L, 0.15, 41)
    lat.import_gauge(u)
    yield lat, u, L
    lat.close()


def clover_op(kappa, csw=CLOVER_CSW):
    return C.make_op(C.OP_CLOVER, mass=1.0 / (2 * kappa) - 4.0, csw=csw)


def test_clover_build_matches_oracle(oracle, clov4):
    """A_oo and A_ee^-1 built on the device (two kappas at once) against
    CloverMatChkb of the oracle; the inverse differs only by the rounding of
    the factorisation."""
    lat, u, L = clov4
    lat.clover_build(CLOVER_KAPPAS, CLOVER_CSW)
    for k, kappa in enumerate(CLOVER_KAPPAS):
        for cb in (0, 1):
            ref = oracle.clover_mat_chkb(u, cb, 1 if cb == 0 else 0, kappa, CLOVER_CSW, L)
            got = lat.clover_export(k, cb)
            print("clover build", k, cb, rel(got, ref), flush=True)
            assert rel(got, ref) < 1e-13


def test_clover_matpc_and_mdagm(oracle, clov4):
    lat, u, L = clov4
    kappa = 0.1365
    lat.clover_build([kappa], CLOVER_CSW)
    op = clover_op(kappa)
    A_oo = oracle.clover_mat_chkb(u, 1, 0, kappa, CLOVER_CSW, L)
    A_ee_inv = oracle.clover_mat_chkb(u, 0, 1, kappa, CLOVER_CSW, L)
    inp = F.gaussian_fermion(lat.vcb, 42)
    for dag in (0, 1):
        ref = oracle.clover_matpc(u, A_oo, A_ee_inv, inp, kappa, dag, L)
        assert rel(lat.matpc(op, inp, dag), ref) < 1e-13
    ref, dp_ref = oracle.clover_mdagm(u, A_oo, A_ee_inv, inp, kappa, L)
    got, dp = lat.mdagm(op, inp)
    assert rel(got, ref) < 1e-13
    assert abs(dp[0] - dp_ref) / dp_ref < 1e-13


def test_clover_cg(oracle, clov4):
    lat, u, L = clov4
    kappa = 0.1365
    lat.clover_build([kappa], CLOVER_CSW)
    op = clover_op(kappa)
    A_oo = oracle.clover_mat_chkb(u, 1, 0, kappa, CLOVER_CSW, L)
    A_ee_inv = oracle.clover_mat_chkb(u, 0, 1, kappa, CLOVER_CSW, L)
    src = F.gaussian_fermion(lat.vcb, 43)
    ref, it_ref, _ = oracle.clover_invcg(u, A_oo, A_ee_inv, src, kappa, 1e-10, 5000, L)
    got, res = lat.invcg(op, src, C.make_cg_arg(1e-10, 5000))
    print("clover cg", it_ref, res[0].iterations, res[0].true_res, rel(got, ref), flush=True)
    assert abs(res[0].iterations - it_ref) <= 1
    assert res[0].true_res < 1e-10
    assert rel(got, ref) < 1e-10
    for prec in (C.PREC_MIXED_SINGLE, C.PREC_MIXED_HALF):
        got2, res2 = lat.invcg(op, src, C.make_cg_arg(1e-10, 5000, precision=prec))
        print("clover mixed", prec, res2[0].iterations, res2[0].true_res, rel(got2, ref), flush=True)
        assert res2[0].true_res <= 1e-10
        assert rel(got2, ref) < 1e-8


def test_clover_multi_kappa_batch(clov4):
    """Sources for two kappas in one batched solve (rhs r uses kappa r // 2)
    give exactly the single-kappa solves."""
    lat, u, L = clov4
    srcs = [F.gaussian_fermion(lat.vcb, 50 + k) for k in range(4)]
    arg = C.make_cg_arg(1e-10, 5000)
    lat.clover_build(CLOVER_KAPPAS, CLOVER_CSW)
    batched, res = lat.invcg(clover_op(CLOVER_KAPPAS[0]), np.concatenate(srcs), arg, nrhs=4)
    n = 24 * lat.vcb
    for r in range(4):
        kappa = CLOVER_KAPPAS[r // 2]
        lat.clover_build([kappa], CLOVER_CSW)
        single, r1 = lat.invcg(clover_op(kappa), srcs[r], arg)
        assert res[r].iterations == r1[0].iterations
        assert np.array_equal(batched[r * n:(r + 1) * n], single)


def test_clover_rejects_mismatched_kappa(clov4):
    lat, u, L = clov4
    lat.clover_build([0.1365], CLOVER_CSW)
    with pytest.raises(C.CpsHipError):
        lat.mdagm(clover_op(0.1370), F.gaussian_fermion(lat.vcb, 1))


# ---------------------------------------------------------------------------
# hipGraph replay of the CG iteration


@pytest.mark.parametrize("prec", [C.PREC_DOUBLE, C.PREC_MIXED_HALF])
@pytest.mark.parametrize("kind", ["dwf", "mobius", "clover"])
def test_graph_replay_matches_eager(prec, kind):
    """The captured iteration, replayed many times with host polls (and, for
    mixed precision, eager reliable updates) between replays, gives exactly
    the eager solve."""
    L = [4, 4, 4, 8]
    ls = 1 if kind == "clover" else 8
    lat = C.Lattice(L, ls, (0, 0, 0, 1))
    lat.import_gauge(wilson_gauge(L, 0.2, 120))
    if kind == "dwf":
        op = C.make_op(C.OP_DWF, mass=0.02, m5=1.8)
    elif kind == "mobius":
        op = C.make_op(C.OP_MOBIUS, mass=0.02, m5=1.8, b=1.5, c=0.5)
    else:
        lat.clover_build([0.13], 1.0)
        op = C.make_op(C.OP_CLOVER, mass=1.0 / 0.26 - 4.0, csw=1.0)
    src = np.concatenate([F.gaussian_fermion(lat.vcb * ls, 121 + k) for k in range(2)])
    out = {}
    for g in (0, 1):
        arg = C.make_cg_arg(1e-10, 3000, precision=prec, reliable_every=40, check_every=4, use_graph=g)
        out[g] = lat.invcg(op, src, arg, nrhs=2)
    assert [r.iterations for r in out[0][1]] == [r.iterations for r in out[1][1]]
    assert np.array_equal(out[0][0], out[1][0])
    assert max(r.true_res for r in out[1][1]) <= 1e-10
    lat.close()


# ---------------------------------------------------------------------------
# resident solve service (continuous batching)


def test_service_results_same_alone_and_in_mix(oracle):
    """Requests on two resident lattices with several operators and masses,
    nrhs up to 5 with 4 slots per engine (slots are refilled mid-request):
    every solution is bit-identical whether the request ran alone or in the
    mix, meets the residual, and matches the FP64 oracle CG."""
    LW, LD, ls = [8, 8, 8, 8], [4, 4, 4, 8], 8
    uw = wilson_gauge(LW, 0.2, 140)
    ud = wilson_gauge(LD, 0.2, 141)
    svc = C.Service(8 << 30, max_slots=4)
    gw = svc.add_gauge(LW, 1, uw)
    gd = svc.add_gauge(LD, ls, ud)
    vw, vd = 24 * int(np.prod(LW)) // 2, 24 * int(np.prod(LD)) // 2 * ls
    reqs = []
    for k, (g, mass, nrhs) in enumerate([(gw, 0.1, 1), (gd, 0.02, 5), (gw, 0.2, 3), (gd, 0.05, 2),
                                          (gd, 0.02, 2), (gw, 0.1, 4)]):
        if g == gw:
            op = C.make_op(C.OP_WILSON, mass=mass)
            src = np.concatenate([F.gaussian_fermion(vw // 24, 150 + 10 * k + r) for r in range(nrhs)])
        else:
            op = C.make_op(C.OP_DWF, mass=mass, m5=1.8)
            src = np.concatenate([F.gaussian_fermion(vd // 24, 150 + 10 * k + r) for r in range(nrhs)])
        reqs.append((g, op, src, nrhs, mass))
    alone = []
    for g, op, src, nrhs, _ in reqs:
        rid = svc.submit(g, op, src, nrhs)
        alone.append(svc.wait(rid))
    ids = [svc.submit(g, op, src, nrhs) for g, op, src, nrhs, _ in reversed(reqs)]
    mix = [svc.wait(rid) for rid in ids][::-1]
    st = svc.stats()
    print("service stats", st, flush=True)
    for (g, op, src, nrhs, mass), (sa, ra, la), (sm, rm, lm) in zip(reqs, alone, mix):
        assert np.array_equal(sa, sm)
        assert [r.iterations for r in ra] == [r.iterations for r in rm]
        assert max(r.true_res for r in rm) <= 1e-10
        n = len(src) // nrhs
        for r in range(nrhs):
            if g == gw:
                ref, _, _ = oracle.wilson_invcg(uw, src[r * n:(r + 1) * n], mass, 1e-10, 5000, LW)
            else:
                ref, _, _ = oracle.dwf_invcg(ud, src[r * n:(r + 1) * n], mass, 1e-10, 5000,
                                             oracle.dwf_struct(LD, ls, 1.8))
            assert rel(sm[r * n:(r + 1) * n], ref) < 1e-9
    assert st["requests_done"] == 2 * len(reqs)
    assert st["last_error"] == 0
    svc.close()


def test_service_request_beyond_budget_fails_cleanly():
    """Requests whose engine can never fit the HBM budget fail with
    CPS_HIP_ERR_NOMEM through the admission error path (each request
    finished once, none read after wait() has freed it), several of them
    in a row, and the service shuts down normally afterwards."""
    L = [4, 4, 4, 8]
    u = wilson_gauge(L, 0.2, 175)
    svc = C.Service(1 << 20, max_slots=2)   # room for the gauge field, not for an engine
    g = svc.add_gauge(L, 1, u)
    op = C.make_op(C.OP_WILSON, mass=0.1)
    vcb = int(np.prod(L)) // 2
    rids = [svc.submit(g, op, np.concatenate([F.gaussian_fermion(vcb, 176 + k) for k in range(nrhs)]), nrhs)
            for k, nrhs in enumerate((1, 3, 2))]
    for rid in rids:
        with pytest.raises(C.CpsHipError) as err:
            svc.wait(rid)
        assert err.value.code == -5
    st = svc.stats()
    assert st["requests_done"] == 3 and st["engines_created"] == 0
    svc.close()


@pytest.mark.parametrize("prec", [C.PREC_MIXED_HALF, C.PREC_MIXED_SINGLE])
def test_service_mixed_precision_shared_masses(oracle, prec):
    """Mixed-precision requests (sloppy CG with FP64 reliable updates per
    slot): Wilson and DWF requests of different masses share one engine per
    lattice; every rhs meets the FP64 true residual, matches the FP64 oracle
    CG, and agrees alone and in the mix within the solver tolerance (reliable
    updates of an engine fall on the same iterations for all its slots, so
    the mix changes when a slot's updates happen, not what it converges
    to)."""
    LW, LD, ls = [8, 8, 8, 8], [4, 4, 4, 8], 8
    uw = wilson_gauge(LW, 0.2, 160)
    ud = wilson_gauge(LD, 0.2, 161)
    svc = C.Service(8 << 30, max_slots=4)
    gw = svc.add_gauge(LW, 1, uw)
    gd = svc.add_gauge(LD, ls, ud)
    vw, vd = 24 * int(np.prod(LW)) // 2, 24 * int(np.prod(LD)) // 2 * ls
    reqs = []
    for k, (g, mass, nrhs) in enumerate([(gw, 0.1, 2), (gd, 0.02, 3), (gw, 0.2, 3), (gd, 0.05, 2), (gw, 0.05, 1)]):
        kind = C.OP_WILSON if g == gw else C.OP_DWF
        op = C.make_op(kind, mass=mass, m5=1.8)
        nv = vw if g == gw else vd
        src = np.concatenate([F.gaussian_fermion(nv // 24, 170 + 10 * k + r) for r in range(nrhs)])
        reqs.append((g, op, src, nrhs, mass))
    alone = []
    for g, op, src, nrhs, _ in reqs:
        alone.append(svc.wait(svc.submit(g, op, src, nrhs, precision=prec, reliable_every=50)))
    created_alone = svc.stats()["engines_created"]
    ids = [svc.submit(g, op, src, nrhs, precision=prec, reliable_every=50) for g, op, src, nrhs, _ in reqs]
    mix = [svc.wait(rid) for rid in ids]
    st = svc.stats()
    print("service mixed stats", prec, st, flush=True)
    for (g, op, src, nrhs, mass), (sa, ra, la), (sm, rm, lm) in zip(reqs, alone, mix):
        print("service mixed request", g, mass, [r.iterations for r in rm], [r.reliable_updates for r in rm],
              max(r.true_res for r in rm), rel(sm, sa), flush=True)
        assert max(r.true_res for r in ra) <= 1e-10 and max(r.true_res for r in rm) <= 1e-10
        assert all(r.status == 1 for r in rm)
        assert rel(sm, sa) < 1e-8
        n = len(src) // nrhs
        for r in range(nrhs):
            if g == gw:
                ref, _, _ = oracle.wilson_invcg(uw, src[r * n:(r + 1) * n], mass, 1e-12, 5000, LW)
            else:
                ref, _, _ = oracle.dwf_invcg(ud, src[r * n:(r + 1) * n], mass, 1e-12, 5000,
                                             oracle.dwf_struct(LD, ls, 1.8))
            assert rel(sm[r * n:(r + 1) * n], ref) < 1e-8
    # one engine per lattice serves all of its masses
    assert created_alone <= 2 and st["engines_created"] <= 2
    assert st["last_error"] == 0
    svc.close()


def test_half_storage_subnormal_site():
    """A site whose values are FP32 subnormals (as far from a point source on
    a large lattice) is stored as zero in FP16 storage, not as inf/NaN: the
    sloppy operator stays finite and within its tolerance of FP64."""
    L, ls = [4, 4, 4, 8], 8
    lat = C.Lattice(L, ls, (0, 0, 0, 1))
    lat.import_gauge(wilson_gauge(L, 0.2, 160))
    op = C.make_op(C.OP_DWF, mass=0.02, m5=1.8)
    inp = F.gaussian_fermion(lat.vcb * ls, 161)
    inp[24 * 5:24 * 6] *= 1e-40          # one 5-D site of subnormal magnitude
    inp[24 * 77:24 * 78] *= 3e-39
    ref, dref = lat.mdagm(op, inp)
    for prec in (C.PREC_MIXED_HALF, C.PREC_MIXED_BF16):
        got, d = lat.mdagm_prec(op, inp, prec)
        assert np.isfinite(got).all()
        assert rel(got, ref) < SLOPPY_TOL[prec]
    lat.close()


# ---------------------------------------------------------------------------
# MatInv of the Moebius and clover operators against the full textbook
# operators (tests/textbook.py)


def _full_setup(L, seed):
    from tests import textbook as T
    from cps_public_amd import layout as LY
    u = F.gauge_field(L, 0.2, seed)
    H = T.hopping(T.links(u, L), L, antiperiodic=(0, 0, 0, 1))
    uw = LY.gauge_canon_to_wilson(LY.bond_cond(u, L, [0, 0, 0, 1]), L)
    return u, H, uw


def test_mobius_matinv_solves_full_operator():
    """DiracOpMobius::MatInv (ZMOB_PC_ORIG): out solves
    [(b D_W + 1) + (c D_W - 1) P] out = in, vectors in DWF_4D_EOPREC_EE order."""
    from tests import textbook as T
    from tests.test_oracle_pins import prec4_from_canon, canon_from_prec4, rand_cplx
    L, ls, mass, m5, b, c = [4, 4, 4, 4], 4, 0.05, 1.8, 1.5, 0.5
    u, H, uw = _full_setup(L, 170)
    D, _ = T.mobius_operator(H, ls, mass, m5, b, c)
    kb, _ = T.mobius_kappas(m5, b, c)
    V = int(np.prod(L))
    rng = np.random.default_rng(171)
    bc5 = rand_cplx(rng, ls, V, 12)
    vin = np.concatenate([prec4_from_canon(bc5, L, ls, 0), prec4_from_canon(bc5, L, ls, 1)])
    lat = C.Lattice(L, ls, (0, 0, 0, 1))
    lat.import_gauge(uw)
    op = C.make_op(C.OP_MOBIUS, mass=mass, m5=m5, b=b, c=c)
    for prec in (C.PREC_DOUBLE, C.PREC_MIXED_HALF):
        out, res = lat.matinv(op, vin, C.make_cg_arg(1e-12, 5000, precision=prec))
        h = len(out) // 2
        x = (canon_from_prec4(out[:h], L, ls, 0) + canon_from_prec4(out[h:], L, ls, 1)).reshape(-1)
        r = D @ x / (2 * kb) - bc5.reshape(-1)
        print("mobius matinv", prec, res[0].iterations, np.linalg.norm(r) / np.linalg.norm(bc5), flush=True)
        assert np.linalg.norm(r) / np.linalg.norm(bc5) < 1e-8
    lat.close()


def test_clover_matinv_solves_full_operator():
    """DiracOpClover::MatInv: out solves (A - kappa H) out = in, [odd][even]."""
    from tests import textbook as T
    from tests.test_oracle_pins import full_from_canon, canon_from_full, rand_cplx
    from cps_public_amd import layout as LY
    L, kappa, csw = [4, 4, 4, 8], 0.13, 1.2
    u, H, uw = _full_setup(L, 180)
    Ub = T.links(LY.bond_cond(u, L, [0, 0, 0, 1]), L)
    D = (T.block_diag_sparse(T.clover_blocks(Ub, L, kappa, csw)) - kappa * H).tocsr()
    V = int(np.prod(L))
    rng = np.random.default_rng(181)
    bvec = rand_cplx(rng, V, 12)
    lat = C.Lattice(L, 1, (0, 0, 0, 1))
    lat.import_gauge(uw)
    lat.clover_build([kappa], csw)
    op = C.make_op(C.OP_CLOVER, mass=1.0 / (2 * kappa) - 4.0, csw=csw)
    for prec in (C.PREC_DOUBLE, C.PREC_MIXED_SINGLE):
        out, res = lat.matinv(op, full_from_canon(bvec, L), C.make_cg_arg(1e-12, 5000, precision=prec))
        x = canon_from_full(out, L).reshape(-1)
        r = D @ x - bvec.reshape(-1)
        print("clover matinv", prec, res[0].iterations, np.linalg.norm(r) / np.linalg.norm(bvec), flush=True)
        assert np.linalg.norm(r) / np.linalg.norm(bvec) < 1e-8
    lat.close()


# ---------------------------------------------------------------------------
# batched point-source propagator (QPropW::Run's 12 spin-colour solves)


def _four2five(b4, ls):
    """Ffour2five(src_5d, src_4d, 0, Ls-1): spins 0,1 at s = 0, spins 2,3 at
    s = Ls-1 (canonical (V, 12) -> (Ls, V, 12))."""
    b5 = np.zeros((ls,) + b4.shape, dtype=complex)
    b5[0, :, :6] = b4[:, :6]
    b5[ls - 1, :, 6:] = b4[:, 6:]
    return b5


def _five2four(x5):
    """Ffive2four(sol_4d, sol_5d, Ls-1, 0): spins 0,1 from s = Ls-1, spins
    2,3 from s = 0."""
    x4 = np.zeros(x5.shape[1:], dtype=complex)
    x4[:, :6] = x5[-1, :, :6]
    x4[:, 6:] = x5[0, :, 6:]
    return x4


@pytest.mark.parametrize("kind,x0,prec", [
    ("wilson", (0, 0, 0, 0), C.PREC_DOUBLE), ("wilson", (1, 2, 3, 1), C.PREC_DOUBLE),
    ("clover", (1, 2, 3, 1), C.PREC_DOUBLE), ("dwf", (1, 2, 3, 1), C.PREC_DOUBLE),
    ("dwf", (2, 0, 1, 3), C.PREC_MIXED_HALF), ("mobius", (1, 2, 3, 1), C.PREC_DOUBLE),
    ("mobius", (0, 1, 1, 0), C.PREC_MIXED_HALF)])
def test_qpropw_point_propagator(kind, x0, prec):
    """cps_hip_qpropw_point against the full textbook operators: column
    (spin, colour) of the propagator is the 4-D projection of D^-1 applied
    to the point source (D_minus first for Moebius), for all 12 columns."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    from tests import textbook as T
    from cps_public_amd import layout as LY
    L = [4, 4, 4, 4]
    V = int(np.prod(L))
    ls = 1 if kind in ("wilson", "clover") else 4
    u, H, uw = _full_setup(L, 190)
    mass, m5, b, c = 0.1, 1.8, 1.5, 0.5
    lat = C.Lattice(L, ls, (0, 0, 0, 1))
    lat.import_gauge(uw)
    if kind == "wilson":
        op = C.make_op(C.OP_WILSON, mass=mass)
        D = sp.identity(12 * V) - (1.0 / (2 * (mass + 4))) * H
    elif kind == "clover":
        kappa, csw = 0.12, 1.1
        lat.clover_build([kappa], csw)
        op = C.make_op(C.OP_CLOVER, mass=1.0 / (2 * kappa) - 4.0, csw=csw)
        Ub = T.links(LY.bond_cond(u, L, [0, 0, 0, 1]), L)
        D = T.block_diag_sparse(T.clover_blocks(Ub, L, kappa, csw)) - kappa * H
    elif kind == "dwf":
        op = C.make_op(C.OP_DWF, mass=mass, m5=m5)
        D = T.dwf_operator(H, ls, mass, m5)[0]
    else:
        op = C.make_op(C.OP_MOBIUS, mass=mass, m5=m5, b=b, c=c)
        Dm, _ = T.mobius_operator(H, ls, mass, m5, b, c)
        kb, _ = T.mobius_kappas(m5, b, c)
        D = Dm / (2 * kb)
        Dminus = (1 - c * (4 - m5)) * sp.identity(12 * V * ls) + 0.5 * c * sp.kron(sp.identity(ls), H)
    lu = spla.splu(sp.csc_matrix(D))
    prop, res = lat.qpropw_point(op, x0, C.make_cg_arg(1e-12, 5000, precision=prec))
    lat.close()
    site = x0[0] + L[0] * (x0[1] + L[1] * (x0[2] + L[2] * x0[3]))
    got = prop.reshape(V, 12, 12)
    worst = 0.0
    for spin in range(4):
        for col in range(3):
            b4 = np.zeros((V, 12), dtype=complex)
            b4[site, 3 * spin + col] = 1.0
            if ls == 1:
                x4 = lu.solve(b4.reshape(-1)).reshape(V, 12)
            else:
                b5 = _four2five(b4, ls).reshape(-1)
                if kind == "mobius":
                    b5 = Dminus @ b5
                x4 = _five2four(lu.solve(b5).reshape(ls, V, 12))
            col_got = got[:, :, 3 * spin + col]
            worst = max(worst, np.linalg.norm(col_got - x4) / np.linalg.norm(x4))
    print("qpropw", kind, x0, prec, [r.iterations for r in res], worst, flush=True)
    assert max(r.true_res for r in res) <= 1e-10
    assert worst < (1e-9 if prec == C.PREC_DOUBLE else 1e-8)


# ---------------------------------------------------------------------------
# site-per-wavefront thread mapping of the hop (HOP_SITEWAVE)


@pytest.mark.parametrize("prec", [C.PREC_DOUBLE, C.PREC_MIXED_HALF])
def test_sitewave_mapping_bit_identical(oracle, monkeypatch, prec):
    """The Moebius operator with Ls = 16 and 4 rhs (one site per wavefront)
    gives the same bits as the element mapping, and matches the oracle."""
    L, ls, nrhs = [4, 4, 4, 8], 16, 4
    mass, m5, b, c = 0.05, 1.8, 1.5, 0.5
    u = wilson_gauge(L, 0.2, 210)
    op = C.make_op(C.OP_MOBIUS, mass=mass, m5=m5, b=b, c=c)
    vcb = int(np.prod(L)) // 2
    src = np.concatenate([F.gaussian_fermion(vcb * ls, 211 + k) for k in range(nrhs)])
    res = {}
    for mode in ("element", "sitewave"):
        monkeypatch.setenv("CPS_HIP_HOP_MODE", mode)
        lat = C.Lattice(L, ls, (0, 0, 0, 1))
        lat.import_gauge(u)
        if prec == C.PREC_DOUBLE:
            res[mode] = (lat.matpc(op, src, 0, nrhs=nrhs), lat.matpc(op, src, 1, nrhs=nrhs))
        else:
            res[mode] = lat.mdagm_prec(op, src, prec, nrhs=nrhs)
        lat.close()
    for a, b_ in zip(res["element"], res["sitewave"]):
        assert np.array_equal(a, b_)
    if prec == C.PREC_DOUBLE:
        m = oracle.mobius_struct(L, ls, m5, b, c)
        n = 24 * vcb * ls
        for k in range(nrhs):
            ref = oracle.mobius_m(u, src[k * n:(k + 1) * n], mass, m)
            assert rel(res["sitewave"][0][k * n:(k + 1) * n], ref) < 1e-13


@pytest.mark.parametrize("case", ["clover", "mobius", "wilson"])
def test_rhsgroup_mapping_bit_identical(monkeypatch, case):
    """Operators whose hops run with several rhs of one site per wavefront
    (HOP_RHSGROUP; clover with two kappas, rhs r on kappa r // 2) give the
    same bits as the element mapping in FP64 (outputs and |M x|^2).  The
    half-precision kernels of the two mappings are separate compilations of
    the same source: the compiler may pair the FP32 products of a sum
    differently in its FMA contraction (the LLVM IR of the two instances has
    the same operations in the same order), so a few FP32 results differ in
    the last bit and, where that crosses an FP16 rounding boundary, a stored
    element by one FP16 step (237 of 3.1M elements at 16^4 on the clover
    operator; none at this size).  Half precision is therefore compared to
    1e-6 (outputs) and 1e-10 (|M x|^2) rather than bit for bit."""
    L = [4, 4, 4, 8]
    u = wilson_gauge(L, 0.2, 250)
    ls = 8 if case == "mobius" else 1
    nrhs = 8 if case == "wilson" else 4
    vcb = int(np.prod(L)) // 2
    src = np.concatenate([F.gaussian_fermion(vcb * ls, 251 + k) for k in range(nrhs)])
    res = {}
    for mode in ("element", "rhsgroup"):
        monkeypatch.setenv("CPS_HIP_HOP_MODE", mode)
        lat = C.Lattice(L, ls, (0, 0, 0, 1))
        lat.import_gauge(u)
        if case == "clover":
            lat.clover_build(CLOVER_KAPPAS, CLOVER_CSW)
            op = clover_op(CLOVER_KAPPAS[0])
        elif case == "mobius":
            op = C.make_op(C.OP_MOBIUS, mass=0.05, m5=1.8, b=1.5, c=0.5)
        else:
            op = C.make_op(C.OP_WILSON, mass=0.1)
        res[mode] = (lat.matpc(op, src, 0, nrhs=nrhs), lat.matpc(op, src, 1, nrhs=nrhs),
                     *lat.mdagm(op, src, nrhs=nrhs), *lat.mdagm_prec(op, src, C.PREC_MIXED_HALF, nrhs=nrhs))
        lat.close()
    for a, b_ in zip(res["element"][:4], res["rhsgroup"][:4]):
        assert np.array_equal(a, b_)
    (ha, da), (hb, db) = res["element"][4:], res["rhsgroup"][4:]
    assert rel(hb, ha) < 1e-6
    assert np.max(np.abs(db - da) / da) < 1e-10


@pytest.mark.parametrize("paired", ["1", "0"])
def test_clover_mixed_residual_step_in_hop(monkeypatch, paired):
    """Clover with two kappa sets, half precision: r -= alpha M^dag M p and
    |r|^2 in the epilogue of MatPcDag's last hop (CLOV_X with the residual
    update, CPS_HIP_CLOVER_R_IN_HOP=1; M^dag M p never stored), with paired
    and with single solution steps.  It rounds M^dag M p differently from the separate pass, so the
    checks are the FP64 stopping rule, iteration counts within 5% and the
    solutions within what the 1e-10 residual bound allows."""
    monkeypatch.setenv("CPS_HIP_PAIRED_X", paired)
    L = [4, 4, 4, 8]
    u = wilson_gauge(L, 0.2, 270)
    nrhs = 4
    vcb = int(np.prod(L)) // 2
    src = np.concatenate([F.gaussian_fermion(vcb, 271 + k) * (1.0 + 3.0 * k) for k in range(nrhs)])
    arg = C.make_cg_arg(1e-10, 2000, precision=C.PREC_MIXED_HALF, reliable_every=40)
    out = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("CPS_HIP_CLOVER_R_IN_HOP", fused)
        lat = C.Lattice(L, 1, (0, 0, 0, 1))
        lat.import_gauge(u)
        lat.clover_build(CLOVER_KAPPAS, CLOVER_CSW)
        out[fused] = lat.invcg(clover_op(CLOVER_KAPPAS[0]), src, arg, nrhs=nrhs)
        lat.close()
    it1 = [r.iterations for r in out["1"][1]]
    it0 = [r.iterations for r in out["0"][1]]
    print("clover r in hop", it1, it0, flush=True)
    assert max(r.true_res for r in out["1"][1]) <= 1e-10
    assert all(r.status == 1 for r in out["1"][1])
    for a, b_ in zip(it1, it0):
        assert abs(a - b_) <= 0.05 * b_ + 2
    assert rel(out["1"][0], out["0"][0]) < 1e-8


@pytest.mark.parametrize("paired", ["1", "0"])
def test_dwf_mixed_residual_step_in_hop(oracle, monkeypatch, paired):
    """Shamir DWF, half precision: r -= alpha M^dag M p and |r|^2 in the
    epilogue of the second hop of M^dag (CLOV_NRR, the default; M^dag M p
    never stored) against the separate pass (CPS_HIP_R_IN_HOP=0), with paired
    and with single solution steps: the FP64 stopping rule, iteration counts
    within 5%, and both solutions against the oracle's FP64 CG."""
    monkeypatch.setenv("CPS_HIP_PAIRED_X", paired)
    L, ls = [4, 4, 4, 8], 8
    u = wilson_gauge(L, 0.2, 280)
    nrhs = 3
    vcb = int(np.prod(L)) // 2
    n = 24 * vcb * ls
    src = np.concatenate([F.gaussian_fermion(vcb * ls, 281 + k) * (1.0 + 2.0 * k) for k in range(nrhs)])
    op = C.make_op(C.OP_DWF, mass=0.02, m5=1.8)
    arg = C.make_cg_arg(1e-10, 3000, precision=C.PREC_MIXED_HALF, reliable_every=40)
    out = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("CPS_HIP_R_IN_HOP", fused)
        lat = C.Lattice(L, ls, (0, 0, 0, 1))
        lat.import_gauge(u)
        out[fused] = lat.invcg(op, src, arg, nrhs=nrhs)
        lat.close()
    it1 = [r.iterations for r in out["1"][1]]
    it0 = [r.iterations for r in out["0"][1]]
    print("dwf r in hop", it1, it0, flush=True)
    for f in ("1", "0"):
        assert max(r.true_res for r in out[f][1]) <= 1e-10
        assert all(r.status == 1 for r in out[f][1])
    for a, b_ in zip(it1, it0):
        assert abs(a - b_) <= 0.05 * b_ + 2
    d = oracle.dwf_struct(L, ls, 1.8)
    for k in range(nrhs):
        ref, _, _ = oracle.dwf_invcg(u, src[k * n:(k + 1) * n], 0.02, 1e-12, 5000, d)
        assert rel(out["1"][0][k * n:(k + 1) * n], ref) < 1e-8


@pytest.mark.parametrize("case", ["clover", "wilson"])
def test_mixed_cg_mapping_consistent(monkeypatch, case):
    """A mixed-precision CG with the 4-D operators' hops free to use the
    element or the rhs-group mapping (the launch with the |M p|^2 chunk
    partials included) takes the same iterations to the same solution.  The
    partial sums use the same tree in both mappings, but the half-precision
    kernels may round a few FP32 results differently
    (test_rhsgroup_mapping_bit_identical): the two-kappa clover solve first
    differs at iteration 7 (|M p|^2 of one rhs, 3.5e-10 relative) and ends
    1.8e-12 apart, so the solutions are compared to 1e-9, well inside what a
    1e-10 residual bound allows."""
    L = [4, 4, 4, 8]
    u = wilson_gauge(L, 0.2, 260)
    nrhs = 4
    vcb = int(np.prod(L)) // 2
    src = np.concatenate([F.gaussian_fermion(vcb, 261 + k) for k in range(nrhs)])
    arg = C.make_cg_arg(1e-10, 2000, precision=C.PREC_MIXED_HALF, reliable_every=50)
    res = {}
    for mode in ("element", "rhsgroup"):
        monkeypatch.setenv("CPS_HIP_HOP_MODE", mode)
        lat = C.Lattice(L, 1, (0, 0, 0, 1))
        lat.import_gauge(u)
        if case == "clover":
            lat.clover_build(CLOVER_KAPPAS, CLOVER_CSW)
            op = clover_op(CLOVER_KAPPAS[0])
        else:
            op = C.make_op(C.OP_WILSON, mass=0.1)
        x, r = lat.invcg(op, src, arg, nrhs=nrhs)
        res[mode] = (x, [k.iterations for k in r], [k.true_res for k in r])
        lat.close()
    assert res["element"][1] == res["rhsgroup"][1]
    assert max(res["element"][2] + res["rhsgroup"][2]) < 1e-10
    assert rel(res["rhsgroup"][0], res["element"][0]) < 1e-9


@pytest.mark.parametrize("kind,ls,nrhs", [("mobius", 16, 2), ("mobius", 12, 1), ("mobius", 8, 1), ("dwf", 16, 2)])
def test_lds_links_mapping_bit_identical(monkeypatch, kind, ls, nrhs):
    """Half-precision MdagM (both hops, the first with its |M x|^2 partial
    sums) with the links staged in LDS gives the same bits as the element
    mapping that loads them per thread; Ls = 12 puts sites across block
    boundaries (nb = 12 does not divide the block)."""
    L = [4, 4, 4, 8]
    u = wilson_gauge(L, 0.2, 230)
    if kind == "mobius":
        op = C.make_op(C.OP_MOBIUS, mass=0.05, m5=1.8, b=1.5, c=0.5)
    else:
        op = C.make_op(C.OP_DWF, mass=0.05, m5=1.8)
    vcb = int(np.prod(L)) // 2
    src = np.concatenate([F.gaussian_fermion(vcb * ls, 231 + k) for k in range(nrhs)])
    res = {}
    for mode in ("element", "lds"):
        monkeypatch.setenv("CPS_HIP_HOP_MODE", mode)
        lat = C.Lattice(L, ls, (0, 0, 0, 1))
        lat.import_gauge(u)
        res[mode] = lat.mdagm_prec(op, src, C.PREC_MIXED_HALF, nrhs=nrhs)
        lat.close()
    for a, b_ in zip(res["element"], res["lds"]):
        assert np.array_equal(a, b_)
    assert np.abs(res["lds"][0]).max() > 0


@pytest.mark.parametrize("links", [C.PREC_MIXED_HALF, C.PREC_MIXED_BF16])
def test_link_row_mapping(oracle, monkeypatch, links):
    """Half-precision Moebius operator (Ls 16) with each link loaded once per
    16-lane row and shared by row broadcasts (HOP_LINK_ROW,
    CPS_HIP_HOP_MODE=row): M^dag M and |M x|^2 agree with the element mapping
    to FP32 rounding (a separate compilation of the same arithmetic) and with
    the FP64 oracle to half-precision accuracy, for FP16 and BF16 links and 1
    and 2 rhs (18-real links: with 16 slots per site the 12-real form is
    never chosen); the mixed CG takes the same iterations (within one) to the
    same solution."""
    L, ls = [4, 4, 4, 8], 16
    mass, m5, b, c = 0.05, 1.8, 1.5, 0.5
    u = wilson_gauge(L, 0.2, 232)
    op = C.make_op(C.OP_MOBIUS, mass=mass, m5=m5, b=b, c=c)
    vcb = int(np.prod(L)) // 2
    n = 24 * vcb * ls
    arg = C.make_cg_arg(1e-10, 3000, precision=links, reliable_every=50)
    res = {}
    for nrhs in (2, 1):
        src = np.concatenate([F.gaussian_fermion(vcb * ls, 233 + k) for k in range(nrhs)])
        for mode in ("element", "row"):
            monkeypatch.setenv("CPS_HIP_HOP_MODE", mode)
            lat = C.Lattice(L, ls, (0, 0, 0, 1))
            lat.import_gauge(u)
            out, dp = lat.mdagm_prec(op, src, links, nrhs=nrhs)
            x, r = lat.invcg(op, src, arg, nrhs=nrhs)
            res[(nrhs, mode)] = (out, dp, x, [k.iterations for k in r], [k.true_res for k in r])
            lat.close()
        (oe, de, xe, ie, te), (orow, drow, xrow, irow, trow) = res[(nrhs, "element")], res[(nrhs, "row")]
        assert np.abs(orow).max() > 0
        assert rel(orow, oe) < 1e-6 and np.max(np.abs(drow - de) / de) < 1e-10
        m = oracle.mobius_struct(L, ls, m5, b, c)
        for k in range(nrhs):
            ref = oracle.mobius_mdagm(u, src[k * n:(k + 1) * n], mass, m)[0]
            assert rel(orow[k * n:(k + 1) * n], ref) < 2e-2
        assert max(te + trow) <= 1e-10
        assert all(abs(a - b_) <= 1 for a, b_ in zip(ie, irow))
        assert rel(xrow, xe) < 1e-9


# ---------------------------------------------------------------------------
# CANONICAL order on the device: gauge import (Convert + BondCond) and
# Lattice::FmatInv on CANONICAL vectors


def test_gauge_import_canonical_equals_wilson_import():
    """Importing the CANONICAL field (signs applied on the device) gives the
    same operator, bit for bit, as importing the WILSON-order field with the
    signs applied on the host."""
    L, ls = [4, 4, 4, 8], 4
    u = F.gauge_field(L, 0.2, 220)
    src = F.gaussian_fermion(int(np.prod(L)) // 2 * ls, 221)
    op = C.make_op(C.OP_DWF, mass=0.05, m5=1.8)
    outs = []
    for canon in (False, True):
        lat = C.Lattice(L, ls, (0, 0, 0, 1))
        if canon:
            lat.import_gauge_canonical(u)
        else:
            lat.import_gauge(wilson_gauge(L, 0.2, 220))
        outs.append(lat.matpc(op, src, 0))
        lat.close()
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("kind,prec", [("wilson", C.PREC_DOUBLE), ("clover", C.PREC_DOUBLE), ("dwf", C.PREC_DOUBLE),
                                       ("mobius", C.PREC_DOUBLE), ("mobius", C.PREC_MIXED_HALF)])
def test_fmatinv_canonical(kind, prec):
    """Lattice::FmatInv on CANONICAL vectors (two rhs): D out = in for the full
    textbook operator (Moebius: D out = D_minus in, the default dminus)."""
    import scipy.sparse as sp
    from tests import textbook as T
    from cps_public_amd import layout as LY
    L = [4, 4, 4, 4]
    V = int(np.prod(L))
    ls = 1 if kind in ("wilson", "clover") else 4
    u = F.gauge_field(L, 0.2, 230)
    H = T.hopping(T.links(u, L), L, antiperiodic=(0, 0, 0, 1))
    mass, m5, b, c = 0.1, 1.8, 1.5, 0.5
    lat = C.Lattice(L, ls, (0, 0, 0, 1))
    lat.import_gauge_canonical(u)
    lhs = None
    if kind == "wilson":
        op = C.make_op(C.OP_WILSON, mass=mass)
        D = sp.identity(12 * V) - (1.0 / (2 * (mass + 4))) * H
    elif kind == "clover":
        kappa, csw = 0.12, 1.1
        lat.clover_build([kappa], csw)
        op = C.make_op(C.OP_CLOVER, mass=1.0 / (2 * kappa) - 4.0, csw=csw)
        D = T.block_diag_sparse(T.clover_blocks(T.links(LY.bond_cond(u, L, [0, 0, 0, 1]), L), L, kappa, csw)) - kappa * H
    elif kind == "dwf":
        op = C.make_op(C.OP_DWF, mass=mass, m5=m5)
        D = T.dwf_operator(H, ls, mass, m5)[0]
    else:
        op = C.make_op(C.OP_MOBIUS, mass=mass, m5=m5, b=b, c=c)
        Dm, _ = T.mobius_operator(H, ls, mass, m5, b, c)
        kb, _ = T.mobius_kappas(m5, b, c)
        D = Dm / (2 * kb)
        lhs = (1 - c * (4 - m5)) * sp.identity(12 * V * ls) + 0.5 * c * sp.kron(sp.identity(ls), H)
    rng = np.random.default_rng(231)
    nrhs = 2
    bz = rng.standard_normal((nrhs, 12 * V * ls)) + 1j * rng.standard_normal((nrhs, 12 * V * ls))
    breal = np.stack([bz.real, bz.imag], axis=-1).reshape(-1)
    out, res = lat.fmatinv(op, breal, C.make_cg_arg(1e-12, 5000, precision=prec), nrhs=nrhs)
    lat.close()
    o = out.reshape(nrhs, -1, 2)
    xz = o[..., 0] + 1j * o[..., 1]
    for k in range(nrhs):
        rhs = bz[k] if lhs is None else lhs @ bz[k]
        r = D @ xz[k] - rhs
        print("fmatinv", kind, prec, res[k].iterations, np.linalg.norm(r) / np.linalg.norm(rhs), flush=True)
        assert np.linalg.norm(r) / np.linalg.norm(rhs) < 1e-9
