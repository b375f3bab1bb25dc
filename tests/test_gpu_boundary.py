"""GPU: boundary conditions and gauge residency through the C ABI.

* The 12-real link formats rebuild the third row as det(U) conj(row0 x row1),
  with det U of the last-slice links measured from the field at import, so
  they are right whatever bc the handle was created with: CPS's DiracOp
  constructor puts the boundary factors into the links
  (dirac_op_base.C:164, d_op_wilson.C:118-121) and a binding may create its
  handle with bc = {0,0,0,0}.
* Twisted boundary conditions (Lattice::BondCond, lattice_base.C:5313-5369;
  GJP::TwistPhase, gjp.C:475-493) through the CANONICAL import, checked
  against the oracle's BondCond, whose twist is pinned by a twisted plane
  wave (tests/test_oracle_pins.py).
* Re-importing an unchanged gauge (CPS's library calls pass the gauge every
  time) keeps the clover sets and low-precision copies.
* A NERSC file written and read back feeds the device operator and the
  service (ReadLatticeParallel, ReadLatticePar.C:60-286).

Tolerances: FP64 operators 1e-13 relative to the oracle; sloppy operators
their storage rounding (FP32 1e-5, FP16 5e-3, BF16 3e-2 relative), as in
test_gpu_parity.test_sloppy_operator."""
import os
import tempfile

import numpy as np
import pytest

import cps_public_amd as C
from cps_public_amd import fields as F
from cps_public_amd import layout as LY
from tests.helpers import rel, wilson_gauge

pytestmark = pytest.mark.gpu

SLOPPY_TOL = {C.PREC_MIXED_SINGLE: 1e-5, C.PREC_MIXED_HALF: 5e-3, C.PREC_MIXED_BF16: 3e-2}
L = [4, 4, 4, 8]


def _ref_mdagm(oracle, kind, u, inp, ls):
    if kind == "wilson":
        op = C.make_op(C.OP_WILSON, mass=0.1)
        ref, dref = oracle.wilson_mdagm(u, inp, 1.0 / (2 * 4.1), L)
    elif kind == "dwf":
        op = C.make_op(C.OP_DWF, mass=0.01, m5=1.8)
        ref, dref = oracle.dwf_mdagm(u, inp, 0.01, oracle.dwf_struct(L, ls, 1.8))
    else:
        kappa, csw = 0.13, 1.0
        op = C.make_op(C.OP_CLOVER, mass=1.0 / (2 * kappa) - 4.0, csw=csw)
        A_oo = oracle.clover_mat_chkb(u, 1, 0, kappa, csw, L)
        A_ee_inv = oracle.clover_mat_chkb(u, 0, 1, kappa, csw, L)
        ref, dref = oracle.clover_mdagm(u, A_oo, A_ee_inv, inp, kappa, L)
    return op, ref, dref


@pytest.mark.parametrize("prec", [C.PREC_MIXED_SINGLE, C.PREC_MIXED_HALF, C.PREC_MIXED_BF16])
@pytest.mark.parametrize("kind", ["wilson", "clover", "dwf"])
def test_sloppy_operator_handle_bc_zero(oracle, prec, kind):
    """A handle created with bc = {0,0,0,0} on a t-antiperiodic WILSON field:
    the sloppy operators (12-real links for FP32 always, for FP16/BF16 with
    one rhs of a 4-D operator) match the FP64 oracle, and give the same bits
    as a handle created with bc = {0,0,0,1}."""
    ls = 1 if kind != "dwf" else 2   # DWF with Ls 2: one slot per site, 12-real
    u = wilson_gauge(L, 0.2, 301)
    inp = F.gaussian_fermion(L[0] * L[1] * L[2] * L[3] // 2 * ls, 302)
    outs = []
    for bc in ((0, 0, 0, 0), (0, 0, 0, 1)):
        lat = C.Lattice(L, ls, bc)
        lat.import_gauge(u)
        ph, ok = lat.boundary_phase()
        assert ok and np.allclose(ph, [1, 1, 1, -1], rtol=0, atol=0), (ph, ok)
        op, ref, dref = _ref_mdagm(oracle, kind, u, inp, ls)
        if kind == "clover":
            lat.clover_build([0.13], 1.0)
        got, d = lat.mdagm_prec(op, inp, prec)
        print("bc0 sloppy", kind, prec, bc, rel(got, ref), flush=True)
        assert rel(got, ref) < SLOPPY_TOL[prec]
        assert abs(d[0] - dref) / dref < SLOPPY_TOL[prec]
        outs.append(got)
        lat.close()
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("prec", [C.PREC_MIXED_SINGLE, C.PREC_MIXED_HALF])
def test_mixed_cg_handle_bc_zero(oracle, prec):
    """The mixed CG of a bc = {0,0,0,0} handle on a signed field converges as
    with the correct bc (same iterations and bits) and meets the FP64
    residual; its solution agrees with the oracle's FP64 solve."""
    u = wilson_gauge(L, 0.3, 311)
    src = F.gaussian_fermion(L[0] * L[1] * L[2] * L[3] // 2, 312)
    op = C.make_op(C.OP_WILSON, mass=1.0 / (2 * 0.12) - 4.0)
    sols, its = [], []
    for bc in ((0, 0, 0, 0), (0, 0, 0, 1)):
        lat = C.Lattice(L, 1, bc)
        lat.import_gauge(u)
        sol, res = lat.invcg(op, src, C.make_cg_arg(1e-10, 5000, precision=prec))
        assert res[0].true_res <= 1e-10
        sols.append(sol)
        its.append(res[0].iterations)
        lat.close()
    assert its[0] == its[1] and np.array_equal(sols[0], sols[1])
    rsol, rit, _ = oracle.wilson_invcg(u, src, 1.0 / (2 * 0.12) - 4.0, 1e-12, 5000, L)
    print("bc0 mixed", prec, its, rit, rel(sols[0], rsol), flush=True)
    assert rel(sols[0], rsol) < 1e-8


TW_BC = (2, 1, 0, 2)
TW_ANGLE = (0.3, 0.0, 0.0, -0.7)


def _twisted_wilson(oracle, u):
    return oracle.gauge_canon_to_wilson(oracle.bond_cond_twist(u, L, list(TW_BC), list(TW_ANGLE)), L)


def test_twisted_canonical_import(oracle):
    """CANONICAL import with twisted x and t, antiperiodic y: the device field
    equals the oracle's BondCond (FP64 operator within 1e-13 of the oracle on
    the oracle-twisted field), and the measured boundary factor of each
    direction is det U = phase^3."""
    u = F.gauge_field(L, 0.2, 320)
    uw = _twisted_wilson(oracle, u)
    lat = C.Lattice(L, 1, TW_BC, twist=TW_ANGLE)
    lat.import_gauge_canonical(u)
    ph, ok = lat.boundary_phase()
    fac = LY.boundary_factor(TW_BC, TW_ANGLE)
    print("twisted phase", ph, ok, flush=True)
    assert ok and np.allclose(ph, np.array(fac) ** 3, rtol=0, atol=1e-12)
    psi = F.gaussian_fermion(lat.vcb, 321)
    for cb in (0, 1):
        for dag in (0, 1):
            ref = oracle.wilson_dslash(uw, psi, cb, dag, L)
            assert rel(lat.wilson_dslash(psi, cb, dag), ref) < 1e-13
    ref, mp_ref = oracle.wilson_mdagm(uw, psi, 0.12, L)
    got, mp = lat.wilson_mdagm(psi, 0.12)
    assert rel(got, ref) < 1e-13 and abs(mp - mp_ref) / mp_ref < 1e-13
    # the WILSON-order import of the same field gives the same bits
    lat2 = C.Lattice(L, 1, (0, 0, 0, 0))
    lat2.import_gauge(uw)
    assert np.array_equal(lat2.wilson_mdagm(psi, 0.12)[0], lat.wilson_mdagm(psi, 0.12)[0])
    lat2.close()
    lat.close()


@pytest.mark.parametrize("prec", [C.PREC_MIXED_SINGLE, C.PREC_MIXED_HALF, C.PREC_MIXED_BF16])
def test_twisted_sloppy_operator_and_cg(oracle, prec):
    """12-real links on a twisted field: the third row gets the complex
    boundary factor, so the sloppy operator stays within its rounding of the
    FP64 oracle, and the mixed CG meets the FP64 residual."""
    u = F.gauge_field(L, 0.2, 330)
    uw = _twisted_wilson(oracle, u)
    lat = C.Lattice(L, 1, TW_BC, twist=TW_ANGLE)
    lat.import_gauge_canonical(u)
    inp = F.gaussian_fermion(lat.vcb, 331)
    op = C.make_op(C.OP_WILSON, mass=0.1)
    ref, dref = oracle.wilson_mdagm(uw, inp, 1.0 / (2 * 4.1), L)
    got, d = lat.mdagm_prec(op, inp, prec)
    print("twisted sloppy", prec, rel(got, ref), flush=True)
    assert rel(got, ref) < SLOPPY_TOL[prec]
    sol, res = lat.invcg(op, inp, C.make_cg_arg(1e-10, 5000, precision=prec))
    rsol, rit, _ = oracle.wilson_invcg(uw, inp, 0.1, 1e-12, 5000, L)
    print("twisted mixed cg", prec, res[0].iterations, rit, res[0].true_res, rel(sol, rsol), flush=True)
    assert res[0].true_res <= 1e-10 and rel(sol, rsol) < 1e-8
    lat.close()


def test_non_su3_field_keeps_18_real_links(oracle):
    """A field that is not SU(3) times a boundary factor cannot be rebuilt
    from two rows: recon_ok is 0, the FP16 sloppy operator keeps 18-real
    links (and stays within its rounding), and FP32 (12-real only) is refused."""
    u = wilson_gauge(L, 0.2, 340).copy()
    rng = np.random.default_rng(341)
    u *= 1.0 + 0.05 * rng.standard_normal(u.shape)   # no longer unitary
    lat = C.Lattice(L, 1, (0, 0, 0, 1))
    lat.import_gauge(u)
    _, ok = lat.boundary_phase()
    assert not ok
    inp = F.gaussian_fermion(lat.vcb, 342)
    op = C.make_op(C.OP_WILSON, mass=0.1)
    ref, _ = oracle.wilson_mdagm(u, inp, 1.0 / (2 * 4.1), L)
    got, _ = lat.mdagm_prec(op, inp, C.PREC_MIXED_HALF)
    assert rel(got, ref) < SLOPPY_TOL[C.PREC_MIXED_HALF]
    with pytest.raises(C.CpsHipError) as e:
        lat.mdagm_prec(op, inp, C.PREC_MIXED_SINGLE)
    assert e.value.code == -1
    lat.close()


def test_reimport_same_gauge_keeps_clover(oracle):
    """CPS's library calls take the gauge with every call
    (d_op_clover.C:216,230,305,333 pass gauge_field to wilson_dslash): an
    unchanged field re-imported keeps the clover sets, so a clover solve
    after wilson_dslash(u) works and gives the same bits; a changed field
    drops them (the clover solve is then refused)."""
    u = wilson_gauge(L, 0.15, 350)
    lat = C.Lattice(L, 1, (0, 0, 0, 0))
    lat.import_gauge(u)
    kappa, csw = 0.13, 1.0
    lat.clover_build([kappa], csw)
    op = C.make_op(C.OP_CLOVER, mass=1.0 / (2 * kappa) - 4.0, csw=csw)
    src = F.gaussian_fermion(lat.vcb, 351)
    arg = C.make_cg_arg(1e-10, 5000)
    sol0, res0 = lat.invcg(op, src, arg)
    psi = F.gaussian_fermion(lat.vcb, 352)
    lat.wilson_dslash(psi, 0, 0)   # the dslash entry point's own staging fields
    bytes0 = lat.device_bytes()
    chi = lat.wilson_dslash(psi, 0, 0, u=u)
    assert rel(chi, oracle.wilson_dslash(u, psi, 0, 0, L)) < 1e-13
    sol1, res1 = lat.invcg(op, src, arg)
    assert np.array_equal(sol0, sol1) and res0[0].iterations == res1[0].iterations
    # the comparison buffer is the only growth
    assert lat.device_bytes() <= bytes0 + u.nbytes + 4096
    u2 = wilson_gauge(L, 0.15, 353)
    lat.wilson_dslash(psi, 0, 0, u=u2)
    with pytest.raises(C.CpsHipError):
        lat.invcg(op, src, arg)
    lat.close()


@pytest.mark.parametrize("rows,nbytes", [(3, 8), (2, 4)])
def test_nersc_file_to_device_operator_and_service(oracle, rows, nbytes):
    """A NERSC file (written by cps_hip_nersc_write, read back by
    cps_hip_nersc_read) imported CANONICALLY gives the operator of the field
    it holds: with 3x3 IEEE64 storage the same bits as importing the
    original field, with two rows in IEEE32 the operator of the rounded,
    rebuilt field (within 1e-6 of the original).  The field read back also
    serves a solve in the resident service."""
    Ln, ls = [4, 4, 4, 8], 4
    u = F.gauge_field(Ln, 0.2, 360)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "conf.nersc")
        C.nersc_write(path, Ln, u, recon_row_3=1 if rows == 2 else 0, bytes_per_real=nbytes)
        ur, info = C.nersc_read(path)
    assert list(info.dims) == Ln
    op = C.make_op(C.OP_DWF, mass=0.05, m5=1.8)
    src = F.gaussian_fermion(int(np.prod(Ln)) // 2 * ls, 361)
    outs = []
    for field in (u, ur):
        lat = C.Lattice(Ln, ls, (0, 0, 0, 1))
        lat.import_gauge_canonical(field)
        outs.append(lat.matpc(op, src, 0))
        lat.close()
    if rows == 3:
        assert np.array_equal(outs[0], outs[1])
    else:
        assert rel(outs[1], outs[0]) < 1e-6
    # the oracle on the field read back
    uw = LY.gauge_canon_to_wilson(LY.bond_cond(ur, Ln, [0, 0, 0, 1]), Ln)
    ref, _ = oracle.dwf_mdagm(uw, src, 0.05, oracle.dwf_struct(Ln, ls, 1.8))
    lat = C.Lattice(Ln, ls, (0, 0, 0, 1))
    lat.import_gauge_canonical(ur)
    got, _ = lat.mdagm(op, src)
    assert rel(got, ref) < 1e-13
    sol_l, res_l = lat.invcg(op, src, C.make_cg_arg(1e-10, 5000))
    lat.close()
    svc = C.Service(8 << 30, max_slots=4)
    gid = svc.add_gauge(Ln, ls, uw)
    rid = svc.submit(gid, op, src, 1, 1e-10, 5000)
    sol_s, res_s, _ = svc.wait(rid)
    svc.close()
    print("nersc", rows, nbytes, res_l[0].iterations, res_s[0].iterations, rel(sol_s, sol_l), flush=True)
    assert res_s[0].true_res <= 1e-10 and rel(sol_s, sol_l) < 1e-9
