// Even-odd MatInv on device fields, and the batched point-source propagator
// driver built on it.
//
// QPropW::Run (QPropW.C:566-613) loops over the 12 spin-colour sources: for
// each it sets a point source, calls Lattice::FmatInv through QPropW::CG
// (QPropW.C:1077-1267) and stores the solution as one source column of the
// WilsonMatrix propagator (LoadRow, QPropW.C:1343-1364).  Here the 12
// sources are one batched solve: they are written straight into the device
// fields, solved together, and the 4-D propagator is gathered on the device,
// so only the propagator itself crosses PCIe.
#include "solver.h"

#include <chrono>

namespace cpship {

int matinv_device(Lattice& L, const cps_hip_op* op, Field& in_o, Field& in_e, Field& x, Field& out_e, int nrhs,
                  const cps_hip_cg_arg* arg, cps_hip_cg_result* res) {
  OpLayout o;
  int rc = op_layout(op, L.ls, &o, L.g);
  if (rc) return rc;
  const int64_t n = o.per_rhs * nrhs;
  Field& tmp = L.field(S_TMP, STORE_D, n);
  Field& src = L.field(S_SRC, STORE_D, n);
  Field& t1 = L.field(S_T1, STORE_D, n);
  Field& t2 = L.field(S_T2, STORE_D, n);
  Field& t3 = L.field(S_T3, STORE_D, n);
  if (L.last_rc) return L.last_rc;
  const double kappa = op_kappa(op);
  switch (op->kind) {
    case CPS_HIP_OP_MOBIUS:
      rc = mobius_matinv_prep(L, op, src.ref(), in_o.ref(), in_e.ref(), t1.ref(), t2.ref(), tmp.ref(), nrhs);
      break;
    case CPS_HIP_OP_CLOVER:
      rc = clover_matinv_prep(L, op, src.ref(), in_o.ref(), in_e.ref(), t1.ref(), t2.ref(), tmp.ref(), nrhs);
      break;
    default:
      // temp = in_o + kappa * D(cb=0) in_e ; src = MatPcDag temp
      // (d_op_wilson.C:289-304, d_op_dwf.C:366-454)
      rc = hop(L, op, PREC_D_D18, tmp.ref(), in_e.ref(), 1, 0, kappa, in_o.ref(), 1.0, nullptr, nrhs);
      if (!rc) rc = apply_matpc(L, op, PREC_D_D18, src.ref(), tmp.ref(), t1.ref(), t2.ref(), 1, nrhs);
  }
  if (rc) return rc;
  std::vector<int> iters, status, nup;
  std::vector<double> tr, sn;
  auto t0 = std::chrono::steady_clock::now();
  if (arg->precision == CPS_HIP_PREC_DOUBLE) rc = cg_double_device(L, op, x, src, nrhs, arg, iters, status, tr, sn);
  else rc = cg_mixed_device(L, op, x, src, nrhs, arg, iters, status, tr, sn, nup);
  if (rc) return rc;
  switch (op->kind) {
    case CPS_HIP_OP_MOBIUS:
      rc = mobius_matinv_reconstruct(L, op, out_e.ref(), x.ref(), in_e.ref(), t1.ref(), t2.ref(), t3.ref(), nrhs);
      break;
    case CPS_HIP_OP_CLOVER:
      rc = clover_matinv_reconstruct(L, op, out_e.ref(), x.ref(), in_e.ref(), t1.ref(), t2.ref(), t3.ref(), nrhs);
      break;
    default:
      // out_e = in_e + kappa * D(cb=1) out_o (d_op_wilson.C:320-323, d_op_dwf.C:489-492)
      rc = hop(L, op, PREC_D_D18, out_e.ref(), x.ref(), 0, 0, kappa, in_e.ref(), 1.0, nullptr, nrhs);
  }
  if (rc) return rc;
  CPSHIP_CHECK(hipStreamSynchronize(L.stream));
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (res) {
    for (int k = 0; k < nrhs; k++) {
      res[k].iterations = iters[k] + 1;   // inv_cg.C:453
      res[k].status = status[k];
      res[k].true_res = tr[k];
      res[k].src_norm_sq = sn[k];
      res[k].reliable_updates = nup.empty() ? 0 : nup[k];
      res[k].seconds = secs;
    }
  }
  return 0;
}

// Fmobius::FmatInv multiplies the source by D_minus first (dminus = 1 by
// default, lattice.h:1124; f_mobius.C:68-78): DiracOpMobius::Dminus
// (d_op_mobius.C:745-778) = -(c D_W - 1/a5) with D_W = (4 - M5) - H/2, i.e.
// (1/a5 - c (4 - M5)) in + (c/2) H in on each checkerboard; applied in place
// to the odd / even source fields.
int mobius_dminus_fields(Lattice& L, const cps_hip_op* op, Field& f_odd, Field& f_even, int nrhs) {
  const int64_t n = (int64_t)L.g.Vcb * L.ls * nrhs;
  Field& t3 = L.field(S_T3, STORE_D, n);
  Field& tmp = L.field(S_TMP, STORE_D, n);
  if (L.last_rc) return L.last_rc;
  const double xc = op->a5_inv - op->mobius_c * (4 - op->m5), hc = 0.5 * op->mobius_c;
  int rc;
  if ((rc = hop(L, op, PREC_D_D18, t3.ref(), f_even.ref(), 1, 0, hc, f_odd.ref(), xc, nullptr, nrhs))) return rc;
  if ((rc = hop(L, op, PREC_D_D18, tmp.ref(), f_odd.ref(), 0, 0, hc, f_even.ref(), xc, nullptr, nrhs))) return rc;
  CPSHIP_CHECK(convert_spinor(STORE_D, t3.ref(), STORE_D, f_odd.ref(), n, L.stream));
  CPSHIP_CHECK(convert_spinor(STORE_D, tmp.ref(), STORE_D, f_even.ref(), n, L.stream));
  return 0;
}

namespace {

// The 12 source entries (one double2 each), addresses computed on the host.
struct PointTargets {
  double2* at[12];
};

__global__ void k_point_src(PointTargets T) {
  if (threadIdx.x != 0) return;
#pragma unroll
  for (int k = 0; k < 12; k++) *T.at[k] = make_double2(1.0, 0.0);
}

// prop[site][s1][c1][spin][colour][re, im] (CANONICAL site order, the
// WilsonMatrix of LoadRow / WilsonMatrix::load_row, WilsonMatrix.C:92-102)
// from the 12 solutions; for the 5-D operators FdwfBase::Ffive2four(sol_4d,
// sol_5d, Ls-1, 0) (f_dwf_base.C:1068-1168): sink spins 0,1 from s = Ls-1,
// spins 2,3 from s = 0.  One thread per (site, rhs).
template <int KIND>
__global__ void k_prop_gather(const double2* __restrict__ xo, const double2* __restrict__ xe, int64_t plane, Geom g,
                              int64_t per_rhs, int ls, double scale, double* __restrict__ prop) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t V = 2 * (int64_t)g.Vcb;
  if (t >= V * 12) return;
  const int k = (int)(t % 12);
  const int64_t site = t / 12;
  int c[4];
  int64_t r = site;
#pragma unroll
  for (int mu = 0; mu < 4; mu++) {
    c[mu] = (int)(r % g.L[mu]);
    r /= g.L[mu];
  }
  const int p = (c[0] + c[1] + c[2] + c[3]) & 1;
  const int site_cb = (c[0] >> 1) + (g.L[0] >> 1) * (c[1] + g.L[1] * (c[2] + g.L[2] * c[3]));
  double* out = prop + site * 288 + 2 * k;   // 6 spin + 2 colour = 2 k
#pragma unroll
  for (int s1 = 0; s1 < 4; s1++) {
    bool odd;
    int64_t e;
    field_slot<KIND>(p, site_cb, s1 < 2 ? ls - 1 : 0, g.Vcb, ls, &odd, &e);
    const double2* d = (odd ? xo : xe) + (int64_t)k * per_rhs + e;
#pragma unroll
    for (int c1 = 0; c1 < 3; c1++) {
      const double2 v = d[(int64_t)(3 * s1 + c1) * plane];
      out[72 * s1 + 24 * c1] = scale * v.x;
      out[72 * s1 + 24 * c1 + 1] = scale * v.y;
    }
  }
}

template <int KIND>
hipError_t launch_gather(const Field& xo, const Field& xe, const Geom& g, int64_t per_rhs, int ls, double scale,
                         double* prop, hipStream_t st) {
  const int64_t nthreads = (int64_t)2 * g.Vcb * 12;
  hipLaunchKernelGGL(k_prop_gather<KIND>, dim3((unsigned)((nthreads + 255) / 256)), dim3(256), 0, st,
                     (const double2*)xo.buf.p, (const double2*)xe.buf.p, xo.nelem, g, per_rhs, ls, scale, prop);
  return hipGetLastError();
}

}  // namespace
}  // namespace cpship

using namespace cpship;

extern "C" int cps_hip_qpropw_point(cps_hip_lattice* lat, const cps_hip_op* op, const int x0[4],
                                    const cps_hip_cg_arg* arg, double* prop, cps_hip_cg_result* res) {
  if (!lat || !op || !x0 || !arg || !prop) return CPS_HIP_ERR_ARG;
  Lattice& L = lat->L;
  L.last_rc = 0;
  if (!L.has_gauge) return CPS_HIP_ERR_NOGAUGE;
  for (int mu = 0; mu < 4; mu++)
    if (x0[mu] < 0 || x0[mu] >= L.g.L[mu]) return CPS_HIP_ERR_ARG;   // FermionVector.C:1575-1581
  OpLayout o;
  int rc = op_layout(op, L.ls, &o, L.g);
  if (rc) return rc;
  const int nrhs = 12;
  const int64_t n = o.per_rhs * nrhs;
  Field& fo = L.field(S_IO_IN, STORE_D, n);
  Field& fe = L.field(S_EVEN_IN, STORE_D, n);
  Field& x = L.field(S_X, STORE_D, n);
  Field& eo = L.field(S_EVEN_OUT, STORE_D, n);
  if (L.last_rc) return L.last_rc;
  const int bs = pick_block(o.per_rhs);
  CPSHIP_CHECK(blas_zero(STORE_D, fo.ref(), o.per_rhs, nrhs, bs, L.stream));
  CPSHIP_CHECK(blas_zero(STORE_D, fe.ref(), o.per_rhs, nrhs, bs, L.stream));
  CPSHIP_CHECK(blas_zero(STORE_D, x.ref(), o.per_rhs, nrhs, bs, L.stream));   // QPropW.C:571 (zero guess)
  // FermionVectorTp::SetPointSource (FermionVector.C:1567-1644) puts 1 at
  // (spin, colour) of site x0; for the 5-D operators FdwfBase::Ffour2five
  // (src_5d, src_4d, 0, Ls-1) (f_dwf_base.C:961-1050) moves spins 0,1 to
  // s = 0 and spins 2,3 to s = Ls-1.  rhs k = 3 spin + colour (the order of
  // QPropW::Run's loops).
  const int p0 = (x0[0] + x0[1] + x0[2] + x0[3]) & 1;
  const int site_cb0 = (x0[0] >> 1) + (L.g.L[0] >> 1) * (x0[1] + L.g.L[1] * (x0[2] + L.g.L[2] * x0[3]));
  PointTargets T;
  for (int k = 0; k < 12; k++) {
    const int spin = k / 3, s = spin < 2 ? 0 : L.ls - 1;
    bool odd = false;
    int64_t e = 0;
    if (op->kind == CPS_HIP_OP_DWF) field_slot<CPS_HIP_OP_DWF>(p0, site_cb0, s, L.g.Vcb, L.ls, &odd, &e);
    else if (op->kind == CPS_HIP_OP_MOBIUS) field_slot<CPS_HIP_OP_MOBIUS>(p0, site_cb0, s, L.g.Vcb, L.ls, &odd, &e);
    else field_slot<CPS_HIP_OP_WILSON>(p0, site_cb0, s, L.g.Vcb, L.ls, &odd, &e);
    e += (int64_t)k * o.per_rhs;
    const Field& f = odd ? fo : fe;
    if (e < 0 || e >= f.nelem) return CPS_HIP_ERR_ARG;
    T.at[k] = (double2*)f.buf.p + (int64_t)k * f.nelem + e;   // complex component k = 3 spin + colour
  }
  hipLaunchKernelGGL(k_point_src, dim3(1), dim3(64), 0, L.stream, T);
  CPSHIP_CHECK(hipGetLastError());
  if (op->kind == CPS_HIP_OP_MOBIUS && (rc = mobius_dminus_fields(L, op, fo, fe, nrhs))) return rc;
  if ((rc = matinv_device(L, op, fo, fe, x, eo, nrhs, arg, res))) return rc;
  // Moebius: the preconditioned system is 2 kappa_b times the operator
  // (d_op_mobius.C:628-633)
  const double fac =
      op->kind == CPS_HIP_OP_MOBIUS ? 1.0 / (op->mobius_b * (4 - op->m5) + op->a5_inv) : 1.0;
  const size_t pbytes = (size_t)2 * L.g.Vcb * 288 * sizeof(double);
  if ((rc = L.io.ensure(pbytes))) return rc;
  double* pd = (double*)L.io.p;
  if (op->kind == CPS_HIP_OP_DWF) {
    CPSHIP_CHECK(launch_gather<CPS_HIP_OP_DWF>(x, eo, L.g, o.per_rhs, L.ls, fac, pd, L.stream));
  } else if (op->kind == CPS_HIP_OP_MOBIUS) {
    CPSHIP_CHECK(launch_gather<CPS_HIP_OP_MOBIUS>(x, eo, L.g, o.per_rhs, L.ls, fac, pd, L.stream));
  } else {
    CPSHIP_CHECK(launch_gather<CPS_HIP_OP_WILSON>(x, eo, L.g, o.per_rhs, L.ls, fac, pd, L.stream));
  }
  CPSHIP_CHECK(hipMemcpyAsync(prop, L.io.p, pbytes, hipMemcpyDeviceToHost, L.stream));
  CPSHIP_CHECK(hipStreamSynchronize(L.stream));
  return 0;
}
