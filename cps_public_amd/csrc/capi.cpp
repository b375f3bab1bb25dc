// extern "C" entry points declared in include/cps_hip.h.
#include <chrono>
#include <cstring>
#include <new>
#include <vector>

#include "solver.h"

using namespace cpship;

namespace {

int check_lat(cps_hip_lattice* lat) {
  if (!lat) return CPS_HIP_ERR_ARG;
  lat->L.last_rc = 0;   // allocation status of this call's scratch requests
  return 0;
}

int maybe_import(cps_hip_lattice* lat, const double* u) {
  if (u) return lat->L.import_gauge(u);
  if (!lat->L.has_gauge) return CPS_HIP_ERR_NOGAUGE;
  return 0;
}

cps_hip_op make_op(int kind, double mass, double m5, double a5_inv) {
  cps_hip_op op;
  std::memset(&op, 0, sizeof(op));
  op.kind = kind;
  op.mass = mass;
  op.m5 = m5;
  op.a5_inv = a5_inv;
  return op;
}

// Wilson-style single hop on host vectors: out (target parity) from in.
int host_hop(cps_hip_lattice* lat, const cps_hip_op& op, double* out, const double* in, int cb, int dag) {
  Lattice& L = lat->L;
  OpLayout o;
  int rc = op_layout(&op, L.ls, &o, L.g);
  if (rc) return rc;
  Field& fi = L.field(S_IO_IN, STORE_D, o.per_rhs);
  Field& fo = L.field(S_IO_OUT, STORE_D, o.per_rhs);
  if (L.last_rc) return L.last_rc;
  // the input checkerboard is cb; the output is 1-cb (4-D parity or 5-D checkerboard)
  rc = upload(L, in, fi, o, cb, 1, 0);
  if (rc) return rc;
  rc = hop(L, &op, PREC_D_D18, fo.ref(), fi.ref(), 1 - cb, dag, 1.0, null_field(), 0.0, nullptr, 1);
  if (rc) return rc;
  return download(L, fo, out, o, 1 - cb, 1);
}

}  // namespace

extern "C" {

int cps_hip_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

const char* cps_hip_version(void) { return "cps_hip 0.1 gfx950"; }

int cps_hip_device_alloc(void** p, size_t bytes) {
  if (!p || bytes == 0) return CPS_HIP_ERR_ARG;
  *p = nullptr;
  if (hipMalloc(p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return CPS_HIP_ERR_NOMEM;
  }
  return 0;
}

int cps_hip_device_free(void* p) {
  if (!p) return 0;
  return hipFree(p) == hipSuccess ? 0 : CPS_HIP_ERR_RUNTIME;
}

int cps_hip_memcpy(void* dst, const void* src, size_t bytes, int kind) {
  if ((!dst || !src) && bytes) return CPS_HIP_ERR_ARG;
  hipMemcpyKind k;
  switch (kind) {
    case CPS_HIP_H2D: k = hipMemcpyHostToDevice; break;
    case CPS_HIP_D2H: k = hipMemcpyDeviceToHost; break;
    case CPS_HIP_D2D: k = hipMemcpyDeviceToDevice; break;
    default: return CPS_HIP_ERR_ARG;
  }
  if (bytes == 0) return 0;
  return hipMemcpy(dst, src, bytes, k) == hipSuccess ? 0 : CPS_HIP_ERR_RUNTIME;
}

int cps_hip_lattice_create(cps_hip_lattice** lat, const int node_sites[4], int ls, const int bc[4]) {
  if (!lat || !node_sites) return CPS_HIP_ERR_ARG;
  *lat = nullptr;
  cps_hip_lattice* h = new (std::nothrow) cps_hip_lattice;
  if (!h) return CPS_HIP_ERR_NOMEM;
  int rc = h->L.init(node_sites, ls, bc);
  if (rc) {
    delete h;
    return rc;
  }
  *lat = h;
  return 0;
}

int cps_hip_lattice_destroy(cps_hip_lattice* lat) {
  if (!lat) return CPS_HIP_ERR_ARG;
  if (lat->L.stream) (void)hipStreamSynchronize(lat->L.stream);
  delete lat;
  return 0;
}

int cps_hip_lattice_set_twist(cps_hip_lattice* lat, const double twist_angle[4]) {
  if (check_lat(lat)) return CPS_HIP_ERR_ARG;
  return lat->L.set_twist(twist_angle);
}

int cps_hip_gauge_boundary_phase(const cps_hip_lattice* lat, double phase[8], int* recon_ok) {
  if (!lat || !phase) return CPS_HIP_ERR_ARG;
  if (!lat->L.has_gauge) return CPS_HIP_ERR_NOGAUGE;
  for (int mu = 0; mu < 4; mu++) {
    phase[2 * mu] = lat->L.link_phase[mu][0];
    phase[2 * mu + 1] = lat->L.link_phase[mu][1];
  }
  if (recon_ok) *recon_ok = lat->L.recon_ok ? 1 : 0;
  return 0;
}

int cps_hip_gauge_import(cps_hip_lattice* lat, const double* u) {
  if (check_lat(lat) || !u) return CPS_HIP_ERR_ARG;
  return lat->L.import_gauge(u);
}

long long cps_hip_lattice_device_bytes(const cps_hip_lattice* lat) { return lat ? lat->L.device_bytes() : 0; }

// ---- Wilson ------------------------------------------------------------
int cps_hip_wilson_dslash(cps_hip_lattice* lat, double* chi, const double* u, const double* psi, int cb, int dag) {
  if (check_lat(lat) || !chi || !psi) return CPS_HIP_ERR_ARG;
  if ((cb != 0 && cb != 1) || (dag != 0 && dag != 1)) return CPS_HIP_ERR_ARG;  // wilson_dslash.C:226-240
  int rc = maybe_import(lat, u);
  if (rc) return rc;
  cps_hip_op op = make_op(CPS_HIP_OP_WILSON, 0, 0, 1);
  return host_hop(lat, op, chi, psi, cb, dag);
}

static int wilson_m_common(cps_hip_lattice* lat, double* chi, const double* u, const double* psi, double kappa,
                           int mode, double* mp_sq) {
  if (check_lat(lat) || !chi || !psi) return CPS_HIP_ERR_ARG;
  int rc = maybe_import(lat, u);
  if (rc) return rc;
  Lattice& L = lat->L;
  // kappa -> mass so that op_kappa reproduces it: kappa = 1/(2(m+4))
  cps_hip_op op = make_op(CPS_HIP_OP_WILSON, 1.0 / (2.0 * kappa) - 4.0, 0, 1);
  OpLayout o;
  rc = op_layout(&op, L.ls, &o, L.g);
  if (rc) return rc;
  Field& fi = L.field(S_IO_IN, STORE_D, o.per_rhs);
  Field& fo = L.field(S_IO_OUT, STORE_D, o.per_rhs);
  Field& t1 = L.field(S_T1, STORE_D, o.per_rhs);
  Field& t2 = L.field(S_T2, STORE_D, o.per_rhs);
  if (L.last_rc) return L.last_rc;
  rc = upload(L, psi, fi, o, 1, 1, 0);
  if (rc) return rc;
  double* part = nullptr;
  int bpr = hop_blocks_per_rhs(L, &op);
  if (mp_sq) {
    rc = L.ensure_partial(bpr);
    if (rc) return rc;
    part = (double*)L.partial.p;
  }
  if (mode == 2) rc = apply_mdagm(L, &op, PREC_D_D18, fo.ref(), fi.ref(), t1.ref(), t2.ref(), part, 1);
  else rc = apply_m(L, &op, PREC_D_D18, fo.ref(), fi.ref(), t1.ref(), mode, nullptr, 1);
  if (rc) return rc;
  if (mp_sq) {
    CGScalars dummy;
    std::memset(&dummy, 0, sizeof(dummy));
    DevBuf outv;
    rc = outv.ensure(sizeof(double));
    if (rc) return rc;
    CPSHIP_CHECK(reduce_rhs(part, bpr, 1, (double*)outv.p, RED_PLAIN, dummy, L.stream));
    CPSHIP_CHECK(hipMemcpyAsync(mp_sq, outv.p, sizeof(double), hipMemcpyDeviceToHost, L.stream));
  }
  return download(L, fo, chi, o, 1, 1);
}

int cps_hip_wilson_m(cps_hip_lattice* lat, double* chi, const double* u, const double* psi, double kappa) {
  return wilson_m_common(lat, chi, u, psi, kappa, 0, nullptr);
}
int cps_hip_wilson_mdag(cps_hip_lattice* lat, double* chi, const double* u, const double* psi, double kappa) {
  return wilson_m_common(lat, chi, u, psi, kappa, 1, nullptr);
}
int cps_hip_wilson_mdagm(cps_hip_lattice* lat, double* chi, const double* u, const double* psi, double* mp_sq,
                         double kappa) {
  return wilson_m_common(lat, chi, u, psi, kappa, 2, mp_sq);
}

// ---- domain wall -----------------------------------------------------------
int cps_hip_dwf_dslash(cps_hip_lattice* lat, double* out, const double* u, const double* in, double mass, int cb,
                       int dag, double m5, double a5_inv) {
  if (check_lat(lat) || !out || !in) return CPS_HIP_ERR_ARG;
  if ((cb != 0 && cb != 1) || (dag != 0 && dag != 1)) return CPS_HIP_ERR_ARG;
  int rc = maybe_import(lat, u);
  if (rc) return rc;
  cps_hip_op op = make_op(CPS_HIP_OP_DWF, mass, m5, a5_inv);
  return host_hop(lat, op, out, in, cb, dag);
}

static int generic_apply(cps_hip_lattice* lat, const cps_hip_op* op, double* out, const double* in, int mode,
                         double* dot_prd, int nrhs) {
  // mode 0: MatPc, 1: MatPcDag, 2: MatPcDagMatPc
  if (check_lat(lat) || !op || !out || !in || nrhs < 1) return CPS_HIP_ERR_ARG;
  Lattice& L = lat->L;
  if (!L.has_gauge) return CPS_HIP_ERR_NOGAUGE;
  OpLayout o;
  int rc = op_layout(op, L.ls, &o, L.g);
  if (rc) return rc;
  const int64_t n = o.per_rhs * nrhs;
  Field& fi = L.field(S_IO_IN, STORE_D, n);
  Field& fo = L.field(S_IO_OUT, STORE_D, n);
  Field& t1 = L.field(S_T1, STORE_D, n);
  Field& t2 = L.field(S_T2, STORE_D, n);
  if (L.last_rc) return L.last_rc;
  // odd checkerboard (4-D parity 1, or 5-D cb 1) for every preconditioned operator
  rc = upload(L, in, fi, o, 1, nrhs, 0);
  if (rc) return rc;
  double* part = nullptr;
  int bpr = 1;
  if (dot_prd) {
    bpr = (op->kind == CPS_HIP_OP_MOBIUS) ? (int)(o.per_rhs / pick_block(o.per_rhs)) : hop_blocks_per_rhs(L, op);
    rc = L.ensure_partial((int64_t)bpr * nrhs);
    if (rc) return rc;
    part = (double*)L.partial.p;
  }
  if (mode == 2) rc = apply_mdagm(L, op, PREC_D_D18, fo.ref(), fi.ref(), t1.ref(), t2.ref(), part, nrhs);
  else rc = apply_matpc(L, op, PREC_D_D18, fo.ref(), fi.ref(), t1.ref(), t2.ref(), mode, nrhs);
  if (rc) return rc;
  if (dot_prd) {
    CGScalars dummy;
    std::memset(&dummy, 0, sizeof(dummy));
    DevBuf outv;
    rc = outv.ensure(nrhs * sizeof(double));
    if (rc) return rc;
    CPSHIP_CHECK(reduce_rhs(part, bpr, nrhs, (double*)outv.p, RED_PLAIN, dummy, L.stream));
    CPSHIP_CHECK(hipMemcpyAsync(dot_prd, outv.p, nrhs * sizeof(double), hipMemcpyDeviceToHost, L.stream));
  }
  return download(L, fo, out, o, 1, nrhs);
}

int cps_hip_dwf_m(cps_hip_lattice* lat, double* out, const double* u, const double* in, double mass, double m5,
                  double a5_inv) {
  if (check_lat(lat)) return CPS_HIP_ERR_ARG;
  int rc = maybe_import(lat, u);
  if (rc) return rc;
  cps_hip_op op = make_op(CPS_HIP_OP_DWF, mass, m5, a5_inv);
  return generic_apply(lat, &op, out, in, 0, nullptr, 1);
}
int cps_hip_dwf_mdag(cps_hip_lattice* lat, double* out, const double* u, const double* in, double mass, double m5,
                     double a5_inv) {
  if (check_lat(lat)) return CPS_HIP_ERR_ARG;
  int rc = maybe_import(lat, u);
  if (rc) return rc;
  cps_hip_op op = make_op(CPS_HIP_OP_DWF, mass, m5, a5_inv);
  return generic_apply(lat, &op, out, in, 1, nullptr, 1);
}
int cps_hip_dwf_mdagm(cps_hip_lattice* lat, double* out, const double* u, const double* in, double* dot_prd,
                      double mass, double m5, double a5_inv) {
  if (check_lat(lat)) return CPS_HIP_ERR_ARG;
  int rc = maybe_import(lat, u);
  if (rc) return rc;
  cps_hip_op op = make_op(CPS_HIP_OP_DWF, mass, m5, a5_inv);
  return generic_apply(lat, &op, out, in, 2, dot_prd, 1);
}

int cps_hip_matpc_dag_matpc(cps_hip_lattice* lat, const cps_hip_op* op, double* out, const double* in,
                            double* dot_prd, int nrhs) {
  return generic_apply(lat, op, out, in, 2, dot_prd, nrhs);
}

int cps_hip_matpc_dag_matpc_prec(cps_hip_lattice* lat, const cps_hip_op* op, int precision, double* out,
                                 const double* in, double* dot_prd, int nrhs) {
  if (check_lat(lat) || !op || !out || !in || nrhs < 1) return CPS_HIP_ERR_ARG;
  Lattice& L = lat->L;
  if (!L.has_gauge) return CPS_HIP_ERR_NOGAUGE;
  if (precision < CPS_HIP_PREC_DOUBLE || precision > CPS_HIP_PREC_MIXED_SINGLE) return CPS_HIP_ERR_ARG;
  int hprec, st;
  precision_formats(precision, op, L.ls, nrhs, L.recon_ok, &st, &hprec);
  OpLayout o;
  int rc = op_layout(op, L.ls, &o, L.g);
  if (rc) return rc;
  const int64_t n = o.per_rhs * nrhs;
  Field& fi = L.field(S_IO_IN, STORE_D, n);
  Field& fo = L.field(S_IO_OUT, STORE_D, n);
  Field& si = L.field(S_PS, st, n);
  Field& so = L.field(S_APS, st, n);
  Field& t1 = L.field(S_T1S, st, n);
  Field& t2 = L.field(S_T2S, st, n);
  if (L.last_rc) return L.last_rc;
  if ((rc = upload(L, in, fi, o, 1, nrhs, 0))) return rc;
  CPSHIP_CHECK(convert_spinor(STORE_D, fi.ref(), st, si.ref(), n, L.stream));
  const int bpr = (op->kind == CPS_HIP_OP_MOBIUS) ? (int)(o.per_rhs / pick_block(o.per_rhs)) : hop_blocks_per_rhs(L, op);
  if ((rc = L.ensure_partial((int64_t)bpr * nrhs))) return rc;
  double* part = (double*)L.partial.p;
  if ((rc = apply_mdagm(L, op, hprec, so.ref(), si.ref(), t1.ref(), t2.ref(), part, nrhs))) return rc;
  CPSHIP_CHECK(convert_spinor(st, so.ref(), STORE_D, fo.ref(), n, L.stream));
  if (dot_prd) {
    CGScalars dummy;
    std::memset(&dummy, 0, sizeof(dummy));
    DevBuf outv;
    if ((rc = outv.ensure(nrhs * sizeof(double)))) return rc;
    CPSHIP_CHECK(reduce_rhs(part, bpr, nrhs, (double*)outv.p, RED_PLAIN, dummy, L.stream));
    CPSHIP_CHECK(hipMemcpyAsync(dot_prd, outv.p, nrhs * sizeof(double), hipMemcpyDeviceToHost, L.stream));
    CPSHIP_CHECK(hipStreamSynchronize(L.stream));
  }
  return download(L, fo, out, o, 1, nrhs);
}

int cps_hip_matpc(cps_hip_lattice* lat, const cps_hip_op* op, double* out, const double* in, int dag, int nrhs) {
  if (dag != 0 && dag != 1) return CPS_HIP_ERR_ARG;
  return generic_apply(lat, op, out, in, dag, nullptr, nrhs);
}

int cps_hip_clover_build(cps_hip_lattice* lat, const double* kappas, int n_kappa, double csw) {
  if (check_lat(lat) || !kappas || n_kappa < 1) return CPS_HIP_ERR_ARG;
  if (!lat->L.has_gauge) return CPS_HIP_ERR_NOGAUGE;
  return clover_build(lat->L, kappas, n_kappa, csw);
}

int cps_hip_clover_export(cps_hip_lattice* lat, int k, int cb, double* A) {
  if (check_lat(lat) || !A || k < 0 || k >= lat->L.n_clover || (cb != 0 && cb != 1)) return CPS_HIP_ERR_ARG;
  Lattice& L = lat->L;
  CPSHIP_CHECK(hipStreamSynchronize(L.stream));
  return clover_export(L, k, cb, A);
}

// ---- InvCg -----------------------------------------------------------------
static int invcg_common(cps_hip_lattice* lat, const cps_hip_op* op, double* sol, const double* src, int nrhs,
                        const cps_hip_cg_arg* arg, cps_hip_cg_result* res, bool on_device) {
  if (check_lat(lat) || !op || !sol || !src || !arg || nrhs < 1) return CPS_HIP_ERR_ARG;
  Lattice& L = lat->L;
  if (!L.has_gauge) return CPS_HIP_ERR_NOGAUGE;
  OpLayout o;
  int rc = op_layout(op, L.ls, &o, L.g);
  if (rc) return rc;
  const int64_t n = o.per_rhs * nrhs;
  Field& x = L.field(S_X, STORE_D, n);
  Field& b = L.field(S_SRC, STORE_D, n);
  if (L.last_rc) return L.last_rc;
  rc = upload(L, src, b, o, 1, nrhs, 0, on_device);
  if (rc) return rc;
  rc = upload(L, sol, x, o, 1, nrhs, 1, on_device);
  if (rc) return rc;
  CPSHIP_CHECK(hipStreamSynchronize(L.stream));
  std::vector<int> iters, status, nup;
  std::vector<double> tr, sn;
  auto t0 = std::chrono::steady_clock::now();
  if (arg->precision == CPS_HIP_PREC_DOUBLE) rc = cg_double_device(L, op, x, b, nrhs, arg, iters, status, tr, sn);
  else rc = cg_mixed_device(L, op, x, b, nrhs, arg, iters, status, tr, sn, nup);
  if (rc) return rc;
  CPSHIP_CHECK(hipStreamSynchronize(L.stream));
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  rc = download(L, x, sol, o, 1, nrhs, on_device);
  if (rc) return rc;
  bool dzero = false;
  for (int k = 0; k < nrhs; k++) {
    if (res) {
      res[k].iterations = iters[k] + 1;
      res[k].status = status[k];
      res[k].true_res = tr[k];
      res[k].src_norm_sq = sn[k];
      res[k].reliable_updates = nup.empty() ? 0 : nup[k];
      res[k].seconds = secs;
    }
    if (status[k] == CG_STATUS_D_ZERO) dzero = true;
  }
  return dzero ? CPS_HIP_ERR_D_ZERO : 0;
}

int cps_hip_invcg(cps_hip_lattice* lat, const cps_hip_op* op, double* sol, const double* src, int nrhs,
                  const cps_hip_cg_arg* arg, cps_hip_cg_result* res) {
  return invcg_common(lat, op, sol, src, nrhs, arg, res, false);
}

int cps_hip_invcg_device(cps_hip_lattice* lat, const cps_hip_op* op, double* sol, const double* src, int nrhs,
                         const cps_hip_cg_arg* arg, cps_hip_cg_result* res) {
  return invcg_common(lat, op, sol, src, nrhs, arg, res, true);
}

// ---- MatInv ----------------------------------------------------------------
// DiracOpWilson::MatInv (d_op_wilson.C:257-334) / DiracOpDwf::MatInv
// (d_op_dwf.C:316-509): temp = in_o + kappa D in_e ; src = MatPcDag temp ;
// InvCg ; out_e = in_e + kappa D out_o.  DiracOpClover::MatInv and
// DiracOpMobius::MatInv (ZMOB_PC_ORIG, d_op_mobius.C:389-640) add the clover
// inverse / M5^-1 steps (clover.hip, mobius.hip).  Full vectors are
// [odd][even] per rhs, except Moebius, which uses DWF_4D_EOPREC_EE
// ([even][odd], d_op_mobius.C:410-415).  The source is preserved.
int cps_hip_matinv(cps_hip_lattice* lat, const cps_hip_op* op, double* out, const double* in, int nrhs,
                   const cps_hip_cg_arg* arg, cps_hip_cg_result* res) {
  if (check_lat(lat) || !op || !out || !in || !arg || nrhs < 1) return CPS_HIP_ERR_ARG;
  Lattice& L = lat->L;
  if (!L.has_gauge) return CPS_HIP_ERR_NOGAUGE;
  OpLayout o;
  int rc = op_layout(op, L.ls, &o, L.g);
  if (rc) return rc;
  const bool mob = op->kind == CPS_HIP_OP_MOBIUS;
  const int blk_odd = mob ? 1 : 0, blk_even = mob ? 0 : 1;   // block order of the full vector
  const int64_t half = o.per_rhs * 24;   // reals of one checkerboard of one rhs
  const int64_t n = o.per_rhs * nrhs;
  std::vector<double> h_odd(half * nrhs), h_even(half * nrhs), h_x(half * nrhs);
  for (int k = 0; k < nrhs; k++) {
    std::memcpy(&h_odd[k * half], in + (2 * k + blk_odd) * half, half * sizeof(double));
    std::memcpy(&h_even[k * half], in + (2 * k + blk_even) * half, half * sizeof(double));
    std::memcpy(&h_x[k * half], out + (2 * k + blk_odd) * half, half * sizeof(double));
  }
  Field& fo = L.field(S_IO_IN, STORE_D, n);
  Field& fe = L.field(S_EVEN_IN, STORE_D, n);
  Field& x = L.field(S_X, STORE_D, n);
  Field& eo = L.field(S_EVEN_OUT, STORE_D, n);
  if (L.last_rc) return L.last_rc;
  if ((rc = upload(L, h_odd.data(), fo, o, 1, nrhs, 0))) return rc;
  if ((rc = upload(L, h_even.data(), fe, o, 0, nrhs, 1))) return rc;
  CPSHIP_CHECK(hipStreamSynchronize(L.stream));
  if (mob) {
    // the CG solves the 2 kappa_b-normalised system: scale the initial guess
    const double g = op->mobius_b * (4 - op->m5) + op->a5_inv;
    for (double& v : h_x) v *= g;
  }
  if ((rc = upload(L, h_x.data(), x, o, 1, nrhs, 0))) return rc;
  if ((rc = matinv_device(L, op, fo, fe, x, eo, nrhs, arg, res))) return rc;
  if ((rc = download(L, x, h_x.data(), o, 1, nrhs))) return rc;
  if ((rc = download(L, eo, h_even.data(), o, 0, nrhs))) return rc;
  // Moebius: the preconditioned system is 2 kappa_b times the operator
  // (d_op_mobius.C:628-633)
  const double fac = mob ? 1.0 / (op->mobius_b * (4 - op->m5) + op->a5_inv) : 1.0;
  for (int k = 0; k < nrhs; k++) {
    double* po = out + (2 * k + blk_odd) * half;
    double* pe = out + (2 * k + blk_even) * half;
    for (int64_t q = 0; q < half; q++) {
      po[q] = fac * h_x[k * half + q];
      pe[q] = fac * h_even[k * half + q];
    }
  }
  return 0;
}

// ---- measurement -------------------------------------------------------------
int cps_hip_bench_copy(cps_hip_lattice* lat, long long n_elem, int reps, double* seconds) {
  if (check_lat(lat) || n_elem < 1 || reps < 1) return CPS_HIP_ERR_ARG;
  Lattice& L = lat->L;
  Field& a = L.field(S_BENCH_IN, STORE_D, n_elem);
  Field& b = L.field(S_BENCH_OUT, STORE_D, n_elem);
  if (L.last_rc) return L.last_rc;
  CPSHIP_CHECK(fill_spinor(STORE_D, a.ref(), n_elem, 3u, L.stream));
  CPSHIP_CHECK(convert_spinor(STORE_D, a.ref(), STORE_D, b.ref(), n_elem, L.stream));
  hipEvent_t e0, e1;
  CPSHIP_CHECK(hipEventCreate(&e0));
  CPSHIP_CHECK(hipEventCreate(&e1));
  CPSHIP_CHECK(hipEventRecord(e0, L.stream));
  for (int r = 0; r < reps; r++) CPSHIP_CHECK(convert_spinor(STORE_D, a.ref(), STORE_D, b.ref(), n_elem, L.stream));
  CPSHIP_CHECK(hipEventRecord(e1, L.stream));
  CPSHIP_CHECK(hipEventSynchronize(e1));
  float ms = 0;
  CPSHIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (seconds) *seconds = ms * 1e-3 / reps;
  return 0;
}

int cps_hip_bench_hop(cps_hip_lattice* lat, int kind, int precision, int nrhs, int reps, double* seconds,
                      double* bytes) {
  if (check_lat(lat) || nrhs < 1 || reps < 1) return CPS_HIP_ERR_ARG;
  Lattice& L = lat->L;
  if (!L.has_gauge) return CPS_HIP_ERR_NOGAUGE;
  cps_hip_op op;
  std::memset(&op, 0, sizeof(op));
  op.kind = kind;
  op.mass = 0.01;
  op.m5 = 1.8;
  op.a5_inv = 1.0;
  OpLayout o;
  int rc = op_layout(&op, L.ls, &o, L.g);
  if (rc) return rc;
  if (precision < CPS_HIP_PREC_DOUBLE || precision > CPS_HIP_PREC_MIXED_SINGLE) return CPS_HIP_ERR_ARG;
  int hprec, st;
  precision_formats(precision, &op, L.ls, nrhs, L.recon_ok, &st, &hprec);
  const int64_t n = o.per_rhs * nrhs;
  Field& fi = L.field(S_BENCH_IN, st, n);
  Field& fo = L.field(S_BENCH_OUT, st, n);
  if (L.last_rc) return L.last_rc;
  CPSHIP_CHECK(fill_spinor(st, fi.ref(), n, 7u, L.stream));
  // the Moebius 4-D hop is the Wilson kernel with the Ls slices as batch
  cps_hip_op hop_op = op;
  if (kind == CPS_HIP_OP_CLOVER) hop_op.kind = CPS_HIP_OP_WILSON;
  for (int w = 0; w < 2; w++)
    if ((rc = hop(L, &hop_op, hprec, fo.ref(), fi.ref(), 1, 0, 1.0, null_field(), 0.0, nullptr, nrhs))) return rc;
  hipEvent_t e0, e1;
  CPSHIP_CHECK(hipEventCreate(&e0));
  CPSHIP_CHECK(hipEventCreate(&e1));
  CPSHIP_CHECK(hipEventRecord(e0, L.stream));
  for (int r = 0; r < reps; r++)
    if ((rc = hop(L, &hop_op, hprec, fo.ref(), fi.ref(), 1, 0, 1.0, null_field(), 0.0, nullptr, nrhs))) return rc;
  CPSHIP_CHECK(hipEventRecord(e1, L.stream));
  CPSHIP_CHECK(hipEventSynchronize(e1));
  float ms = 0;
  CPSHIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (seconds) *seconds = ms * 1e-3 / reps;
  if (bytes) {
    // compulsory traffic: every link once (4V links), each spinor read and
    // written once per rhs (compulsory-traffic model)
    const double V = 2.0 * L.g.Vcb;
    const double lb = (double)link_bytes(link_storage_of(hprec));
    const double sb = st == STORE_D ? 24 * 8 : st == STORE_F ? 24 * 4 : 24 * 2 + 4;
    *bytes = 4 * V * lb + (double)n * 2 * sb;
  }
  return 0;
}

}  // extern "C"
