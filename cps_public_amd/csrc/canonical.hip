// CPS CANONICAL order on the device: the gauge import that replaces
// Lattice::Convert(WILSON) + BondCond, and the fermion conversion of the
// DiracOp constructors / destructors with CNV_FRM_YES, so a caller holding
// CANONICAL fields never runs CPS's host-side conversions.
//
// CANONICAL orders (convert_func.C:318-494, convert.C:444-760,
// lattice_base.C:5350-5351):
//   gauge   [t][z][y][x][mu][row][col][re,im]
//   fermion [t][z][y][x][spin][colour][re,im]; 5-D fields are Ls such 4-D
//           fields, s-major ([s][t][z][y][x]...).
#include "solver.h"

#include <chrono>
#include <cstring>

namespace cpship {

namespace {

// One thread per (canonical site, mu): U_mu(x) to the WILSON array
// [parity][site_cb][mu][col][row][re,im] (the transpose of convert_func.C:
// 402-404), times the boundary factor g.bf[mu] on the last slice of the
// direction: -1 antiperiodic, the twist phase e^{-i pi theta} twisted
// (Lattice::BondCond, lattice_base.C:5313-5369; GJP::TwistPhase, gjp.C:475).
__global__ void k_gauge_canon_to_wilson(const double* __restrict__ canon, double* __restrict__ wil, Geom g) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t V = 2 * (int64_t)g.Vcb;
  if (t >= 4 * V) return;
  const int mu = (int)(t & 3);
  const int64_t site = t >> 2;
  int c[4];
  int64_t r = site;
#pragma unroll
  for (int d = 0; d < 4; d++) {
    c[d] = (int)(r % g.L[d]);
    r /= g.L[d];
  }
  const int p = (c[0] + c[1] + c[2] + c[3]) & 1;
  const int site_cb = (c[0] >> 1) + (g.L[0] >> 1) * (c[1] + g.L[1] * (c[2] + g.L[2] * c[3]));
  const bool last = c[mu] == g.L[mu] - 1;
  const double fr = last ? g.bf[mu][0] : 1.0, fi = last ? g.bf[mu][1] : 0.0;
  const double* src = canon + 72 * site + 18 * mu;
  double* dst = wil + 72 * ((int64_t)p * g.Vcb + site_cb) + 18 * mu;
#pragma unroll
  for (int row = 0; row < 3; row++)
#pragma unroll
    for (int col = 0; col < 3; col++) {
      const double a = src[2 * (3 * row + col)], b = src[2 * (3 * row + col) + 1];
      // products rounded separately (no fused multiply-add), as a host
      // BondCond without contraction computes the complex product
      {
#pragma clang fp contract(off)
        dst[2 * (row + 3 * col)] = fr * a - fi * b;
        dst[2 * (row + 3 * col) + 1] = fr * b + fi * a;
      }
    }
}

// One thread per (rhs, s, canonical site): the spinor between the CANONICAL
// vector and the device field that holds its checkerboard (FP64 fields).
template <int KIND, bool PACK>
__global__ void k_ferm_canon(double* __restrict__ canon, double2* __restrict__ fo, double2* __restrict__ fe,
                             int64_t plane, Geom g, int ls, int64_t per_rhs, int64_t n, double scale) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int64_t V = 2 * (int64_t)g.Vcb;
  const int64_t site = t % V;
  const int64_t rs = t / V;          // rhs * ls + s
  const int s = (int)(rs % ls);
  const int64_t rhs = rs / ls;
  int c[4];
  int64_t r = site;
#pragma unroll
  for (int d = 0; d < 4; d++) {
    c[d] = (int)(r % g.L[d]);
    r /= g.L[d];
  }
  const int p = (c[0] + c[1] + c[2] + c[3]) & 1;
  const int site_cb = (c[0] >> 1) + (g.L[0] >> 1) * (c[1] + g.L[1] * (c[2] + g.L[2] * c[3]));
  bool odd;
  int64_t e;
  field_slot<KIND>(p, site_cb, s, g.Vcb, ls, &odd, &e);
  double2* f = (odd ? fo : fe) + rhs * per_rhs + e;
  double* v = canon + 24 * t;
#pragma unroll
  for (int q = 0; q < 12; q++) {
    if (PACK) {
      f[q * plane] = make_double2(scale * v[2 * q], scale * v[2 * q + 1]);
    } else {
      const double2 w = f[q * plane];
      v[2 * q] = scale * w.x;
      v[2 * q + 1] = scale * w.y;
    }
  }
}

template <bool PACK>
int launch_ferm_canon(Lattice& L, int kind, double* canon, const Field& fo, const Field& fe, int nrhs, double scale) {
  OpLayout o;
  cps_hip_op op;
  std::memset(&op, 0, sizeof(op));
  op.kind = kind;
  int rc = op_layout(&op, L.ls, &o, L.g);
  if (rc) return rc;
  const int ls = (kind == CPS_HIP_OP_DWF || kind == CPS_HIP_OP_MOBIUS) ? L.ls : 1;
  const int64_t n = (int64_t)2 * L.g.Vcb * ls * nrhs;
  if (fo.nelem < o.per_rhs * nrhs || fe.nelem < o.per_rhs * nrhs || fo.nelem != fe.nelem) return CPS_HIP_ERR_ARG;
  const dim3 grid((unsigned)((n + 255) / 256)), block(256);
  double2* po = (double2*)fo.buf.p;
  double2* pe = (double2*)fe.buf.p;
  switch (kind) {
    case CPS_HIP_OP_DWF:
      hipLaunchKernelGGL((k_ferm_canon<CPS_HIP_OP_DWF, PACK>), grid, block, 0, L.stream, canon, po, pe, fo.nelem, L.g,
                         ls, o.per_rhs, n, scale);
      break;
    case CPS_HIP_OP_MOBIUS:
      hipLaunchKernelGGL((k_ferm_canon<CPS_HIP_OP_MOBIUS, PACK>), grid, block, 0, L.stream, canon, po, pe, fo.nelem,
                         L.g, ls, o.per_rhs, n, scale);
      break;
    default:
      hipLaunchKernelGGL((k_ferm_canon<CPS_HIP_OP_WILSON, PACK>), grid, block, 0, L.stream, canon, po, pe, fo.nelem,
                         L.g, ls, o.per_rhs, n, scale);
  }
  CPSHIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace

int pack_canonical(Lattice& L, int kind, const double* canon, Field& f_odd, Field& f_even, int nrhs, double scale) {
  return launch_ferm_canon<true>(L, kind, const_cast<double*>(canon), f_odd, f_even, nrhs, scale);
}

int unpack_canonical(Lattice& L, int kind, const Field& f_odd, const Field& f_even, double* canon, int nrhs,
                     double scale) {
  return launch_ferm_canon<false>(L, kind, canon, f_odd, f_even, nrhs, scale);
}

int Lattice::import_gauge_canonical(const double* u_canon) {
  const size_t n = (size_t)72 * 2 * g.Vcb;
  int rc = io.ensure(n * sizeof(double));
  if (rc) return rc;
  if ((rc = io2.ensure(n * sizeof(double)))) return rc;
  CPSHIP_CHECK(hipMemcpyAsync(io.p, u_canon, n * sizeof(double), hipMemcpyHostToDevice, stream));
  const int64_t nt = (int64_t)8 * g.Vcb;
  hipLaunchKernelGGL(k_gauge_canon_to_wilson, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, stream,
                     (const double*)io.p, (double*)io2.p, g);
  CPSHIP_CHECK(hipGetLastError());
  return import_gauge_from((const double*)io2.p, hipMemcpyDeviceToDevice);
}

}  // namespace cpship

using namespace cpship;

extern "C" {

int cps_hip_gauge_import_canonical(cps_hip_lattice* lat, const double* u_canon) {
  if (!lat || !u_canon) return CPS_HIP_ERR_ARG;
  lat->L.last_rc = 0;
  return lat->L.import_gauge_canonical(u_canon);
}

int cps_hip_fmatinv(cps_hip_lattice* lat, const cps_hip_op* op, double* out, const double* in, int nrhs,
                    const cps_hip_cg_arg* arg, cps_hip_cg_result* res) {
  if (!lat || !op || !out || !in || !arg || nrhs < 1) return CPS_HIP_ERR_ARG;
  Lattice& L = lat->L;
  L.last_rc = 0;
  if (!L.has_gauge) return CPS_HIP_ERR_NOGAUGE;
  OpLayout o;
  int rc = op_layout(op, L.ls, &o, L.g);
  if (rc) return rc;
  const int64_t n = o.per_rhs * nrhs;
  const size_t bytes = (size_t)2 * n * 24 * sizeof(double);   // both checkerboards
  Field& fo = L.field(S_IO_IN, STORE_D, n);
  Field& fe = L.field(S_EVEN_IN, STORE_D, n);
  Field& x = L.field(S_X, STORE_D, n);
  Field& eo = L.field(S_EVEN_OUT, STORE_D, n);
  if (L.last_rc) return L.last_rc;
  if ((rc = L.io.ensure(bytes)) || (rc = L.io2.ensure(bytes))) return rc;
  const bool mob = op->kind == CPS_HIP_OP_MOBIUS;
  // the Moebius CG solves the 2 kappa_b-normalised system
  const double g = mob ? op->mobius_b * (4 - op->m5) + op->a5_inv : 1.0;
  CPSHIP_CHECK(hipMemcpyAsync(L.io.p, in, bytes, hipMemcpyHostToDevice, L.stream));
  CPSHIP_CHECK(hipMemcpyAsync(L.io2.p, out, bytes, hipMemcpyHostToDevice, L.stream));
  if ((rc = pack_canonical(L, op->kind, (const double*)L.io.p, fo, fe, nrhs, 1.0))) return rc;
  // the initial guess: its odd half starts the CG (the even half is rebuilt)
  if ((rc = pack_canonical(L, op->kind, (const double*)L.io2.p, x, eo, nrhs, g))) return rc;
  if (mob && (rc = mobius_dminus_fields(L, op, fo, fe, nrhs))) return rc;
  if ((rc = matinv_device(L, op, fo, fe, x, eo, nrhs, arg, res))) return rc;
  if ((rc = unpack_canonical(L, op->kind, x, eo, (double*)L.io2.p, nrhs, 1.0 / g))) return rc;
  CPSHIP_CHECK(hipMemcpyAsync(out, L.io2.p, bytes, hipMemcpyDeviceToHost, L.stream));
  CPSHIP_CHECK(hipStreamSynchronize(L.stream));
  return 0;
}

}  // extern "C"
