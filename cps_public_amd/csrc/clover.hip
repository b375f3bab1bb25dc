// Clover-improved Wilson operator: the clover field is built on the device
// from the resident gauge field, and the preconditioned operator
//   M_pc = A_oo - kappa^2 D_oe A_ee^-1 D_eo     (d_op_clover.C:194-248)
// runs as two hops whose epilogues apply the clover matrices (hop_clover.hip).
//
// Clover matrix (include/util/dirac_op.h:930-951):
//   A(x) = 1 - 1/2 kappa c_sw sum_{mu<nu} [gamma_mu, gamma_nu] F_munu(x),
//   F_munu = 1/8 (P_munu - P_munu^dag), P_munu the sum of the four plaquette
//   leaves around x in the (mu, nu) plane.
// In the chiral basis [gamma_mu, gamma_nu] is block diagonal (it commutes
// with gamma_5), so A is two 6x6 Hermitian blocks per site, stored in the
// packing of clover_mat_mlt (cps_hip_internal.h).  Even sites hold A^-1
// (CloverMatChkb with the inverse flag, d_op_clover_supp.C:492-540), here
// computed by a Cholesky factorisation in FP64.
#include "hop_kernel.h"
#include "solver.h"

namespace cpship {

namespace {

typedef cplx<double> zd;

__device__ __forceinline__ void m3_mul(const zd a[9], const zd b[9], zd c[9]) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      zd s = cmul(a[3 * i], b[j]);
      s = cadd(s, cmul(a[3 * i + 1], b[3 + j]));
      c[3 * i + j] = cadd(s, cmul(a[3 * i + 2], b[6 + j]));
    }
}

__device__ __forceinline__ void m3_dag(const zd a[9], zd c[9]) {
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) c[3 * i + j] = cmk<double>(a[3 * j + i].re, -a[3 * j + i].im);
}

// U_mu at coordinates x + shift (periodic wrap); the boundary signs already
// in the links cancel around every closed leaf.
__device__ __forceinline__ void link_at(const GaugeRef& u, const Geom& g, const int x[4], const int sh[4], int mu,
                                        zd m[9]) {
  int y[4];
#pragma unroll
  for (int d = 0; d < 4; d++) {
    int v = x[d] + sh[d];
    v = v < 0 ? v + g.L[d] : (v >= g.L[d] ? v - g.L[d] : v);
    y[d] = v;
  }
  const int par = (y[0] + y[1] + y[2] + y[3]) & 1;
  link_load<LinkD18, double>(u, par, mu, cb_index(g, y), g.Vcb, make_float2(1.f, 0.f), m);
}

// 2x2 spin block (chirality beta: 0 spins 0,1; 1 spins 2,3) of
// [gamma_mu, gamma_nu] for the plane index pl = (01,02,03,12,13,23) in the
// chiral basis of dirac_op.h:955-978, as {re, im} of entries (0,0),(0,1),
// (1,0),(1,1).
__device__ __forceinline__ void commutator_block(int pl, int beta, double c[8]) {
  const double s = beta == 0 ? 1.0 : -1.0;
#pragma unroll
  for (int k = 0; k < 8; k++) c[k] = 0.0;
  switch (pl) {
    case 0: c[1] = -2; c[7] = 2; break;                  // (x,y): diag(-2i, 2i)
    case 1: c[2] = -2; c[4] = 2; break;                  // (x,z): [[0,-2],[2,0]]
    case 2: c[3] = 2 * s; c[5] = 2 * s; break;           // (x,t): +-[[0,2i],[2i,0]]
    case 3: c[3] = -2; c[5] = -2; break;                 // (y,z): [[0,-2i],[-2i,0]]
    case 4: c[2] = -2 * s; c[4] = 2 * s; break;          // (y,t): +-[[0,-2],[2,0]]
    default: c[1] = 2 * s; c[7] = -2 * s; break;         // (z,t): +-diag(2i,-2i)
  }
}

// One thread per site of the full lattice (e = parity*Vcb + site_cb):
// acc[e*72 ..] receives the packed A of that site.
__global__ void k_clover_build(GaugeRef u, Geom g, double omega, double* __restrict__ acc) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 2 * (int64_t)g.Vcb) return;
  const int parity = (int)(e / g.Vcb);
  const int site = (int)(e - (int64_t)parity * g.Vcb);
  int x[4];
  {
    const int lh = g.L[0] >> 1;
    int r = site / lh;
    const int xh = site - r * lh;
    x[1] = r % g.L[1];
    r /= g.L[1];
    x[2] = r % g.L[2];
    x[3] = r / g.L[2];
    x[0] = 2 * xh + ((x[1] + x[2] + x[3] + parity) & 1);
  }
  double* a = acc + e * 72;
  for (int k = 0; k < 72; k++) a[k] = 0.0;
  for (int b = 0; b < 2; b++)
    for (int i = 0; i < 6; i++) a[36 * b + i * i + 2 * i] = 1.0;
  int pl = 0;
  for (int mu = 0; mu < 4; mu++)
    for (int nu = mu + 1; nu < 4; nu++, pl++) {
      int z[4] = {0, 0, 0, 0}, em[4] = {0, 0, 0, 0}, en[4] = {0, 0, 0, 0};
      em[mu] = 1;
      en[nu] = 1;
      int mm[4], mn[4], mmn[4], mpn[4], pmn[4];   // -mu, -nu, -mu-nu, -mu+nu, mu-nu
      for (int d = 0; d < 4; d++) {
        mm[d] = -em[d];
        mn[d] = -en[d];
        mmn[d] = -em[d] - en[d];
        mpn[d] = -em[d] + en[d];
        pmn[d] = em[d] - en[d];
      }
      // link (or its adjoint) at x + sh in direction m
      auto get = [&](const int sh[4], int m, bool dg, zd out[9]) {
        if (dg) {
          zd t[9];
          link_at(u, g, x, sh, m, t);
          m3_dag(t, out);
        } else {
          link_at(u, g, x, sh, m, out);
        }
      };
      // P += M1 M2 M3 M4
      zd P[9], M1[9], M2[9], M3[9], M4[9], T1[9], T2[9], Lf[9];
      for (int k = 0; k < 9; k++) P[k] = cmk<double>(0.0, 0.0);
      auto leaf = [&]() {
        m3_mul(M1, M2, T1);
        m3_mul(M3, M4, T2);
        m3_mul(T1, T2, Lf);
        for (int k = 0; k < 9; k++) P[k] = cadd(P[k], Lf[k]);
      };
      // U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag
      get(z, mu, false, M1); get(em, nu, false, M2); get(en, mu, true, M3); get(z, nu, true, M4); leaf();
      // U_nu(x) U_mu(x-mu+nu)^dag U_nu(x-mu)^dag U_mu(x-mu)
      get(z, nu, false, M1); get(mpn, mu, true, M2); get(mm, nu, true, M3); get(mm, mu, false, M4); leaf();
      // U_mu(x-mu)^dag U_nu(x-mu-nu)^dag U_mu(x-mu-nu) U_nu(x-nu)
      get(mm, mu, true, M1); get(mmn, nu, true, M2); get(mmn, mu, false, M3); get(mn, nu, false, M4); leaf();
      // U_nu(x-nu)^dag U_mu(x-nu) U_nu(x+mu-nu) U_mu(x)^dag
      get(mn, nu, true, M1); get(mn, mu, false, M2); get(pmn, nu, false, M3); get(z, mu, true, M4); leaf();
      // F = (P - P^dag) / 8
      zd F[9];
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
          F[3 * i + j] = cmk<double>((P[3 * i + j].re - P[3 * j + i].re) * 0.125,
                                     (P[3 * i + j].im + P[3 * j + i].im) * 0.125);
      // A_beta[(s,c),(s',c')] -= omega/2 C_beta[s][s'] F[c][c'], lower triangle
      for (int b = 0; b < 2; b++) {
        double c[8];
        commutator_block(pl, b, c);
        for (int s = 0; s < 2; s++)
          for (int sp = 0; sp <= s; sp++) {
            const double cr = c[2 * (2 * s + sp)], ci = c[2 * (2 * s + sp) + 1];
            if (cr == 0.0 && ci == 0.0) continue;
            for (int cc = 0; cc < 3; cc++)
              for (int cp = 0; cp < 3; cp++) {
                const int i = 3 * s + cc, j = 3 * sp + cp;
                if (j > i) continue;
                const zd f = F[3 * cc + cp];
                const double vr = -0.5 * omega * (cr * f.re - ci * f.im);
                const double vi = -0.5 * omega * (cr * f.im + ci * f.re);
                if (i == j) {
                  a[36 * b + i * i + 2 * i] += vr;
                } else {
                  a[36 * b + i * i + 2 * j] += vr;
                  a[36 * b + i * i + 2 * j + 1] += vi;
                }
              }
          }
      }
    }
}

// One thread per (site, chiral block): even sites are inverted (Cholesky
// A = L L^dag, A^-1 = L^-dag L^-1), then the block goes to the FP64 SoA set
// of (kappa index k, checkerboard = parity).
__global__ void k_clover_pack(const double* __restrict__ acc, double2* __restrict__ out, int Vcb, int k) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4 * (int64_t)Vcb) return;
  const int b = (int)(t & 1);
  const int64_t e = t >> 1;
  const int parity = (int)(e / Vcb);
  const int site = (int)(e - (int64_t)parity * Vcb);
  const double* a = acc + e * 72 + 36 * b;
  double r[36];
  for (int q = 0; q < 36; q++) r[q] = a[q];
  if (parity == 0) {
    zd Lm[21];   // lower triangle incl. diagonal, (i,j) at i*(i+1)/2 + j
    auto A = [&](int i, int j) { return i == j ? cmk<double>(r[i * i + 2 * i], 0.0)
                                               : cmk<double>(r[i * i + 2 * j], r[i * i + 2 * j + 1]); };
    for (int j = 0; j < 6; j++) {
      double d = A(j, j).re;
      for (int q = 0; q < j; q++) {
        const zd l = Lm[j * (j + 1) / 2 + q];
        d -= l.re * l.re + l.im * l.im;
      }
      const double ljj = sqrt(d);
      Lm[j * (j + 1) / 2 + j] = cmk<double>(ljj, 0.0);
      for (int i = j + 1; i < 6; i++) {
        zd s = A(i, j);
        for (int q = 0; q < j; q++) s = csub(s, cmul(Lm[i * (i + 1) / 2 + q], cconj(Lm[j * (j + 1) / 2 + q])));
        Lm[i * (i + 1) / 2 + j] = cmk<double>(s.re / ljj, s.im / ljj);
      }
    }
    zd W[21];    // W = L^-1, lower
    for (int j = 0; j < 6; j++) {
      W[j * (j + 1) / 2 + j] = cmk<double>(1.0 / Lm[j * (j + 1) / 2 + j].re, 0.0);
      for (int i = j + 1; i < 6; i++) {
        zd s = cmk<double>(0.0, 0.0);
        for (int q = j; q < i; q++) s = cadd(s, cmul(Lm[i * (i + 1) / 2 + q], W[q * (q + 1) / 2 + j]));
        const double lii = Lm[i * (i + 1) / 2 + i].re;
        W[i * (i + 1) / 2 + j] = cmk<double>(-s.re / lii, -s.im / lii);
      }
    }
    for (int i = 0; i < 6; i++)
      for (int j = 0; j <= i; j++) {
        zd s = cmk<double>(0.0, 0.0);
        for (int q = i; q < 6; q++) s = cadd(s, cmul(cconj(W[q * (q + 1) / 2 + i]), W[q * (q + 1) / 2 + j]));
        if (i == j) {
          r[i * i + 2 * i] = s.re;
        } else {
          r[i * i + 2 * j] = s.re;
          r[i * i + 2 * j + 1] = s.im;
        }
      }
  }
  double2* p = out + ((int64_t)(k * 2 + parity) * 36 + 18 * b) * Vcb + site;
  for (int q = 0; q < 18; q++) p[(int64_t)q * Vcb] = make_double2(r[2 * q], r[2 * q + 1]);
}

// FP64 SoA clover sets -> FP32 (18 planes of float4) or FP16 (9 planes of
// 8 halves); one thread per site of one set.
template <int CS>
__global__ void k_clover_convert(const double2* __restrict__ in, void* __restrict__ out, int64_t nsets, int Vcb) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nsets * Vcb) return;
  const int64_t set = t / Vcb;
  const int site = (int)(t - set * Vcb);
  const double2* p = in + set * 36 * Vcb + site;
  double r[72];
  for (int q = 0; q < 36; q++) {
    const double2 h = p[(int64_t)q * Vcb];
    r[2 * q] = h.x;
    r[2 * q + 1] = h.y;
  }
  if (CS == CLOVER_F) {
    float4* o = reinterpret_cast<float4*>(out) + set * 18 * Vcb + site;
    for (int q = 0; q < 18; q++)
      o[(int64_t)q * Vcb] = make_float4(float(r[4 * q]), float(r[4 * q + 1]), float(r[4 * q + 2]), float(r[4 * q + 3]));
  } else {
    uint4* o = reinterpret_cast<uint4*>(out) + set * 9 * Vcb + site;
    for (int q = 0; q < 9; q++) {
      const double* v = r + 8 * q;
      o[(int64_t)q * Vcb] = make_uint4(f2h(float(v[0]), float(v[1])), f2h(float(v[2]), float(v[3])),
                                       f2h(float(v[4]), float(v[5])), f2h(float(v[6]), float(v[7])));
    }
  }
}

int nblk(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

int storage_of_hprec(int hprec) {
  return hprec == PREC_D_D18 ? CLOVER_D : hprec == PREC_F_F12 ? CLOVER_F : CLOVER_H;
}

int ensure_clover(Lattice& L, int cs) {
  if (L.n_clover == 0) return CPS_HIP_ERR_ARG;
  if (L.clover_valid[cs]) return 0;
  const int64_t nsets = 2 * (int64_t)L.n_clover;
  int rc = L.clover[cs].ensure((size_t)nsets * clover_nplanes(cs) * L.g.Vcb * 16);
  if (rc) return rc;
  const int64_t n = nsets * L.g.Vcb;
  if (cs == CLOVER_F)
    hipLaunchKernelGGL(k_clover_convert<CLOVER_F>, dim3(nblk(n, 128)), dim3(128), 0, L.stream,
                       (const double2*)L.clover[CLOVER_D].p, L.clover[cs].p, nsets, L.g.Vcb);
  else
    hipLaunchKernelGGL(k_clover_convert<CLOVER_H>, dim3(nblk(n, 128)), dim3(128), 0, L.stream,
                       (const double2*)L.clover[CLOVER_D].p, L.clover[cs].p, nsets, L.g.Vcb);
  CPSHIP_CHECK(hipGetLastError());
  L.clover_valid[cs] = true;
  return 0;
}

// Which clover sets a solve of nrhs right-hand sides uses: with one kappa
// every rhs uses it (and the operator's mass must give that kappa); with
// n kappas, nrhs = n * k and rhs r uses kappa index r / k.
int clover_sets(Lattice& L, const cps_hip_op* op, int nrhs, int* rhs_per_set) {
  if (L.n_clover == 0) return CPS_HIP_ERR_ARG;
  // inside a CG with an active-rhs window the set follows the global index
  if (L.win_total > 0) nrhs = L.win_total;
  if (op->csw != L.clover_csw) return CPS_HIP_ERR_ARG;
  if (L.n_clover == 1) {
    const double k = op_kappa(op);
    if (std::fabs(k - L.clover_kappas[0]) > 1e-12 * k) return CPS_HIP_ERR_ARG;
    *rhs_per_set = nrhs;
    return 0;
  }
  if (nrhs % L.n_clover) return CPS_HIP_ERR_ARG;
  *rhs_per_set = nrhs / L.n_clover;
  return 0;
}

}  // namespace

int clover_build(Lattice& L, const double* kappas, int n_kappa, double csw) {
  if (!kappas || n_kappa < 1) return CPS_HIP_ERR_ARG;
  for (int k = 0; k < n_kappa; k++)
    if (!(kappas[k] > 0.0)) return CPS_HIP_ERR_ARG;
  int rc = L.ensure_gauge(LINK_D18);
  if (rc) return rc;
  const int Vcb = L.g.Vcb;
  rc = L.clover[CLOVER_D].ensure((size_t)2 * n_kappa * 36 * Vcb * 16);
  if (rc) return rc;
  DevBuf acc;
  rc = acc.ensure((size_t)2 * Vcb * 72 * sizeof(double));
  if (rc) return rc;
  for (int k = 0; k < n_kappa; k++) {
    hipLaunchKernelGGL(k_clover_build, dim3(nblk(2 * (int64_t)Vcb, 64)), dim3(64), 0, L.stream,
                       L.gauge_ref(LINK_D18), L.g, kappas[k] * csw, (double*)acc.p);
    CPSHIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_clover_pack, dim3(nblk(4 * (int64_t)Vcb, 64)), dim3(64), 0, L.stream,
                       (const double*)acc.p, (double2*)L.clover[CLOVER_D].p, Vcb, k);
    CPSHIP_CHECK(hipGetLastError());
  }
  std::vector<double> ksq(n_kappa);
  for (int k = 0; k < n_kappa; k++) ksq[k] = kappas[k] * kappas[k];
  rc = L.clover_ksq.ensure(n_kappa * sizeof(double));
  if (rc) return rc;
  CPSHIP_CHECK(hipMemcpyAsync(L.clover_ksq.p, ksq.data(), n_kappa * sizeof(double), hipMemcpyHostToDevice, L.stream));
  CPSHIP_CHECK(hipStreamSynchronize(L.stream));
  L.n_clover = n_kappa;
  L.clover_kappas.assign(kappas, kappas + n_kappa);
  L.clover_csw = csw;
  L.clover_valid[CLOVER_D] = true;
  L.clover_valid[CLOVER_F] = L.clover_valid[CLOVER_H] = false;
  return 0;
}

int clover_export(Lattice& L, int k, int cb, double* A) {
  if (k < 0 || k >= L.n_clover || (cb != 0 && cb != 1) || !A) return CPS_HIP_ERR_ARG;
  const int Vcb = L.g.Vcb;
  std::vector<double> soa((size_t)36 * Vcb * 2);
  const double2* src = (const double2*)L.clover[CLOVER_D].p + (int64_t)(k * 2 + cb) * 36 * Vcb;
  CPSHIP_CHECK(hipMemcpy(soa.data(), src, soa.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (int q = 0; q < 36; q++)
    for (int s = 0; s < Vcb; s++) {
      A[(size_t)s * 72 + 2 * q] = soa[((size_t)q * Vcb + s) * 2];
      A[(size_t)s * 72 + 2 * q + 1] = soa[((size_t)q * Vcb + s) * 2 + 1];
    }
  return 0;
}

// MatPc (dag 0) / MatPcDag (dag 1): out = A_oo in - kappa^2 D A_ee^-1 D in,
// with D the hop of the given dag (A is Hermitian).  partial: per-block
// |out|^2 of the second hop.
static int clover_matpc_ru(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& out, const FieldRef& in,
                           const FieldRef& t1, int dag, double* partial, int nrhs, const FieldRef* R,
                           const double* ralpha, const int* ractive) {
  int per = 0;
  int rc = clover_sets(L, op, nrhs, &per);
  if (rc) return rc;
  const int cs = storage_of_hprec(hprec);
  if ((rc = ensure_clover(L, cs))) return rc;
  HopClover hc;
  hc.ref.data = L.clover[cs].p;
  hc.ref.plane = L.g.Vcb;
  hc.rhs_per_set = per;
  hc.rhs0 = L.win_total > 0 ? L.win_lo : 0;
  hc.mode = CLOV_POST;
  hc.kscale = nullptr;
  rc = hop_ex(L, op, hprec, t1, in, 0, dag, 1.0, null_field(), 0.0, nullptr, nrhs, &hc);
  if (rc) return rc;
  hc.mode = CLOV_X;
  hc.kscale = (const double*)L.clover_ksq.p;
  hc.R = R;
  hc.ralpha = ralpha;
  hc.ractive = ractive;
  return hop_ex(L, op, hprec, out, t1, 1, dag, -1.0, in, 1.0, partial, nrhs, &hc);
}

int clover_matpc(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& out, const FieldRef& in,
                 const FieldRef& t1, int dag, double* partial, int nrhs) {
  return clover_matpc_ru(L, op, hprec, out, in, t1, dag, partial, nrhs, nullptr, nullptr, nullptr);
}

// MatPcDag of t2 with the CG residual step in the epilogue of its last hop:
// r -= alpha[rhs] (MatPcDag t2) for the rhs with active[rhs], partial =
// |r|^2 (FP16 r and hops; M^dag M p is never stored)
int clover_mdag_rupdate(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& r, const FieldRef& t2,
                        const FieldRef& t1, const double* alpha, const int* active, double* partial, int nrhs) {
  if (storage_of_hprec(hprec) != CLOVER_H) return CPS_HIP_ERR_ARG;
  return clover_matpc_ru(L, op, hprec, r, t2, t1, 1, partial, nrhs, &r, alpha, active);
}

// MatPcDagMatPc (d_op_clover.C:156-181): partial = |MatPc in|^2 per block.
int clover_mdagm(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& out, const FieldRef& in,
                 const FieldRef& t1, const FieldRef& t2, double* partial, int nrhs) {
  int rc = clover_matpc(L, op, hprec, t2, in, t1, 0, partial, nrhs);
  if (rc) return rc;
  return clover_matpc(L, op, hprec, out, t2, t1, 1, nullptr, nrhs);
}

namespace {

// out = A in for every site of a single-checkerboard field (A the clover set
// of kappa index rhs / rhs_per_set at checkerboard cb)
template <typename S>
__global__ void __launch_bounds__(256) k_clover_apply(FieldRef out, FieldRef in, CloverRef c, int cb, int rhs_per_set,
                                                      int Vcb) {
  using R = typename SpinTraits<S>::R;
  const int rhs = blockIdx.y;
  const int site = blockIdx.x * blockDim.x + threadIdx.x;
  if (site >= Vcb) return;
  const int64_t e = (int64_t)rhs * Vcb + site;
  const int set = rhs / rhs_per_set;
  const int64_t c0 = (int64_t)(set * 2 + cb) * CloverOf<S>::nplanes * c.plane + site;
  cplx<R> x[12], y[12];
  spin_load<S, R>(in, e, x);
  clover_mul<S, R>(c, c0, x, y);
  spin_store<S, R>(out, e, y);
}

int clover_apply_d(Lattice& L, const FieldRef& out, const FieldRef& in, int cb, int rhs_per_set, int nrhs) {
  CloverRef c;
  c.data = L.clover[CLOVER_D].p;
  c.plane = L.g.Vcb;
  dim3 grid((unsigned)((L.g.Vcb + 255) / 256), (unsigned)nrhs), block(256);
  hipLaunchKernelGGL(k_clover_apply<SpinD>, grid, block, 0, L.stream, out, in, c, cb, rhs_per_set, L.g.Vcb);
  CPSHIP_CHECK(hipGetLastError());
  return 0;
}

}  // namespace

// DiracOpClover::MatInv (d_op_clover.C, MatInv(Vector*, Vector*, Float*,
// PreserveType)): full vectors [odd][even];
//   t = in_o + kappa D(cb even) A_ee^-1 in_e ; src = MatPcDag t ; InvCg ;
//   out_e = A_ee^-1 (in_e + kappa D(cb odd) out_o).
// One kappa (the clover field must hold a single set).
int clover_matinv_prep(Lattice& L, const cps_hip_op* op, const FieldRef& src, const FieldRef& in_o,
                       const FieldRef& in_e, const FieldRef& t1, const FieldRef& t2, const FieldRef& t3, int nrhs) {
  int per = 0;
  int rc = clover_sets(L, op, nrhs, &per);
  if (rc) return rc;
  if (L.n_clover != 1) return CPS_HIP_ERR_ARG;
  const double kappa = L.clover_kappas[0];
  if ((rc = clover_apply_d(L, t3, in_e, 0, per, nrhs))) return rc;
  if ((rc = hop(L, op, PREC_D_D18, t2, t3, 1, 0, kappa, in_o, 1.0, nullptr, nrhs))) return rc;
  return clover_matpc(L, op, PREC_D_D18, src, t2, t1, 1, nullptr, nrhs);
}

int clover_matinv_reconstruct(Lattice& L, const cps_hip_op* op, const FieldRef& out_e, const FieldRef& out_o,
                              const FieldRef& in_e, const FieldRef& t1, const FieldRef& t2, const FieldRef& t3,
                              int nrhs) {
  (void)t2;
  (void)t3;
  int per = 0;
  int rc = clover_sets(L, op, nrhs, &per);
  if (rc) return rc;
  const double kappa = L.clover_kappas[0];
  if ((rc = hop(L, op, PREC_D_D18, t1, out_o, 0, 0, kappa, in_e, 1.0, nullptr, nrhs))) return rc;
  return clover_apply_d(L, out_e, t1, 0, per, nrhs);
}

}  // namespace cpship
