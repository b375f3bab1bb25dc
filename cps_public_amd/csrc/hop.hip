// Wilson / domain-wall hopping kernels for gfx950.
//
// One thread per output element (rhs, site, j).  The 4-D hop gathers the 8
// neighbour spinors, spin-projects each to a half spinor (2 spins x 3
// colours), multiplies by U or U^dagger and reconstructs the lower spins;
// links are shared by every j (and every rhs) of a site, so lanes of one
// wavefront that belong to one site read them with one broadcast load.  The
// domain-wall kernel adds the fifth-dimension hop from the same site's
// opposite-chirality slots.  An epilogue forms out = xcoef*X + scale*hop and
// optionally the per-rhs |out|^2 partial sums (wavefront shuffles, then one
// LDS pass per block, one value per block written; a second kernel sums the
// block values in a fixed order so results are reproducible).
//
// Operator conventions follow the CPS noarch code:
//   chi(x) = sum_mu U_mu(x) (1 - s gamma_mu) psi(x+mu)
//          + U_mu(x-mu)^dagger (1 + s gamma_mu) psi(x-mu),  s = +1 (dag 0), -1 (dag 1)
//   (wilson_dslash.C:173-515; gamma matrices of OneMinusPlusGammaTimesPsi,
//   wilson_dslash.C:64-128), and for domain wall
//   out(s) += (2/a5) P+ in(s-1) + (2/a5) P- in(s+1), -mf at the walls,
//   P+ and P- exchanged for dag 1 (dwf_dslash_5_plus.C:111-340).
#include "cps_hip_internal.h"
#include "hop.h"

namespace cpship {

// (1 + A gamma_mu) psi projected to spins (0,1): h[0..2] spin 0, h[3..5] spin 1
template <int MU, int A, typename R>
__device__ __forceinline__ void project(const cplx<R> p[12], cplx<R> h[6]) {
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const cplx<R> p0 = p[c], p1 = p[3 + c], p2 = p[6 + c], p3 = p[9 + c];
    if constexpr (MU == 0) {        // h0 = p0 + A i p3, h1 = p1 + A i p2
      h[c] = cmk<R>(p0.re - A * p3.im, p0.im + A * p3.re);
      h[3 + c] = cmk<R>(p1.re - A * p2.im, p1.im + A * p2.re);
    } else if constexpr (MU == 1) { // h0 = p0 - A p3, h1 = p1 + A p2
      h[c] = cmk<R>(p0.re - A * p3.re, p0.im - A * p3.im);
      h[3 + c] = cmk<R>(p1.re + A * p2.re, p1.im + A * p2.im);
    } else if constexpr (MU == 2) { // h0 = p0 + A i p2, h1 = p1 - A i p3
      h[c] = cmk<R>(p0.re - A * p2.im, p0.im + A * p2.re);
      h[3 + c] = cmk<R>(p1.re + A * p3.im, p1.im - A * p3.re);
    } else {                        // h0 = p0 + A p2, h1 = p1 + A p3
      h[c] = cmk<R>(p0.re + A * p2.re, p0.im + A * p2.im);
      h[3 + c] = cmk<R>(p1.re + A * p3.re, p1.im + A * p3.im);
    }
  }
}

// acc += full spinor rebuilt from the colour-rotated half spinor w
template <int MU, int A, typename R>
__device__ __forceinline__ void reconstruct(cplx<R> acc[12], const cplx<R> w[6]) {
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const cplx<R> w0 = w[c], w1 = w[3 + c];
    acc[c].re += w0.re; acc[c].im += w0.im;
    acc[3 + c].re += w1.re; acc[3 + c].im += w1.im;
    if constexpr (MU == 0) {        // r2 = -A i w1, r3 = -A i w0
      acc[6 + c].re += A * w1.im; acc[6 + c].im -= A * w1.re;
      acc[9 + c].re += A * w0.im; acc[9 + c].im -= A * w0.re;
    } else if constexpr (MU == 1) { // r2 = A w1, r3 = -A w0
      acc[6 + c].re += A * w1.re; acc[6 + c].im += A * w1.im;
      acc[9 + c].re -= A * w0.re; acc[9 + c].im -= A * w0.im;
    } else if constexpr (MU == 2) { // r2 = -A i w0, r3 = A i w1
      acc[6 + c].re += A * w0.im; acc[6 + c].im -= A * w0.re;
      acc[9 + c].re -= A * w1.im; acc[9 + c].im += A * w1.re;
    } else {                        // r2 = A w0, r3 = A w1
      acc[6 + c].re += A * w0.re; acc[6 + c].im += A * w0.im;
      acc[9 + c].re += A * w1.re; acc[9 + c].im += A * w1.im;
    }
  }
}

template <typename R>
__device__ __forceinline__ void su3_mul(const cplx<R> u[9], const cplx<R> h[6], cplx<R> w[6]) {
#pragma unroll
  for (int s = 0; s < 2; s++)
#pragma unroll
    for (int r = 0; r < 3; r++) {
      cplx<R> a = cmul(u[3 * r + 0], h[3 * s + 0]);
      cplx<R> b = cmul(u[3 * r + 1], h[3 * s + 1]);
      cplx<R> c = cmul(u[3 * r + 2], h[3 * s + 2]);
      w[3 * s + r] = cmk<R>(a.re + b.re + c.re, a.im + b.im + c.im);
    }
}

template <typename R>
__device__ __forceinline__ void su3_dag_mul(const cplx<R> u[9], const cplx<R> h[6], cplx<R> w[6]) {
#pragma unroll
  for (int s = 0; s < 2; s++)
#pragma unroll
    for (int r = 0; r < 3; r++) {
      cplx<R> a = cmulc(u[0 * 3 + r], h[3 * s + 0]);
      cplx<R> b = cmulc(u[1 * 3 + r], h[3 * s + 1]);
      cplx<R> c = cmulc(u[2 * 3 + r], h[3 * s + 2]);
      w[3 * s + r] = cmk<R>(a.re + b.re + c.re, a.im + b.im + c.im);
    }
}

struct Coord {
  int x[4];
};

__device__ __forceinline__ Coord site_coords(const Geom& g, int site_cb, int parity) {
  Coord c;
  const int lh = g.L[0] >> 1;
  int r = site_cb / lh;
  const int xh = site_cb - r * lh;
  c.x[1] = r % g.L[1];
  r /= g.L[1];
  c.x[2] = r % g.L[2];
  c.x[3] = r / g.L[2];
  c.x[0] = 2 * xh + ((c.x[1] + c.x[2] + c.x[3] + parity) & 1);
  return c;
}

__device__ __forceinline__ int cb_index(const Geom& g, const int x[4]) {
  return (x[0] >> 1) + (g.L[0] >> 1) * (x[1] + g.L[1] * (x[2] + g.L[2] * x[3]));
}

template <int MU, int DIR /* +1 forward, -1 backward */, typename SIN, typename LK, typename R, int DAG>
__device__ __forceinline__ void hop_dir(const HopParams& P, const Coord& c, int out_par, int site_cb,
                                        int64_t in_base, int j, cplx<R> acc[12]) {
  const Geom& g = P.geom;
  int n[4] = {c.x[0], c.x[1], c.x[2], c.x[3]};
  if (DIR > 0) n[MU] = (n[MU] + 1 == g.L[MU]) ? 0 : n[MU] + 1;
  else n[MU] = (n[MU] == 0) ? g.L[MU] - 1 : n[MU] - 1;
  const int nbr = cb_index(g, n);
  cplx<R> psi[12];
  spin_load<SIN, R>(P.in, in_base + (int64_t)nbr * P.nb + j, psi);
  cplx<R> h[6], w[6], u[9];
  constexpr int S = (DAG == 0) ? 1 : -1;
  if constexpr (DIR > 0) {
    // U_mu(x) (1 - s gamma_mu) psi(x+mu); link of the output site
    R sign = (P.link_bc_sign && g.bc[MU] && c.x[MU] == g.L[MU] - 1) ? R(-1) : R(1);
    project<MU, -S, R>(psi, h);
    link_load<LK, R>(P.gauge, out_par, MU, site_cb, g.Vcb, sign, u);
    su3_mul<R>(u, h, w);
    reconstruct<MU, -S, R>(acc, w);
  } else {
    // U_mu(x-mu)^dagger (1 + s gamma_mu) psi(x-mu); link of the neighbour
    R sign = (P.link_bc_sign && g.bc[MU] && c.x[MU] == 0) ? R(-1) : R(1);
    project<MU, S, R>(psi, h);
    link_load<LK, R>(P.gauge, 1 - out_par, MU, nbr, g.Vcb, sign, u);
    su3_dag_mul<R>(u, h, w);
    reconstruct<MU, S, R>(acc, w);
  }
}

template <typename R>
__device__ __forceinline__ R wave_sum(R v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum v over the block (all threads of a block share one rhs) and write one
// partial per block.
template <typename R>
__device__ __forceinline__ void block_partial(R v, double* partial, int64_t slot) {
  __shared__ double red[16];
  double d = wave_sum<double>(double(v));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = d;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    const int nw = (blockDim.x + 63) >> 6;
    for (int k = 0; k < nw; k++) s += red[k];
    partial[slot] = s;
  }
}

// Maps the hardware block id so that each XCD (blocks are dealt round-robin
// over the 8 XCDs) works on a contiguous range of sites, which keeps
// neighbour reuse inside one XCD's L2.
__device__ __forceinline__ int64_t logical_block(int64_t bid, int64_t nblk) {
  if ((nblk & 7) != 0) return bid;
  return (bid & 7) * (nblk >> 3) + (bid >> 3);
}

template <typename SIN, typename SOUT, typename LK, typename R, int DAG, int KIND>
__global__ void __launch_bounds__(256) k_hop(HopParams P) {
  const int64_t nblk = (int64_t)gridDim.x;
  const int64_t lb = logical_block(blockIdx.x, nblk);
  const int64_t e = lb * blockDim.x + threadIdx.x;   // element within this launch
  const int64_t per_rhs = (int64_t)P.nsites_out * P.nb;
  const int rhs = (int)(e / per_rhs);
  const int64_t rem = e - (int64_t)rhs * per_rhs;
  const int site = (int)(rem / P.nb);
  const int j = (int)(rem - (int64_t)site * P.nb);
  const bool valid = rhs < P.nrhs;
  R nrm = R(0);
  if (valid) {
    const Geom& g = P.geom;
    int out_par, site_cb;
    if (KIND == KIND_DWF) {       // fields hold both 4-D parities
      out_par = site >= g.Vcb ? 1 : 0;
      site_cb = site - out_par * g.Vcb;
    } else {
      out_par = P.out_parity;
      site_cb = site;
    }
    const Coord c = site_coords(g, site_cb, out_par);
    int64_t in_base;
    if (KIND == KIND_DWF) in_base = ((int64_t)rhs * 2 * g.Vcb + (int64_t)(1 - out_par) * g.Vcb) * P.nb;
    else in_base = (int64_t)rhs * g.Vcb * P.nb;

    cplx<R> acc[12];
#pragma unroll
    for (int k = 0; k < 12; k++) acc[k] = cmk<R>(R(0), R(0));
    hop_dir<0, +1, SIN, LK, R, DAG>(P, c, out_par, site_cb, in_base, j, acc);
    hop_dir<1, +1, SIN, LK, R, DAG>(P, c, out_par, site_cb, in_base, j, acc);
    hop_dir<2, +1, SIN, LK, R, DAG>(P, c, out_par, site_cb, in_base, j, acc);
    hop_dir<3, +1, SIN, LK, R, DAG>(P, c, out_par, site_cb, in_base, j, acc);
    hop_dir<0, -1, SIN, LK, R, DAG>(P, c, out_par, site_cb, in_base, j, acc);
    hop_dir<1, -1, SIN, LK, R, DAG>(P, c, out_par, site_cb, in_base, j, acc);
    hop_dir<2, -1, SIN, LK, R, DAG>(P, c, out_par, site_cb, in_base, j, acc);
    hop_dir<3, -1, SIN, LK, R, DAG>(P, c, out_par, site_cb, in_base, j, acc);

    if constexpr (KIND == KIND_DWF) {
      // fifth dimension: input slots at the same 4-D site (same parity block)
      const int ls = P.ls;
      const int cls_out = (out_par + P.out_cb5) & 1;        // s = 2j + cls_out
      const int s = 2 * j + cls_out;
      const int64_t site_base = ((int64_t)rhs * 2 * g.Vcb + (int64_t)out_par * g.Vcb + site_cb) * P.nb;
      // in slots at this site hold s' = 2j' + (1 - cls_out)
      const int sm = (s == 0) ? ls - 1 : s - 1;
      const int sp = (s == ls - 1) ? 0 : s + 1;
      const R cm = (s == 0) ? R(P.neg_mass_two_over_a5) : R(P.two_over_a5);
      const R cp = (s == ls - 1) ? R(P.neg_mass_two_over_a5) : R(P.two_over_a5);
      const int jm = sm >> 1, jp = sp >> 1;
      // dag 0: upper from s-1, lower from s+1; dag 1: lower from s-1, upper from s+1
      const int c_from_m = (DAG == 0) ? 0 : 6;
      const int c_from_p = (DAG == 0) ? 6 : 0;
      cplx<R> hm[6], hp[6];
      spin_load_half<SIN, R>(P.in, site_base + jm, c_from_m, hm);
      spin_load_half<SIN, R>(P.in, site_base + jp, c_from_p, hp);
#pragma unroll
      for (int k = 0; k < 6; k++) {
        acc[c_from_m + k].re += cm * hm[k].re;
        acc[c_from_m + k].im += cm * hm[k].im;
        acc[c_from_p + k].re += cp * hp[k].re;
        acc[c_from_p + k].im += cp * hp[k].im;
      }
    }

    const int64_t eo = (int64_t)rhs * per_rhs + rem;
    cplx<R> res[12];
    const R sc = R(P.scale);
    if (P.X.data) {
      cplx<R> xv[12];
      spin_load<SOUT, R>(P.X, eo, xv);
      const R xc = R(P.xcoef);
#pragma unroll
      for (int k = 0; k < 12; k++) res[k] = cmk<R>(xc * xv[k].re + sc * acc[k].re, xc * xv[k].im + sc * acc[k].im);
    } else {
#pragma unroll
      for (int k = 0; k < 12; k++) res[k] = cmk<R>(sc * acc[k].re, sc * acc[k].im);
    }
    spin_store<SOUT, R>(P.out, eo, res);
    if (P.partial) {
#pragma unroll
      for (int k = 0; k < 12; k++) nrm += res[k].re * res[k].re + res[k].im * res[k].im;
    }
  }
  if (P.partial) {
    // blocks never straddle two rhs (block size divides nsites_out*nb)
    const int64_t bpr = per_rhs / blockDim.x;
    block_partial<R>(nrm, P.partial, lb);
    (void)bpr;
  }
}

template <typename SIN, typename SOUT, typename LK, typename R>
static hipError_t launch_t(const HopParams& P, hipStream_t st) {
  const int64_t nelem = (int64_t)P.nrhs * P.nsites_out * P.nb;
  const int bs = P.block;
  const int64_t nblk = (nelem + bs - 1) / bs;
  dim3 grid((unsigned)nblk), block(bs);
  if (P.kind == KIND_DWF) {
    if (P.dag) hipLaunchKernelGGL((k_hop<SIN, SOUT, LK, R, 1, KIND_DWF>), grid, block, 0, st, P);
    else hipLaunchKernelGGL((k_hop<SIN, SOUT, LK, R, 0, KIND_DWF>), grid, block, 0, st, P);
  } else {
    if (P.dag) hipLaunchKernelGGL((k_hop<SIN, SOUT, LK, R, 1, KIND_WILSON>), grid, block, 0, st, P);
    else hipLaunchKernelGGL((k_hop<SIN, SOUT, LK, R, 0, KIND_WILSON>), grid, block, 0, st, P);
  }
  return hipGetLastError();
}

hipError_t launch_hop(const HopParams& P, hipStream_t st) {
  switch (P.prec) {
    case PREC_D_D18: return launch_t<SpinD, SpinD, LinkD18, double>(P, st);
    case PREC_F_F12: return launch_t<SpinF, SpinF, LinkF12, float>(P, st);
    case PREC_H_H12: return launch_t<SpinH, SpinH, LinkH12, float>(P, st);
    case PREC_H_B12: return launch_t<SpinH, SpinH, LinkB12, float>(P, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace cpship
