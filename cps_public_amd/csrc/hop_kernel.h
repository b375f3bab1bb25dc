// Wilson / domain-wall hopping kernel template for gfx950 (instantiated in
// hop.hip, and with the clover epilogues in hop_clover.hip).
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
#pragma once
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

// Neighbours of one output site: direction d = 0..3 is +mu (mu = d), d = 4..7
// is -mu (mu = d-4), the CPS summation order (wilson_dslash.C:437-447).
struct Nbr {
  int idx[8];      // checkerboard index of the neighbour site
  unsigned neg;    // bit d: the link of direction d lies on the last slice of
                   // direction d & 3 (it carries the boundary factor)
};

__device__ __forceinline__ Nbr neighbours(const Geom& g, const Coord& c) {
  Nbr N;
  N.neg = 0;
#pragma unroll
  for (int d = 0; d < 8; d++) {
    const int mu = d & 3;
    int n[4] = {c.x[0], c.x[1], c.x[2], c.x[3]};
    if (d < 4) {
      n[mu] = (n[mu] + 1 == g.L[mu]) ? 0 : n[mu] + 1;
      if (c.x[mu] == g.L[mu] - 1) N.neg |= 1u << d;
    } else {
      n[mu] = (n[mu] == 0) ? g.L[mu] - 1 : n[mu] - 1;
      if (c.x[mu] == 0) N.neg |= 1u << d;
    }
    N.idx[d] = cb_index(g, n);
  }
  return N;
}

// Third-row factor of the link of direction d (bit d of the site's boundary
// mask: the link lies on the last slice of direction d & 3)
__device__ __forceinline__ float2 link_phase(const HopParams& P, unsigned bmask, int d) {
  const int mu = d & 3;
  return ((bmask >> d) & 1u) ? make_float2(P.bph[mu][0], P.bph[mu][1]) : make_float2(1.f, 0.f);
}

// ---- packed half-precision path ----------------------------------------------
// FP16 spinors are projected in packed f16 arithmetic on their normalised
// values, and the SU(3) product uses v_dot2 (two f16 or bf16 products added
// into an f32 accumulator per instruction); the site norm is applied once to
// the colour-rotated half spinor.  This keeps the sloppy hop within the
// vector-ALU budget of the HBM-bound roofline.
typedef _Float16 hv2 __attribute__((ext_vector_type(2)));
typedef __bf16 bv2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ hv2 as_hv2(uint32_t w) { return __builtin_bit_cast(hv2, w); }
__device__ __forceinline__ bv2 as_bv2(uint32_t w) { return __builtin_bit_cast(bv2, w); }
__device__ __forceinline__ hv2 hswap(hv2 a) { return __builtin_shufflevector(a, a, 1, 0); }

template <int MU, int A>
__device__ __forceinline__ void project_h(const SpinRaw<SpinH>& r, hv2 h[6]) {
  const uint32_t p[12] = {r.q[0].x, r.q[0].y, r.q[0].z, r.q[0].w, r.q[1].x, r.q[1].y,
                          r.q[1].z, r.q[1].w, r.q[2].x, r.q[2].y, r.q[2].z, r.q[2].w};
  const hv2 ia = {(_Float16)(-A), (_Float16)(A)};   // A*i*q = swap(q) * (-A, A)
  const hv2 aa = {(_Float16)(A), (_Float16)(A)};
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const hv2 p0 = as_hv2(p[c]), p1 = as_hv2(p[3 + c]), p2 = as_hv2(p[6 + c]), p3 = as_hv2(p[9 + c]);
    if constexpr (MU == 0) {        // h0 = p0 + A i p3, h1 = p1 + A i p2
      h[c] = p0 + hswap(p3) * ia;
      h[3 + c] = p1 + hswap(p2) * ia;
    } else if constexpr (MU == 1) { // h0 = p0 - A p3, h1 = p1 + A p2
      h[c] = p0 - p3 * aa;
      h[3 + c] = p1 + p2 * aa;
    } else if constexpr (MU == 2) { // h0 = p0 + A i p2, h1 = p1 - A i p3
      h[c] = p0 + hswap(p2) * ia;
      h[3 + c] = p1 - hswap(p3) * ia;
    } else {                        // h0 = p0 + A p2, h1 = p1 + A p3
      h[c] = p0 + p2 * aa;
      h[3 + c] = p1 + p3 * aa;
    }
  }
}

// acc += n * (full spinor rebuilt from w)
template <int MU, int A>
__device__ __forceinline__ void reconstruct_scaled(cplx<float> acc[12], const cplx<float> w[6], float n) {
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const cplx<float> w0 = w[c], w1 = w[3 + c];
    acc[c].re += n * w0.re; acc[c].im += n * w0.im;
    acc[3 + c].re += n * w1.re; acc[3 + c].im += n * w1.im;
    const float an = A * n;
    if constexpr (MU == 0) {
      acc[6 + c].re += an * w1.im; acc[6 + c].im -= an * w1.re;
      acc[9 + c].re += an * w0.im; acc[9 + c].im -= an * w0.re;
    } else if constexpr (MU == 1) {
      acc[6 + c].re += an * w1.re; acc[6 + c].im += an * w1.im;
      acc[9 + c].re -= an * w0.re; acc[9 + c].im -= an * w0.im;
    } else if constexpr (MU == 2) {
      acc[6 + c].re += an * w0.im; acc[6 + c].im -= an * w0.re;
      acc[9 + c].re -= an * w1.im; acc[9 + c].im += an * w1.re;
    } else {
      acc[6 + c].re += an * w0.re; acc[6 + c].im += an * w0.im;
      acc[9 + c].re += an * w1.re; acc[9 + c].im += an * w1.im;
    }
  }
}

// third row of a 12-real link, ph * conj(row0 x row1), from f32 values (ph
// = det U, see link_row2)
__device__ __forceinline__ void row2_f32(const float2 u[6], float2 ph, float2 r[3]) {
  cplx<float> m[9];
#pragma unroll
  for (int k = 0; k < 6; k++) m[k] = cmk<float>(u[k].x, u[k].y);
  link_row2<float>(m, ph);
#pragma unroll
  for (int k = 0; k < 3; k++) r[k] = make_float2(m[6 + k].re, m[6 + k].im);
}

// link elements as packed 32-bit words (f16 or bf16 pairs), third row rebuilt for 12-real storage
template <typename LK>
__device__ __forceinline__ void link_words(const LinkRaw<LK>& r, float2 ph, uint32_t U[9]) {
  if constexpr (LinkIs18Half<LK>::value) {
    U[0] = r.a.x; U[1] = r.a.y; U[2] = r.a.z; U[3] = r.a.w;
    U[4] = r.b.x; U[5] = r.b.y; U[6] = r.b.z; U[7] = r.b.w; U[8] = r.c;
    (void)ph;
  } else {
    U[0] = r.a.x; U[1] = r.a.y; U[2] = r.a.z; U[3] = r.a.w; U[4] = r.b.x; U[5] = r.b.y;
    float2 f[6], r2[3];
#pragma unroll
    for (int k = 0; k < 6; k++) f[k] = LinkIsBf16<LK>::value ? bf2f(U[k]) : h2f(U[k]);
    row2_f32(f, ph, r2);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      if constexpr (LinkIsBf16<LK>::value) {
        const bv2 b = {(__bf16)r2[k].x, (__bf16)r2[k].y};
        U[6 + k] = __builtin_bit_cast(uint32_t, b);
      } else {
        U[6 + k] = f2h(r2[k].x, r2[k].y);
      }
    }
  }
}

// One v_pk_add_f16 per needed operand: h = p + alpha q (alpha = +1, -1, +i, -i
// for ALPHA = 0..3) delivered directly in the form the dot2 product wants:
//   VAR 0: (h.re, -h.im)   VAR 1: (h.im, h.re)      (U h)
//   VAR 2: (h.re,  h.im)   VAR 3: (h.im, -h.re)     (U^dagger h)
// The half selection and negations are instruction modifiers, so the
// projection costs 12 packed instructions per direction.
#define CPSHIP_PKADD(OS0, OS1, OH0, OH1, NL0, NL1, NH0, NH1)                                                  \
  asm("v_pk_add_f16 %0, %1, %2 op_sel:[" #OS0 "," #OS1 "] op_sel_hi:[" #OH0 "," #OH1 "] neg_lo:[" #NL0 "," #NL1 \
      "] neg_hi:[" #NH0 "," #NH1 "]"                                                                            \
      : "=v"(r)                                                                                                \
      : "v"(p), "v"(q))

template <int ALPHA, int VAR>
__device__ __forceinline__ uint32_t proj_word(uint32_t p, uint32_t q) {
  uint32_t r;
  if constexpr (ALPHA == 0) {          // h = (p.x + q.x, p.y + q.y)
    if constexpr (VAR == 0) CPSHIP_PKADD(0, 0, 1, 1, 0, 0, 1, 1);
    else if constexpr (VAR == 1) CPSHIP_PKADD(1, 1, 0, 0, 0, 0, 0, 0);
    else if constexpr (VAR == 2) CPSHIP_PKADD(0, 0, 1, 1, 0, 0, 0, 0);
    else CPSHIP_PKADD(1, 1, 0, 0, 0, 0, 1, 1);
  } else if constexpr (ALPHA == 1) {   // h = (p.x - q.x, p.y - q.y)
    if constexpr (VAR == 0) CPSHIP_PKADD(0, 0, 1, 1, 0, 1, 1, 0);
    else if constexpr (VAR == 1) CPSHIP_PKADD(1, 1, 0, 0, 0, 1, 0, 1);
    else if constexpr (VAR == 2) CPSHIP_PKADD(0, 0, 1, 1, 0, 1, 0, 1);
    else CPSHIP_PKADD(1, 1, 0, 0, 0, 1, 1, 0);
  } else if constexpr (ALPHA == 2) {   // h = (p.x - q.y, p.y + q.x)
    if constexpr (VAR == 0) CPSHIP_PKADD(0, 1, 1, 0, 0, 1, 1, 1);
    else if constexpr (VAR == 1) CPSHIP_PKADD(1, 0, 0, 1, 0, 0, 0, 1);
    else if constexpr (VAR == 2) CPSHIP_PKADD(0, 1, 1, 0, 0, 1, 0, 0);
    else CPSHIP_PKADD(1, 0, 0, 1, 0, 0, 1, 0);
  } else {                             // h = (p.x + q.y, p.y - q.x)
    if constexpr (VAR == 0) CPSHIP_PKADD(0, 1, 1, 0, 0, 0, 1, 0);
    else if constexpr (VAR == 1) CPSHIP_PKADD(1, 0, 0, 1, 0, 1, 0, 0);
    else if constexpr (VAR == 2) CPSHIP_PKADD(0, 1, 1, 0, 0, 0, 0, 1);
    else CPSHIP_PKADD(1, 0, 0, 1, 0, 1, 1, 1);
  }
  return r;
}
#undef CPSHIP_PKADD

// alpha of h0 = p0 + alpha0 q0 and h1 = p1 + alpha1 q1 for (1 + A gamma_mu),
// with q0, q1 the spin-3/2 or spin-2/3 partners (see project_h)
template <int MU, int A> struct ProjAlpha;
template <int A> struct ProjAlpha<0, A> { static constexpr int a0 = A > 0 ? 2 : 3, a1 = A > 0 ? 2 : 3, q0 = 3, q1 = 2; };
template <int A> struct ProjAlpha<1, A> { static constexpr int a0 = A > 0 ? 1 : 0, a1 = A > 0 ? 0 : 1, q0 = 3, q1 = 2; };
template <int A> struct ProjAlpha<2, A> { static constexpr int a0 = A > 0 ? 2 : 3, a1 = A > 0 ? 3 : 2, q0 = 2, q1 = 3; };
template <int A> struct ProjAlpha<3, A> { static constexpr int a0 = A > 0 ? 0 : 1, a1 = A > 0 ? 0 : 1, q0 = 2, q1 = 3; };

__device__ __forceinline__ float dot2_first(uint32_t a, uint32_t b) {
  float r;
  asm("v_dot2_f32_f16 %0, %1, %2, 0" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

// link words as f16 pairs (bf16 storage is widened to f16 here)
template <typename LK>
__device__ __forceinline__ void link_words_f16(const LinkRaw<LK>& r, float2 ph, uint32_t U[9]) {
  link_words<LK>(r, ph, U);
  if constexpr (LinkIsBf16<LK>::value) {
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const float2 f = bf2f(U[k]);
      uint32_t h;
      asm("v_cvt_pk_f16_f32 %0, %1, %2" : "=v"(h) : "v"(f.x), "v"(f.y));
      U[k] = h;
    }
  }
}

// one direction from the link as f16 words (U[3 row + col], third row
// already rebuilt)
template <int D, int DAG>
__device__ __forceinline__ void dir_compute_h2w(const SpinRaw<SpinH>& sr, const uint32_t U[9], cplx<float> acc[12]) {
  constexpr int MU = D & 3;
  constexpr int S = (DAG == 0) ? 1 : -1;
  constexpr int A = (D < 4) ? -S : S;
  using PA = ProjAlpha<MU, A>;
  constexpr int V1 = (D < 4) ? 0 : 2, V2 = (D < 4) ? 1 : 3;
  const uint32_t p[12] = {sr.q[0].x, sr.q[0].y, sr.q[0].z, sr.q[0].w, sr.q[1].x, sr.q[1].y,
                          sr.q[1].z, sr.q[1].w, sr.q[2].x, sr.q[2].y, sr.q[2].z, sr.q[2].w};
  uint32_t hb[6], hx[6];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    hb[c] = proj_word<PA::a0, V1>(p[c], p[3 * PA::q0 + c]);
    hx[c] = proj_word<PA::a0, V2>(p[c], p[3 * PA::q0 + c]);
    hb[3 + c] = proj_word<PA::a1, V1>(p[3 + c], p[3 * PA::q1 + c]);
    hx[3 + c] = proj_word<PA::a1, V2>(p[3 + c], p[3 * PA::q1 + c]);
  }
  cplx<float> w[6];
#pragma unroll
  for (int s = 0; s < 2; s++)
#pragma unroll
    for (int r = 0; r < 3; r++) {
      const uint32_t u0 = D < 4 ? U[3 * r + 0] : U[0 * 3 + r];
      const uint32_t u1 = D < 4 ? U[3 * r + 1] : U[1 * 3 + r];
      const uint32_t u2 = D < 4 ? U[3 * r + 2] : U[2 * 3 + r];
      float re = dot2_first(u0, hb[3 * s + 0]);
      float im = dot2_first(u0, hx[3 * s + 0]);
      re = __builtin_amdgcn_fdot2(as_hv2(u1), as_hv2(hb[3 * s + 1]), re, false);
      im = __builtin_amdgcn_fdot2(as_hv2(u1), as_hv2(hx[3 * s + 1]), im, false);
      re = __builtin_amdgcn_fdot2(as_hv2(u2), as_hv2(hb[3 * s + 2]), re, false);
      im = __builtin_amdgcn_fdot2(as_hv2(u2), as_hv2(hx[3 * s + 2]), im, false);
      w[3 * s + r] = cmk<float>(re, im);
    }
  reconstruct_scaled<MU, A>(acc, w, sr.n);
}

template <int D, typename LK, int DAG>
__device__ __forceinline__ void dir_compute_h2(const SpinRaw<SpinH>& sr, const LinkRaw<LK>& lr, float2 ph,
                                               cplx<float> acc[12]) {
  uint32_t U[9];
  link_words_f16<LK>(lr, ph, U);
  dir_compute_h2w<D, DAG>(sr, U, acc);
}

// forward:  U_mu(x) (1 - s gamma_mu) psi(x+mu)        (link of the output site)
// backward: U_mu(x-mu)^dagger (1 + s gamma_mu) psi(x-mu) (link of the neighbour)
template <int D, typename SIN, typename LK, typename R, int DAG>
__device__ __forceinline__ void dir_compute(const SpinRaw<SIN>& sr, const LinkRaw<LK>& lr, float2 ph, cplx<R> acc[12]) {
  if constexpr (std::is_same<SIN, SpinH>::value) {
    dir_compute_h2<D, LK, DAG>(sr, lr, ph, acc);
    return;
  }
  constexpr int MU = D & 3;
  constexpr int S = (DAG == 0) ? 1 : -1;
  cplx<R> psi[12], h[6], w[6], u[9];
  raw_decode<SIN, R>(sr, psi);
  link_decode<LK, R>(lr, ph, u);
  if constexpr (D < 4) {
    project<MU, -S, R>(psi, h);
    su3_mul<R>(u, h, w);
    reconstruct<MU, -S, R>(acc, w);
  } else {
    project<MU, S, R>(psi, h);
    su3_dag_mul<R>(u, h, w);
    reconstruct<MU, S, R>(acc, w);
  }
}

// Directions D0 .. D0+NB-1: all their spinor and link loads are issued
// before the first of them is decoded.
template <int D0, int NB, typename SIN, typename LK, typename R, int DAG>
__device__ __forceinline__ void hop_batch(const HopParams& P, const Nbr& N, int out_par, int site_cb,
                                          int64_t in_base, int j, cplx<R> acc[12]) {
  const int Vcb = P.geom.Vcb;
  SpinRaw<SIN> sr[NB];
  LinkRaw<LK> lr[NB];
#pragma unroll
  for (int k = 0; k < NB; k++) {
    const int d = D0 + k;
    raw_load<SIN>(sr[k], P.in, in_base + (int64_t)N.idx[d] * P.nb + j);
    const int mu = d & 3;
    const int64_t lidx = (d < 4) ? (int64_t)(out_par * 4 + mu) * Vcb + site_cb
                                 : (int64_t)((1 - out_par) * 4 + mu) * Vcb + N.idx[d];
    link_raw_load<LK>(lr[k], P.gauge, lidx);
  }
#define CPSHIP_DIR(K)                                                                             \
  if constexpr ((K) < NB) {                                                                       \
    const float2 ph = link_phase(P, N.neg, D0 + (K));                                             \
    dir_compute<D0 + (K), SIN, LK, R, DAG>(sr[K], lr[K], ph, acc);                                \
  }
  CPSHIP_DIR(0) CPSHIP_DIR(1) CPSHIP_DIR(2) CPSHIP_DIR(3) CPSHIP_DIR(4) CPSHIP_DIR(5) CPSHIP_DIR(6) CPSHIP_DIR(7)
#undef CPSHIP_DIR
}

// HOP_LDS_LINKS: the spinor loads of directions D0 .. D0+NB-1 ...
template <int D0, int NB>
__device__ __forceinline__ void hop_lds_load(const HopParams& P, const Nbr& N, int64_t in_base, int j,
                                             SpinRaw<SpinH> sr[NB]) {
#pragma unroll
  for (int k = 0; k < NB; k++) raw_load<SpinH>(sr[k], P.in, in_base + (int64_t)N.idx[D0 + k] * P.nb + j);
}

// ... and their products with the links staged in LDS (block site sl)
template <int D0, int NB, int DAG>
__device__ __forceinline__ void hop_lds_compute(const HopParams& P, const uint32_t* lds, int sl,
                                                const SpinRaw<SpinH> sr[NB], cplx<float> acc[12]) {
#define CPSHIP_DIR(K)                                                                             \
  if constexpr ((K) < NB) {                                                                       \
    constexpr int D = D0 + (K);                                                                   \
    const uint4* w = reinterpret_cast<const uint4*>(lds + sl * LDS_SITE_WORDS + 8 * D);           \
    const uint4 a = w[0], b = w[1];                                                               \
    const uint32_t U[9] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w,                                \
                           lds[P.lds_sites * LDS_SITE_WORDS + 8 * sl + D]};                       \
    dir_compute_h2w<D, DAG>(sr[K], U, acc);                                                       \
  }
  CPSHIP_DIR(0) CPSHIP_DIR(1) CPSHIP_DIR(2) CPSHIP_DIR(3) CPSHIP_DIR(4) CPSHIP_DIR(5) CPSHIP_DIR(6) CPSHIP_DIR(7)
#undef CPSHIP_DIR
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
  __syncthreads();   // red[] is reused by the block's next sub-block
}

// Chunk partials (HopParams::chunk_part, one element per site): the sum of
// v over 64 consecutive sites of one rhs as the pairwise tree of the xor
// butterfly (offsets 1, 2, .., 32; lane i and lane i^m form the same sum).
// Element order: a wavefront is one chunk.
__device__ __forceinline__ void chunk_partial_elem(const HopParams& P, int rhs, int site, double v) {
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) v += __shfl_xor(v, m);
  if ((threadIdx.x & 63) == 0) P.partial[(int64_t)rhs * (P.nsites_out >> 6) + (site >> 6)] = v;
}

// HOP_RHSGROUP (lane = g S + q, S = 64 / G sites of rhs g): the butterfly
// over the S lanes of each rhs gives the first log2 S levels of that tree;
// the G wavefronts of a chunk (in one block, G <= W) then combine pairwise,
// (a0 + a1) + (a2 + a3), as the remaining levels.
__device__ __forceinline__ void chunk_partial_rhsgroup(const HopParams& P, int64_t sb64, double v) {
  __shared__ double cred[16];
  const int s_shift = 6 - P.rg_shift, S = 1 << s_shift, G = 1 << P.rg_shift;
  for (int m = 1; m < S; m <<= 1) v += __shfl_xor(v, m);
  const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6), W = (int)(blockDim.x >> 6);
  if ((lane & (S - 1)) == 0) cred[w * G + (lane >> s_shift)] = v;
  __syncthreads();
  const int t = (int)threadIdx.x;
  if (t < W) {
    const int c = t >> P.rg_shift, g = t & (G - 1);   // chunk c of the block, rhs g of the group
    // G is 2 or 4 (G <= W <= 4)
    const double* cw = cred + c * G * G + g;
    const double sum = G == 2 ? cw[0] + cw[G] : (cw[0] + cw[G]) + (cw[2 * G] + cw[3 * G]);
    const int sb = (int)sb64;
    const int per_g = P.nsites_out / (W << s_shift);
    const int rg = sb / per_g, sblk = sb - rg * per_g;
    const int64_t rhs = ((int64_t)rg << P.rg_shift) + g;
    P.partial[rhs * (P.nsites_out >> 6) + (int64_t)sblk * (W >> P.rg_shift) + c] = sum;
  }
}

// Maps the hardware block id so that each XCD (blocks are dealt round-robin
// over the 8 XCDs) works on a contiguous range of sites, which keeps
// neighbour reuse inside one XCD's L2.
__device__ __forceinline__ int64_t logical_block(int64_t bid, int64_t nblk, int remap) {
  if (!remap || (nblk & 7) != 0) return bid;
  return (bid & 7) * (nblk >> 3) + (bid >> 3);
}

// Register budget of the hop kernels: at least this many waves per SIMD
// (512 VGPRs / waves).  With the load batch depth (HopBatch) it trades loads
// in flight per wave against waves per SIMD; the defaults are the measured
// best on MI355X (half: batch 4, 3 waves, 47 us vs 56 us for batch 8 with 2
// waves on the 16^3x32x16 domain-wall hop; FP64: batch 2, 2 waves).
#ifndef CPSHIP_HOP_WAVES_D
#define CPSHIP_HOP_WAVES_D 2
#endif
#ifndef CPSHIP_HOP_WAVES_F
#define CPSHIP_HOP_WAVES_F 3
#endif
#ifndef CPSHIP_HOP_WAVES_H
#define CPSHIP_HOP_WAVES_H 3
#endif
// spinor load batch of the LDS-link mapping (the links no longer take
// registers, so all 8 directions can be in flight)
#ifndef CPSHIP_HOP_BATCH_LDS
#define CPSHIP_HOP_BATCH_LDS 8
#endif
template <typename S> struct HopWaves;
template <> struct HopWaves<SpinD> { static constexpr int value = CPSHIP_HOP_WAVES_D; };
template <> struct HopWaves<SpinF> { static constexpr int value = CPSHIP_HOP_WAVES_F; };
template <> struct HopWaves<SpinH> { static constexpr int value = CPSHIP_HOP_WAVES_H; };

// Which output element thread threadIdx.x of logical sub-block `sb` (of
// nsub) computes, and the partial-sum slot of that sub-block.
struct HopElem {
  int rhs, site, j;
  int64_t pslot;   // rhs * blocks_per_rhs + sub-block within the rhs
  // site range of the block (element mapping): sites qb .. qb+nsl-1 in site
  // order (through the spatial block b's permutation for HOP_SPATIAL); the
  // element's site is qb + sl
  int qb, nsl, sl, b;
};

// HOP_SITEWAVE decode (Wilson kind): sub-block sb = (spatial block b | -,
// rhs group G, site block); wave w of the block computes site
// (site block) * W + w for rhs G*g + lane / nb, g = 64 / nb.  The site is
// wave-uniform.
__device__ __forceinline__ HopElem hop_decode_sitewave(const HopParams& P, int64_t sb, int64_t nsub) {
  HopElem E;
  const int W = blockDim.x >> 6;
  const int g = 64 / P.nb;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  E.rhs = 0;
  int site;
  if (P.flags & HOP_SPATIAL) {
    const int64_t per_b = nsub >> 3;
    const int b = (int)(sb / per_b);
    const int64_t n_in = sb - (int64_t)b * per_b;
    const int64_t spb = P.perm_S / W;   // site blocks per spatial block
    const int G = (int)(n_in / spb);
    const int64_t sblk = n_in - (int64_t)G * spb;
    site = (int)P.perm[(int64_t)b * P.perm_S + sblk * W + w];
    E.rhs = G * g + lane / P.nb;
  } else {
    const int64_t spb = (int64_t)P.nsites_out / W;
    const int G = (int)(sb / spb);
    const int64_t sblk = sb - (int64_t)G * spb;
    site = (int)(sblk * W + w);
    E.rhs = G * g + lane / P.nb;
  }
  E.site = __builtin_amdgcn_readfirstlane(site);
  E.j = lane - (lane / P.nb) * P.nb;
  E.pslot = 0;   // no partials in this mode
  E.qb = E.nsl = E.sl = E.b = 0;
  return E;
}

// the block's site range: first element el0, the element's own site index q
__device__ __forceinline__ void hop_block_sites(const HopParams& P, int el0, int q, HopElem& E) {
  const int el1 = el0 + (int)blockDim.x - 1;
  if (P.nb_shift >= 0) {
    E.qb = el0 >> P.nb_shift;
    E.nsl = (el1 >> P.nb_shift) - E.qb + 1;
  } else {
    E.qb = el0 / P.nb;
    E.nsl = el1 / P.nb - E.qb + 1;
  }
  E.sl = q - E.qb;
}

// HOP_RHSGROUP decode (Wilson kind, nb a power of two, no partials): lane =
// (g S + q) nb + j with g < G rhs and q < S = 64 / (G nb) consecutive sites;
// logical block sb = (rhs group, block of W S sites), W wavefronts per block
__device__ __forceinline__ HopElem hop_decode_rhsgroup(const HopParams& P, int64_t sb64) {
  HopElem E;
  const int sb = (int)sb64;
  const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6), W = (int)(blockDim.x >> 6);
  const int s_shift = 6 - P.rg_shift - P.nb_shift;   // log2 S
  const int per_g = P.nsites_out / (W << s_shift);
  const int rg = sb / per_g;
  const int sblk = sb - rg * per_g;
  const int gq = lane >> P.nb_shift;
  const int g = gq >> s_shift;
  E.j = lane & (P.nb - 1);
  E.rhs = (rg << P.rg_shift) + g;
  E.site = ((sblk * W + w) << s_shift) + (gq - (g << s_shift));
  E.pslot = 0;   // no partials in this mode
  E.qb = E.nsl = E.sl = E.b = 0;
  return E;
}

// Block-uniform quantities (rhs, sub-block) are computed with 32-bit
// arithmetic on the block index (scalar registers); per thread only the
// element within the sub-block is split into (site, j), with a shift when
// nb is a power of two.  Counts per rhs stay below 2^31 (checked on the
// host).
template <int KIND>
__device__ __forceinline__ HopElem hop_decode(const HopParams& P, int64_t sb64, int64_t nsub64) {
  HopElem E;
  const int sb = (int)sb64, nsub = (int)nsub64;
  const int el_base = threadIdx.x;
  if (KIND == KIND_WILSON && (P.flags & HOP_SPATIAL)) {
    // sb = (spatial block b = XCD, rhs, sub-block within (b, rhs))
    const int per_b = nsub >> 3;
    const int b = sb / per_b;
    const int n_in = sb - b * per_b;
    const int bpr = P.bpr_b;   // sub-blocks per (b, rhs)
    E.rhs = n_in / bpr;
    const int blk = n_in - E.rhs * bpr;
    const int el0 = blk * (int)blockDim.x;
    const int el = el0 + el_base;
    const int q = P.nb_shift >= 0 ? (el >> P.nb_shift) : el / P.nb;
    E.j = el - q * P.nb;
    E.site = (int)P.perm[(int64_t)b * P.perm_S + q];
    E.pslot = (int64_t)E.rhs * (bpr << 3) + (int64_t)b * bpr + blk;
    E.b = b;
    hop_block_sites(P, el0, q, E);
  } else {
    const int bpr = P.bpr;     // blocks per rhs
    E.rhs = sb / bpr;
    const int blk = sb - E.rhs * bpr;
    const int el0 = blk * (int)blockDim.x;
    const int el = el0 + el_base;
    E.site = P.nb_shift >= 0 ? (el >> P.nb_shift) : el / P.nb;
    E.j = el - E.site * P.nb;
    E.pslot = sb;
    E.b = 0;
    hop_block_sites(P, el0, E.site, E);
  }
  return E;
}

// LDS staging of the links (HOP_LDS_LINKS): thread t < 8 nsl loads the link
// of direction d = t % 8 of block site sl = t / 8, as hop_batch would, and
// stores it in the f16 word form of the dot2 product (third row rebuilt,
// boundary factor applied).  Every thread of the block calls it; the caller
// synchronises the block before reading.
template <typename LK, int KIND>
__device__ __forceinline__ void lds_fill_links(const HopParams& P, const HopElem& E, uint32_t* lds) {
  const int t = threadIdx.x;
  if (t >= 8 * E.nsl) return;
  const int sl = t >> 3, d = t & 7, mu = d & 3;
  const int q = E.qb + sl;
  const Geom& g = P.geom;
  const int site = (KIND == KIND_WILSON && (P.flags & HOP_SPATIAL)) ? (int)P.perm[(int64_t)E.b * P.perm_S + q] : q;
  int out_par, site_cb;
  if (KIND == KIND_DWF) {
    out_par = site >= g.Vcb ? 1 : 0;
    site_cb = site - out_par * g.Vcb;
  } else {
    out_par = P.out_parity;
    site_cb = site;
  }
  const int64_t ns = (int64_t)out_par * g.Vcb + site_cb;
  int64_t lidx;
  if (d < 4) {
    lidx = (int64_t)(out_par * 4 + mu) * g.Vcb + site_cb;
  } else {
    const uint32_t nidx = reinterpret_cast<const uint32_t*>(P.nbr)[(2 * (int64_t)g.Vcb + ns) * 4 + mu];
    lidx = (int64_t)((1 - out_par) * 4 + mu) * g.Vcb + nidx;
  }
  const float2 ph = link_phase(P, P.nbr_neg[ns], d);
  LinkRaw<LK> lr;
  link_raw_load<LK>(lr, P.gauge, lidx);
  uint32_t U[9];
  link_words_f16<LK>(lr, ph, U);
  uint4* w = reinterpret_cast<uint4*>(lds + sl * LDS_SITE_WORDS + 8 * d);
  w[0] = make_uint4(U[0], U[1], U[2], U[3]);
  w[1] = make_uint4(U[4], U[5], U[6], U[7]);
  lds[P.lds_sites * LDS_SITE_WORDS + t] = U[8];
}

// neighbour indices and boundary bits of the element's site, precomputed per
// (parity, site) by k_build_nbr (the j lanes of a site read one address)
template <int KIND>
__device__ __forceinline__ Nbr hop_nbr(const HopParams& P, const HopElem& E) {
  const Geom& g = P.geom;
  int out_par, site_cb;
  if (KIND == KIND_DWF) {
    out_par = E.site >= g.Vcb ? 1 : 0;
    site_cb = E.site - out_par * g.Vcb;
  } else {
    out_par = P.out_parity;
    site_cb = E.site;
  }
  Nbr N;
  const int64_t ns = (int64_t)out_par * g.Vcb + site_cb;
  const uint4 a = P.nbr[ns], b = P.nbr[2 * (int64_t)g.Vcb + ns];
  N.idx[0] = (int)a.x; N.idx[1] = (int)a.y; N.idx[2] = (int)a.z; N.idx[3] = (int)a.w;
  N.idx[4] = (int)b.x; N.idx[5] = (int)b.y; N.idx[6] = (int)b.z; N.idx[7] = (int)b.w;
  N.neg = P.nbr_neg[ns];
  return N;
}

// ---- Moebius fifth-dimension stage across a 16-lane row -----------------
// With nb = Ls = 16 the slots s = 0..15 of a site are lanes 16r .. 16r+15 of
// the wavefront (every thread mapping keeps them so), and the M5^-1
// recurrence (m5inv_regs) becomes log-step scans over the row: DPP row
// shifts (slots from outside the row read 0), rotations and the broadcast of
// the row's lane 15.
enum { DPP_ROW_SHL = 0x100, DPP_ROW_SHR = 0x110, DPP_ROW_ROR = 0x120, DPP_ROW_NEWBCAST = 0x150 };

template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
}

// ---- HOP_LINK_ROW: links shared by a DPP row ------------------------------
// With nb = 16 in element order the lanes 16r .. 16r+15 of a wavefront are
// the 16 slots of one site, and every direction's link is the same for all
// of them.  Lane j of the row loads 32-bit word min(j, NW - 1) of the link
// (one 4-byte load and one register per link instead of 24 or 36 bytes), and
// the row's lanes exchange the words by row broadcasts (v_mov_b32 with
// row_newbcast); the words then take the path of a per-lane load.
template <typename LK> struct LinkRowWords { static constexpr int n = 6; };   // 12-real: rows 0, 1
template <> struct LinkRowWords<LinkH18> { static constexpr int n = 9; };
template <> struct LinkRowWords<LinkB18> { static constexpr int n = 9; };

// byte offset of lane j's word of link 0 and the stride between links: the
// 16-byte plane holds words 0-3 (and the second 16-byte plane words 4-7 of
// 18-real links); the last plane holds word 8 (18-real, 4 bytes per link)
// or words 4, 5 (12-real, 8 bytes per link)
template <typename LK>
__device__ __forceinline__ void link_row_addr(const GaugeRef& g, int j, uint32_t& off, uint32_t& stride) {
  constexpr int NW = LinkRowWords<LK>::n;
  const int w = j < NW ? j : NW - 1;
  const uint32_t pb = (uint32_t)(g.plane * 16);
  if (w < 4) {
    off = 4u * w;
    stride = 16;
  } else if (NW == 9 && w < 8) {
    off = pb + 4u * (w - 4);
    stride = 16;
  } else if (NW == 9) {
    off = 2 * pb;
    stride = 4;
  } else {
    off = pb + 4u * (w - 4);
    stride = 8;
  }
}

template <int W>
__device__ __forceinline__ uint32_t row_word(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, DPP_ROW_NEWBCAST + W, 0xF, 0xF, true);
}

// the link as a per-lane load would have delivered it, from the row's words
template <typename LK>
__device__ __forceinline__ LinkRaw<LK> link_from_row(uint32_t v) {
  LinkRaw<LK> r;
  r.a = make_uint4(row_word<0>(v), row_word<1>(v), row_word<2>(v), row_word<3>(v));
  if constexpr (LinkRowWords<LK>::n == 9) {
    r.b = make_uint4(row_word<4>(v), row_word<5>(v), row_word<6>(v), row_word<7>(v));
    r.c = row_word<8>(v);
  } else {
    r.b = make_uint2(row_word<4>(v), row_word<5>(v));
  }
  return r;
}

// hop_batch with the links of HOP_LINK_ROW (half-precision spinors)
template <int D0, int NB, typename LK, int DAG>
__device__ __forceinline__ void hop_batch_row(const HopParams& P, const Nbr& N, int out_par, int site_cb,
                                              int64_t in_base, int j, uint32_t loff, uint32_t lstride,
                                              cplx<float> acc[12]) {
  const int Vcb = P.geom.Vcb;
  SpinRaw<SpinH> sr[NB];
  uint32_t lw[NB];
#pragma unroll
  for (int k = 0; k < NB; k++) {
    const int d = D0 + k;
    raw_load<SpinH>(sr[k], P.in, in_base + (int64_t)N.idx[d] * P.nb + j);
    const int mu = d & 3;
    const uint32_t lidx = (d < 4) ? (uint32_t)((out_par * 4 + mu) * Vcb + site_cb)
                                  : (uint32_t)(((1 - out_par) * 4 + mu) * Vcb + N.idx[d]);
    lw[k] = ld_off<uint32_t>(P.gauge.data, 0, loff + lidx * lstride);
  }
#define CPSHIP_DIR(K)                                                                             \
  if constexpr ((K) < NB) {                                                                       \
    const float2 ph = link_phase(P, N.neg, D0 + (K));                                             \
    uint32_t U[9];                                                                                \
    link_words_f16<LK>(link_from_row<LK>(lw[K]), ph, U);                                          \
    dir_compute_h2w<D0 + (K), DAG>(sr[K], U, acc);                                                \
  }
  CPSHIP_DIR(0) CPSHIP_DIR(1) CPSHIP_DIR(2) CPSHIP_DIR(3) CPSHIP_DIR(4) CPSHIP_DIR(5) CPSHIP_DIR(6) CPSHIP_DIR(7)
#undef CPSHIP_DIR
}

// forward type: y_s = sum_{s' <= s} a^(s-s') x_s' (scan), then
// y_s += fF_s y_15 (s < 15), y_15 *= inv_d
__device__ __forceinline__ float m5_fwd(float x, bool last, const float (&ap)[4], float fF, float inv_d) {
  float y = x;
  y = fmaf(ap[0], dpp_f<DPP_ROW_SHR + 1>(y), y);
  y = fmaf(ap[1], dpp_f<DPP_ROW_SHR + 2>(y), y);
  y = fmaf(ap[2], dpp_f<DPP_ROW_SHR + 4>(y), y);
  y = fmaf(ap[3], dpp_f<DPP_ROW_SHR + 8>(y), y);
  const float y15 = dpp_f<DPP_ROW_NEWBCAST + 15>(y);
  return last ? y * inv_d : fmaf(fF, y15, y);
}

// backward type: x_15 <- inv_d x_15 + sum_{s < 15} fB_s x_s, then the
// suffix scan y_s = sum_{s' >= s} a^(s'-s) x_s'
__device__ __forceinline__ float m5_bwd(float x, bool last, const float (&ap)[4], float fB, float inv_d) {
  float z = fB * x;   // fB_15 = 0
  z += dpp_f<DPP_ROW_SHR + 1>(z);
  z += dpp_f<DPP_ROW_SHR + 2>(z);
  z += dpp_f<DPP_ROW_SHR + 4>(z);
  z += dpp_f<DPP_ROW_SHR + 8>(z);
  float y = last ? fmaf(inv_d, x, z) : x;
  y = fmaf(ap[0], dpp_f<DPP_ROW_SHL + 1>(y), y);
  y = fmaf(ap[1], dpp_f<DPP_ROW_SHL + 2>(y), y);
  y = fmaf(ap[2], dpp_f<DPP_ROW_SHL + 4>(y), y);
  y = fmaf(ap[3], dpp_f<DPP_ROW_SHL + 8>(y), y);
  return y;
}

// b5 v + c5 (P v) for one real of the upper (UP) or lower spins; FROM_M: the
// shifted term comes from slot s-1 (else s+1), -mass at the wall
template <bool FROM_M>
__device__ __forceinline__ float s5_row(float v, float b5, float cw) {
  const float sh = FROM_M ? dpp_f<DPP_ROW_ROR + 1>(v) : dpp_f<DPP_ROW_ROR + 15>(v);
  return fmaf(cw, sh, b5 * v);
}

template <int DAG>
__device__ __forceinline__ void m5_row(const HopParams& P, int s, cplx<float> (&res)[12]) {
  const float a = float(P.m5a), inv_d = float(P.m5inv_d), mass = float(P.m5mass);
  const float b5 = float(P.m5b), c5 = float(P.m5c);
  float ap[4];
  ap[0] = a;
  ap[1] = a * a;
  ap[2] = ap[1] * ap[1];
  ap[3] = ap[2] * ap[2];
  float as = 1.f;   // a^s
  if (s & 1) as *= ap[0];
  if (s & 2) as *= ap[1];
  if (s & 4) as *= ap[2];
  if (s & 8) as *= ap[3];
  const bool last = s == 15;
  const float fF = last ? 0.f : -mass * inv_d * a * as;   // - mass inv_d a^(s+1)
  const float fB = last ? 0.f : -a * mass * inv_d * as;   // K mass inv_d a^s
  const float cm = c5 * (s == 0 ? -mass : 1.f), cp = c5 * (last ? -mass : 1.f);
  if constexpr (DAG == 0) {
    // M5^-1: upper spins forward type, lower backward type; then the s-term
    // with P of dag 0 (upper from s-1, lower from s+1)
#pragma unroll
    for (int k = 0; k < 6; k++) {
      res[k].re = s5_row<true>(m5_fwd(res[k].re, last, ap, fF, inv_d), b5, cm);
      res[k].im = s5_row<true>(m5_fwd(res[k].im, last, ap, fF, inv_d), b5, cm);
    }
#pragma unroll
    for (int k = 6; k < 12; k++) {
      res[k].re = s5_row<false>(m5_bwd(res[k].re, last, ap, fB, inv_d), b5, cp);
      res[k].im = s5_row<false>(m5_bwd(res[k].im, last, ap, fB, inv_d), b5, cp);
    }
  } else {
    // the s-term with P of dag 1 (upper from s+1, lower from s-1), then
    // M5^-dag: upper spins backward type, lower forward type
#pragma unroll
    for (int k = 0; k < 6; k++) {
      res[k].re = m5_bwd(s5_row<false>(res[k].re, b5, cp), last, ap, fB, inv_d);
      res[k].im = m5_bwd(s5_row<false>(res[k].im, b5, cp), last, ap, fB, inv_d);
    }
#pragma unroll
    for (int k = 6; k < 12; k++) {
      res[k].re = m5_fwd(s5_row<true>(res[k].re, b5, cm), last, ap, fF, inv_d);
      res[k].im = m5_fwd(s5_row<true>(res[k].im, b5, cm), last, ap, fF, inv_d);
    }
  }
}

// One output element: the hop, the fifth-dimension term (DWF), the epilogue
// and the store; returns the element's |out|^2 when partials are requested
// (|r|^2 of the updated residual for CLOV_S5XRR), summed in FP64.
template <typename SIN, typename SOUT, typename LK, typename R, int DAG, int KIND, bool NT, int CLOV, int LDSL,
          bool ROWL = false>
__device__ __forceinline__ double hop_element(const HopParams& P, const HopElem& E, const Nbr& N, uint32_t* lds) {
  const int rhs = E.rhs, site = E.site, j = E.j;
  const int64_t per_rhs = (int64_t)P.nsites_out * P.nb;
  const int64_t rem = (int64_t)site * P.nb + j;
  const bool valid = rhs < P.nrhs;
  double nrm = 0.0;
  const Geom& g = P.geom;
  int out_par, site_cb;
  if (KIND == KIND_DWF) {       // fields hold both 4-D parities
    out_par = site >= g.Vcb ? 1 : 0;
    site_cb = site - out_par * g.Vcb;
  } else {
    out_par = P.out_parity;
    site_cb = site;
  }
  int64_t in_base;
  if (KIND == KIND_DWF) in_base = ((int64_t)rhs * 2 * g.Vcb + (int64_t)(1 - out_par) * g.Vcb) * P.nb;
  else in_base = (int64_t)rhs * g.Vcb * P.nb;
  cplx<R> acc[12];
#pragma unroll
  for (int k = 0; k < 12; k++) acc[k] = cmk<R>(R(0), R(0));
  if constexpr (LDSL) {
    // the spinor loads go out before the block waits for its links
    constexpr int NBL = CPSHIP_HOP_BATCH_LDS;
    SpinRaw<SpinH> sr[NBL];
    if (valid) hop_lds_load<0, NBL>(P, N, in_base, j, sr);
    lds_fill_links<LK, KIND>(P, E, lds);
    __syncthreads();
    if (valid) {
      hop_lds_compute<0, NBL, DAG>(P, lds, E.sl, sr, acc);
      if constexpr (NBL == 4) {
        hop_lds_load<4, 4>(P, N, in_base, j, sr);
        hop_lds_compute<4, 4, DAG>(P, lds, E.sl, sr, acc);
      }
    }
  }
  if (valid) {
    constexpr int NB = (LDSL || ROWL) ? 0 : HopBatch<SIN>::value;
    if constexpr (ROWL && std::is_same<SIN, SpinH>::value) {
      uint32_t loff, lstride;
      link_row_addr<LK>(P.gauge, j, loff, lstride);
      hop_batch_row<0, 4, LK, DAG>(P, N, out_par, site_cb, in_base, j, loff, lstride, acc);
      hop_batch_row<4, 4, LK, DAG>(P, N, out_par, site_cb, in_base, j, loff, lstride, acc);
    } else if constexpr (NB == 0) {
    } else if constexpr (NB == 8) {
      hop_batch<0, 8, SIN, LK, R, DAG>(P, N, out_par, site_cb, in_base, j, acc);
    } else if constexpr (NB == 4) {
      hop_batch<0, 4, SIN, LK, R, DAG>(P, N, out_par, site_cb, in_base, j, acc);
      hop_batch<4, 4, SIN, LK, R, DAG>(P, N, out_par, site_cb, in_base, j, acc);
    } else if constexpr (NB == 2) {
      hop_batch<0, 2, SIN, LK, R, DAG>(P, N, out_par, site_cb, in_base, j, acc);
      hop_batch<2, 2, SIN, LK, R, DAG>(P, N, out_par, site_cb, in_base, j, acc);
      hop_batch<4, 2, SIN, LK, R, DAG>(P, N, out_par, site_cb, in_base, j, acc);
      hop_batch<6, 2, SIN, LK, R, DAG>(P, N, out_par, site_cb, in_base, j, acc);
    } else {
      hop_batch<0, 1, SIN, LK, R, DAG>(P, N, out_par, site_cb, in_base, j, acc);
      hop_batch<1, 1, SIN, LK, R, DAG>(P, N, out_par, site_cb, in_base, j, acc);
      hop_batch<2, 1, SIN, LK, R, DAG>(P, N, out_par, site_cb, in_base, j, acc);
      hop_batch<3, 1, SIN, LK, R, DAG>(P, N, out_par, site_cb, in_base, j, acc);
      hop_batch<4, 1, SIN, LK, R, DAG>(P, N, out_par, site_cb, in_base, j, acc);
      hop_batch<5, 1, SIN, LK, R, DAG>(P, N, out_par, site_cb, in_base, j, acc);
      hop_batch<6, 1, SIN, LK, R, DAG>(P, N, out_par, site_cb, in_base, j, acc);
      hop_batch<7, 1, SIN, LK, R, DAG>(P, N, out_par, site_cb, in_base, j, acc);
    }

    if constexpr (KIND == KIND_DWF) {
      // fifth dimension: input slots at the same 4-D site (same parity block)
      const int ls = P.ls;
      const int cls_out = (out_par + P.out_cb5) & 1;        // s = 2j + cls_out
      const int s = 2 * j + cls_out;
      const int64_t site_base = ((int64_t)rhs * 2 * g.Vcb + (int64_t)out_par * g.Vcb + site_cb) * P.nb;
      // in slots at this site hold s' = 2j' + (1 - cls_out)
      const int sm = (s == 0) ? ls - 1 : s - 1;
      const int sp = (s == ls - 1) ? 0 : s + 1;
      // wall coefficient -2 mf / a5: one for the launch, or per rhs (CLOV_KS:
      // kwall, a kernel of its own so that no conditional load enters the
      // stencil)
      const R cw = CLOV == CLOV_KS ? R(P.kwall[P.rhs0 + rhs]) : R(P.neg_mass_two_over_a5);
      const R cm = (s == 0) ? cw : R(P.two_over_a5);
      const R cp = (s == ls - 1) ? cw : R(P.two_over_a5);
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
    if constexpr (CLOV == CLOV_S5X || CLOV == CLOV_S5XR || CLOV == CLOV_S5XRR) {
      // j = s of the Moebius layout; the Ls slots of this site are adjacent
      const int ls = P.nb;
      const int64_t sbase = eo - j;
      const int sm = (j == 0) ? ls - 1 : j - 1;
      const int sp = (j == ls - 1) ? 0 : j + 1;
      const R cm = (j == 0) ? R(-P.x5mass) : R(1);
      const R cp = (j == ls - 1) ? R(-P.x5mass) : R(1);
      // P of the hop's own dag (compile-time component offsets keep res[]
      // in registers)
      constexpr int c_from_m = DAG ? 6 : 0;
      constexpr int c_from_p = DAG ? 0 : 6;
      cplx<R> xv[12], hm[6], hp[6];
      spin_load<SOUT, R>(P.X, eo, xv);
      constexpr bool ROW = (CLOV == CLOV_S5XR || CLOV == CLOV_S5XRR) && std::is_same<R, float>::value;
      if constexpr (ROW) {
#pragma unroll
        for (int k = 0; k < 6; k++) {
          hm[k] = cmk<R>(dpp_f<DPP_ROW_ROR + 1>(xv[c_from_m + k].re), dpp_f<DPP_ROW_ROR + 1>(xv[c_from_m + k].im));
          hp[k] = cmk<R>(dpp_f<DPP_ROW_ROR + 15>(xv[c_from_p + k].re), dpp_f<DPP_ROW_ROR + 15>(xv[c_from_p + k].im));
        }
      } else {
        spin_load_half<SOUT, R>(P.X, sbase + sm, c_from_m, hm);
        spin_load_half<SOUT, R>(P.X, sbase + sp, c_from_p, hp);
      }
      const R xc = R(P.xcoef), x5 = R(P.x5coef), sc = R(P.scale);
      const R sb = R(P.scale * P.hb5);
#pragma unroll
      for (int k = 0; k < 12; k++) res[k] = cmk<R>(xc * xv[k].re + sb * acc[k].re, xc * xv[k].im + sb * acc[k].im);
#pragma unroll
      for (int k = 0; k < 6; k++) {
        res[c_from_m + k].re += x5 * (cm * hm[k].re);
        res[c_from_m + k].im += x5 * (cm * hm[k].im);
        res[c_from_p + k].re += x5 * (cp * hp[k].re);
        res[c_from_p + k].im += x5 * (cp * hp[k].im);
      }
      if (P.hc5 != 0.0) {
        // P applied to the hop itself: the hop of slot s-1 (s+1) is the one
        // of the lane below (above) in the site's aligned group of nb lanes
        const int lane = (int)(threadIdx.x & 63);
        const int lm = (j == 0) ? lane + ls - 1 : lane - 1;
        const int lp = (j == ls - 1) ? lane - (ls - 1) : lane + 1;
        const R scm = R(P.scale * P.hc5) * cm, scp = R(P.scale * P.hc5) * cp;
#pragma unroll
        for (int k = 0; k < 6; k++) {
          R mre, mim, pre, pim;
          if constexpr (ROW) {
            mre = dpp_f<DPP_ROW_ROR + 1>(acc[c_from_m + k].re);
            mim = dpp_f<DPP_ROW_ROR + 1>(acc[c_from_m + k].im);
            pre = dpp_f<DPP_ROW_ROR + 15>(acc[c_from_p + k].re);
            pim = dpp_f<DPP_ROW_ROR + 15>(acc[c_from_p + k].im);
          } else {
            mre = __shfl(acc[c_from_m + k].re, lm, 64);
            mim = __shfl(acc[c_from_m + k].im, lm, 64);
            pre = __shfl(acc[c_from_p + k].re, lp, 64);
            pim = __shfl(acc[c_from_p + k].im, lp, 64);
          }
          res[c_from_m + k].re += scm * mre;
          res[c_from_m + k].im += scm * mim;
          res[c_from_p + k].re += scp * pre;
          res[c_from_p + k].im += scp * pim;
        }
      }
    } else if constexpr (CLOV != CLOV_NONE && CLOV != CLOV_KS && CLOV != CLOV_NRR) {
      const int set = (P.rhs0 + rhs) / P.rhs_per_set;
      // (no branch here: a conditional load splits the block and the stencil
      // above it then spills)
      R sc;
      if constexpr (CLOV == CLOV_X || CLOV == CLOV_XRR) sc = R(P.scale * P.kscale[set]);
      else sc = R(P.scale);
      const int64_t c0 = (int64_t)(set * 2 + out_par) * CloverOf<SIN>::nplanes * P.clover.plane + site_cb;
      if constexpr (CLOV == CLOV_POST) {
#pragma unroll
        for (int k = 0; k < 12; k++) acc[k] = cmk<R>(sc * acc[k].re, sc * acc[k].im);
        clover_mul<SIN, R>(P.clover, c0, acc, res);
      } else {
        cplx<R> xv[12], ax[12];
        spin_load<SOUT, R>(P.X, eo, xv);
        clover_mul<SIN, R>(P.clover, c0, xv, ax);
        const R xc = R(P.xcoef);
#pragma unroll
        for (int k = 0; k < 12; k++)
          res[k] = cmk<R>(xc * ax[k].re + sc * acc[k].re, xc * ax[k].im + sc * acc[k].im);
      }
    } else {
      // scale of the hop: one for the launch, or (CLOV_KS, Wilson kind)
      // times kscale[rhs0 + rhs], the per-rhs kappa^2 of the Wilson M
      R sc;
      if constexpr (CLOV == CLOV_KS && KIND == KIND_WILSON) sc = R(P.scale * P.kscale[P.rhs0 + rhs]);
      else sc = R(P.scale);
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
      if constexpr (std::is_same<SIN, SpinH>::value && KIND == KIND_WILSON) {
        if (P.m5post) m5_row<DAG>(P, j, res);
      }
    }
    if constexpr (CLOV == CLOV_S5XRR || CLOV == CLOV_XRR || CLOV == CLOV_NRR) {
      // the CG step r -= alpha (M^dag M p) and |r|^2, with M^dag M p in FP32
      if (P.ractive[rhs]) {
        const R a = R(P.ralpha[rhs]);
        cplx<R> rv[12];
        spin_load<SpinH, R>(P.R, eo, rv);
#pragma unroll
        for (int k = 0; k < 12; k++) {
          rv[k].re -= a * res[k].re;
          rv[k].im -= a * res[k].im;
          nrm += double(rv[k].re) * rv[k].re + double(rv[k].im) * rv[k].im;
        }
        spin_store<SpinH, R>(P.R, eo, rv);
      }
    } else {
      spin_store<SOUT, R, NT>(P.out, eo, res);
      if (P.partial) {
#pragma unroll
        for (int k = 0; k < 12; k++) nrm += double(res[k].re) * res[k].re + double(res[k].im) * res[k].im;
      }
    }
  }
  return nrm;
}


// SW = 0: one output element per thread in element order; SW = 1
// (HOP_SITEWAVE): one site per wavefront, so the site's neighbour entries
// and links are read with scalar loads; SW = 2 (HOP_LDS_LINKS): element
// order, links staged per block in LDS; SW = 3 (HOP_RHSGROUP): S sites x G
// rhs per wavefront; SW = 4 (HOP_LINK_ROW): element order, links shared
// across each 16-lane row.
template <typename SIN, typename SOUT, typename LK, typename R, int DAG, int KIND, bool NT, int CLOV, int SW>
__global__ void __launch_bounds__(256, HopWaves<SIN>::value) k_hop(HopParams P) {
  extern __shared__ uint4 hop_lds_area[];   // HOP_LDS_LINKS (SW = 2): the block's links
  const int64_t nsub = (int64_t)gridDim.x;
  const int64_t hb = logical_block(blockIdx.x, gridDim.x, !(P.flags & HOP_NO_XCD_REMAP));
  const HopElem E = SW == 1 ? hop_decode_sitewave(P, hb, nsub)
                   : SW == 3 ? hop_decode_rhsgroup(P, hb) : hop_decode<KIND>(P, hb, nsub);
  const Nbr N = hop_nbr<KIND>(P, E);
  const double nrm = hop_element<SIN, SOUT, LK, R, DAG, KIND, NT, CLOV, SW == 2, SW == 4>(
      P, E, N, reinterpret_cast<uint32_t*>(hop_lds_area));
  if (P.partial) {
    if (!P.chunk_part) block_partial<double>(nrm, P.partial, E.pslot);
    else if constexpr (SW == 3) chunk_partial_rhsgroup(P, hb, nrm);
    else chunk_partial_elem(P, E.rhs, E.site, nrm);
  }
}

}  // namespace cpship
