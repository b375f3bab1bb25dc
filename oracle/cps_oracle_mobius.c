/*
 * cps_oracle_mobius.c -- CPU restatement of the CPS 4-D even-odd
 * preconditioned Moebius operator (DiracOpMobius with ZMOB_PC_ORIG and the
 * constant real b, c that d_op_mobius.C:110-122 installs) and of the clover
 * operator (DiracOpClover).
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP path; see cps_oracle.h).
 *
 * Moebius fields use DWF_4D_EOPREC order restricted to one 4-D parity: Ls
 * consecutive slices of V/2 sites, slice stride 24*V/2.
 */
#include "cps_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define SPINOR 24

static int vol4(const int L[4]) { return L[0] * L[1] * L[2] * L[3]; }

/* d_op_mobius.C:106-109 */
void ora_mobius_kappas(const ora_mobius *m, double *kappa_b, double *kappa_c) {
  *kappa_b = 1.0 / (2 * (m->b * (4 - m->m5) + m->a5_inv));
  *kappa_c = 1.0 / (2 * (m->c * (4 - m->m5) - m->a5_inv));
}

/* out += z * in over n reals (zTimesV1PlusV2 with a real factor,
 * zmobius.h:257-263). */
static void axpy(double *out, double z, const double *in, int n) {
  int i;
  for (i = 0; i < n; i++) out[i] = z * in[i] + out[i];
}

/* zmobius_kappa_dslash_5_plus_cmplx, zmobius_dslash_5_plus.C:125-391 with a
 * constant real factor fact[s] = coef (s_nodes = 1).  Accumulates into out.
 * Pinned by: tests/test_oracle_pins.py::test_mobius_preconditioned_ops (the
 * Schur complement of an independently built sparse Moebius operator, b,c =
 * (1.5,0.5), (1,0), (2,1)) and ::test_mobius_m5inv. */
void ora_mobius_dslash_5_plus(double *out, const double *in, double mass,
                              int dag, double coef, const ora_mobius *m) {
  int ls = m->ls, vol_4d_cb = vol4(m->L) / 2;
  size_t ls_stride = (size_t)SPINOR * vol_4d_cb;
  double neg_mass_fact = -mass * coef;
  int s, x;
  int up = (dag == 1) ? 12 : 0; /* half hopping from s-1 */
  int dn = (dag == 0) ? 12 : 0; /* half hopping from s+1 */
  for (s = 1; s < ls; s++)
    for (x = 0; x < vol_4d_cb; x++)
      axpy(out + up + 24 * x + s * ls_stride, coef, in + up + 24 * x + (s - 1) * ls_stride, 12);
  for (x = 0; x < vol_4d_cb; x++)
    axpy(out + up + 24 * x, neg_mass_fact, in + up + 24 * x + (ls - 1) * ls_stride, 12);
  for (s = 0; s < ls - 1; s++)
    for (x = 0; x < vol_4d_cb; x++)
      axpy(out + dn + 24 * x + s * ls_stride, coef, in + dn + 24 * x + (s + 1) * ls_stride, 12);
  for (x = 0; x < vol_4d_cb; x++)
    axpy(out + dn + 24 * x + (ls - 1) * ls_stride, neg_mass_fact, in + dn + 24 * x, 12);
}

/* zmobius_dslash_4, zmobius_dslash_4.C:34-224: 4-D hop on every slice of one
 * 4-D parity, preceded (dag 0) or followed (dag 1) by b psi + c P(+-) psi(s-+1). */
void ora_mobius_dslash_4(double *out, const double *u, const double *in,
                         int cb, int dag, const ora_mobius *m, double mass) {
  int ls = m->ls;
  size_t f_size = (size_t)SPINOR * vol4(m->L) / 2;
  double *frm = (double *)malloc(f_size * ls * sizeof(double));
  size_t k;
  int i;
  if (!dag) {
    for (i = 0; i < ls; i++)
      for (k = 0; k < f_size; k++) frm[i * f_size + k] = m->b * in[i * f_size + k];
    ora_mobius_dslash_5_plus(frm, in, mass, dag, m->c, m);
    for (i = 0; i < ls; i++)
      ora_wilson_dslash(out + i * f_size, u, frm + i * f_size, cb, dag, m->L);
  } else {
    for (i = 0; i < ls; i++)
      ora_wilson_dslash(frm + i * f_size, u, in + i * f_size, cb, dag, m->L);
    for (i = 0; i < ls; i++)
      for (k = 0; k < f_size; k++) out[i * f_size + k] = m->b * frm[i * f_size + k];
    ora_mobius_dslash_5_plus(out, frm, mass, dag, m->c, m);
  }
  free(frm);
}

/* zmobius_m5inv_cmplx_dag0 / _dag1, zmobius_m5inv.C:247-686, for a constant
 * real K = kappa_b/kappa_c.  In place.
 * Pinned by: tests/test_oracle_pins.py::test_mobius_m5inv (M5 applied to the
 * result, with M5 built independently, gives the input back; even Ls). */
void ora_mobius_m5inv(double *f, double mass, int dag, const ora_mobius *m) {
  int ls = m->ls, vol = vol4(m->L) / 2;
  double K, kb, kc, prod, inv_d_last, factL, factR;
  int i, s, x;
  ora_mobius_kappas(m, &kb, &kc);
  K = kb / kc;
#define F(s_, x_) (f + 24 * ((x_) + vol * (s_)))
  if (dag == 0) {
    prod = 1.0;
    for (i = 0; i < ls; i++) prod *= -K;
    prod *= mass;
    inv_d_last = 1.0 / (1.0 + prod);
    for (x = 0; x < vol; x++)
      for (i = 0; i < 12; i++) F(ls - 1, x)[12 + i] *= inv_d_last;
    factL = K * mass * inv_d_last;
    for (s = 0; s <= ls - 2; ++s) {
      for (x = 0; x < vol; x++) {
        axpy(F(s + 1, x), -K, F(s, x), 12);
        axpy(F(ls - 1, x) + 12, factL, F(s, x) + 12, 12);
      }
      factL *= -K;
    }
    prod = 1.0;
    for (i = 0; i < ls - 1; i++) prod *= -K;
    factR = -prod * mass * inv_d_last;
    for (s = ls - 2; s >= 0; --s) {
      for (x = 0; x < vol; x++) {
        axpy(F(s, x), factR, F(ls - 1, x), 12);
        axpy(F(s, x) + 12, -K, F(s + 1, x) + 12, 12);
      }
      factR /= -K;
    }
    for (x = 0; x < vol; x++)
      for (i = 0; i < 12; i++) F(ls - 1, x)[i] *= inv_d_last;
  } else {
    prod = 1.0;
    for (i = 0; i < ls; i++) prod *= K;
    prod *= mass;
    inv_d_last = 1.0 / (1.0 + prod);
    for (x = 0; x < vol; x++)
      for (i = 0; i < 12; i++) F(ls - 1, x)[i] *= inv_d_last;
    factR = K * mass * inv_d_last;
    for (s = 0; s <= ls - 2; ++s) {
      for (x = 0; x < vol; x++) {
        axpy(F(ls - 1, x), factR, F(s, x), 12);
        axpy(F(s + 1, x) + 12, -K, F(s, x) + 12, 12);
      }
      factR *= -K;
    }
    prod = 1.0;
    for (i = 0; i < ls - 2; i++) prod *= -K;
    prod *= -K;
    factL = -prod * mass * inv_d_last;
    for (s = ls - 2; s >= 0; --s) {
      for (x = 0; x < vol; x++) {
        axpy(F(s, x), -K, F(s + 1, x), 12);
        axpy(F(s, x) + 12, factL, F(ls - 1, x) + 12, 12);
      }
      if (s > 0) factL /= -K;
    }
    for (x = 0; x < vol; x++)
      for (i = 0; i < 12; i++) F(ls - 1, x)[12 + i] *= inv_d_last;
  }
#undef F
}

/* zmobius_m_orig, zmobius_m-orig.h:3-122 */
void ora_mobius_m(double *out, const double *u, const double *in, double mass,
                  const ora_mobius *m, double *tmp2) {
  int ls = m->ls;
  size_t ls_stride = (size_t)SPINOR * vol4(m->L) / 2, f_size = ls_stride * ls, k;
  double kb, kc;
  int s;
  ora_mobius_kappas(m, &kb, &kc);
  ora_mobius_dslash_4(out, u, in, 1, 0, m, mass);
  for (s = 0; s < ls; s++)
    for (k = 0; k < ls_stride; k++) out[s * ls_stride + k] *= kb;
  ora_mobius_m5inv(out, mass, 0, m);
  ora_mobius_dslash_4(tmp2, u, out, 0, 0, m, mass);
  for (s = 0; s < ls; s++)
    for (k = 0; k < ls_stride; k++) tmp2[s * ls_stride + k] *= -kb;
  memcpy(out, in, f_size * sizeof(double));
  for (k = 0; k < f_size; k++) out[k] += tmp2[k];
  ora_mobius_dslash_5_plus(out, in, mass, 0, kb / kc, m);
}

/* zmobius_mdag_orig, zmobius_mdag-orig.h */
void ora_mobius_mdag(double *out, const double *u, const double *in,
                     double mass, const ora_mobius *m, double *tmp2,
                     double *tmp3) {
  int ls = m->ls;
  size_t ls_stride = (size_t)SPINOR * vol4(m->L) / 2, f_size = ls_stride * ls, k;
  double kb, kc;
  int s;
  (void)tmp3;
  ora_mobius_kappas(m, &kb, &kc);
  memcpy(out, in, f_size * sizeof(double));
  for (s = 0; s < ls; s++)
    for (k = 0; k < ls_stride; k++) out[s * ls_stride + k] *= -kb;
  ora_mobius_dslash_4(tmp2, u, out, 1, 1, m, mass);
  ora_mobius_m5inv(tmp2, mass, 1, m);
  for (s = 0; s < ls; s++)
    for (k = 0; k < ls_stride; k++) tmp2[s * ls_stride + k] *= kb;
  ora_mobius_dslash_4(out, u, tmp2, 0, 1, m, mass);
  for (k = 0; k < f_size; k++) out[k] += in[k];
  ora_mobius_dslash_5_plus(out, in, mass, 1, kb / kc, m);
}

/* mobius_mdagm, mobius_mdagm.C:33-60 */
void ora_mobius_mdagm(double *out, const double *u, const double *in,
                      double *dot_prd, double mass, const ora_mobius *m,
                      double *tmp1, double *tmp2, double *tmp3) {
  size_t f_size = (size_t)SPINOR * vol4(m->L) / 2 * m->ls;
  ora_mobius_m(tmp1, u, in, mass, m, tmp2);
  if (dot_prd) *dot_prd = ora_dot(tmp1, tmp1, f_size);
  ora_mobius_mdag(out, u, tmp1, mass, m, tmp2, tmp3);
}

typedef struct {
  const double *u;
  double mass;
  const ora_mobius *m;
  double *t1, *t2, *t3;
} mobius_ctx;

static void mobius_mdagm_cb(void *c, double *out, const double *in, double *dp) {
  mobius_ctx *w = (mobius_ctx *)c;
  ora_mobius_mdagm(out, w->u, in, dp, w->mass, w->m, w->t1, w->t2, w->t3);
}

int ora_mobius_invcg(double *sol, const double *src, const double *u,
                     double mass, double stop_rsd, int max_num_iter,
                     const ora_mobius *m, double *true_res) {
  size_t f = (size_t)SPINOR * vol4(m->L) / 2 * m->ls;
  mobius_ctx w;
  int it;
  w.u = u;
  w.mass = mass;
  w.m = m;
  w.t1 = (double *)malloc(f * sizeof(double));
  w.t2 = (double *)malloc(f * sizeof(double));
  w.t3 = (double *)malloc(f * sizeof(double));
  it = ora_invcg(sol, src, f, stop_rsd, max_num_iter, 0.0, true_res,
                 mobius_mdagm_cb, &w, NULL);
  free(w.t1);
  free(w.t2);
  free(w.t3);
  return it;
}

/* ======================================================================
 * Clover.  DiracOpClover::SiteFuv / SiteCloverMat / CloverMatChkb,
 * d_op_clover_supp.C:188-540 (isotropic lattice, so omega_xi = omega,
 * gjp.C:408); clover_mat_mlt_C, clover.C:127-172; MatPcDagOrNot and
 * MatPcDagMatPc, d_op_clover.C:156-248.  The inverse of each 6x6 block is
 * formed by an LDL^dagger factorisation of the packed Hermitian matrix as
 * mat_inv(MAT_INV_ALG_LDL_CMPR) does (dense_matrix.C:188-262); the rounding
 * of that step is this restatement's own.
 * Pinned by: tests/test_oracle_pins.py::test_clover_matrix_textbook (every
 * site's 12x12 matrix against the documented formula of dirac_op.h:930-951
 * built independently from the four leaves, A A^{-1} = 1) and
 * ::test_clover_preconditioned_ops (Schur complement of A - kappa H, c_sw = 0
 * reduces to Wilson).
 * ====================================================================== */
typedef struct { double re, im; } cplx;
typedef struct { cplx e[3][3]; } mat3;

static cplx cmul(cplx a, cplx b) { cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; return r; }
static cplx cadd(cplx a, cplx b) { cplx r = {a.re + b.re, a.im + b.im}; return r; }
static cplx cconj(cplx a) { cplx r = {a.re, -a.im}; return r; }

static mat3 mmul(const mat3 *a, const mat3 *b) {
  mat3 r;
  int i, j, k;
  for (i = 0; i < 3; i++)
    for (j = 0; j < 3; j++) {
      cplx s = {0, 0};
      for (k = 0; k < 3; k++) s = cadd(s, cmul(a->e[i][k], b->e[k][j]));
      r.e[i][j] = s;
    }
  return r;
}
static mat3 mdag(const mat3 *a) {
  mat3 r;
  int i, j;
  for (i = 0; i < 3; i++)
    for (j = 0; j < 3; j++) r.e[i][j] = cconj(a->e[j][i]);
  return r;
}

/* Link U_mu(site) in canonical orientation from a WILSON-order gauge field
 * (element (row,col) at 2*(row+3*col), convert_func.C:402-404). */
static mat3 get_link(const double *u, const int *site_in, int mu, const int L[4]) {
  int s[4], i, par = 0, off;
  mat3 r;
  int vol = vol4(L);
  for (i = 0; i < 4; i++) {
    s[i] = ((site_in[i] % L[i]) + L[i]) % L[i];
    par += s[i];
  }
  off = s[0] + L[0] * (s[1] + L[1] * (s[2] + L[2] * s[3]));
  off = (off - off % 2) / 2 + (par % 2) * (vol / 2);
  {
    const double *p = u + 72 * off + 18 * mu;
    int row, col;
    for (row = 0; row < 3; row++)
      for (col = 0; col < 3; col++) {
        r.e[row][col].re = p[2 * (row + 3 * col)];
        r.e[row][col].im = p[2 * (row + 3 * col) + 1];
      }
  }
  return r;
}

/* F_{mu nu}(x) = (Q - Q^dagger)/8, Q the sum of the four clover leaves
 * (dirac_op.h:940-952 and SiteFuv). */
static mat3 site_fuv(const double *u, const int *x, int mu, int nu, const int L[4]) {
  int a[4], i, j;
  mat3 q, t, r;
#define SITE(dmu, dnu) (a[0] = x[0], a[1] = x[1], a[2] = x[2], a[3] = x[3], a[mu] += (dmu), a[nu] += (dnu), a)
  mat3 Umu_x = get_link(u, SITE(0, 0), mu, L);
  mat3 Unu_x = get_link(u, SITE(0, 0), nu, L);
  mat3 Unu_xpmu = get_link(u, SITE(1, 0), nu, L);
  mat3 Umu_xpnu = get_link(u, SITE(0, 1), mu, L);
  mat3 Umu_xmmupnu = get_link(u, SITE(-1, 1), mu, L);
  mat3 Unu_xmmu = get_link(u, SITE(-1, 0), nu, L);
  mat3 Umu_xmmu = get_link(u, SITE(-1, 0), mu, L);
  mat3 Unu_xmmumnu = get_link(u, SITE(-1, -1), nu, L);
  mat3 Umu_xmmumnu = get_link(u, SITE(-1, -1), mu, L);
  mat3 Unu_xmnu = get_link(u, SITE(0, -1), nu, L);
  mat3 Umu_xmnu = get_link(u, SITE(0, -1), mu, L);
  mat3 Unu_xpmumnu = get_link(u, SITE(1, -1), nu, L);
#undef SITE
  mat3 d1, d2;
  /* leaf 1: U_mu(x) U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+ */
  t = mmul(&Umu_x, &Unu_xpmu);
  d1 = mdag(&Umu_xpnu); t = mmul(&t, &d1);
  d1 = mdag(&Unu_x); q = mmul(&t, &d1);
  /* leaf 2: U_nu(x) U_mu(x-mu+nu)^+ U_nu(x-mu)^+ U_mu(x-mu) */
  d1 = mdag(&Umu_xmmupnu); t = mmul(&Unu_x, &d1);
  d2 = mdag(&Unu_xmmu); t = mmul(&t, &d2);
  t = mmul(&t, &Umu_xmmu);
  for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) q.e[i][j] = cadd(q.e[i][j], t.e[i][j]);
  /* leaf 3: U_mu(x-mu)^+ U_nu(x-mu-nu)^+ U_mu(x-mu-nu) U_nu(x-nu) */
  d1 = mdag(&Umu_xmmu); d2 = mdag(&Unu_xmmumnu); t = mmul(&d1, &d2);
  t = mmul(&t, &Umu_xmmumnu);
  t = mmul(&t, &Unu_xmnu);
  for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) q.e[i][j] = cadd(q.e[i][j], t.e[i][j]);
  /* leaf 4: U_nu(x-nu)^+ U_mu(x-nu) U_nu(x+mu-nu) U_mu(x)^+ */
  d1 = mdag(&Unu_xmnu); t = mmul(&d1, &Umu_xmnu);
  t = mmul(&t, &Unu_xpmumnu);
  d2 = mdag(&Umu_x); t = mmul(&t, &d2);
  for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) q.e[i][j] = cadd(q.e[i][j], t.e[i][j]);
  for (i = 0; i < 3; i++)
    for (j = 0; j < 3; j++) {
      r.e[i][j].re = (q.e[i][j].re - q.e[j][i].re) * 0.125;
      r.e[i][j].im = (q.e[i][j].im + q.e[j][i].im) * 0.125;
    }
  return r;
}

static mat3 mscale_i(const mat3 *a, double w) { /* i*w*a */
  mat3 r;
  int i, j;
  for (i = 0; i < 3; i++)
    for (j = 0; j < 3; j++) {
      r.e[i][j].re = -w * a->e[i][j].im;
      r.e[i][j].im = w * a->e[i][j].re;
    }
  return r;
}

/* Packed 6x6 Hermitian layout (clover.C:113-121): diagonal (i,i) real at
 * i*i+2*i, lower element (i,j), i>j, complex at i*i+2*j. */
static void pack6(double *A, cplx M[6][6]) {
  int i, j;
  for (i = 0; i < 6; i++) {
    for (j = 0; j < i; j++) {
      A[i * i + 2 * j] = M[i][j].re;
      A[i * i + 2 * j + 1] = M[i][j].im;
    }
    A[i * i + 2 * i] = M[i][i].re;
  }
}
static void unpack6(cplx M[6][6], const double *A) {
  int i, j;
  for (i = 0; i < 6; i++) {
    for (j = 0; j < i; j++) {
      M[i][j].re = A[i * i + 2 * j];
      M[i][j].im = A[i * i + 2 * j + 1];
      M[j][i].re = M[i][j].re;
      M[j][i].im = -M[i][j].im;
    }
    M[i][i].re = A[i * i + 2 * i];
    M[i][i].im = 0;
  }
}

static void site_clover(double *A, const double *u, const int *x, double omega,
                        const int L[4]) {
  mat3 F01 = site_fuv(u, x, 0, 1, L), F02 = site_fuv(u, x, 0, 2, L);
  mat3 F03 = site_fuv(u, x, 0, 3, L), F12 = site_fuv(u, x, 1, 2, L);
  mat3 F13 = site_fuv(u, x, 1, 3, L), F23 = site_fuv(u, x, 2, 3, L);
  mat3 As[6], t1, t2;
  int i, j, blk;
  /* As[1] = omega (i F12 - F02) - omega (i F03 + F13)
   * As[4] = omega (i F12 - F02) + omega (i F03 + F13) */
  t1 = mscale_i(&F12, omega);
  t2 = mscale_i(&F03, omega);
  for (i = 0; i < 3; i++)
    for (j = 0; j < 3; j++) {
      cplx a, b;
      a.re = t1.e[i][j].re - omega * F02.e[i][j].re;
      a.im = t1.e[i][j].im - omega * F02.e[i][j].im;
      b.re = t2.e[i][j].re + omega * F13.e[i][j].re;
      b.im = t2.e[i][j].im + omega * F13.e[i][j].im;
      As[1].e[i][j].re = a.re - b.re;
      As[1].e[i][j].im = a.im - b.im;
      As[4].e[i][j].re = b.re + a.re;
      As[4].e[i][j].im = b.im + a.im;
    }
  /* As[0] = i omega (F01 - F23), As[3] = i omega (F01 + F23) */
  t1 = mscale_i(&F01, omega);
  t2 = mscale_i(&F23, omega);
  for (i = 0; i < 3; i++)
    for (j = 0; j < 3; j++) {
      As[0].e[i][j].re = t1.e[i][j].re - t2.e[i][j].re;
      As[0].e[i][j].im = t1.e[i][j].im - t2.e[i][j].im;
      As[3].e[i][j].re = t2.e[i][j].re + t1.e[i][j].re;
      As[3].e[i][j].im = t2.e[i][j].im + t1.e[i][j].im;
    }
  for (i = 0; i < 3; i++)
    for (j = 0; j < 3; j++) {
      double d = (i == j) ? 1.0 : 0.0;
      As[2].e[i][j].re = d - As[0].e[i][j].re;
      As[2].e[i][j].im = -As[0].e[i][j].im;
      As[5].e[i][j].re = d - As[3].e[i][j].re;
      As[5].e[i][j].im = -As[3].e[i][j].im;
      As[0].e[i][j].re += d;
      As[3].e[i][j].re += d;
    }
  /* block = [[As0, As1^+], [As1, As2]] (upper spins), [[As3,As4^+],[As4,As5]] */
  for (blk = 0; blk < 2; blk++) {
    cplx M[6][6];
    const mat3 *a = &As[3 * blk], *b = &As[3 * blk + 1], *c = &As[3 * blk + 2];
    for (i = 0; i < 3; i++)
      for (j = 0; j < 3; j++) {
        M[i][j] = a->e[i][j];
        M[i + 3][j] = b->e[i][j];
        M[i][j + 3] = cconj(b->e[j][i]);
        M[i + 3][j + 3] = c->e[i][j];
      }
    pack6(A + 36 * blk, M);
  }
}

/* Inverse of a packed Hermitian positive 6x6 block through L D L^dagger. */
static void inv6_ldl(double *A) {
  cplx M[6][6], Lm[6][6], X[6][6];
  double D[6];
  int i, j, k;
  unpack6(M, A);
  for (i = 0; i < 6; i++) {
    for (j = 0; j < i; j++) {
      cplx s = M[i][j];
      for (k = 0; k < j; k++) {
        cplx t = cmul(Lm[i][k], cconj(Lm[j][k]));
        s.re -= t.re * D[k];
        s.im -= t.im * D[k];
      }
      Lm[i][j].re = s.re / D[j];
      Lm[i][j].im = s.im / D[j];
    }
    {
      double s = M[i][i].re;
      for (k = 0; k < i; k++)
        s -= (Lm[i][k].re * Lm[i][k].re + Lm[i][k].im * Lm[i][k].im) * D[k];
      D[i] = s;
    }
    Lm[i][i].re = 1;
    Lm[i][i].im = 0;
    for (j = i + 1; j < 6; j++) Lm[i][j].re = Lm[i][j].im = 0;
  }
  /* solve A X = I column by column */
  for (j = 0; j < 6; j++) {
    cplx y[6];
    for (i = 0; i < 6; i++) {
      cplx s;
      s.re = (i == j) ? 1.0 : 0.0;
      s.im = 0;
      for (k = 0; k < i; k++) {
        cplx t = cmul(Lm[i][k], y[k]);
        s.re -= t.re;
        s.im -= t.im;
      }
      y[i] = s;
    }
    for (i = 0; i < 6; i++) { y[i].re /= D[i]; y[i].im /= D[i]; }
    for (i = 5; i >= 0; i--) {
      cplx s = y[i];
      for (k = i + 1; k < 6; k++) {
        cplx t = cmul(cconj(Lm[k][i]), X[k][j]);
        s.re -= t.re;
        s.im -= t.im;
      }
      X[i][j] = s;
    }
  }
  pack6(A, X);
}

void ora_clover_mat_chkb(double *A, const double *u, int cb, int inverse,
                         double kappa, double csw, const int L[4]) {
  double omega = kappa * csw;
  int t, vol = vol4(L);
  /* the sites of parity cb in the order x/2 + L0/2 (y + L1 (z + L2 t)) */
#pragma omp parallel for schedule(static)
  for (t = 0; t < L[3]; ++t) {
    int site[4];
    site[3] = t;
    for (site[2] = 0; site[2] < L[2]; ++site[2])
      for (site[1] = 0; site[1] < L[1]; ++site[1])
        for (site[0] = (site[3] + site[2] + site[1] + cb) % 2; site[0] < L[0]; site[0] += 2) {
          size_t idx = (size_t)(site[0] / 2) + (size_t)(L[0] / 2) * (site[1] + (size_t)L[1] * (site[2] + (size_t)L[2] * t));
          site_clover(A + 72 * idx, u, site, omega, L);
        }
  }
  if (inverse) {
    int i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < vol; i++) inv6_ldl(A + 36 * (size_t)i);
  }
}

/* clover_mat_mlt_C, clover.C:127-172: Y = A X per 6-component half spinor. */
void ora_clover_mat_mlt(double *out, const double *mat, const double *in, int n) {
  int i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < n * 2; ++i) {
    double *Y = out + 12 * (size_t)i;
    const double *A = mat + 36 * (size_t)i, *X = in + 12 * (size_t)i;
    Y[0] = A[0]*X[0] + A[1]*X[2] + A[2]*X[3] + A[4]*X[4] + A[5]*X[5] + A[9]*X[6] + A[10]*X[7] + A[16]*X[8] + A[17]*X[9] + A[25]*X[10] + A[26]*X[11];
    Y[1] = A[0]*X[1] + A[1]*X[3] - A[2]*X[2] + A[4]*X[5] - A[5]*X[4] + A[9]*X[7] - A[10]*X[6] + A[16]*X[9] - A[17]*X[8] + A[25]*X[11] - A[26]*X[10];
    Y[2] = A[1]*X[0] - A[2]*X[1] + A[3]*X[2] + A[6]*X[4] + A[7]*X[5] + A[11]*X[6] + A[12]*X[7] + A[18]*X[8] + A[19]*X[9] + A[27]*X[10] + A[28]*X[11];
    Y[3] = A[1]*X[1] + A[2]*X[0] + A[3]*X[3] + A[6]*X[5] - A[7]*X[4] + A[11]*X[7] - A[12]*X[6] + A[18]*X[9] - A[19]*X[8] + A[27]*X[11] - A[28]*X[10];
    Y[4] = A[4]*X[0] - A[5]*X[1] + A[6]*X[2] - A[7]*X[3] + A[8]*X[4] + A[13]*X[6] + A[14]*X[7] + A[20]*X[8] + A[21]*X[9] + A[29]*X[10] + A[30]*X[11];
    Y[5] = A[4]*X[1] + A[5]*X[0] + A[6]*X[3] + A[7]*X[2] + A[8]*X[5] + A[13]*X[7] - A[14]*X[6] + A[20]*X[9] - A[21]*X[8] + A[29]*X[11] - A[30]*X[10];
    Y[6] = A[9]*X[0] - A[10]*X[1] + A[11]*X[2] - A[12]*X[3] + A[13]*X[4] - A[14]*X[5] + A[15]*X[6] + A[22]*X[8] + A[23]*X[9] + A[31]*X[10] + A[32]*X[11];
    Y[7] = A[9]*X[1] + A[10]*X[0] + A[11]*X[3] + A[12]*X[2] + A[13]*X[5] + A[14]*X[4] + A[15]*X[7] + A[22]*X[9] - A[23]*X[8] + A[31]*X[11] - A[32]*X[10];
    Y[8] = A[16]*X[0] - A[17]*X[1] + A[18]*X[2] - A[19]*X[3] + A[20]*X[4] - A[21]*X[5] + A[22]*X[6] - A[23]*X[7] + A[24]*X[8] + A[33]*X[10] + A[34]*X[11];
    Y[9] = A[16]*X[1] + A[17]*X[0] + A[18]*X[3] + A[19]*X[2] + A[20]*X[5] + A[21]*X[4] + A[22]*X[7] + A[23]*X[6] + A[24]*X[9] + A[33]*X[11] - A[34]*X[10];
    Y[10] = A[25]*X[0] - A[26]*X[1] + A[27]*X[2] - A[28]*X[3] + A[29]*X[4] - A[30]*X[5] + A[31]*X[6] - A[32]*X[7] + A[33]*X[8] - A[34]*X[9] + A[35]*X[10];
    Y[11] = A[25]*X[1] + A[26]*X[0] + A[27]*X[3] + A[28]*X[2] + A[29]*X[5] + A[30]*X[4] + A[31]*X[7] + A[32]*X[6] + A[33]*X[9] + A[34]*X[8] + A[35]*X[11];
  }
}

/* DiracOpClover::MatPcDagOrNot, d_op_clover.C:194-248 */
void ora_clover_matpc(double *out, const double *u, const double *A_oo,
                      const double *A_ee_inv, const double *in, double kappa,
                      int dag, const int L[4], double *buf0, double *unused) {
  int half = vol4(L) / 2;
  (void)unused;
  ora_wilson_dslash(buf0, u, in, 1, dag, L);
  ora_clover_mat_mlt(out, A_ee_inv, buf0, half);
  ora_wilson_dslash(buf0, u, out, 0, dag, L);
  ora_clover_mat_mlt(out, A_oo, in, half);
  ora_ftimesv1plusv2(out, -kappa * kappa, buf0, out, (size_t)SPINOR * half);
}

/* DiracOpClover::MatPcDagMatPc, d_op_clover.C:156-181 */
void ora_clover_mdagm(double *out, const double *u, const double *A_oo,
                      const double *A_ee_inv, const double *in,
                      double *dot_prd, double kappa, const int L[4],
                      double *tmp1, double *tmp2, double *tmp3) {
  size_t f = (size_t)SPINOR * vol4(L) / 2;
  (void)tmp3;
  ora_clover_matpc(tmp1, u, A_oo, A_ee_inv, in, kappa, 0, L, tmp2, NULL);
  if (dot_prd) *dot_prd = ora_dot(tmp1, tmp1, f);
  ora_clover_matpc(out, u, A_oo, A_ee_inv, tmp1, kappa, 1, L, tmp2, NULL);
}

typedef struct {
  const double *u, *Aoo, *Aei;
  double kappa;
  const int *L;
  double *t1, *t2;
} clover_ctx;

static void clover_mdagm_cb(void *c, double *out, const double *in, double *dp) {
  clover_ctx *w = (clover_ctx *)c;
  ora_clover_mdagm(out, w->u, w->Aoo, w->Aei, in, dp, w->kappa, w->L, w->t1, w->t2, NULL);
}

int ora_clover_invcg(double *sol, const double *src, const double *u,
                     const double *A_oo, const double *A_ee_inv, double kappa,
                     double stop_rsd, int max_num_iter, const int L[4],
                     double *true_res) {
  size_t f = (size_t)SPINOR * vol4(L) / 2;
  clover_ctx w;
  int it;
  w.u = u;
  w.Aoo = A_oo;
  w.Aei = A_ee_inv;
  w.kappa = kappa;
  w.L = L;
  w.t1 = (double *)malloc(f * sizeof(double));
  w.t2 = (double *)malloc(f * sizeof(double));
  it = ora_invcg(sol, src, f, stop_rsd, max_num_iter, 0.0, true_res,
                 clover_mdagm_cb, &w, NULL);
  free(w.t1);
  free(w.t2);
  return it;
}
