/*
 * cps_oracle_wilson.c -- CPU restatement of CPS noarch layout conversion,
 * boundary conditions, the Wilson hopping term, the Shamir domain-wall
 * operator, the vector BLAS and DiracOp::InvCg.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the HIP path; see cps_oracle.h).
 * Parity status of each routine is recorded next to it ("pinned by").
 */
#include "cps_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define SPINOR 24
#define GAUGE 72

static int vol4(const int L[4]) { return L[0] * L[1] * L[2] * L[3]; }

/* Thread count of the OpenMP site loops (results do not depend on it). */
void ora_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int ora_get_threads(void) { return omp_get_max_threads(); }

/* ora_invcg prints |r|^2 to stderr every `every` iterations (0: silent), so
 * that a long CPU solve reports its own progress. */
static int progress_every = 0;
void ora_set_progress(int every) { progress_every = every > 0 ? every : 0; }

/* ======================================================================
 * Layout.  Follows convert_func.C:318-415 (gauge CANONICAL->WILSON,
 * site number (x - x%2 + lx*(y+ly*(z+lz*t)) + vol*parity)/2 and the link
 * reals moved from r+2*(col+3*(row+3*mu)) to r+2*(row+3*(col+3*mu))), and
 * convert_func.C:429-494 (fermion, odd sites first).
 * Pinned by: tests/test_oracle_pins.py::test_fermion_layouts,
 * ::test_gauge_layout, ::test_bond_cond (the index formulas restated
 * independently, round trips).
 * ====================================================================== */
static int wilson_gauge_site(const int L[4], int x, int y, int z, int t) {
  int vol = vol4(L);
  int par = (x + y + z + t) % 2;
  return (x - x % 2 + L[0] * (y + L[1] * (z + L[2] * t)) + vol * par) / 2;
}

static int wilson_ferm_site(const int L[4], int x, int y, int z, int t) {
  int vol = vol4(L);
  int par = (x + y + z + t + 1) % 2; /* odd sites first */
  return (x - x % 2 + L[0] * (y + L[1] * (z + L[2] * t)) + vol * par) / 2;
}

void ora_gauge_canon_to_wilson(double *dst, const double *src, const int L[4]) {
  int x, y, z, t, mu, row, col, r;
  for (t = 0; t < L[3]; t++)
    for (z = 0; z < L[2]; z++)
      for (y = 0; y < L[1]; y++)
        for (x = 0; x < L[0]; x++) {
          int cur = x + L[0] * (y + L[1] * (z + L[2] * t));
          int nw = wilson_gauge_site(L, x, y, z, t);
          for (mu = 0; mu < 4; mu++)
            for (row = 0; row < 3; row++)
              for (col = 0; col < 3; col++)
                for (r = 0; r < 2; r++)
                  dst[GAUGE * nw + r + 2 * (row + 3 * (col + 3 * mu))] =
                      src[GAUGE * cur + r + 2 * (col + 3 * (row + 3 * mu))];
        }
}

void ora_gauge_wilson_to_canon(double *dst, const double *src, const int L[4]) {
  int x, y, z, t, mu, row, col, r;
  for (t = 0; t < L[3]; t++)
    for (z = 0; z < L[2]; z++)
      for (y = 0; y < L[1]; y++)
        for (x = 0; x < L[0]; x++) {
          int cur = x + L[0] * (y + L[1] * (z + L[2] * t));
          int nw = wilson_gauge_site(L, x, y, z, t);
          for (mu = 0; mu < 4; mu++)
            for (row = 0; row < 3; row++)
              for (col = 0; col < 3; col++)
                for (r = 0; r < 2; r++)
                  dst[GAUGE * cur + r + 2 * (col + 3 * (row + 3 * mu))] =
                      src[GAUGE * nw + r + 2 * (row + 3 * (col + 3 * mu))];
        }
}

void ora_ferm_canon_to_wilson(double *dst, const double *src, const int L[4]) {
  int x, y, z, t;
  for (t = 0; t < L[3]; t++)
    for (z = 0; z < L[2]; z++)
      for (y = 0; y < L[1]; y++)
        for (x = 0; x < L[0]; x++) {
          int cur = x + L[0] * (y + L[1] * (z + L[2] * t));
          int nw = wilson_ferm_site(L, x, y, z, t);
          memcpy(dst + SPINOR * nw, src + SPINOR * cur, SPINOR * sizeof(double));
        }
}

void ora_ferm_wilson_to_canon(double *dst, const double *src, const int L[4]) {
  int x, y, z, t;
  for (t = 0; t < L[3]; t++)
    for (z = 0; z < L[2]; z++)
      for (y = 0; y < L[1]; y++)
        for (x = 0; x < L[0]; x++) {
          int cur = x + L[0] * (y + L[1] * (z + L[2] * t));
          int nw = wilson_ferm_site(L, x, y, z, t);
          memcpy(dst + SPINOR * cur, src + SPINOR * nw, SPINOR * sizeof(double));
        }
}

/* 5-D DWF WILSON order, convert.C:635-700: each canonical s-slice is first
 * converted to 4-D WILSON order (odd half then even half); then the odd 5-D
 * checkerboard collects, for s = 0..Ls-1, the half of slice s with 4-D
 * parity (s+1)%2, followed by the even 5-D checkerboard (4-D parity s%2). */
void ora_ferm5_canon_to_wilson(double *dst, const double *src, const int L[4],
                               int ls) {
  size_t stride = (size_t)SPINOR * vol4(L);
  size_t half = stride / 2;
  double *slice = (double *)malloc(stride * sizeof(double));
  int s;
  for (s = 0; s < ls; s++) {
    ora_ferm_canon_to_wilson(slice, src + s * stride, L);
    /* slice: [odd 4-D][even 4-D] */
    int par_odd5 = (s + 1) % 2; /* 4-D parity of this slice in the odd 5-D cb */
    int par_even5 = s % 2;
    /* WILSON 4-D block for 4-D parity p: odd (p=1) is block 0, even is block 1 */
    memcpy(dst + s * half, slice + (par_odd5 == 1 ? 0 : half), half * sizeof(double));
    memcpy(dst + ls * half + s * half, slice + (par_even5 == 1 ? 0 : half),
           half * sizeof(double));
  }
  free(slice);
}

void ora_ferm5_wilson_to_canon(double *dst, const double *src, const int L[4],
                               int ls) {
  size_t stride = (size_t)SPINOR * vol4(L);
  size_t half = stride / 2;
  double *slice = (double *)malloc(stride * sizeof(double));
  int s;
  for (s = 0; s < ls; s++) {
    int par_odd5 = (s + 1) % 2;
    int par_even5 = s % 2;
    memcpy(slice + (par_odd5 == 1 ? 0 : half), src + s * half, half * sizeof(double));
    memcpy(slice + (par_even5 == 1 ? 0 : half), src + ls * half + s * half,
           half * sizeof(double));
    ora_ferm_wilson_to_canon(dst + s * stride, slice, L);
  }
  free(slice);
}

/* DWF_4D_EOPREC order (convert.C:700-760): the odd 4-D half of every slice,
 * s = 0..Ls-1, followed by the even 4-D half of every slice. */
void ora_ferm5_canon_to_4dprec(double *dst, const double *src, const int L[4],
                               int ls) {
  size_t stride = (size_t)SPINOR * vol4(L);
  size_t half = stride / 2;
  double *slice = (double *)malloc(stride * sizeof(double));
  int s;
  for (s = 0; s < ls; s++) {
    ora_ferm_canon_to_wilson(slice, src + s * stride, L);
    memcpy(dst + s * half, slice, half * sizeof(double));
    memcpy(dst + ls * half + s * half, slice + half, half * sizeof(double));
  }
  free(slice);
}

void ora_ferm5_4dprec_to_canon(double *dst, const double *src, const int L[4],
                               int ls) {
  size_t stride = (size_t)SPINOR * vol4(L);
  size_t half = stride / 2;
  double *slice = (double *)malloc(stride * sizeof(double));
  int s;
  for (s = 0; s < ls; s++) {
    memcpy(slice, src + s * half, half * sizeof(double));
    memcpy(slice + half, src + ls * half + s * half, half * sizeof(double));
    ora_ferm_wilson_to_canon(dst + s * stride, slice, L);
  }
  free(slice);
}

/* Lattice::BondCond, lattice_base.C:5313-5369, antiperiodic directions:
 * multiply the mu-link of every site with x_mu = L_mu - 1 by -1.  CANONICAL
 * order.  (Twisted directions: ora_bond_cond_twist.) */
void ora_bond_cond(double *u, const int L[4], const int bc[4]) {
  int mu, x[4], i;
  for (mu = 0; mu < 4; mu++) {
    if (!bc[mu]) continue;
    for (x[3] = 0; x[3] < L[3]; x[3]++)
      for (x[2] = 0; x[2] < L[2]; x[2]++)
        for (x[1] = 0; x[1] < L[1]; x[1]++)
          for (x[0] = 0; x[0] < L[0]; x[0]++) {
            if (x[mu] != L[mu] - 1) continue;
            int off = mu + 4 * (x[0] + L[0] * (x[1] + L[1] * (x[2] + L[2] * x[3])));
            for (i = 0; i < 18; i++) u[18 * off + i] = -u[18 * off + i];
          }
  }
}

/* Lattice::BondCond, lattice_base.C:5313-5369, with the twisted kind: for
 * bc[mu] = 1 (BND_CND_APRD) the last-slice mu-links are multiplied by -1,
 * for bc[mu] = 2 (BND_CND_TWISTED) by GJP::TwistPhase(mu) = cos(pi theta) -
 * i sin(pi theta) (gjp.C:475-493, theta = twist[mu] in units of pi), or by
 * its conjugate when undo = 1 (the odd calls of :5327-5331, which remove the
 * twist again).  CANONICAL order. */
void ora_bond_cond_twist(double *u, const int L[4], const int bc[4], const double twist[4], int undo) {
  const double pi = 3.1415926535897932384626433832795;
  int mu, x[4], i;
  for (mu = 0; mu < 4; mu++) {
    double fr, fi;
    if (bc[mu] == 1) {
      fr = -1.0;
      fi = 0.0;
    } else if (bc[mu] == 2) {
      fr = cos(twist[mu] * pi);
      fi = -sin(twist[mu] * pi);
      if (undo) fi = -fi;
    } else {
      continue;
    }
    for (x[3] = 0; x[3] < L[3]; x[3]++)
      for (x[2] = 0; x[2] < L[2]; x[2]++)
        for (x[1] = 0; x[1] < L[1]; x[1]++)
          for (x[0] = 0; x[0] < L[0]; x[0]++) {
            if (x[mu] != L[mu] - 1) continue;
            int off = mu + 4 * (x[0] + L[0] * (x[1] + L[1] * (x[2] + L[2] * x[3])));
            double *m = u + 18 * off;
            for (i = 0; i < 9; i++) {
              double a = m[2 * i], b = m[2 * i + 1];
              m[2 * i] = fr * a - fi * b;
              m[2 * i + 1] = fr * b + fi * a;
            }
          }
  }
}

/* ======================================================================
 * Wilson hopping term.  Follows wilson_dslash.C:64-167 (spin projector
 * OneMinusPlusGammaTimesPsi and the link multiply UmuTimes, kept as full
 * 4-spinor operations as CPS does) and wilson_dslash.C:173-515 (site loop,
 * neighbour addressing, link offsets, order of the 8-term sum).
 * Pinned by: tests/test_oracle_pins.py::test_wilson_dslash_textbook (both
 * cb and dag against an independently built sparse hopping matrix with the
 * gamma basis of dirac_op.h:955-978, dag 1 as its adjoint),
 * ::test_wilson_plane_wave (closed form), ::test_wilson_gauge_covariance,
 * ::test_wilson_preconditioned_ops (Schur complement).
 * ====================================================================== */
#define UELEM(u, r, row, col, d) (u)[(r) + 2 * ((row) + 3 * ((col) + 3 * (d)))]
#define FERM(f, r, c, s) (f)[(r) + 2 * ((c) + 3 * (s))]

/* mp = 0: (1 - sdag*gamma_mu) psi ; mp = 1: (1 + sdag*gamma_mu) psi */
static void one_minus_plus_gamma(double *into, const double *psi, int sdag,
                                 int mu, int mp) {
  double sign = (mp == 1) ? 1.0 : -1.0;
  double k = sign * sdag;
  int c;
  for (c = 0; c < 3; c++) {
    if (mu == 0) {
      FERM(into, 0, c, 0) = FERM(psi, 0, c, 0) + k * (-FERM(psi, 1, c, 3));
      FERM(into, 1, c, 0) = FERM(psi, 1, c, 0) + k * (FERM(psi, 0, c, 3));
      FERM(into, 0, c, 1) = FERM(psi, 0, c, 1) + k * (-FERM(psi, 1, c, 2));
      FERM(into, 1, c, 1) = FERM(psi, 1, c, 1) + k * (FERM(psi, 0, c, 2));
      FERM(into, 0, c, 2) = FERM(psi, 0, c, 2) + k * (FERM(psi, 1, c, 1));
      FERM(into, 1, c, 2) = FERM(psi, 1, c, 2) + k * (-FERM(psi, 0, c, 1));
      FERM(into, 0, c, 3) = FERM(psi, 0, c, 3) + k * (FERM(psi, 1, c, 0));
      FERM(into, 1, c, 3) = FERM(psi, 1, c, 3) + k * (-FERM(psi, 0, c, 0));
    } else if (mu == 1) {
      FERM(into, 0, c, 0) = FERM(psi, 0, c, 0) + k * (-FERM(psi, 0, c, 3));
      FERM(into, 1, c, 0) = FERM(psi, 1, c, 0) + k * (-FERM(psi, 1, c, 3));
      FERM(into, 0, c, 1) = FERM(psi, 0, c, 1) + k * (FERM(psi, 0, c, 2));
      FERM(into, 1, c, 1) = FERM(psi, 1, c, 1) + k * (FERM(psi, 1, c, 2));
      FERM(into, 0, c, 2) = FERM(psi, 0, c, 2) + k * (FERM(psi, 0, c, 1));
      FERM(into, 1, c, 2) = FERM(psi, 1, c, 2) + k * (FERM(psi, 1, c, 1));
      FERM(into, 0, c, 3) = FERM(psi, 0, c, 3) + k * (-FERM(psi, 0, c, 0));
      FERM(into, 1, c, 3) = FERM(psi, 1, c, 3) + k * (-FERM(psi, 1, c, 0));
    } else if (mu == 2) {
      FERM(into, 0, c, 0) = FERM(psi, 0, c, 0) + k * (-FERM(psi, 1, c, 2));
      FERM(into, 1, c, 0) = FERM(psi, 1, c, 0) + k * (FERM(psi, 0, c, 2));
      FERM(into, 0, c, 1) = FERM(psi, 0, c, 1) + k * (FERM(psi, 1, c, 3));
      FERM(into, 1, c, 1) = FERM(psi, 1, c, 1) + k * (-FERM(psi, 0, c, 3));
      FERM(into, 0, c, 2) = FERM(psi, 0, c, 2) + k * (FERM(psi, 1, c, 0));
      FERM(into, 1, c, 2) = FERM(psi, 1, c, 2) + k * (-FERM(psi, 0, c, 0));
      FERM(into, 0, c, 3) = FERM(psi, 0, c, 3) + k * (-FERM(psi, 1, c, 1));
      FERM(into, 1, c, 3) = FERM(psi, 1, c, 3) + k * (FERM(psi, 0, c, 1));
    } else {
      FERM(into, 0, c, 0) = FERM(psi, 0, c, 0) + k * (FERM(psi, 0, c, 2));
      FERM(into, 1, c, 0) = FERM(psi, 1, c, 0) + k * (FERM(psi, 1, c, 2));
      FERM(into, 0, c, 1) = FERM(psi, 0, c, 1) + k * (FERM(psi, 0, c, 3));
      FERM(into, 1, c, 1) = FERM(psi, 1, c, 1) + k * (FERM(psi, 1, c, 3));
      FERM(into, 0, c, 2) = FERM(psi, 0, c, 2) + k * (FERM(psi, 0, c, 0));
      FERM(into, 1, c, 2) = FERM(psi, 1, c, 2) + k * (FERM(psi, 1, c, 0));
      FERM(into, 0, c, 3) = FERM(psi, 0, c, 3) + k * (FERM(psi, 0, c, 1));
      FERM(into, 1, c, 3) = FERM(psi, 1, c, 3) + k * (FERM(psi, 1, c, 1));
    }
  }
}

/* rowcol = 0: into = U psi ; rowcol = 1: into = U^dagger psi */
static void umu_times(double *into, const double *u, const double *psi, int mu,
                      int rowcol) {
  int s, c;
  for (s = 0; s < 4; s++)
    for (c = 0; c < 3; c++) {
      if (rowcol == 0) {
        FERM(into, 0, c, s) =
            (UELEM(u, 0, c, 0, mu) * FERM(psi, 0, 0, s) + UELEM(u, 0, c, 1, mu) * FERM(psi, 0, 1, s) +
             UELEM(u, 0, c, 2, mu) * FERM(psi, 0, 2, s) - UELEM(u, 1, c, 0, mu) * FERM(psi, 1, 0, s) -
             UELEM(u, 1, c, 1, mu) * FERM(psi, 1, 1, s) - UELEM(u, 1, c, 2, mu) * FERM(psi, 1, 2, s));
        FERM(into, 1, c, s) =
            (UELEM(u, 0, c, 0, mu) * FERM(psi, 1, 0, s) + UELEM(u, 0, c, 1, mu) * FERM(psi, 1, 1, s) +
             UELEM(u, 0, c, 2, mu) * FERM(psi, 1, 2, s) + UELEM(u, 1, c, 0, mu) * FERM(psi, 0, 0, s) +
             UELEM(u, 1, c, 1, mu) * FERM(psi, 0, 1, s) + UELEM(u, 1, c, 2, mu) * FERM(psi, 0, 2, s));
      } else {
        FERM(into, 0, c, s) =
            (UELEM(u, 0, 0, c, mu) * FERM(psi, 0, 0, s) + UELEM(u, 0, 1, c, mu) * FERM(psi, 0, 1, s) +
             UELEM(u, 0, 2, c, mu) * FERM(psi, 0, 2, s) + UELEM(u, 1, 0, c, mu) * FERM(psi, 1, 0, s) +
             UELEM(u, 1, 1, c, mu) * FERM(psi, 1, 1, s) + UELEM(u, 1, 2, c, mu) * FERM(psi, 1, 2, s));
        FERM(into, 1, c, s) =
            (UELEM(u, 0, 0, c, mu) * FERM(psi, 1, 0, s) + UELEM(u, 0, 1, c, mu) * FERM(psi, 1, 1, s) +
             UELEM(u, 0, 2, c, mu) * FERM(psi, 1, 2, s) - UELEM(u, 1, 0, c, mu) * FERM(psi, 0, 0, s) -
             UELEM(u, 1, 1, c, mu) * FERM(psi, 0, 1, s) - UELEM(u, 1, 2, c, mu) * FERM(psi, 0, 2, s));
      }
    }
}

void ora_wilson_dslash(double *chi_p, const double *u_p, const double *psi_p,
                       int cb, int dag, const int L[4]) {
  int lx = L[0], ly = L[1], lz = L[2], lt = L[3];
  int vol = vol4(L) / 2; /* checkerboard volume, Wilson.vol[0] */
  int sdag = (dag == 0) ? 1 : -1;
  int cbn = (cb == 0) ? 1 : 0;
  int x, y;
  /* Sites are independent (each writes its own chi): OpenMP threads only
   * split the loop, the arithmetic of every site is unchanged. */
#pragma omp parallel for collapse(2) schedule(static)
  for (x = 0; x < lx; x++)
    for (y = 0; y < ly; y++) {
      double tmp[SPINOR], temps[4][2][SPINOR];
      int z, t, mu, i;
      for (z = 0; z < lz; z++)
        for (t = 0; t < lt; t++) {
          if ((x + y + z + t) % 2 != cbn) continue;
          int xyzt = (x / 2) + (lx / 2) * (y + ly * (z + lz * t));
          int pos[4] = {x, y, z, t};
          for (mu = 0; mu < 4; mu++) {
            int posp[4] = {x, y, z, t};
            posp[mu] = (posp[mu] + 1) % L[mu];
            int posp_xyzt = (posp[0] / 2) + (lx / 2) * (posp[1] + ly * (posp[2] + lz * posp[3]));
            const double *u = u_p + GAUGE * (xyzt + vol * cbn);
            const double *psi = psi_p + SPINOR * posp_xyzt;
            one_minus_plus_gamma(tmp, psi, sdag, mu, 0);
            umu_times(temps[mu][0], u, tmp, mu, 0);
          }
          for (mu = 0; mu < 4; mu++) {
            int posm[4] = {x, y, z, t};
            posm[mu] = posm[mu] - 1 + ((L[mu] - posm[mu]) / L[mu]) * L[mu];
            int posm_xyzt = (posm[0] / 2) + (lx / 2) * (posm[1] + ly * (posm[2] + lz * posm[3]));
            const double *u = u_p + GAUGE * (posm_xyzt + vol * cb);
            const double *psi = psi_p + SPINOR * posm_xyzt;
            one_minus_plus_gamma(tmp, psi, sdag, mu, 1);
            umu_times(temps[mu][1], u, tmp, mu, 1);
          }
          (void)pos;
          double *chi = chi_p + SPINOR * xyzt;
          /* tmp1..tmp4 are the forward terms mu=0..3, tmp5..tmp8 the backward */
          for (i = 0; i < SPINOR; i++)
            chi[i] = (temps[0][0][i] + temps[1][0][i] + temps[2][0][i] + temps[3][0][i] +
                      temps[0][1][i] + temps[1][1][i] + temps[2][1][i] + temps[3][1][i]);
        }
    }
}

/* wilson_m.C:49-93 */
void ora_wilson_m(double *chi, const double *u, const double *psi, double kappa,
                  const int L[4], double *tmp1) {
  int n, vol = vol4(L) / 2;
  ora_wilson_dslash(tmp1, u, psi, 1, 0, L);
  ora_wilson_dslash(chi, u, tmp1, 0, 0, L);
  for (n = 0; n < SPINOR * vol; n++) chi[n] = (psi[n] - kappa * kappa * chi[n]);
}

/* wilson_mdag.C:50-95 */
void ora_wilson_mdag(double *chi, const double *u, const double *psi,
                     double kappa, const int L[4], double *tmp1) {
  int n, vol = vol4(L) / 2;
  ora_wilson_dslash(tmp1, u, psi, 1, 1, L);
  ora_wilson_dslash(chi, u, tmp1, 0, 1, L);
  for (n = 0; n < SPINOR * vol; n++) chi[n] = (psi[n] - kappa * kappa * chi[n]);
}

/* wilson_mdagm.C:83-164; mp_sq = |M psi|^2 summed in site order */
void ora_wilson_mdagm(double *chi, const double *u, const double *psi,
                      double *mp_sq, double kappa, const int L[4], double *tmp1,
                      double *tmp2) {
  int n, vol = vol4(L) / 2;
  double sum = 0.0;
  ora_wilson_dslash(tmp1, u, psi, 1, 0, L);
  ora_wilson_dslash(tmp2, u, tmp1, 0, 0, L);
  for (n = 0; n < SPINOR * vol; n++) {
    tmp2[n] = (psi[n] - kappa * kappa * tmp2[n]);
    if (mp_sq) sum = sum + tmp2[n] * tmp2[n];
  }
  if (mp_sq) *mp_sq = sum;
  ora_wilson_dslash(tmp1, u, tmp2, 1, 1, L);
  ora_wilson_dslash(chi, u, tmp1, 0, 1, L);
  for (n = 0; n < SPINOR * vol; n++) chi[n] = (tmp2[n] - kappa * kappa * chi[n]);
}

/* ======================================================================
 * Vector BLAS, vector_util.C:727-835 (single thread: sequential sum).
 * ====================================================================== */
double ora_dot(const double *a, const double *b, size_t n) {
  double sum = 0.0;
  size_t i;
  for (i = 0; i < n; ++i) sum += a[i] * b[i];
  return sum;
}

void ora_ftimesv1plusv2(double *a, double b, const double *c, const double *d,
                        size_t n) {
  size_t i;
  for (i = 0; i < n; ++i) a[i] = b * c[i] + d[i];
}

/* ======================================================================
 * Shamir domain wall.  dwf_init.C:97-98 (kappa), dwf_dslash.C:32-50,
 * dwf_dslash_4.C:68-141, dwf_dslash_5_plus.C:111-340 (s_nodes = 1),
 * dwf_m.C:31-62, dwf_mdag.C:32-62, dwf_mdagm.C:32-61.
 * Pinned by: tests/test_oracle_pins.py::test_dwf_dslash_textbook (both cb
 * and dag against an independently built sparse 5-D operator) and
 * ::test_dwf_preconditioned_ops_and_solvers (Schur complement, kappa_5,
 * CG and MatInv residuals through the sparse operator).
 * ====================================================================== */
double ora_dwf_kappa(const ora_dwf *d) {
  return 1.0 / (2 * (4 + d->a5_inv - d->m5));
}

void ora_dwf_dslash_4(double *out, const double *u, const double *in, int cb,
                      int dag, const ora_dwf *d) {
  size_t size_cb = (size_t)SPINOR * (vol4(d->L) / 2);
  int i;
  for (i = 0; i < d->ls; i++) {
    int parity = (i + cb) % 2;
    ora_wilson_dslash(out + i * size_cb, u, in + i * size_cb, parity, dag, d->L);
  }
}

void ora_dwf_dslash_5_plus(double *out, const double *in, double mass, int dag,
                           const ora_dwf *d) {
  int ls = d->ls;
  int vol_4d_cb = vol4(d->L) / 2;
  size_t ls_stride = (size_t)SPINOR * vol_4d_cb;
  double two_over_a5 = 2.0 * d->a5_inv;
  double neg_mass_two_over_a5 = -2.0 * mass * d->a5_inv;
  const double *f_in;
  double *f_out;
  int s, x;

  /* out[s] upper (lower if dag) += 2/a5 * in[s-1], s = 1..ls-1 */
  f_in = in + (dag == 1 ? 12 : 0);
  f_out = out + (dag == 1 ? 12 : 0) + ls_stride;
  for (s = 1; s < ls; s++)
    for (x = 0; x < vol_4d_cb; x++) {
      ora_ftimesv1plusv2(f_out, two_over_a5, f_in, f_out, 12);
      f_in += 24;
      f_out += 24;
    }
  /* out[0] += -mf*2/a5 * in[ls-1] */
  f_in = in + (ls - 1) * ls_stride + (dag == 1 ? 12 : 0);
  f_out = out + (dag == 1 ? 12 : 0);
  for (x = 0; x < vol_4d_cb; x++) {
    ora_ftimesv1plusv2(f_out, neg_mass_two_over_a5, f_in, f_out, 12);
    f_in += 24;
    f_out += 24;
  }
  /* out[s] lower (upper if dag) += 2/a5 * in[s+1], s = 0..ls-2 */
  f_in = in + (dag == 0 ? 12 : 0) + ls_stride;
  f_out = out + (dag == 0 ? 12 : 0);
  for (s = 0; s < ls - 1; s++)
    for (x = 0; x < vol_4d_cb; x++) {
      ora_ftimesv1plusv2(f_out, two_over_a5, f_in, f_out, 12);
      f_in += 24;
      f_out += 24;
    }
  /* out[ls-1] += -mf*2/a5 * in[0] */
  f_in = in + (dag == 0 ? 12 : 0);
  f_out = out + (dag == 0 ? 12 : 0) + (ls - 1) * ls_stride;
  for (x = 0; x < vol_4d_cb; x++) {
    ora_ftimesv1plusv2(f_out, neg_mass_two_over_a5, f_in, f_out, 12);
    f_in += 24;
    f_out += 24;
  }
}

void ora_dwf_dslash(double *out, const double *u, const double *in, double mass,
                    int cb, int dag, const ora_dwf *d) {
  ora_dwf_dslash_4(out, u, in, cb, dag, d);
  ora_dwf_dslash_5_plus(out, in, mass, dag, d);
}

void ora_dwf_m(double *out, const double *u, const double *in, double mass,
               const ora_dwf *d, double *tmp2) {
  size_t f_size = (size_t)SPINOR * vol4(d->L) * d->ls / 2;
  double kappa = ora_dwf_kappa(d);
  double minus_kappa_sq = -kappa * kappa;
  ora_dwf_dslash(tmp2, u, in, mass, 1, 0, d);
  ora_dwf_dslash(out, u, tmp2, mass, 0, 0, d);
  ora_ftimesv1plusv2(out, minus_kappa_sq, out, in, f_size);
}

void ora_dwf_mdag(double *out, const double *u, const double *in, double mass,
                  const ora_dwf *d, double *tmp2) {
  size_t f_size = (size_t)SPINOR * vol4(d->L) * d->ls / 2;
  double kappa = ora_dwf_kappa(d);
  double minus_kappa_sq = -kappa * kappa;
  ora_dwf_dslash(tmp2, u, in, mass, 1, 1, d);
  ora_dwf_dslash(out, u, tmp2, mass, 0, 1, d);
  ora_ftimesv1plusv2(out, minus_kappa_sq, out, in, f_size);
}

void ora_dwf_mdagm(double *out, const double *u, const double *in,
                   double *dot_prd, double mass, const ora_dwf *d, double *tmp1,
                   double *tmp2) {
  size_t f_size = (size_t)SPINOR * vol4(d->L) * d->ls / 2;
  ora_dwf_m(tmp1, u, in, mass, d, tmp2);
  if (dot_prd) *dot_prd = ora_dot(tmp1, tmp1, f_size);
  ora_dwf_mdag(out, u, tmp1, mass, d, tmp2);
}

/* ======================================================================
 * DiracOp::InvCgShift with shift = NULL, inv_cg.C:84-448.
 * Returns itr+1 exactly as the reference does (inv_cg.C:446).
 * res_hist (optional, max_num_iter entries) receives |res|^2 after each
 * iteration.  d == 0 makes the reference exit(5) (inv_cg.C:326-329); here it
 * returns -1 so a test can observe it.
 * Pinned by: tests/test_oracle_pins.py::test_wilson_matinv_and_cg (the CPS
 * self-check max|Mx-b| of tests/f_wilson_inv/main.C:183-246 through the
 * independently built operator), ::test_wilson_cg_iteration_scale (the
 * iteration count logged in tests/f_wilson_inv/qcdlinux1.log:69), and the
 * solver residual checks of the DWF, Moebius and clover tests.
 * ====================================================================== */
int ora_invcg(double *sol, const double *src, size_t f_size_cb, double stop_rsd,
              int max_num_iter, double src_norm_sq, double *true_res,
              ora_mdagm_fn mdagm, void *ctx, double *res_hist) {
  double stp_cnd, res_norm_sq_prv, res_norm_sq_cur, a, b, d;
  int itr, max_itr, i;
  double *res = (double *)malloc(f_size_cb * sizeof(double));
  double *dir = (double *)malloc(f_size_cb * sizeof(double));
  double *mmp = (double *)malloc(f_size_cb * sizeof(double));
  size_t k;

  if (src_norm_sq == 0) src_norm_sq = ora_dot(src, src, f_size_cb);
  stp_cnd = src_norm_sq * stop_rsd * stop_rsd;

  mdagm(ctx, mmp, sol, NULL);
  memcpy(res, src, f_size_cb * sizeof(double));
  for (k = 0; k < f_size_cb; k++) res[k] -= mmp[k];
  memcpy(dir, res, f_size_cb * sizeof(double));
  res_norm_sq_cur = ora_dot(res, res, f_size_cb);

  itr = 0;
  max_itr = max_num_iter - 1;
  if (res_norm_sq_cur <= stp_cnd) max_itr = 0;

  for (i = 0; i < max_itr; i++) {
    itr = itr + 1;
    res_norm_sq_prv = res_norm_sq_cur;
    mdagm(ctx, mmp, dir, &d);
    if (d == 0.0) {
      itr = -2;
      break;
    }
    a = res_norm_sq_prv / d;
    ora_ftimesv1plusv2(sol, a, dir, sol, f_size_cb);
    ora_ftimesv1plusv2(res, -a, mmp, res, f_size_cb);
    res_norm_sq_cur = ora_dot(res, res, f_size_cb);
    if (res_hist) res_hist[itr - 1] = res_norm_sq_cur;
    if (progress_every && itr % progress_every == 0) {
      fprintf(stderr, "oracle CG iteration %d: |r|^2 %.3e (stop at %.3e)\n", itr, res_norm_sq_cur, stp_cnd);
      fflush(stderr);
    }
    if (res_norm_sq_cur <= stp_cnd) break;
    b = res_norm_sq_cur / res_norm_sq_prv;
    ora_ftimesv1plusv2(dir, b, dir, res, f_size_cb);
  }

  if (itr >= 0) {
    mdagm(ctx, mmp, sol, NULL);
    memcpy(res, src, f_size_cb * sizeof(double));
    for (k = 0; k < f_size_cb; k++) res[k] -= mmp[k];
    res_norm_sq_cur = ora_dot(res, res, f_size_cb);
    if (true_res) *true_res = sqrt(res_norm_sq_cur / src_norm_sq);
  }
  free(res);
  free(dir);
  free(mmp);
  return itr + 1;
}

/* ---- drivers ----------------------------------------------------------- */
typedef struct {
  const double *u;
  double kappa;
  const int *L;
  double *tmp1, *tmp2;
} wilson_ctx;

static void wilson_mdagm_cb(void *c, double *out, const double *in, double *dp) {
  wilson_ctx *w = (wilson_ctx *)c;
  ora_wilson_mdagm(out, w->u, in, dp, w->kappa, w->L, w->tmp1, w->tmp2);
}

int ora_wilson_invcg(double *sol, const double *src, const double *u,
                     double mass, double stop_rsd, int max_num_iter,
                     const int L[4], double *true_res) {
  size_t f = (size_t)SPINOR * vol4(L) / 2;
  wilson_ctx w;
  int it;
  w.u = u;
  w.kappa = 1.0 / (2.0 * (mass + 4.0)); /* d_op_wilson.C:134-137 */
  w.L = L;
  w.tmp1 = (double *)malloc(f * sizeof(double));
  w.tmp2 = (double *)malloc(f * sizeof(double));
  it = ora_invcg(sol, src, f, stop_rsd, max_num_iter, 0.0, true_res,
                 wilson_mdagm_cb, &w, NULL);
  free(w.tmp1);
  free(w.tmp2);
  return it;
}

/* DiracOpWilson::MatInv, d_op_wilson.C:257-334 with prs_in = PRESERVE_YES.
 * in/out: full WILSON-order vectors [odd][even]; out holds the initial guess
 * on its odd half. */
int ora_wilson_matinv(double *out, double *in, const double *u, double mass,
                      double stop_rsd, int max_num_iter, const int L[4],
                      double *true_res) {
  size_t f = (size_t)SPINOR * vol4(L) / 2;
  double kappa = 1.0 / (2.0 * (mass + 4.0));
  double *temp = (double *)malloc(f * sizeof(double));
  double *temp2 = (double *)malloc(f * sizeof(double));
  double *tmp1 = (double *)malloc(f * sizeof(double));
  double *even_in = in + f, *even_out = out + f;
  int iter;
  ora_wilson_dslash(temp, u, even_in, 0, 0, L);
  ora_ftimesv1plusv2(temp, kappa, temp, in, f);
  memcpy(temp2, in, f * sizeof(double));
  ora_wilson_mdag(in, u, temp, kappa, L, tmp1);
  iter = ora_wilson_invcg(out, in, u, mass, stop_rsd, max_num_iter, L, true_res);
  memcpy(in, temp2, f * sizeof(double));
  ora_wilson_dslash(temp, u, out, 1, 0, L);
  ora_ftimesv1plusv2(even_out, kappa, temp, even_in, f);
  free(temp);
  free(temp2);
  free(tmp1);
  return iter;
}

typedef struct {
  const double *u;
  double mass;
  const ora_dwf *d;
  double *tmp1, *tmp2;
} dwf_ctx;

static void dwf_mdagm_cb(void *c, double *out, const double *in, double *dp) {
  dwf_ctx *w = (dwf_ctx *)c;
  ora_dwf_mdagm(out, w->u, in, dp, w->mass, w->d, w->tmp1, w->tmp2);
}

int ora_dwf_invcg(double *sol, const double *src, const double *u, double mass,
                  double stop_rsd, int max_num_iter, const ora_dwf *d,
                  double *true_res) {
  size_t f = (size_t)SPINOR * vol4(d->L) * d->ls / 2;
  dwf_ctx w;
  int it;
  w.u = u;
  w.mass = mass;
  w.d = d;
  w.tmp1 = (double *)malloc(f * sizeof(double));
  w.tmp2 = (double *)malloc(f * sizeof(double));
  it = ora_invcg(sol, src, f, stop_rsd, max_num_iter, 0.0, true_res,
                 dwf_mdagm_cb, &w, NULL);
  free(w.tmp1);
  free(w.tmp2);
  return it;
}

/* DiracOpDwf::MatInv, d_op_dwf.C:316-509 (CG branch, PRESERVE_YES; the two
 * debug checksums of d_op_dwf.C:368-389,428-449 are diagnostics only). */
int ora_dwf_matinv(double *out, double *in, const double *u, double mass,
                   double stop_rsd, int max_num_iter, const ora_dwf *d,
                   double *true_res) {
  size_t f = (size_t)SPINOR * vol4(d->L) * d->ls / 2;
  double kappa = ora_dwf_kappa(d);
  double *temp = (double *)malloc(f * sizeof(double));
  double *temp2 = (double *)malloc(f * sizeof(double));
  double *tmp2 = (double *)malloc(f * sizeof(double));
  double *even_in = in + f, *even_out = out + f;
  int iter;
  ora_dwf_dslash(temp, u, even_in, mass, 0, 0, d);
  ora_ftimesv1plusv2(temp, kappa, temp, in, f);
  memcpy(temp2, in, f * sizeof(double));
  ora_dwf_mdag(in, u, temp, mass, d, tmp2);
  iter = ora_dwf_invcg(out, in, u, mass, stop_rsd, max_num_iter, d, true_res);
  memcpy(in, temp2, f * sizeof(double));
  ora_dwf_dslash(temp, u, out, mass, 1, 0, d);
  ora_ftimesv1plusv2(even_out, kappa, temp, even_in, f);
  free(temp);
  free(temp2);
  free(tmp2);
  return iter;
}
