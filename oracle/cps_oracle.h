/*
 * cps_oracle.h -- CPU restatement of the CPS noarch Wilson / domain-wall
 * even-odd preconditioned operators and their CG inverter.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in cps_public_amd/.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product never does.
 *
 * Every routine states the reference file:line it follows (paths relative to
 * the CPS source tree cps_pp/).  Operation order inside each routine follows
 * the reference so that, compiled without FP contraction, results agree with
 * CPS to the last bit or within a few ulp.
 *
 * Layouts (all arrays are double):
 *   CANONICAL gauge : [t][z][y][x][mu][row][col][re,im]      (lattice_base.C:5350)
 *   WILSON gauge    : [parity(even first)][site_cb][mu][col][row][re,im]
 *                     i.e. element U(row,col) at 2*(row+3*col)   (convert_func.C:318-415)
 *   CANONICAL fermion: [t][z][y][x][spin][colour][re,im]
 *   WILSON fermion  : [odd block][even block], each [site_cb][spin][colour][re,im]
 *                     site_cb = x/2 + lx/2*(y+ly*(z+lz*t))      (convert_func.C:429-494)
 *   DWF 5-D (WILSON): [odd 5-D cb: s=0..Ls-1 slices][even 5-D cb: slices],
 *                     slice s of 5-D parity cb has 4-D parity (s+cb)%2
 *                                                               (convert.C:635-700)
 */
#ifndef CPS_ORACLE_H
#define CPS_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- layout ------------------------------------------------------------ */
/* thread count of the OpenMP site loops; results do not depend on it */
void ora_set_threads(int n);
int ora_get_threads(void);
/* ora_invcg prints |r|^2 to stderr every `every` iterations (0: silent) */
void ora_set_progress(int every);

void ora_gauge_canon_to_wilson(double *dst, const double *src, const int L[4]);
void ora_gauge_wilson_to_canon(double *dst, const double *src, const int L[4]);
void ora_ferm_canon_to_wilson(double *dst, const double *src, const int L[4]);
void ora_ferm_wilson_to_canon(double *dst, const double *src, const int L[4]);
void ora_ferm5_canon_to_wilson(double *dst, const double *src, const int L[4], int ls);
void ora_ferm5_wilson_to_canon(double *dst, const double *src, const int L[4], int ls);
void ora_ferm5_canon_to_4dprec(double *dst, const double *src, const int L[4], int ls);
void ora_ferm5_4dprec_to_canon(double *dst, const double *src, const int L[4], int ls);
/* bc[mu] = 1 for anti-periodic; toggles the sign of links with x_mu = L_mu-1 */
void ora_bond_cond(double *u_canon, const int L[4], const int bc[4]);
/* bc[mu] = 1 antiperiodic (x -1), 2 twisted (x e^{-i pi twist[mu]}, its
 * conjugate with undo = 1) on the links with x_mu = L_mu-1 */
void ora_bond_cond_twist(double *u_canon, const int L[4], const int bc[4], const double twist[4], int undo);

/* ---- Wilson ------------------------------------------------------------ */
void ora_wilson_dslash(double *chi, const double *u, const double *psi,
                       int cb, int dag, const int L[4]);
void ora_wilson_m(double *chi, const double *u, const double *psi,
                  double kappa, const int L[4], double *tmp1);
void ora_wilson_mdag(double *chi, const double *u, const double *psi,
                     double kappa, const int L[4], double *tmp1);
void ora_wilson_mdagm(double *chi, const double *u, const double *psi,
                      double *mp_sq, double kappa, const int L[4],
                      double *tmp1, double *tmp2);

/* ---- Shamir domain wall ------------------------------------------------ */
typedef struct {
  int L[4];
  int ls;
  double m5;      /* DwfHeight */
  double a5_inv;  /* DwfA5Inv  */
} ora_dwf;

double ora_dwf_kappa(const ora_dwf *d);
void ora_dwf_dslash_4(double *out, const double *u, const double *in,
                      int cb, int dag, const ora_dwf *d);
void ora_dwf_dslash_5_plus(double *out, const double *in, double mass,
                           int dag, const ora_dwf *d);
void ora_dwf_dslash(double *out, const double *u, const double *in,
                    double mass, int cb, int dag, const ora_dwf *d);
void ora_dwf_m(double *out, const double *u, const double *in, double mass,
               const ora_dwf *d, double *tmp2);
void ora_dwf_mdag(double *out, const double *u, const double *in, double mass,
                  const ora_dwf *d, double *tmp2);
void ora_dwf_mdagm(double *out, const double *u, const double *in,
                   double *dot_prd, double mass, const ora_dwf *d,
                   double *tmp1, double *tmp2);

/* ---- Moebius (4-D even-odd preconditioned, ZMOB_PC_ORIG, real b,c) ------ */
typedef struct {
  int L[4];
  int ls;
  double m5;
  double a5_inv;
  double b, c;
} ora_mobius;

void ora_mobius_kappas(const ora_mobius *m, double *kappa_b, double *kappa_c);
void ora_mobius_dslash_4(double *out, const double *u, const double *in,
                         int cb, int dag, const ora_mobius *m, double mass);
void ora_mobius_m5inv(double *inout, double mass, int dag, const ora_mobius *m);
void ora_mobius_dslash_5_plus(double *out, const double *in, double mass,
                              int dag, double coef, const ora_mobius *m);
void ora_mobius_m(double *out, const double *u, const double *in, double mass,
                  const ora_mobius *m, double *tmp2);
void ora_mobius_mdag(double *out, const double *u, const double *in,
                     double mass, const ora_mobius *m, double *tmp2,
                     double *tmp3);
void ora_mobius_mdagm(double *out, const double *u, const double *in,
                      double *dot_prd, double mass, const ora_mobius *m,
                      double *tmp1, double *tmp2, double *tmp3);

/* ---- clover ------------------------------------------------------------ */
/* Build the packed per-site clover matrices (72 reals: two 6x6 Hermitian
 * chiral blocks, 36 reals each) for the sites of checkerboard `cb`
 * (0 even, 1 odd) from a WILSON-order gauge field.  inverse != 0 stores
 * A^{-1}. */
void ora_clover_mat_chkb(double *A, const double *u, int cb, int inverse,
                         double kappa, double csw, const int L[4]);
void ora_clover_mat_mlt(double *out, const double *A, const double *in,
                        int nsites);
void ora_clover_matpc(double *out, const double *u, const double *A_oo,
                      const double *A_ee_inv, const double *in, double kappa,
                      int dag, const int L[4], double *tmp1, double *tmp2);
void ora_clover_mdagm(double *out, const double *u, const double *A_oo,
                      const double *A_ee_inv, const double *in,
                      double *dot_prd, double kappa, const int L[4],
                      double *tmp1, double *tmp2, double *tmp3);

/* ---- vector BLAS ------------------------------------------------------- */
double ora_dot(const double *a, const double *b, size_t n);
void ora_ftimesv1plusv2(double *a, double b, const double *c, const double *d,
                        size_t n);

/* ---- CG ---------------------------------------------------------------- */
typedef void (*ora_mdagm_fn)(void *ctx, double *out, const double *in,
                             double *dot_prd);
int ora_invcg(double *sol, const double *src, size_t f_size, double stop_rsd,
              int max_num_iter, double src_norm_sq, double *true_res,
              ora_mdagm_fn mdagm, void *ctx, double *res_hist);

/* Convenience drivers used by tests and the CPU baseline. */
int ora_wilson_invcg(double *sol, const double *src, const double *u,
                     double mass, double stop_rsd, int max_num_iter,
                     const int L[4], double *true_res);
int ora_wilson_matinv(double *out, double *in, const double *u, double mass,
                      double stop_rsd, int max_num_iter, const int L[4],
                      double *true_res);
int ora_dwf_invcg(double *sol, const double *src, const double *u,
                  double mass, double stop_rsd, int max_num_iter,
                  const ora_dwf *d, double *true_res);
int ora_dwf_matinv(double *out, double *in, const double *u, double mass,
                   double stop_rsd, int max_num_iter, const ora_dwf *d,
                   double *true_res);
int ora_mobius_invcg(double *sol, const double *src, const double *u,
                     double mass, double stop_rsd, int max_num_iter,
                     const ora_mobius *m, double *true_res);
int ora_clover_invcg(double *sol, const double *src, const double *u,
                     const double *A_oo, const double *A_ee_inv, double kappa,
                     double stop_rsd, int max_num_iter, const int L[4],
                     double *true_res);

#ifdef __cplusplus
}
#endif
#endif
