// Vector BLAS and CG scalar kernels (see blas.hip).
#pragma once
#include "cps_hip_internal.h"
#include "layout.h"

namespace cpship {

enum ReduceMode { RED_PLAIN = 0, RED_CG_ALPHA = 1, RED_CG_BETA = 2, RED_CG_INIT = 3, RED_SRC_NORM = 4 };
enum CGStatus {
  CG_STATUS_RUNNING = 0, CG_STATUS_CONVERGED = 1, CG_STATUS_MAXITER = 2, CG_STATUS_D_ZERO = 3,
  CG_STATUS_NEED_UPDATE = 4   // mixed precision: sloppy residual fell below delta * last true residual
};

// per-rhs CG scalars, all device arrays of length nrhs
struct CGScalars {
  double *rr, *rr_prev, *d, *alpha, *beta, *stp, *src_norm, *rsd;
  int *active, *iters, *status;
  int max_itr;        // max_num_iter - 1 (inv_cg.C:275)
  double* rr_upd;     // mixed: true |r|^2 at the last reliable update (nullptr for FP64 CG)
  double delta2;      // mixed: reliable-update trigger |r_s|^2 < delta2 * rr_upd
  double* hist;       // optional |r|^2 history ring [nrhs][hist_len]
  int hist_len;
  double* alpha_prev; // mixed CG with paired solution steps: alpha of the
                      // iteration before (set by RED_CG_ALPHA), or nullptr
};

hipError_t blas_norm(int st, const FieldRef& x, int64_t per_rhs, int nrhs, int bs, double* partial, hipStream_t s);
hipError_t blas_redot(int st, const FieldRef& x, const FieldRef& y, int64_t per_rhs, int nrhs, int bs,
                      double* partial, hipStream_t s);
hipError_t blas_sub_norm_d(const FieldRef& b, const FieldRef& a, const FieldRef& r, int64_t per_rhs, int nrhs, int bs,
                           double* partial, hipStream_t s);
hipError_t blas_cg_xr_d(const FieldRef& x, const FieldRef& p, const FieldRef& r, const FieldRef& ap, const CGScalars& cg,
                        int64_t per_rhs, int nrhs, int bs, double* partial, hipStream_t s);
hipError_t blas_xpay(int st, const FieldRef& p, const FieldRef& r, const double* beta, const int* active,
                     int64_t per_rhs, int nrhs, int bs, hipStream_t s);
hipError_t blas_redot_mixed(int sx, const FieldRef& x, int sy, const FieldRef& y, int64_t per_rhs, int nrhs, int bs,
                            double* partial, hipStream_t s);
hipError_t blas_axpy(int sx, const FieldRef& x, int sy, const FieldRef& y, const double* a, double a_const,
                     int64_t per_rhs, int nrhs, int bs, hipStream_t s);
hipError_t reduce_rhs(const double* partial, int bpr, int nrhs, double* out, int mode, const CGScalars& cg,
                      hipStream_t s);

hipError_t blas_cg_xr_mixed(int sx, const FieldRef& x, int sp, const FieldRef& p, int sr, const FieldRef& r, int sa,
                            const FieldRef& ap, const CGScalars& cg, int64_t per_rhs, int nrhs, int bs,
                            double* partial, hipStream_t s);
// the mixed CG with a lagged solution step: x += alpha p; p = r + beta p
// (active rhs), and r -= alpha A p with |r|^2
hipError_t blas_xpay_x_mixed(int sp, const FieldRef& p, int sr, const FieldRef& r, int sx, const FieldRef& x,
                             const double* alpha, const double* beta, const int* active, int64_t per_rhs, int nrhs,
                             int bs, hipStream_t s);
// FP16 p, FP32 x, FP16 or FP32 r, Moebius layout with Ls = 16: also
// t5 = b5 p + c5 P p
hipError_t blas_xpay_x_s5(const FieldRef& p, int sr, const FieldRef& r, const FieldRef& x, const FieldRef& t5,
                          const double* alpha, const double* beta, const int* active, int64_t per_rhs, int nrhs, int bs,
                          double b5, double c5, double mass, hipStream_t s);
// paired iterations (FP16 p in two buffers, FP32 x): p_out = r + beta p_in;
// with xupd first x += a_old p_out(before) + a_new p_in; t5 (optional,
// Moebius Ls 16): b5 p_out + c5 P p_out
hipError_t blas_xpay_pair(const FieldRef& pin, const FieldRef& pout, int sr, const FieldRef& r, const FieldRef& x,
                          const FieldRef* t5, bool xupd, const double* a_old, const double* a_new, const double* beta,
                          const int* active, int64_t per_rhs, int nrhs, int bs, double b5, double c5, double mass,
                          hipStream_t s);
hipError_t blas_cg_r_mixed(int sr, const FieldRef& r, int sa, const FieldRef& ap, const CGScalars& cg,
                           int64_t per_rhs, int nrhs, int bs, double* partial, hipStream_t s);
hipError_t blas_xpay_mixed(int sp, const FieldRef& p, int sr, const FieldRef& r, const double* beta, const int* active,
                           int64_t per_rhs, int nrhs, int bs, hipStream_t s);
hipError_t blas_zero(int st, const FieldRef& x, int64_t per_rhs, int nrhs, int bs, hipStream_t s);

}  // namespace cpship
