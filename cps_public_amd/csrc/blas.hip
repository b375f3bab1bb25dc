// Vector BLAS and CG scalar kernels for the device fields.
//
// Every kernel takes one thread per field element (12 complex components) and
// per-rhs coefficients; blocks never straddle two right-hand sides, so each
// block's reduction is a per-rhs partial (wavefront shuffle, then LDS), and
// k_reduce_rhs adds the partials of one rhs in a fixed order.  Reductions are
// therefore reproducible run to run.
//
// The CG scalar logic follows DiracOp::InvCgShift (inv_cg.C:293-364):
//   a = |r|^2_prev / d ; x += a p ; r -= a Ap ; stop if |r|^2 <= stp_cnd ;
//   b = |r|^2 / |r|^2_prev ; p = b p + r.
#include "cps_hip_internal.h"
#include "blas.h"

namespace cpship {

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ void block_partial_d(double v, double* partial, int64_t slot) {
  __shared__ double red[16];
  v = wsum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    const int nw = (blockDim.x + 63) >> 6;
    for (int k = 0; k < nw; k++) s += red[k];
    partial[slot] = s;
  }
}

// grid: (blocks_per_rhs, nrhs); element e = rhs*per_rhs + blockIdx.x*blockDim.x + tid
#define ELEM_SETUP                                                        \
  const int rhs = blockIdx.y;                                             \
  const int64_t e = (int64_t)rhs * per_rhs + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; \
  const int64_t slot = (int64_t)rhs * gridDim.x + blockIdx.x;

template <typename S>
__global__ void __launch_bounds__(256) k_norm(FieldRef x, int64_t per_rhs, double* partial) {
  ELEM_SETUP
  using R = typename SpinTraits<S>::R;
  cplx<R> v[12];
  spin_load<S, R>(x, e, v);
  double n = 0;
#pragma unroll
  for (int c = 0; c < 12; c++) n += double(v[c].re) * v[c].re + double(v[c].im) * v[c].im;
  block_partial_d(n, partial, slot);
}

template <typename S>
__global__ void __launch_bounds__(256) k_redot(FieldRef x, FieldRef y, int64_t per_rhs, double* partial) {
  ELEM_SETUP
  using R = typename SpinTraits<S>::R;
  cplx<R> a[12], b[12];
  spin_load<S, R>(x, e, a);
  spin_load<S, R>(y, e, b);
  double n = 0;
#pragma unroll
  for (int c = 0; c < 12; c++) n += double(a[c].re) * b[c].re + double(a[c].im) * b[c].im;
  block_partial_d(n, partial, slot);
}

// Re <x, y> with x and y in different storages
template <typename SX, typename SY>
__global__ void __launch_bounds__(256) k_redot_t(FieldRef x, FieldRef y, int64_t per_rhs, double* partial) {
  ELEM_SETUP
  using RX = typename SpinTraits<SX>::R;
  using RY = typename SpinTraits<SY>::R;
  cplx<RX> a[12];
  cplx<RY> b[12];
  spin_load<SX, RX>(x, e, a);
  spin_load<SY, RY>(y, e, b);
  double n = 0;
#pragma unroll
  for (int c = 0; c < 12; c++) n += double(a[c].re) * double(b[c].re) + double(a[c].im) * double(b[c].im);
  block_partial_d(n, partial, slot);
}

// r = b - a (FP64), optional |r|^2
__global__ void __launch_bounds__(256) k_sub_norm_d(FieldRef b, FieldRef a, FieldRef r, int64_t per_rhs, double* partial) {
  ELEM_SETUP
  cplx<double> vb[12], va[12];
  spin_load<SpinD, double>(b, e, vb);
  spin_load<SpinD, double>(a, e, va);
  double n = 0;
#pragma unroll
  for (int c = 0; c < 12; c++) {
    va[c] = cmk<double>(vb[c].re - va[c].re, vb[c].im - va[c].im);
    n += va[c].re * va[c].re + va[c].im * va[c].im;
  }
  spin_store<SpinD, double>(r, e, va);
  if (partial) block_partial_d(n, partial, slot);
}

// x += a p ; r -= a Ap ; |r|^2   (inv_cg.C:340-349), skipped for inactive rhs
__global__ void __launch_bounds__(256) k_cg_xr_d(FieldRef x, FieldRef p, FieldRef r, FieldRef ap,
                                                const double* alpha, const int* active, int64_t per_rhs,
                                                double* partial) {
  ELEM_SETUP
  double n = 0;
  if (active[rhs]) {
    const double a = alpha[rhs];
    {   // x += a p
      cplx<double> vx[12], vp[12];
      spin_load<SpinD, double>(x, e, vx);
      spin_load<SpinD, double>(p, e, vp);
#pragma unroll
      for (int c = 0; c < 12; c++) vx[c] = cmk<double>(a * vp[c].re + vx[c].re, a * vp[c].im + vx[c].im);
      spin_store<SpinD, double>(x, e, vx);
    }
    {   // r -= a Ap, |r|^2
      cplx<double> vr[12], va[12];
      spin_load<SpinD, double>(r, e, vr);
      spin_load<SpinD, double>(ap, e, va);
#pragma unroll
      for (int c = 0; c < 12; c++) {
        vr[c] = cmk<double>(-a * va[c].re + vr[c].re, -a * va[c].im + vr[c].im);
        n += vr[c].re * vr[c].re + vr[c].im * vr[c].im;
      }
      spin_store<SpinD, double>(r, e, vr);
    }
  }
  block_partial_d(n, partial, slot);
}

// p = b p + r for active rhs (inv_cg.C:360)
template <typename S>
__global__ void __launch_bounds__(256) k_xpay(FieldRef p, FieldRef r, const double* beta, const int* active, int64_t per_rhs) {
  const int rhs = blockIdx.y;
  if (!active[rhs]) return;
  const int64_t e = (int64_t)rhs * per_rhs + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  using R = typename SpinTraits<S>::R;
  const R b = R(beta[rhs]);
  cplx<R> vp[12], vr[12];
  spin_load<S, R>(p, e, vp);
  spin_load<S, R>(r, e, vr);
#pragma unroll
  for (int c = 0; c < 12; c++) vp[c] = cmk<R>(b * vp[c].re + vr[c].re, b * vp[c].im + vr[c].im);
  spin_store<S, R>(p, e, vp);
}

// y = a*x + y (per-rhs a), any precision pair handled in compute precision of Y
template <typename SX, typename SY>
__global__ void __launch_bounds__(256) k_axpy(FieldRef x, FieldRef y, const double* a, double a_const, int64_t per_rhs) {
  const int rhs = blockIdx.y;
  if (a && a[rhs] == 0.0) return;   // a per-rhs coefficient of 0 leaves y as it is
  const int64_t e = (int64_t)rhs * per_rhs + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  using RX = typename SpinTraits<SX>::R;
  using RY = typename SpinTraits<SY>::R;
  const RY c = RY(a ? a[rhs] : a_const);
  cplx<RX> vx[12];
  cplx<RY> vy[12];
  spin_load<SX, RX>(x, e, vx);
  spin_load<SY, RY>(y, e, vy);
#pragma unroll
  for (int k = 0; k < 12; k++) vy[k] = cmk<RY>(fma_r(c, RY(vx[k].re), vy[k].re), fma_r(c, RY(vx[k].im), vy[k].im));
  spin_store<SY, RY>(y, e, vy);
}


// xs += a p ; rs -= a ap ; |rs|^2 (sloppy CG step, any storage combination)
template <typename SX, typename SP, typename SR, typename SA>
__global__ void __launch_bounds__(256) k_cg_xr_t(FieldRef x, FieldRef p, FieldRef r, FieldRef ap, const double* alpha,
                          const int* active, int64_t per_rhs, double* partial) {
  ELEM_SETUP
  double n = 0;
  if (active[rhs]) {
    using R = float;
    const R a = R(alpha[rhs]);
    cplx<typename SpinTraits<SX>::R> vx[12];
    cplx<typename SpinTraits<SP>::R> vp[12];
    cplx<typename SpinTraits<SR>::R> vr[12];
    cplx<typename SpinTraits<SA>::R> va[12];
    spin_load<SX, typename SpinTraits<SX>::R>(x, e, vx);
    spin_load<SP, typename SpinTraits<SP>::R>(p, e, vp);
    spin_load<SR, typename SpinTraits<SR>::R>(r, e, vr);
    spin_load<SA, typename SpinTraits<SA>::R>(ap, e, va);
#pragma unroll
    for (int c = 0; c < 12; c++) {
      vx[c].re = fma_r(a, R(vp[c].re), vx[c].re); vx[c].im = fma_r(a, R(vp[c].im), vx[c].im);
      vr[c].re -= a * R(va[c].re); vr[c].im -= a * R(va[c].im);
      n += double(vr[c].re) * vr[c].re + double(vr[c].im) * vr[c].im;
    }
    spin_store<SX, typename SpinTraits<SX>::R>(x, e, vx);
    spin_store<SR, typename SpinTraits<SR>::R>(r, e, vr);
  }
  block_partial_d(n, partial, slot);
}

// The mixed CG's lagged solution update: the step x += alpha p of the
// previous iteration is applied by the kernel that reads p next, the
// direction update p = r + beta p (alpha[rhs] = 0 when no step is pending;
// a rhs that stopped still gets its last step, but no new direction).
template <typename SP, typename SR, typename SX>
__global__ void __launch_bounds__(256) k_xpay_x_t(FieldRef p, FieldRef r, FieldRef x, const double* alpha,
                                                  const double* beta, const int* active, int64_t per_rhs) {
  const int rhs = blockIdx.y;
  const double a = alpha[rhs];
  const bool act = active[rhs] != 0;
  if (!act && a == 0.0) return;
  const int64_t e = (int64_t)rhs * per_rhs + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  using R = float;
  cplx<typename SpinTraits<SP>::R> vp[12];
  spin_load<SP, typename SpinTraits<SP>::R>(p, e, vp);
  if (a != 0.0) {
    const R ar = R(a);
    cplx<typename SpinTraits<SX>::R> vx[12];
    spin_load<SX, typename SpinTraits<SX>::R>(x, e, vx);
#pragma unroll
    for (int c = 0; c < 12; c++) {
      vx[c].re = fma_r(ar, R(vp[c].re), vx[c].re);
      vx[c].im = fma_r(ar, R(vp[c].im), vx[c].im);
    }
    spin_store<SX, typename SpinTraits<SX>::R>(x, e, vx);
  }
  if (act) {
    const R b = R(beta[rhs]);
    cplx<typename SpinTraits<SR>::R> vr[12];
    spin_load<SR, typename SpinTraits<SR>::R>(r, e, vr);
    cplx<R> o[12];
#pragma unroll
    for (int c = 0; c < 12; c++) o[c] = cmk<R>(fma_r(b, R(vp[c].re), R(vr[c].re)), fma_r(b, R(vp[c].im), R(vr[c].im)));
    spin_store<SP, R>(p, e, o);
  }
}

// The same for the Moebius layout with Ls = 16, writing also
// t5 = b5 p + c5 P p (P of dag 0: upper spins from s-1, lower from s+1, -mass
// at the walls), the first step of the Moebius M, from the stored (FP16) p:
// the 16 slots of a site are one 16-lane row, so the slots s-1, s+1 come
// from the neighbouring lanes.  The arithmetic is that of k_s5.
__device__ __forceinline__ float ror_f(float v, int n) {
  return n == 1 ? __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x121, 0xF, 0xF, true))
                : __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x12F, 0xF, 0xF, true));
}

template <typename SR>
__global__ void __launch_bounds__(256) k_xpay_x_s5(FieldRef p, FieldRef r, FieldRef x, FieldRef t5, const double* alpha,
                                                   const double* beta, const int* active, int64_t per_rhs, float b5,
                                                   float c5, float mass) {
  const int rhs = blockIdx.y;
  const double a = alpha[rhs];
  const bool act = active[rhs] != 0;
  if (!act && a == 0.0) return;
  const int64_t e = (int64_t)rhs * per_rhs + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  using R = float;
  cplx<R> vp[12];
  spin_load<SpinH, R>(p, e, vp);
  if (a != 0.0) {
    const R ar = R(a);
    cplx<R> vx[12];
    spin_load<SpinF, R>(x, e, vx);
#pragma unroll
    for (int c = 0; c < 12; c++) {
      vx[c].re = fma_r(ar, R(vp[c].re), vx[c].re);
      vx[c].im = fma_r(ar, R(vp[c].im), vx[c].im);
    }
    spin_store<SpinF, R>(x, e, vx);
  }
  if (act) {
    const R b = R(beta[rhs]);
    cplx<R> vr[12];
    spin_load<SR, R>(r, e, vr);
    cplx<R> o[12], pr[12], t[12];
#pragma unroll
    for (int c = 0; c < 12; c++) o[c] = cmk<R>(fma_r(b, R(vp[c].re), R(vr[c].re)), fma_r(b, R(vp[c].im), R(vr[c].im)));
    spin_store<SpinH, R>(p, e, o);
    half_roundtrip(o, pr);
    const int s = (int)(e & 15);
    const R cm = (s == 0) ? -mass : R(1), cp = (s == 15) ? -mass : R(1);
    const R one = R(1);
#pragma unroll
    for (int k = 0; k < 6; k++) {
      t[k] = cmk<R>(one * fma_r(b5, pr[k].re, c5 * (cm * ror_f(pr[k].re, 1))),
                    one * fma_r(b5, pr[k].im, c5 * (cm * ror_f(pr[k].im, 1))));
      t[6 + k] = cmk<R>(one * fma_r(b5, pr[6 + k].re, c5 * (cp * ror_f(pr[6 + k].re, 15))),
                        one * fma_r(b5, pr[6 + k].im, c5 * (cp * ror_f(pr[6 + k].im, 15))));
    }
    spin_store<SpinH, R>(t5, e, t);
  }
}

// Paired iterations of the mixed CG with FP16 directions in two buffers: an
// iteration writes its new direction p_out = r + beta p_in into the buffer
// that held the direction before p_in; with XUPD (every second iteration)
// the two pending solution steps are applied first, in order,
// x += a_old p_out(before) ; x += a_new p_in -- the same fused multiply-adds
// as one step per iteration, with one pass over x for two iterations.  S5:
// also t5 = b5 p_out + c5 P p_out (as k_xpay_x_s5).  A rhs that stopped in
// the first iteration of a pair gets its last direction moved into p_out.
template <typename SR, bool XUPD, bool S5>
__global__ void __launch_bounds__(256) k_xpay_pair(FieldRef pin, FieldRef pout, FieldRef r, FieldRef x, FieldRef t5,
                                                   const double* a_old, const double* a_new, const double* beta,
                                                   const int* active, int64_t per_rhs, float b5, float c5,
                                                   float mass) {
  const int rhs = blockIdx.y;
  const double ao = XUPD ? a_old[rhs] : 0.0, an = XUPD ? a_new[rhs] : 0.0;
  const bool act = active[rhs] != 0;
  if (!act && ao == 0.0 && an == 0.0) return;
  const int64_t e = (int64_t)rhs * per_rhs + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  using R = float;
  cplx<R> vi[12];
  spin_load<SpinH, R>(pin, e, vi);
  if (XUPD && (ao != 0.0 || an != 0.0)) {
    cplx<R> vx[12];
    spin_load<SpinF, R>(x, e, vx);
    if (ao != 0.0) {
      const R a = R(ao);
      cplx<R> vo[12];
      spin_load<SpinH, R>(pout, e, vo);
#pragma unroll
      for (int c = 0; c < 12; c++) {
        vx[c].re = fma_r(a, vo[c].re, vx[c].re);
        vx[c].im = fma_r(a, vo[c].im, vx[c].im);
      }
    }
    if (an != 0.0) {
      const R a = R(an);
#pragma unroll
      for (int c = 0; c < 12; c++) {
        vx[c].re = fma_r(a, vi[c].re, vx[c].re);
        vx[c].im = fma_r(a, vi[c].im, vx[c].im);
      }
    }
    spin_store<SpinF, R>(x, e, vx);
  }
  if (XUPD && !act && an != 0.0) {
    // the rhs stopped at the previous iteration (its step a_new is pending):
    // its latest direction p_in is copied, unchanged, into the buffer that
    // holds the direction at the end of a pair
    const uint4* src = reinterpret_cast<const uint4*>(pin.data) + e;
    uint4* dst = reinterpret_cast<uint4*>(pout.data) + e;
#pragma unroll
    for (int k = 0; k < 3; k++) dst[k * pout.plane] = src[k * pin.plane];
    pout.norm[e] = pin.norm[e];
  }
  if (act) {
    const R b = R(beta[rhs]);
    cplx<R> vr[12];
    spin_load<SR, R>(r, e, vr);
    cplx<R> o[12];
#pragma unroll
    for (int c = 0; c < 12; c++) o[c] = cmk<R>(fma_r(b, vi[c].re, R(vr[c].re)), fma_r(b, vi[c].im, R(vr[c].im)));
    spin_store<SpinH, R>(pout, e, o);
    if constexpr (S5) {
      cplx<R> pr[12], t[12];
      half_roundtrip(o, pr);
      const int s = (int)(e & 15);
      const R cm = (s == 0) ? -mass : R(1), cp = (s == 15) ? -mass : R(1);
      const R one = R(1);
#pragma unroll
      for (int k = 0; k < 6; k++) {
        t[k] = cmk<R>(one * fma_r(b5, pr[k].re, c5 * (cm * ror_f(pr[k].re, 1))),
                      one * fma_r(b5, pr[k].im, c5 * (cm * ror_f(pr[k].im, 1))));
        t[6 + k] = cmk<R>(one * fma_r(b5, pr[6 + k].re, c5 * (cp * ror_f(pr[6 + k].re, 15))),
                          one * fma_r(b5, pr[6 + k].im, c5 * (cp * ror_f(pr[6 + k].im, 15))));
      }
      spin_store<SpinH, R>(t5, e, t);
    }
  }
}

// r -= alpha A p and the per-block |r|^2 (the solution step is lagged)
template <typename SR, typename SA>
__global__ void __launch_bounds__(256) k_cg_r_t(FieldRef r, FieldRef ap, const double* alpha, const int* active,
                                                int64_t per_rhs, double* partial) {
  ELEM_SETUP
  double n = 0;
  if (active[rhs]) {
    using R = float;
    const R a = R(alpha[rhs]);
    cplx<typename SpinTraits<SR>::R> vr[12];
    cplx<typename SpinTraits<SA>::R> va[12];
    spin_load<SR, typename SpinTraits<SR>::R>(r, e, vr);
    spin_load<SA, typename SpinTraits<SA>::R>(ap, e, va);
#pragma unroll
    for (int c = 0; c < 12; c++) {
      vr[c].re -= a * R(va[c].re); vr[c].im -= a * R(va[c].im);
      n += double(vr[c].re) * vr[c].re + double(vr[c].im) * vr[c].im;
    }
    spin_store<SR, typename SpinTraits<SR>::R>(r, e, vr);
  }
  block_partial_d(n, partial, slot);
}

// p = b p + r with p and r of different storage
template <typename SP, typename SR>
__global__ void __launch_bounds__(256) k_xpay_t(FieldRef p, FieldRef r, const double* beta, const int* active, int64_t per_rhs) {
  const int rhs = blockIdx.y;
  if (!active[rhs]) return;
  const int64_t e = (int64_t)rhs * per_rhs + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  using R = float;
  const R b = R(beta[rhs]);
  cplx<typename SpinTraits<SP>::R> vp[12];
  cplx<typename SpinTraits<SR>::R> vr[12];
  spin_load<SP, typename SpinTraits<SP>::R>(p, e, vp);
  spin_load<SR, typename SpinTraits<SR>::R>(r, e, vr);
  cplx<R> o[12];
#pragma unroll
  for (int c = 0; c < 12; c++) o[c] = cmk<R>(fma_r(b, R(vp[c].re), R(vr[c].re)), fma_r(b, R(vp[c].im), R(vr[c].im)));
  spin_store<SP, R>(p, e, o);
}

// fill a field with zeros (used instead of memset so it can sit in captured graphs)
template <typename S>
__global__ void __launch_bounds__(256) k_zero(FieldRef x, int64_t per_rhs) {
  const int64_t e = (int64_t)blockIdx.y * per_rhs + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  using R = typename SpinTraits<S>::R;
  cplx<R> v[12];
#pragma unroll
  for (int c = 0; c < 12; c++) v[c] = cmk<R>(R(0), R(0));
  spin_store<S, R>(x, e, v);
}

// One block per rhs: out[rhs] = sum of that rhs's partials (fixed order),
// then the CG scalar step selected by mode.
__global__ void k_reduce_rhs(const double* partial, int bpr, double* out, int mode, CGScalars cg) {
  const int rhs = blockIdx.x;
  __shared__ double red[256];
  double s = 0.0;
  for (int k = threadIdx.x; k < bpr; k += blockDim.x) s += partial[(int64_t)rhs * bpr + k];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x >> 1; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const double v = red[0];
  if (out) out[rhs] = v;
  if (mode == RED_CG_ALPHA) {
    // d = <p, A p>  ->  a = rr / d
    cg.d[rhs] = v;
    if (cg.alpha_prev) cg.alpha_prev[rhs] = cg.alpha[rhs];
    if (cg.active[rhs]) {
      cg.iters[rhs] += 1;
      cg.rr_prev[rhs] = cg.rr[rhs];
      if (v == 0.0) {           // inv_cg.C:326-329
        cg.active[rhs] = 0;
        cg.status[rhs] = CG_STATUS_D_ZERO;
        cg.alpha[rhs] = 0.0;
      } else {
        cg.alpha[rhs] = cg.rr[rhs] / v;
      }
    } else {
      cg.alpha[rhs] = 0.0;
    }
  } else if (mode == RED_CG_BETA) {
    if (cg.active[rhs]) {
      cg.rr[rhs] = v;
      if (cg.hist) cg.hist[(int64_t)rhs * cg.hist_len + (cg.iters[rhs] - 1) % cg.hist_len] = v;
      if (v <= cg.stp[rhs]) {
        cg.active[rhs] = 0;
        cg.status[rhs] = CG_STATUS_CONVERGED;
        cg.beta[rhs] = 0.0;
      } else if (cg.rr_upd && v < cg.delta2 * cg.rr_upd[rhs]) {
        cg.active[rhs] = 0;
        cg.status[rhs] = CG_STATUS_NEED_UPDATE;
        cg.beta[rhs] = v / cg.rr_prev[rhs];
      } else if (cg.iters[rhs] >= cg.max_itr) {
        cg.active[rhs] = 0;
        cg.status[rhs] = CG_STATUS_MAXITER;
        cg.beta[rhs] = v / cg.rr_prev[rhs];
      } else {
        cg.beta[rhs] = v / cg.rr_prev[rhs];
      }
    }
  } else if (mode == RED_CG_INIT) {
    // initial |r|^2: decides whether any iteration runs (inv_cg.C:267-276)
    cg.rr[rhs] = v;
    cg.iters[rhs] = 0;
    if (v <= cg.stp[rhs] || cg.max_itr <= 0) {
      cg.active[rhs] = 0;
      cg.status[rhs] = CG_STATUS_CONVERGED;
    } else {
      cg.active[rhs] = 1;
      cg.status[rhs] = CG_STATUS_RUNNING;
    }
  } else if (mode == RED_SRC_NORM) {
    // stp_cnd = |src|^2 rsd^2 (inv_cg.C:166-174)
    cg.src_norm[rhs] = v;
    cg.stp[rhs] = v * cg.rsd[rhs] * cg.rsd[rhs];
  }
}

static inline dim3 egrid(int64_t per_rhs, int bs, int nrhs) {
  return dim3((unsigned)(per_rhs / bs), (unsigned)nrhs);
}

hipError_t blas_norm(int st, const FieldRef& x, int64_t per_rhs, int nrhs, int bs, double* partial, hipStream_t s) {
  if (st == STORE_D) hipLaunchKernelGGL(k_norm<SpinD>, egrid(per_rhs, bs, nrhs), dim3(bs), 0, s, x, per_rhs, partial);
  else if (st == STORE_F) hipLaunchKernelGGL(k_norm<SpinF>, egrid(per_rhs, bs, nrhs), dim3(bs), 0, s, x, per_rhs, partial);
  else hipLaunchKernelGGL(k_norm<SpinH>, egrid(per_rhs, bs, nrhs), dim3(bs), 0, s, x, per_rhs, partial);
  return hipGetLastError();
}

hipError_t blas_redot(int st, const FieldRef& x, const FieldRef& y, int64_t per_rhs, int nrhs, int bs,
                      double* partial, hipStream_t s) {
  if (st == STORE_D) hipLaunchKernelGGL(k_redot<SpinD>, egrid(per_rhs, bs, nrhs), dim3(bs), 0, s, x, y, per_rhs, partial);
  else if (st == STORE_F) hipLaunchKernelGGL(k_redot<SpinF>, egrid(per_rhs, bs, nrhs), dim3(bs), 0, s, x, y, per_rhs, partial);
  else hipLaunchKernelGGL(k_redot<SpinH>, egrid(per_rhs, bs, nrhs), dim3(bs), 0, s, x, y, per_rhs, partial);
  return hipGetLastError();
}

hipError_t blas_sub_norm_d(const FieldRef& b, const FieldRef& a, const FieldRef& r, int64_t per_rhs, int nrhs, int bs,
                           double* partial, hipStream_t s) {
  hipLaunchKernelGGL(k_sub_norm_d, egrid(per_rhs, bs, nrhs), dim3(bs), 0, s, b, a, r, per_rhs, partial);
  return hipGetLastError();
}

hipError_t blas_cg_xr_d(const FieldRef& x, const FieldRef& p, const FieldRef& r, const FieldRef& ap, const CGScalars& cg,
                        int64_t per_rhs, int nrhs, int bs, double* partial, hipStream_t s) {
  hipLaunchKernelGGL(k_cg_xr_d, egrid(per_rhs, bs, nrhs), dim3(bs), 0, s, x, p, r, ap, cg.alpha, cg.active, per_rhs,
                     partial);
  return hipGetLastError();
}

hipError_t blas_xpay(int st, const FieldRef& p, const FieldRef& r, const double* beta, const int* active,
                     int64_t per_rhs, int nrhs, int bs, hipStream_t s) {
  if (st == STORE_D) hipLaunchKernelGGL(k_xpay<SpinD>, egrid(per_rhs, bs, nrhs), dim3(bs), 0, s, p, r, beta, active, per_rhs);
  else if (st == STORE_F) hipLaunchKernelGGL(k_xpay<SpinF>, egrid(per_rhs, bs, nrhs), dim3(bs), 0, s, p, r, beta, active, per_rhs);
  else hipLaunchKernelGGL(k_xpay<SpinH>, egrid(per_rhs, bs, nrhs), dim3(bs), 0, s, p, r, beta, active, per_rhs);
  return hipGetLastError();
}

hipError_t blas_redot_mixed(int sx, const FieldRef& x, int sy, const FieldRef& y, int64_t per_rhs, int nrhs, int bs,
                            double* partial, hipStream_t s) {
  dim3 g = egrid(per_rhs, bs, nrhs), b(bs);
  if (sx == STORE_F && sy == STORE_H) hipLaunchKernelGGL((k_redot_t<SpinF, SpinH>), g, b, 0, s, x, y, per_rhs, partial);
  else if (sx == STORE_F && sy == STORE_F) hipLaunchKernelGGL((k_redot_t<SpinF, SpinF>), g, b, 0, s, x, y, per_rhs, partial);
  else if (sx == STORE_D && sy == STORE_D) hipLaunchKernelGGL((k_redot_t<SpinD, SpinD>), g, b, 0, s, x, y, per_rhs, partial);
  else if (sx == STORE_H && sy == STORE_H) hipLaunchKernelGGL((k_redot_t<SpinH, SpinH>), g, b, 0, s, x, y, per_rhs, partial);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t blas_axpy(int sx, const FieldRef& x, int sy, const FieldRef& y, const double* a, double a_const,
                     int64_t per_rhs, int nrhs, int bs, hipStream_t s) {
  dim3 g = egrid(per_rhs, bs, nrhs), b(bs);
#define CPSHIP_AXPY(A, TA, B, TB) \
  if (sx == A && sy == B) { hipLaunchKernelGGL((k_axpy<TA, TB>), g, b, 0, s, x, y, a, a_const, per_rhs); return hipGetLastError(); }
  CPSHIP_AXPY(STORE_D, SpinD, STORE_D, SpinD)
  CPSHIP_AXPY(STORE_F, SpinF, STORE_D, SpinD)
  CPSHIP_AXPY(STORE_H, SpinH, STORE_D, SpinD)
  CPSHIP_AXPY(STORE_F, SpinF, STORE_F, SpinF)
  CPSHIP_AXPY(STORE_H, SpinH, STORE_F, SpinF)
  CPSHIP_AXPY(STORE_H, SpinH, STORE_H, SpinH)
  CPSHIP_AXPY(STORE_F, SpinF, STORE_H, SpinH)
#undef CPSHIP_AXPY
  return hipErrorInvalidValue;
}

hipError_t reduce_rhs(const double* partial, int bpr, int nrhs, double* out, int mode, const CGScalars& cg,
                      hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_rhs, dim3(nrhs), dim3(256), 0, s, partial, bpr, out, mode, cg);
  return hipGetLastError();
}

hipError_t blas_cg_xr_mixed(int sx, const FieldRef& x, int sp, const FieldRef& p, int sr, const FieldRef& r, int sa,
                            const FieldRef& ap, const CGScalars& cg, int64_t per_rhs, int nrhs, int bs,
                            double* partial, hipStream_t s) {
  dim3 g = egrid(per_rhs, bs, nrhs), b(bs);
#define CPSHIP_XR(X, TX, P, TP, RR, TR, A, TA)                                                          \
  if (sx == X && sp == P && sr == RR && sa == A) {                                                      \
    hipLaunchKernelGGL((k_cg_xr_t<TX, TP, TR, TA>), g, b, 0, s, x, p, r, ap, cg.alpha, cg.active, per_rhs, partial); \
    return hipGetLastError();                                                                           \
  }
  CPSHIP_XR(STORE_F, SpinF, STORE_H, SpinH, STORE_F, SpinF, STORE_H, SpinH)
  CPSHIP_XR(STORE_H, SpinH, STORE_H, SpinH, STORE_F, SpinF, STORE_H, SpinH)
  CPSHIP_XR(STORE_F, SpinF, STORE_H, SpinH, STORE_H, SpinH, STORE_H, SpinH)
  CPSHIP_XR(STORE_F, SpinF, STORE_F, SpinF, STORE_F, SpinF, STORE_F, SpinF)
  CPSHIP_XR(STORE_H, SpinH, STORE_H, SpinH, STORE_H, SpinH, STORE_H, SpinH)
#undef CPSHIP_XR
  return hipErrorInvalidValue;
}

hipError_t blas_xpay_mixed(int sp, const FieldRef& p, int sr, const FieldRef& r, const double* beta, const int* active,
                           int64_t per_rhs, int nrhs, int bs, hipStream_t s) {
  dim3 g = egrid(per_rhs, bs, nrhs), b(bs);
  if (sp == STORE_H && sr == STORE_F) { hipLaunchKernelGGL((k_xpay_t<SpinH, SpinF>), g, b, 0, s, p, r, beta, active, per_rhs); return hipGetLastError(); }
  if (sp == STORE_H && sr == STORE_H) { hipLaunchKernelGGL((k_xpay_t<SpinH, SpinH>), g, b, 0, s, p, r, beta, active, per_rhs); return hipGetLastError(); }
  if (sp == STORE_F && sr == STORE_F) { hipLaunchKernelGGL((k_xpay_t<SpinF, SpinF>), g, b, 0, s, p, r, beta, active, per_rhs); return hipGetLastError(); }
  return hipErrorInvalidValue;
}

hipError_t blas_xpay_x_mixed(int sp, const FieldRef& p, int sr, const FieldRef& r, int sx, const FieldRef& x,
                             const double* alpha, const double* beta, const int* active, int64_t per_rhs, int nrhs,
                             int bs, hipStream_t s) {
  dim3 g = egrid(per_rhs, bs, nrhs), b(bs);
#define CPSHIP_XPX(P, TP, RR, TR, X, TX)                                                                     \
  if (sp == P && sr == RR && sx == X) {                                                                      \
    hipLaunchKernelGGL((k_xpay_x_t<TP, TR, TX>), g, b, 0, s, p, r, x, alpha, beta, active, per_rhs);         \
    return hipGetLastError();                                                                                \
  }
  CPSHIP_XPX(STORE_H, SpinH, STORE_F, SpinF, STORE_F, SpinF)
  CPSHIP_XPX(STORE_H, SpinH, STORE_F, SpinF, STORE_H, SpinH)
  CPSHIP_XPX(STORE_H, SpinH, STORE_H, SpinH, STORE_H, SpinH)
  CPSHIP_XPX(STORE_H, SpinH, STORE_H, SpinH, STORE_F, SpinF)
  CPSHIP_XPX(STORE_F, SpinF, STORE_F, SpinF, STORE_F, SpinF)
#undef CPSHIP_XPX
  return hipErrorInvalidValue;
}

hipError_t blas_xpay_x_s5(const FieldRef& p, int sr, const FieldRef& r, const FieldRef& x, const FieldRef& t5,
                          const double* alpha, const double* beta, const int* active, int64_t per_rhs, int nrhs, int bs,
                          double b5, double c5, double mass, hipStream_t s) {
  if (per_rhs % 16 || bs % 64) return hipErrorInvalidValue;   // rows of 16 slots within whole wavefronts
  dim3 g = egrid(per_rhs, bs, nrhs), b(bs);
  if (sr == STORE_H)
    hipLaunchKernelGGL(k_xpay_x_s5<SpinH>, g, b, 0, s, p, r, x, t5, alpha, beta, active, per_rhs, float(b5),
                       float(c5), float(mass));
  else if (sr == STORE_F)
    hipLaunchKernelGGL(k_xpay_x_s5<SpinF>, g, b, 0, s, p, r, x, t5, alpha, beta, active, per_rhs, float(b5),
                       float(c5), float(mass));
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t blas_xpay_pair(const FieldRef& pin, const FieldRef& pout, int sr, const FieldRef& r, const FieldRef& x,
                          const FieldRef* t5, bool xupd, const double* a_old, const double* a_new, const double* beta,
                          const int* active, int64_t per_rhs, int nrhs, int bs, double b5, double c5, double mass,
                          hipStream_t s) {
  if (t5 && (per_rhs % 16 || bs % 64)) return hipErrorInvalidValue;   // rows of 16 slots in whole wavefronts
  dim3 g = egrid(per_rhs, bs, nrhs), b(bs);
  const FieldRef tt = t5 ? *t5 : pout;
  const float fb = float(b5), fc = float(c5), fm = float(mass);
#define CPSHIP_PAIR(TR, XU, S5)                                                                               \
  hipLaunchKernelGGL((k_xpay_pair<TR, XU, S5>), g, b, 0, s, pin, pout, r, x, tt, a_old, a_new, beta, active,  \
                     per_rhs, fb, fc, fm)
  if (sr == STORE_H) {
    if (t5) { if (xupd) CPSHIP_PAIR(SpinH, true, true); else CPSHIP_PAIR(SpinH, false, true); }
    else { if (xupd) CPSHIP_PAIR(SpinH, true, false); else CPSHIP_PAIR(SpinH, false, false); }
  } else if (sr == STORE_F) {
    if (t5) { if (xupd) CPSHIP_PAIR(SpinF, true, true); else CPSHIP_PAIR(SpinF, false, true); }
    else { if (xupd) CPSHIP_PAIR(SpinF, true, false); else CPSHIP_PAIR(SpinF, false, false); }
  } else {
    return hipErrorInvalidValue;
  }
#undef CPSHIP_PAIR
  return hipGetLastError();
}

hipError_t blas_cg_r_mixed(int sr, const FieldRef& r, int sa, const FieldRef& ap, const CGScalars& cg,
                           int64_t per_rhs, int nrhs, int bs, double* partial, hipStream_t s) {
  dim3 g = egrid(per_rhs, bs, nrhs), b(bs);
  if (sr == STORE_F && sa == STORE_H) hipLaunchKernelGGL((k_cg_r_t<SpinF, SpinH>), g, b, 0, s, r, ap, cg.alpha, cg.active, per_rhs, partial);
  else if (sr == STORE_H && sa == STORE_H) hipLaunchKernelGGL((k_cg_r_t<SpinH, SpinH>), g, b, 0, s, r, ap, cg.alpha, cg.active, per_rhs, partial);
  else if (sr == STORE_F && sa == STORE_F) hipLaunchKernelGGL((k_cg_r_t<SpinF, SpinF>), g, b, 0, s, r, ap, cg.alpha, cg.active, per_rhs, partial);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t blas_zero(int st, const FieldRef& x, int64_t per_rhs, int nrhs, int bs, hipStream_t s) {
  dim3 g = egrid(per_rhs, bs, nrhs), b(bs);
  if (st == STORE_D) hipLaunchKernelGGL(k_zero<SpinD>, g, b, 0, s, x, per_rhs);
  else if (st == STORE_F) hipLaunchKernelGGL(k_zero<SpinF>, g, b, 0, s, x, per_rhs);
  else hipLaunchKernelGGL(k_zero<SpinH>, g, b, 0, s, x, per_rhs);
  return hipGetLastError();
}

}  // namespace cpship
