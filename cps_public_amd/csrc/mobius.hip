// Moebius domain wall (4-D even-odd preconditioned, constant real b and c).
//
// Follows zmobius_m_orig / zmobius_mdag_orig (zmobius_m-orig.h:3-122,
// zmobius_mdag-orig.h) of the CPS noarch Moebius operator:
//   M   = 1 + (kb/kc) D5 - kb D4(b,c) M5^-1 kb D4(b,c)
// with D4(b,c) = Wilson hop of (b + c P(+-) shift) and M5^-1 the LDU
// recurrence of zmobius_m5inv.C:247-686.  Fields are single 4-D parity with
// the Ls slices as the inner batch index, so the 4-D hop is the Wilson
// kernel with nb = Ls (links loaded once per site for all s).
//
// Kernels here:
//   k_s5     out = xc*X + a1*(b1*A + c1*P A) + a2*(b2*B + c2*P B)
//            P the CPS fifth-dimension shift (dwf_dslash_5_plus / the
//            zmobius kappa_dslash_5_plus) with -mass at the walls
//   k_m5inv  in-place M5^-1; a block stages S sites x Ls slices in LDS, one
//            thread per (site, component) runs the recurrence over s
#include "hop.h"
#include "solver.h"

#include <cstring>

namespace cpship {

struct S5Params {
  FieldRef out, A, B, X;
  double a1, b1, c1, a2, b2, c2, xc;
  double mass;
  int dag, ls, nrhs;
  int64_t per_rhs;  // Vcb*Ls
  double* partial;
};

// value of the CPS shift term (P+ from s-1, P- from s+1 for dag 0; exchanged
// for dag 1) at slot s of a site whose Ls slots start at element `base`
template <typename S, typename R, int DAG>
__device__ __forceinline__ void s5_term(const FieldRef& f, int64_t base, int s, int ls, R mass, cplx<R> t[12]) {
  const int sm = (s == 0) ? ls - 1 : s - 1;
  const int sp = (s == ls - 1) ? 0 : s + 1;
  const R cm = (s == 0) ? -mass : R(1);
  const R cp = (s == ls - 1) ? -mass : R(1);
  constexpr int c_from_m = DAG ? 6 : 0;
  constexpr int c_from_p = DAG ? 0 : 6;
  cplx<R> hm[6], hp[6];
  spin_load_half<S, R>(f, base + sm, c_from_m, hm);
  spin_load_half<S, R>(f, base + sp, c_from_p, hp);
#pragma unroll
  for (int k = 0; k < 6; k++) {
    t[c_from_m + k] = cmk<R>(cm * hm[k].re, cm * hm[k].im);
    t[c_from_p + k] = cmk<R>(cp * hp[k].re, cp * hp[k].im);
  }
}

template <typename S, int DAG>
__global__ void __launch_bounds__(256) k_s5(S5Params P) {
  using R = typename SpinTraits<S>::R;
  const int rhs = blockIdx.y;
  const int64_t e = (int64_t)rhs * P.per_rhs + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int s = (int)(e % P.ls);
  const int64_t base = e - s;
  const R mass = R(P.mass);
  cplx<R> res[12];
  {
    cplx<R> v[12], t[12];
    spin_load<S, R>(P.A, e, v);
    s5_term<S, R, DAG>(P.A, base, s, P.ls, mass, t);
    const R a = R(P.a1), b = R(P.b1), c = R(P.c1);
    // b v + c t as one fused multiply-add, written out (the fused direction
    // update of the mixed CG computes the same expression)
#pragma unroll
    for (int k = 0; k < 12; k++)
      res[k] = cmk<R>(a * fma_r(b, v[k].re, c * t[k].re), a * fma_r(b, v[k].im, c * t[k].im));
  }
  if (P.B.data) {
    cplx<R> v[12], t[12];
    spin_load<S, R>(P.B, e, v);
    s5_term<S, R, DAG>(P.B, base, s, P.ls, mass, t);
    const R a = R(P.a2), b = R(P.b2), c = R(P.c2);
#pragma unroll
    for (int k = 0; k < 12; k++) {
      res[k].re += a * (b * v[k].re + c * t[k].re);
      res[k].im += a * (b * v[k].im + c * t[k].im);
    }
  }
  if (P.X.data) {
    cplx<R> v[12];
    spin_load<S, R>(P.X, e, v);
    const R xc = R(P.xc);
#pragma unroll
    for (int k = 0; k < 12; k++) {
      res[k].re += xc * v[k].re;
      res[k].im += xc * v[k].im;
    }
  }
  spin_store<S, R>(P.out, e, res);
  if (P.partial) {
    double n = 0;
#pragma unroll
    for (int k = 0; k < 12; k++) n += double(res[k].re) * res[k].re + double(res[k].im) * res[k].im;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, 64);
    __shared__ double red[16];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0;
      for (int w = 0; w < (int)((blockDim.x + 63) >> 6); w++) t += red[w];
      P.partial[(int64_t)rhs * gridDim.x + blockIdx.x] = t;
    }
  }
}

struct M5Params {
  FieldRef f;       // input
  FieldRef out;     // output (may be f)
  int ls, dag;
  double K, mass;
  int64_t nsites;   // nrhs * Vcb
  // fused fifth-dimension term t = b5 v + c5 P v (P of this dag, -mass at
  // the walls): S5_NONE, S5_AFTER (t of the M5^-1 result is stored) or
  // S5_BEFORE (M5^-1 is applied to t of the input)
  int s5;
  double b5, c5;
};
enum { S5_NONE = 0, S5_AFTER = 1, S5_BEFORE = 2 };

// b5 v_s + c5 (P v)_s from the LDS copy of one site's Ls slots
template <typename R, int DAG>
__device__ __forceinline__ void s5_lds_t(const R* site_lds, int s, int ls, R mass, R b5, R c5, R t[24]) {
  const int sm = (s == 0) ? ls - 1 : s - 1;
  const int sp = (s == ls - 1) ? 0 : s + 1;
  const R cm = (s == 0) ? -mass : R(1);
  const R cp = (s == ls - 1) ? -mass : R(1);
  // dag 0: reals 0-11 (spins 0,1) from s-1, 12-23 from s+1; dag 1 exchanged
  constexpr int r_from_m = DAG ? 12 : 0, r_from_p = DAG ? 0 : 12;
  const R* v = site_lds + s * 24;
  const R* vm = site_lds + sm * 24;
  const R* vp = site_lds + sp * 24;
#pragma unroll
  for (int k = 0; k < 12; k++) {
    t[r_from_m + k] = b5 * v[r_from_m + k] + c5 * (cm * vm[r_from_m + k]);
    t[r_from_p + k] = b5 * v[r_from_p + k] + c5 * (cp * vp[r_from_p + k]);
  }
}


// zmobius_m5inv_cmplx_dag0 / _dag1 with constant real K, one component
template <typename R>
__device__ __forceinline__ void m5inv_component(R* v /* [s] stride 12 */, int stride, int ls, int dag, int lower,
                                                R K, R mass) {
#define V(s_) v[(s_) * stride]
  if (dag == 0) {
    R prod = 1;
    for (int i = 0; i < ls; i++) prod *= -K;
    prod *= mass;
    const R inv_d_last = R(1) / (R(1) + prod);
    if (!lower) {
      for (int s = 0; s <= ls - 2; ++s) V(s + 1) += -K * V(s);
      R prd = 1;
      for (int i = 0; i < ls - 1; i++) prd *= -K;
      R factR = -prd * mass * inv_d_last;
      for (int s = ls - 2; s >= 0; --s) {
        V(s) += factR * V(ls - 1);
        factR /= -K;
      }
      V(ls - 1) *= inv_d_last;
    } else {
      V(ls - 1) *= inv_d_last;
      R factL = K * mass * inv_d_last;
      for (int s = 0; s <= ls - 2; ++s) {
        V(ls - 1) += factL * V(s);
        factL *= -K;
      }
      for (int s = ls - 2; s >= 0; --s) V(s) += -K * V(s + 1);
    }
  } else {
    R prod = 1;
    for (int i = 0; i < ls; i++) prod *= K;
    prod *= mass;
    const R inv_d_last = R(1) / (R(1) + prod);
    if (!lower) {
      V(ls - 1) *= inv_d_last;
      R factR = K * mass * inv_d_last;
      for (int s = 0; s <= ls - 2; ++s) {
        V(ls - 1) += factR * V(s);
        factR *= -K;
      }
      for (int s = ls - 2; s >= 0; --s) V(s) += -K * V(s + 1);
    } else {
      for (int s = 0; s <= ls - 2; ++s) V(s + 1) += -K * V(s);
      R prd = 1;
      for (int i = 0; i < ls - 2; i++) prd *= -K;
      prd *= -K;
      R factL = -prd * mass * inv_d_last;
      for (int s = ls - 2; s >= 0; --s) {
        V(s) += factL * V(ls - 1);
        if (s > 0) factL /= -K;
      }
      V(ls - 1) *= inv_d_last;
    }
  }
#undef V
}

// block: SITES sites x ls slots; LDS holds the block's fields in compute precision
template <typename S, int DAG>
__global__ void k_m5inv(M5Params P) {
  using R = typename SpinTraits<S>::R;
  extern __shared__ unsigned char smem_raw[];
  R* lds = reinterpret_cast<R*>(smem_raw);   // [site_local][s][24]
  const int ls = P.ls;
  const int sites_per_block = blockDim.x / ls;
  const int64_t site0 = (int64_t)blockIdx.x * sites_per_block;
  // phase 1: element (site_local, s) -> LDS
  const int sl_t = threadIdx.x / ls, s_t = threadIdx.x - (threadIdx.x / ls) * ls;
  const int64_t site_t = site0 + sl_t;
  const bool mine = sl_t < sites_per_block && site_t < P.nsites;
  if (mine) {
    cplx<R> v[12];
    spin_load<S, R>(P.f, site_t * ls + s_t, v);
    R* d = lds + ((int64_t)sl_t * ls + s_t) * 24;
#pragma unroll
    for (int c = 0; c < 12; c++) {
      d[2 * c] = v[c].re;
      d[2 * c + 1] = v[c].im;
    }
  }
  __syncthreads();
  if (P.s5 == S5_BEFORE) {
    R t[24];
    if (mine) s5_lds_t<R, DAG>(lds + (int64_t)sl_t * ls * 24, s_t, ls, R(P.mass), R(P.b5), R(P.c5), t);
    __syncthreads();
    if (mine) {
      R* d = lds + ((int64_t)sl_t * ls + s_t) * 24;
#pragma unroll
      for (int k = 0; k < 24; k++) d[k] = t[k];
    }
    __syncthreads();
  }
  // phase 2: (site_local, real component r in 0..23) runs the recurrence
  for (int w = threadIdx.x; w < sites_per_block * 24; w += blockDim.x) {
    const int sl = w / 24, r = w - sl * 24;
    if (site0 + sl < P.nsites) {
      const int lower = r >= 12;   // spins 2,3
      m5inv_component<R>(lds + (int64_t)sl * ls * 24 + r, 24, ls, DAG, lower, R(P.K), R(P.mass));
    }
  }
  __syncthreads();
  if (mine) {
    cplx<R> v[12];
    if (P.s5 == S5_AFTER) {
      R t[24];
      s5_lds_t<R, DAG>(lds + (int64_t)sl_t * ls * 24, s_t, ls, R(P.mass), R(P.b5), R(P.c5), t);
#pragma unroll
      for (int c = 0; c < 12; c++) v[c] = cmk<R>(t[2 * c], t[2 * c + 1]);
    } else {
      const R* d = lds + ((int64_t)sl_t * ls + s_t) * 24;
#pragma unroll
      for (int c = 0; c < 12; c++) v[c] = cmk<R>(d[2 * c], d[2 * c + 1]);
    }
    spin_store<S, R>(P.out, site_t * ls + s_t, v);
  }
}

// The same recurrence with Ls a compile-time constant: the Ls values of one
// complex component live in registers (fully unrolled, no LDS round trip
// per step) and the divisions by -K become products with 1/(-K).
template <typename R, int LS, int DAG, int LOWER>
__device__ __forceinline__ void m5inv_regs(cplx<R> (&v)[LS], R K, R mass) {
  const R mK = -K, mKinv = R(1) / (-K);
  if constexpr (DAG == 0) {
    R prod = mass;
#pragma unroll
    for (int i = 0; i < LS; i++) prod *= mK;
    const R inv_d_last = R(1) / (R(1) + prod);
    if constexpr (!LOWER) {
#pragma unroll
      for (int s = 0; s <= LS - 2; ++s) {
        v[s + 1].re += mK * v[s].re;
        v[s + 1].im += mK * v[s].im;
      }
      R prd = 1;
#pragma unroll
      for (int i = 0; i < LS - 1; i++) prd *= mK;
      R f = -prd * mass * inv_d_last;
#pragma unroll
      for (int s = LS - 2; s >= 0; --s) {
        v[s].re += f * v[LS - 1].re;
        v[s].im += f * v[LS - 1].im;
        f *= mKinv;
      }
      v[LS - 1].re *= inv_d_last;
      v[LS - 1].im *= inv_d_last;
    } else {
      v[LS - 1].re *= inv_d_last;
      v[LS - 1].im *= inv_d_last;
      R f = K * mass * inv_d_last;
#pragma unroll
      for (int s = 0; s <= LS - 2; ++s) {
        v[LS - 1].re += f * v[s].re;
        v[LS - 1].im += f * v[s].im;
        f *= mK;
      }
#pragma unroll
      for (int s = LS - 2; s >= 0; --s) {
        v[s].re += mK * v[s + 1].re;
        v[s].im += mK * v[s + 1].im;
      }
    }
  } else {
    R prod = mass;
#pragma unroll
    for (int i = 0; i < LS; i++) prod *= K;
    const R inv_d_last = R(1) / (R(1) + prod);
    if constexpr (!LOWER) {
      v[LS - 1].re *= inv_d_last;
      v[LS - 1].im *= inv_d_last;
      R f = K * mass * inv_d_last;
#pragma unroll
      for (int s = 0; s <= LS - 2; ++s) {
        v[LS - 1].re += f * v[s].re;
        v[LS - 1].im += f * v[s].im;
        f *= mK;
      }
#pragma unroll
      for (int s = LS - 2; s >= 0; --s) {
        v[s].re += mK * v[s + 1].re;
        v[s].im += mK * v[s + 1].im;
      }
    } else {
#pragma unroll
      for (int s = 0; s <= LS - 2; ++s) {
        v[s + 1].re += mK * v[s].re;
        v[s + 1].im += mK * v[s].im;
      }
      R prd = 1;
#pragma unroll
      for (int i = 0; i < LS - 1; i++) prd *= mK;
      R f = -prd * mass * inv_d_last;
#pragma unroll
      for (int s = LS - 2; s >= 0; --s) {
        v[s].re += f * v[LS - 1].re;
        v[s].im += f * v[LS - 1].im;
        if (s > 0) f *= mKinv;
      }
      v[LS - 1].re *= inv_d_last;
      v[LS - 1].im *= inv_d_last;
    }
  }
}

// M5^-1 for a compile-time Ls.  A block stages M5_SITES sites x LS slots in
// LDS (per-site stride padded by 4 reals against bank conflicts); phase 2
// gives each (site, complex component) pair to one thread, which runs the
// recurrence on registers.
template <int LS> struct M5Sites { static constexpr int value = LS <= 16 ? 256 / 16 : 256 / LS; };

template <typename S, int DAG, int LS>
__global__ void __launch_bounds__(256) k_m5inv_t(M5Params P) {
  using R = typename SpinTraits<S>::R;
  constexpr int SITES = M5Sites<LS>::value;
  constexpr int STRIDE = LS * 24 + 4;   // reals per site in LDS
  __shared__ R lds[SITES * STRIDE];
  const int64_t site0 = (int64_t)blockIdx.x * SITES;
  const int sl_t = threadIdx.x / LS, s_t = threadIdx.x - (threadIdx.x / LS) * LS;
  const int64_t site_t = site0 + sl_t;
  const bool mine = sl_t < SITES && site_t < P.nsites;
  if (mine) {
    cplx<R> v[12];
    spin_load<S, R>(P.f, site_t * LS + s_t, v);
    R* d = lds + sl_t * STRIDE + s_t * 24;
#pragma unroll
    for (int c = 0; c < 12; c++) {
      d[2 * c] = v[c].re;
      d[2 * c + 1] = v[c].im;
    }
  }
  __syncthreads();
  if (P.s5 == S5_BEFORE) {
    R t[24];
    if (mine) s5_lds_t<R, DAG>(lds + sl_t * STRIDE, s_t, LS, R(P.mass), R(P.b5), R(P.c5), t);
    __syncthreads();
    if (mine) {
      R* d = lds + sl_t * STRIDE + s_t * 24;
#pragma unroll
      for (int k = 0; k < 24; k++) d[k] = t[k];
    }
    __syncthreads();
  }
  // phase 2: (site, complex component c) -> registers; components 0-5 are
  // spins 0,1 (upper), 6-11 spins 2,3 (lower)
  for (int w = threadIdx.x; w < SITES * 12; w += blockDim.x) {
    const int sl = w / 12, c = w - sl * 12;
    if (site0 + sl < P.nsites) {
      R* col = lds + sl * STRIDE + 2 * c;
      cplx<R> v[LS];
#pragma unroll
      for (int s = 0; s < LS; s++) v[s] = cmk<R>(col[s * 24], col[s * 24 + 1]);
      if (c < 6) m5inv_regs<R, LS, DAG, 0>(v, R(P.K), R(P.mass));
      else m5inv_regs<R, LS, DAG, 1>(v, R(P.K), R(P.mass));
#pragma unroll
      for (int s = 0; s < LS; s++) {
        col[s * 24] = v[s].re;
        col[s * 24 + 1] = v[s].im;
      }
    }
  }
  __syncthreads();
  if (mine) {
    cplx<R> v[12];
    if (P.s5 == S5_AFTER) {
      R t[24];
      s5_lds_t<R, DAG>(lds + sl_t * STRIDE, s_t, LS, R(P.mass), R(P.b5), R(P.c5), t);
#pragma unroll
      for (int c = 0; c < 12; c++) v[c] = cmk<R>(t[2 * c], t[2 * c + 1]);
    } else {
      const R* d = lds + sl_t * STRIDE + s_t * 24;
#pragma unroll
      for (int c = 0; c < 12; c++) v[c] = cmk<R>(d[2 * c], d[2 * c + 1]);
    }
    spin_store<S, R>(P.out, site_t * LS + s_t, v);
  }
}

template <int LS>
static bool launch_m5inv_t(const M5Params& P, int st, int dag, hipStream_t stream) {
  constexpr int SITES = M5Sites<LS>::value;
  dim3 grid((unsigned)((P.nsites + SITES - 1) / SITES)), block(SITES * LS);
  switch (st) {
#define CPSHIP_M5T(S)                                                                  \
  if (dag) hipLaunchKernelGGL((k_m5inv_t<S, 1, LS>), grid, block, 0, stream, P);       \
  else hipLaunchKernelGGL((k_m5inv_t<S, 0, LS>), grid, block, 0, stream, P);
    case STORE_D: CPSHIP_M5T(SpinD) break;
    case STORE_F: CPSHIP_M5T(SpinF) break;
    default: CPSHIP_M5T(SpinH) break;
#undef CPSHIP_M5T
  }
  return true;
}

static int store_of(int hprec) { return hprec == PREC_D_D18 ? STORE_D : hprec == PREC_F_F12 ? STORE_F : STORE_H; }

static int launch_s5(Lattice& L, int hprec, S5Params& P) {
  const int bs = pick_block(P.per_rhs);
  dim3 grid((unsigned)(P.per_rhs / bs), (unsigned)P.nrhs), block(bs);
  switch (store_of(hprec)) {
#define CPSHIP_S5(S)                                                                   \
  if (P.dag) hipLaunchKernelGGL((k_s5<S, 1>), grid, block, 0, L.stream, P);            \
  else hipLaunchKernelGGL((k_s5<S, 0>), grid, block, 0, L.stream, P);
    case STORE_D: CPSHIP_S5(SpinD) break;
    case STORE_F: CPSHIP_S5(SpinF) break;
    default: CPSHIP_S5(SpinH) break;
#undef CPSHIP_S5
  }
  CPSHIP_CHECK(hipGetLastError());
  return 0;
}

static int launch_m5inv(Lattice& L, int hprec, const FieldRef& f, int dag, double K, double mass, int nrhs,
                        const FieldRef* out = nullptr, int s5 = S5_NONE, double b5 = 0.0, double c5 = 0.0) {
  M5Params P;
  P.f = f;
  P.out = out ? *out : f;
  P.s5 = s5;
  P.b5 = b5;
  P.c5 = c5;
  P.ls = L.ls;
  P.dag = dag;
  P.K = K;
  P.mass = mass;
  P.nsites = (int64_t)nrhs * L.g.Vcb;
  const int st = store_of(hprec);
  bool done = false;
  switch (L.ls) {
    case 4: done = launch_m5inv_t<4>(P, st, dag, L.stream); break;
    case 6: done = launch_m5inv_t<6>(P, st, dag, L.stream); break;
    case 8: done = launch_m5inv_t<8>(P, st, dag, L.stream); break;
    case 10: done = launch_m5inv_t<10>(P, st, dag, L.stream); break;
    case 12: done = launch_m5inv_t<12>(P, st, dag, L.stream); break;
    case 16: done = launch_m5inv_t<16>(P, st, dag, L.stream); break;
    case 20: done = launch_m5inv_t<20>(P, st, dag, L.stream); break;
    case 24: done = launch_m5inv_t<24>(P, st, dag, L.stream); break;
    case 32: done = launch_m5inv_t<32>(P, st, dag, L.stream); break;
    default: break;
  }
  if (done) {
    CPSHIP_CHECK(hipGetLastError());
    return 0;
  }
  // other Ls: the generic kernel (Ls at run time)
  const int sites_per_block = std::max(1, 256 / L.ls);
  const int bs = sites_per_block * L.ls;
  const size_t rbytes = st == STORE_D ? sizeof(double) : sizeof(float);
  const size_t lds = (size_t)sites_per_block * L.ls * 24 * rbytes;
  dim3 grid((unsigned)((P.nsites + sites_per_block - 1) / sites_per_block)), block(bs);
  switch (st) {
#define CPSHIP_M5(S)                                                                   \
  if (dag) hipLaunchKernelGGL((k_m5inv<S, 1>), grid, block, lds, L.stream, P);         \
  else hipLaunchKernelGGL((k_m5inv<S, 0>), grid, block, lds, L.stream, P);
    case STORE_D: CPSHIP_M5(SpinD) break;
    case STORE_F: CPSHIP_M5(SpinF) break;
    default: CPSHIP_M5(SpinH) break;
#undef CPSHIP_M5
  }
  CPSHIP_CHECK(hipGetLastError());
  return 0;
}

static void mobius_kappas(const cps_hip_op* op, double* kb, double* kc) {
  // d_op_mobius.C:106-109
  *kb = 1.0 / (2 * (op->mobius_b * (4 - op->m5) + op->a5_inv));
  *kc = 1.0 / (2 * (op->mobius_c * (4 - op->m5) - op->a5_inv));
}

// 4-D hop of every slice (Wilson kernel, nb = Ls), output parity `target`;
// with s5x the epilogue adds X + x5coef P X (P of direction dag, -mass walls)
// and writes the per-block |out|^2 partials when partial is given
// the Moebius fifth-dimension stage fused into a half-precision hop's
// epilogue (HopParams::m5post): ratio K = kappa_b / kappa_c, the walls' mass,
// and the coefficients of b5 + c5 P
struct M5Post {
  double K, mass, b, c;
};

// the CG residual update fused into the last hop of M^dag (HopParams::rupd)
struct RUpdate {
  FieldRef r;
  const double* alpha;
  const int* active;
};

// CPS_HIP_M5_FUSED=0 keeps M5^-1 as its own pass
static bool m5_fused_enabled() { return knob(&cps_hip_tuning::m5_fused, "CPS_HIP_M5_FUSED") != 0; }

// CPS_HIP_S5_FUSED=0 keeps the dag-1 s-term as its own pass
static bool s5_fused_enabled() { return knob(&cps_hip_tuning::s5_fused, "CPS_HIP_S5_FUSED") != 0; }

// s5x: out = X + x5coef (P X) + scale (hb5 hop + hc5 P hop) (CLOV_S5X)
static int hop4(Lattice& L, int hprec, const FieldRef& out, const FieldRef& in, int target, int dag, double scale,
                int nrhs, bool s5x = false, const FieldRef* X = nullptr, double x5coef = 0.0, double mass = 0.0,
                double* partial = nullptr, double hb5 = 1.0, double hc5 = 0.0, const M5Post* m5 = nullptr,
                const RUpdate* ru = nullptr) {
  cps_hip_op w;
  std::memset(&w, 0, sizeof(w));
  w.kind = CPS_HIP_OP_MOBIUS;   // single-parity layout with nb = Ls
  HopParams P;
  std::memset(&P, 0, sizeof(P));
  P.geom = L.g;
  P.kind = KIND_WILSON;
  P.prec = hprec;
  P.dag = dag;
  P.nb = L.ls;
  P.nrhs = nrhs;
  P.nsites_out = L.g.Vcb;
  P.out_parity = target;
  P.block = pick_block((long long)L.g.Vcb * L.ls);
  if (int rc = L.set_hop_phase(P, hprec)) return rc;
  P.scale = scale;
  P.in = in;
  P.out = out;
  P.X = null_field();
  P.partial = partial;
  if (s5x) {
    P.clov = CLOV_S5X;
    P.X = *X;
    P.xcoef = 1.0;
    P.x5coef = x5coef;
    P.x5mass = mass;
    P.hb5 = hb5;
    P.hc5 = hc5;
    // CPS_HIP_S5X_ROW=0 loads the slots s-1, s+1 of X instead
    P.s5_row = (store_of(hprec) == STORE_H && L.ls == 16 && knob(&cps_hip_tuning::s5x_row, "CPS_HIP_S5X_ROW") != 0) ? 1 : 0;
  }
  if (ru) {
    P.s5_row = 1;   // the residual step exists in the row variant only
    P.rupd = 1;
    P.R = ru->r;
    P.ralpha = ru->alpha;
    P.ractive = ru->active;
  }
  if (m5) {
    P.m5post = 1;
    P.m5a = -m5->K;
    P.m5inv_d = 1.0 / (1.0 + m5->mass * std::pow(m5->K, 16));
    P.m5mass = m5->mass;
    P.m5b = m5->b;
    P.m5c = m5->c;
  }
  const int lstore = link_storage_of(hprec);
  int rc = L.ensure_gauge(lstore);
  if (rc) return rc;
  P.gauge = L.gauge_ref(lstore);
  return tuned_launch(L, P);
}

static S5Params s5_base(Lattice& L, const cps_hip_op* op, int dag, int nrhs) {
  S5Params P;
  std::memset(&P, 0, sizeof(P));
  P.mass = op->mass;
  P.dag = dag;
  P.ls = L.ls;
  P.nrhs = nrhs;
  P.per_rhs = (int64_t)L.g.Vcb * L.ls;
  P.A = P.B = P.X = null_field();
  return P;
}

// zmobius_m_orig (dag 0) / zmobius_mdag_orig (dag 1); t1, t2 scratch
int mobius_m(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& out, const FieldRef& in,
             const FieldRef& t1, const FieldRef& t2, int dag, double* partial, int nrhs, bool t1_ready) {
  double kb, kc;
  mobius_kappas(op, &kb, &kc);
  const double ratio = kb / kc;
  const double b = op->mobius_b, c = op->mobius_c;
  int rc;
  // half precision with Ls = 16: M5^-1 and its s-term run in the epilogue of
  // the first hop (the 16 slots of a site are one 16-lane row)
  const bool m5f = store_of(hprec) == STORE_H && L.ls == 16 && m5_fused_enabled();
  const M5Post m5{ratio, op->mass, b, c};
  if (dag == 0) {
    // t1 = b in + c P in ; out = kb * D4(cb=1) t1
    S5Params P = s5_base(L, op, 0, nrhs);
    P.out = t1; P.A = in; P.a1 = 1; P.b1 = b; P.c1 = c;
    if (!t1_ready && (rc = launch_s5(L, hprec, P))) return rc;
    if (m5f) {
      // t2 = b M5^-1 (kb D4 t1) + c P M5^-1 (kb D4 t1) ;
      // out = in + (kb/kc) P in - kb * D4(cb=0) t2
      if ((rc = hop4(L, hprec, t2, t1, 0, 0, kb, nrhs, false, nullptr, 0.0, 0.0, nullptr, 1.0, 0.0, &m5))) return rc;
      return hop4(L, hprec, out, t2, 1, 0, -kb, nrhs, true, &in, ratio, op->mass, partial);
    }
    if ((rc = hop4(L, hprec, out, t1, 0, 0, kb, nrhs))) return rc;
    // t1 = b M5^-1 out + c P M5^-1 out (one kernel)
    if ((rc = launch_m5inv(L, hprec, out, 0, ratio, op->mass, nrhs, &t1, S5_AFTER, b, c))) return rc;
    // out = in + (kb/kc) P in - kb * D4(cb=0) t1 (the hop's epilogue)
    return hop4(L, hprec, out, t1, 1, 0, -kb, nrhs, true, &in, ratio, op->mass, partial);
  }
  // dag 1: t1 = -kb * D4dag(cb=1) in ; t2 = M5^-dag (b t1 + c Pdag t1)
  if (m5f) {
    if ((rc = hop4(L, hprec, t2, in, 0, 1, -kb, nrhs, false, nullptr, 0.0, 0.0, nullptr, 1.0, 0.0, &m5))) return rc;
  } else {
    if ((rc = hop4(L, hprec, t1, in, 0, 1, -kb, nrhs))) return rc;
    if ((rc = launch_m5inv(L, hprec, t1, 1, ratio, op->mass, nrhs, &t2, S5_BEFORE, b, c))) return rc;
  }
  // t1 = kb * D4dag(cb=0) t2 ; out = (b t1 + c Pdag t1) + (in + ratio Pdag in):
  // one hop with the s-term fused into its epilogue when the Ls slots of a
  // site fit a wavefront's lane groups, else a hop and a k_s5 pass
  if (L.ls <= 64 && 64 % L.ls == 0 && s5_fused_enabled())
    return hop4(L, hprec, out, t2, 1, 1, kb, nrhs, true, &in, ratio, op->mass, partial, b, c);
  if ((rc = hop4(L, hprec, t1, t2, 1, 1, kb, nrhs))) return rc;
  S5Params P = s5_base(L, op, 1, nrhs);
  P.out = out; P.A = t1; P.a1 = 1; P.b1 = b; P.c1 = c;
  P.B = in; P.a2 = 1; P.b2 = 1; P.c2 = ratio;
  P.partial = partial;
  return launch_s5(L, hprec, P);
}

// r -= alpha M^dag t3 for the rhs with active set, and the partials |r|^2:
// the second half of a CG iteration's M^dag M with the residual step in the
// epilogue of M^dag's last hop (half precision, Ls 16: both fifth-dimension
// stages run in hop epilogues, so M^dag is two hops)
int mobius_mdag_rupdate(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& r, const FieldRef& t3,
                        const FieldRef& t1, const FieldRef& t2, const double* alpha, const int* active,
                        double* partial, int nrhs) {
  (void)t1;
  if (store_of(hprec) != STORE_H || L.ls != 16 || !m5_fused_enabled() || !s5_fused_enabled()) return CPS_HIP_ERR_ARG;
  double kb, kc;
  mobius_kappas(op, &kb, &kc);
  const double ratio = kb / kc;
  const M5Post m5{ratio, op->mass, op->mobius_b, op->mobius_c};
  int rc = hop4(L, hprec, t2, t3, 0, 1, -kb, nrhs, false, nullptr, 0.0, 0.0, nullptr, 1.0, 0.0, &m5);
  if (rc) return rc;
  const RUpdate ru{r, alpha, active};
  return hop4(L, hprec, r, t2, 1, 1, kb, nrhs, true, &t3, ratio, op->mass, partial, op->mobius_b, op->mobius_c,
              nullptr, &ru);
}

// mobius_mdagm (mobius_mdagm.C:33-60): t3 = M in, |t3|^2, out = M^dag t3.
// t1, t2 are the operator scratch; the intermediate uses scratch slot S_T3/S_T3S.
int mobius_mdagm(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& out, const FieldRef& in,
                 const FieldRef& t1, const FieldRef& t2, double* partial, int nrhs, bool t1_ready) {
  const int st = store_of(hprec);
  Field& t3 = L.field_min(st == STORE_D ? S_T3 : S_T3S, st, (int64_t)nrhs * L.g.Vcb * L.ls);
  if (L.last_rc) return L.last_rc;
  int rc = mobius_m(L, op, hprec, t3.ref(), in, t1, t2, 0, partial, nrhs, t1_ready);
  if (rc) return rc;
  return mobius_m(L, op, hprec, out, t3.ref(), t1, t2, 1, nullptr, nrhs);
}

// DiracOpMobius::MatInv, ZMOB_PC_ORIG (d_op_mobius.C:389-640), for the
// full operator D = (b D_W + 1) + (c D_W - 1) P on vectors in
// DWF_4D_EOPREC_EE order (even 4-D parity first).
// Source preparation: t = kb D4(b,c)(cb even) M5^-1 in_e + in_o, and the
// CG source src = MatPcDag t.
int mobius_matinv_prep(Lattice& L, const cps_hip_op* op, const FieldRef& src, const FieldRef& in_o,
                       const FieldRef& in_e, const FieldRef& t1, const FieldRef& t2, const FieldRef& t3, int nrhs) {
  double kb, kc;
  mobius_kappas(op, &kb, &kc);
  int rc;
  if ((rc = launch_m5inv(L, PREC_D_D18, in_e, 0, kb / kc, op->mass, nrhs, &t3))) return rc;
  S5Params P = s5_base(L, op, 0, nrhs);
  P.out = t1; P.A = t3; P.a1 = 1; P.b1 = op->mobius_b; P.c1 = op->mobius_c;
  if ((rc = launch_s5(L, PREC_D_D18, P))) return rc;
  if ((rc = hop(L, op, PREC_D_D18, t3, t1, 1, 0, kb, in_o, 1.0, nullptr, nrhs))) return rc;
  return mobius_m(L, op, PREC_D_D18, src, t3, t1, t2, 1, nullptr, nrhs);
}

// Reconstruction: out_e = M5^-1 (kb D4(b,c)(cb odd) out_o) + M5^-1 in_e
// (the overall factor 2 kb of the reference is applied by the caller).
int mobius_matinv_reconstruct(Lattice& L, const cps_hip_op* op, const FieldRef& out_e, const FieldRef& out_o,
                              const FieldRef& in_e, const FieldRef& t1, const FieldRef& t2, const FieldRef& t3,
                              int nrhs) {
  double kb, kc;
  mobius_kappas(op, &kb, &kc);
  int rc;
  S5Params P = s5_base(L, op, 0, nrhs);
  P.out = t1; P.A = out_o; P.a1 = 1; P.b1 = op->mobius_b; P.c1 = op->mobius_c;
  if ((rc = launch_s5(L, PREC_D_D18, P))) return rc;
  if ((rc = hop(L, op, PREC_D_D18, t2, t1, 0, 0, kb, null_field(), 0.0, nullptr, nrhs))) return rc;
  if ((rc = launch_m5inv(L, PREC_D_D18, t2, 0, kb / kc, op->mass, nrhs, &out_e))) return rc;
  if ((rc = launch_m5inv(L, PREC_D_D18, in_e, 0, kb / kc, op->mass, nrhs, &t3))) return rc;
  CPSHIP_CHECK(blas_axpy(STORE_D, t3, STORE_D, out_e, nullptr, 1.0, (int64_t)L.g.Vcb * L.ls, nrhs,
                         pick_block((int64_t)L.g.Vcb * L.ls), L.stream));
  return 0;
}

}  // namespace cpship
