// Layout transfer between the CPS WILSON-order arrays (array of structures,
// FP64) and the device structure-of-arrays fields, plus precision
// conversion between device fields.
//
// CPS orders followed (convert_func.C:318-494, convert.C:635-760):
//   4-D fermion checkerboard : [site_cb][spin][colour][re,im]
//   5-D DWF checkerboard cb  : slices s = 0..Ls-1 of Vcb sites, slice s at
//                              4-D parity (s+cb)%2
//   4-D-preconditioned 5-D   : slices s = 0..Ls-1 of Vcb sites of one parity
//   gauge                    : [parity][site_cb][mu][col][row][re,im]
#include "cps_hip_internal.h"
#include "layout.h"

#include <algorithm>

namespace cpship {

__device__ __forceinline__ int64_t soa_index(int map, int64_t k, int Vcb, int ls, int cb5, int nrhs_unused) {
  // k enumerates CPS order: rhs-major, then (s, site_cb) or site_cb
  (void)nrhs_unused;
  if (map == MAP_4D) return k;
  const int64_t per = (int64_t)Vcb * ls;
  const int64_t rhs = k / per;
  const int64_t rem = k - rhs * per;
  const int s = (int)(rem / Vcb);
  const int i = (int)(rem - (int64_t)s * Vcb);
  if (map == MAP_DWF5) {
    const int p = (s + cb5) & 1;
    const int nb = ls >> 1;
    return ((rhs * 2 + p) * Vcb + i) * nb + (s >> 1);
  }
  // MAP_4DPREC
  return (rhs * Vcb + i) * ls + s;
}

template <typename S>
__global__ void k_pack_spinor(const double* __restrict__ aos, FieldRef out, int map, int Vcb, int ls, int cb5,
                              int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  using R = typename SpinTraits<S>::R;
  cplx<R> v[12];
  const double* p = aos + 24 * k;
#pragma unroll
  for (int c = 0; c < 12; c++) v[c] = cmk<R>(R(p[2 * c]), R(p[2 * c + 1]));
  spin_store<S, R>(out, soa_index(map, k, Vcb, ls, cb5, 0), v);
}

template <typename S>
__global__ void k_unpack_spinor(FieldRef in, double* __restrict__ aos, int map, int Vcb, int ls, int cb5,
                                int64_t n) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  using R = typename SpinTraits<S>::R;
  cplx<R> v[12];
  spin_load<S, R>(in, soa_index(map, k, Vcb, ls, cb5, 0), v);
  double* p = aos + 24 * k;
#pragma unroll
  for (int c = 0; c < 12; c++) {
    p[2 * c] = double(v[c].re);
    p[2 * c + 1] = double(v[c].im);
  }
}

template <typename SIN, typename SOUT>
__global__ void k_convert(FieldRef in, FieldRef out, int64_t n) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  using R = float;
  if constexpr (std::is_same<SIN, SpinD>::value && std::is_same<SOUT, SpinD>::value) {
    cplx<double> v[12];
    spin_load<SIN, double>(in, e, v);
    spin_store<SOUT, double>(out, e, v);
  } else {
    using RI = typename SpinTraits<SIN>::R;
    cplx<RI> v[12];
    spin_load<SIN, RI>(in, e, v);
    using RO = typename SpinTraits<SOUT>::R;
    cplx<RO> w[12];
#pragma unroll
    for (int c = 0; c < 12; c++) w[c] = cmk<RO>(RO(v[c].re), RO(v[c].im));
    spin_store<SOUT, RO>(out, e, w);
  }
}

__device__ __forceinline__ uint32_t pack_bf16(double re, double im) {
  __hip_bfloat16 a = __float2bfloat16(float(re)), b = __float2bfloat16(float(im));
  return (uint32_t)(*reinterpret_cast<uint16_t*>(&a)) | ((uint32_t)(*reinterpret_cast<uint16_t*>(&b)) << 16);
}

// one thread per (parity, site_cb, mu)
template <typename LK>
__global__ void k_pack_gauge(const double* __restrict__ aos, GaugeRef out, int Vcb) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)8 * Vcb;
  if (k >= n) return;
  const int mu = (int)(k & 3);
  const int64_t ps = k >> 2;                 // parity*Vcb + site
  const int parity = (int)(ps / Vcb);
  const int site = (int)(ps - (int64_t)parity * Vcb);
  const double* src = aos + 72 * ps + 18 * mu;
  const int64_t idx = (int64_t)(parity * 4 + mu) * Vcb + site;
  // element c = 3*row + col; WILSON order holds U(row,col) at 2*(row+3*col)
  double re[9], im[9];
  for (int c = 0; c < 9; c++) {
    const int w = 2 * ((c / 3) + 3 * (c % 3));
    re[c] = src[w];
    im[c] = src[w + 1];
  }
  void* base = const_cast<void*>(out.data);
  if constexpr (std::is_same<LK, LinkD18>::value) {
    double2* p = reinterpret_cast<double2*>(base) + idx;
    for (int c = 0; c < 9; c++) p[c * out.plane] = make_double2(re[c], im[c]);
  } else if constexpr (std::is_same<LK, LinkF12>::value) {
    float4* p = reinterpret_cast<float4*>(base) + idx;
    for (int q = 0; q < 3; q++)
      p[q * out.plane] = make_float4(float(re[2 * q]), float(im[2 * q]), float(re[2 * q + 1]), float(im[2 * q + 1]));
  } else if constexpr (std::is_same<LK, LinkH18>::value || std::is_same<LK, LinkB18>::value) {
    uint32_t w[9];
    for (int c = 0; c < 9; c++)
      w[c] = std::is_same<LK, LinkH18>::value ? f2h(float(re[c]), float(im[c])) : pack_bf16(re[c], im[c]);
    uint4* a = reinterpret_cast<uint4*>(base);
    a[idx] = make_uint4(w[0], w[1], w[2], w[3]);
    a[out.plane + idx] = make_uint4(w[4], w[5], w[6], w[7]);
    reinterpret_cast<uint32_t*>(a + 2 * out.plane)[idx] = w[8];
  } else {
    uint32_t w[6];
    for (int c = 0; c < 6; c++)
      w[c] = std::is_same<LK, LinkH12>::value ? f2h(float(re[c]), float(im[c])) : pack_bf16(re[c], im[c]);
    reinterpret_cast<uint4*>(base)[idx] = make_uint4(w[0], w[1], w[2], w[3]);
    reinterpret_cast<uint2*>(reinterpret_cast<uint4*>(base) + out.plane)[idx] = make_uint2(w[4], w[5]);
  }
}

static inline unsigned nblocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

hipError_t pack_spinor(int storage, const double* aos, const FieldRef& out, int map, int Vcb, int ls, int cb5,
                       int64_t n, hipStream_t st) {
  const int bs = 256;
  switch (storage) {
    case STORE_D: hipLaunchKernelGGL(k_pack_spinor<SpinD>, dim3(nblocks(n, bs)), dim3(bs), 0, st, aos, out, map, Vcb, ls, cb5, n); break;
    case STORE_F: hipLaunchKernelGGL(k_pack_spinor<SpinF>, dim3(nblocks(n, bs)), dim3(bs), 0, st, aos, out, map, Vcb, ls, cb5, n); break;
    case STORE_H: hipLaunchKernelGGL(k_pack_spinor<SpinH>, dim3(nblocks(n, bs)), dim3(bs), 0, st, aos, out, map, Vcb, ls, cb5, n); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t unpack_spinor(int storage, const FieldRef& in, double* aos, int map, int Vcb, int ls, int cb5,
                         int64_t n, hipStream_t st) {
  const int bs = 256;
  switch (storage) {
    case STORE_D: hipLaunchKernelGGL(k_unpack_spinor<SpinD>, dim3(nblocks(n, bs)), dim3(bs), 0, st, in, aos, map, Vcb, ls, cb5, n); break;
    case STORE_F: hipLaunchKernelGGL(k_unpack_spinor<SpinF>, dim3(nblocks(n, bs)), dim3(bs), 0, st, in, aos, map, Vcb, ls, cb5, n); break;
    case STORE_H: hipLaunchKernelGGL(k_unpack_spinor<SpinH>, dim3(nblocks(n, bs)), dim3(bs), 0, st, in, aos, map, Vcb, ls, cb5, n); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t convert_spinor(int sin, const FieldRef& in, int sout, const FieldRef& out, int64_t n, hipStream_t st) {
  const int bs = 256;
  dim3 g(nblocks(n, bs)), b(bs);
#define CPSHIP_CONV(A, TA, B, TB) \
  if (sin == A && sout == B) { hipLaunchKernelGGL((k_convert<TA, TB>), g, b, 0, st, in, out, n); return hipGetLastError(); }
  CPSHIP_CONV(STORE_D, SpinD, STORE_D, SpinD)
  CPSHIP_CONV(STORE_D, SpinD, STORE_F, SpinF)
  CPSHIP_CONV(STORE_D, SpinD, STORE_H, SpinH)
  CPSHIP_CONV(STORE_F, SpinF, STORE_D, SpinD)
  CPSHIP_CONV(STORE_F, SpinF, STORE_F, SpinF)
  CPSHIP_CONV(STORE_F, SpinF, STORE_H, SpinH)
  CPSHIP_CONV(STORE_H, SpinH, STORE_D, SpinD)
  CPSHIP_CONV(STORE_H, SpinH, STORE_F, SpinF)
  CPSHIP_CONV(STORE_H, SpinH, STORE_H, SpinH)
#undef CPSHIP_CONV
  return hipErrorInvalidValue;
}

hipError_t pack_gauge(int link_storage, const double* aos, const GaugeRef& out, int Vcb, hipStream_t st) {
  const int bs = 256;
  const int64_t n = (int64_t)8 * Vcb;
  switch (link_storage) {
    case LINK_D18: hipLaunchKernelGGL(k_pack_gauge<LinkD18>, dim3(nblocks(n, bs)), dim3(bs), 0, st, aos, out, Vcb); break;
    case LINK_F12: hipLaunchKernelGGL(k_pack_gauge<LinkF12>, dim3(nblocks(n, bs)), dim3(bs), 0, st, aos, out, Vcb); break;
    case LINK_H12: hipLaunchKernelGGL(k_pack_gauge<LinkH12>, dim3(nblocks(n, bs)), dim3(bs), 0, st, aos, out, Vcb); break;
    case LINK_B12: hipLaunchKernelGGL(k_pack_gauge<LinkB12>, dim3(nblocks(n, bs)), dim3(bs), 0, st, aos, out, Vcb); break;
    case LINK_H18: hipLaunchKernelGGL(k_pack_gauge<LinkH18>, dim3(nblocks(n, bs)), dim3(bs), 0, st, aos, out, Vcb); break;
    case LINK_B18: hipLaunchKernelGGL(k_pack_gauge<LinkB18>, dim3(nblocks(n, bs)), dim3(bs), 0, st, aos, out, Vcb); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// det of the 3x3 complex matrix m (18 doubles, element (i, j) at 2 (i + 3 j);
// the determinant does not depend on the storage transpose)
__device__ __forceinline__ double2 det3(const double* m) {
  double ar[9], ai[9];
#pragma unroll
  for (int k = 0; k < 9; k++) {
    ar[k] = m[2 * k];
    ai[k] = m[2 * k + 1];
  }
  // cofactor expansion along the first index: sum_j a(0,j) C(0,j)
  auto mulr = [](double xr, double xi, double yr, double yi) { return xr * yr - xi * yi; };
  auto muli = [](double xr, double xi, double yr, double yi) { return xr * yi + xi * yr; };
  // element (i, j) at k = i + 3 j
  double dr = 0, di = 0;
#pragma unroll
  for (int j = 0; j < 3; j++) {
    const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    // minor (1, j1)(2, j2) - (1, j2)(2, j1)
    const double pr = mulr(ar[1 + 3 * j1], ai[1 + 3 * j1], ar[2 + 3 * j2], ai[2 + 3 * j2]) -
                      mulr(ar[1 + 3 * j2], ai[1 + 3 * j2], ar[2 + 3 * j1], ai[2 + 3 * j1]);
    const double pi = muli(ar[1 + 3 * j1], ai[1 + 3 * j1], ar[2 + 3 * j2], ai[2 + 3 * j2]) -
                      muli(ar[1 + 3 * j2], ai[1 + 3 * j2], ar[2 + 3 * j1], ai[2 + 3 * j1]);
    dr += mulr(ar[3 * j], ai[3 * j], pr, pi);
    di += muli(ar[3 * j], ai[3 * j], pr, pi);
  }
  return make_double2(dr, di);
}

// 4 threads: det of the reference link of each direction; also clears dev
__global__ void k_link_det_ref(const double* __restrict__ aos, Geom g, double* __restrict__ ref,
                               unsigned* __restrict__ dev) {
  const int mu = threadIdx.x;
  if (mu >= 4) return;
  dev[2 * mu] = 0u;
  dev[2 * mu + 1] = 0u;
  int x[4] = {0, 0, 0, 0};
  x[mu] = g.L[mu] - 1;
  const int p = (x[0] + x[1] + x[2] + x[3]) & 1;
  const int64_t site_cb = (x[0] >> 1) + (int64_t)(g.L[0] >> 1) * (x[1] + g.L[1] * (x[2] + (int64_t)g.L[2] * x[3]));
  const double2 d = det3(aos + 72 * ((int64_t)p * g.Vcb + site_cb) + 18 * mu);
  ref[2 * mu] = d.x;
  ref[2 * mu + 1] = d.y;
}

// one thread per link (parity, site_cb, mu): the largest deviation of det U
// from 1 (interior) or from the reference (last slice), per direction
__global__ void k_link_det_dev(const double* __restrict__ aos, Geom g, const double* __restrict__ ref,
                               unsigned* __restrict__ dev) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 8 * (int64_t)g.Vcb) return;
  const int mu = (int)(t & 3);
  const int64_t ps = t >> 2;   // parity * Vcb + site_cb
  const int p = ps >= g.Vcb ? 1 : 0;
  const int64_t site_cb = ps - (int64_t)p * g.Vcb;
  const int lh = g.L[0] >> 1;
  int64_t r = site_cb / lh;
  const int xh = (int)(site_cb - r * lh);
  int x[4];
  x[1] = (int)(r % g.L[1]);
  r /= g.L[1];
  x[2] = (int)(r % g.L[2]);
  x[3] = (int)(r / g.L[2]);
  x[0] = 2 * xh + ((x[1] + x[2] + x[3] + p) & 1);
  const bool last = x[mu] == g.L[mu] - 1;
  const double2 d = det3(aos + 72 * ps + 18 * mu);
  const double er = last ? ref[2 * mu] : 1.0, ei = last ? ref[2 * mu + 1] : 0.0;
  float e = (float)sqrt((d.x - er) * (d.x - er) + (d.y - ei) * (d.y - ei));
  if (!(e <= 3.0e38f)) e = 3.0e38f;   // NaN or inf: the largest finite value
  atomicMax(&dev[2 * mu + (last ? 1 : 0)], __float_as_uint(e));
}

hipError_t link_det_check(const double* aos, const Geom& g, double* ref, unsigned* dev, hipStream_t st) {
  hipLaunchKernelGGL(k_link_det_ref, dim3(1), dim3(64), 0, st, aos, g, ref, dev);
  const int64_t n = (int64_t)8 * g.Vcb;
  hipLaunchKernelGGL(k_link_det_dev, dim3(nblocks(n, 256)), dim3(256), 0, st, aos, g, (const double*)ref, dev);
  return hipGetLastError();
}

__global__ void k_flag_clear(unsigned* flag) {
  if (threadIdx.x == 0) *flag = 0u;
}

__global__ void k_words_differ(const uint64_t* __restrict__ a, const uint64_t* __restrict__ b, int64_t n,
                               unsigned* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  bool d = false;
  for (int64_t k = i; k < n; k += stride) d |= a[k] != b[k];
  if (__any(d) && (threadIdx.x & 63) == 0) atomicOr(flag, 1u);
}

hipError_t words_differ(const uint64_t* a, const uint64_t* b, int64_t n, unsigned* flag, hipStream_t st) {
  hipLaunchKernelGGL(k_flag_clear, dim3(1), dim3(64), 0, st, flag);
  const int64_t nb = std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_words_differ, dim3((unsigned)std::max<int64_t>(nb, 1)), dim3(256), 0, st, a, b, n, flag);
  return hipGetLastError();
}

__device__ __forceinline__ float hash_unit(uint64_t k) {
  k ^= k >> 33; k *= 0xff51afd7ed558ccdULL; k ^= k >> 33; k *= 0xc4ceb9fe1a85ec53ULL; k ^= k >> 33;
  return float((k >> 40) & 0xFFFFFF) * (2.0f / 16777216.0f) - 1.0f;
}

template <typename S>
__global__ void k_fill(FieldRef f, int64_t n, unsigned seed) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  using R = typename SpinTraits<S>::R;
  cplx<R> v[12];
#pragma unroll
  for (int c = 0; c < 12; c++) {
    const uint64_t k = ((uint64_t)seed << 40) ^ ((uint64_t)e * 24 + 2 * c);
    v[c] = cmk<R>(R(hash_unit(k)), R(hash_unit(k + 1)));
  }
  spin_store<S, R>(f, e, v);
}

hipError_t fill_spinor(int storage, const FieldRef& f, int64_t n, unsigned seed, hipStream_t st) {
  const int bs = 256;
  if (storage == STORE_D) hipLaunchKernelGGL(k_fill<SpinD>, dim3(nblocks(n, bs)), dim3(bs), 0, st, f, n, seed);
  else if (storage == STORE_F) hipLaunchKernelGGL(k_fill<SpinF>, dim3(nblocks(n, bs)), dim3(bs), 0, st, f, n, seed);
  else hipLaunchKernelGGL(k_fill<SpinH>, dim3(nblocks(n, bs)), dim3(bs), 0, st, f, n, seed);
  return hipGetLastError();
}

}  // namespace cpship
