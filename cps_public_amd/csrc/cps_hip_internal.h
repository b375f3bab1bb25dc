// Internal definitions shared by the HIP kernels and the host driver.
//
// Device layout (structure of arrays, one "plane" per complex component):
//   spinor field element e = (rhs * nsites + site) * nb + j
//     site : checkerboard site index (for fields that hold both 4-D parities,
//            site = parity * Vcb + site_cb); site_cb = x/2 + Lx/2*(y+Ly*(z+Lz*t))
//     j    : inner batch index (the Ls/2 fifth-dimension slots of a 5-D
//            checkerboarded DWF field, the Ls slices of a 4-D-preconditioned
//            Moebius field, 1 for a 4-D Wilson field)
//     plane c (0..11) holds complex component c = 3*spin + colour.
//   half-precision spinors add one float norm per element (max |real|).
//   gauge: plane k (complex element k = 3*row+col, canonical orientation)
//     at index (parity*4 + mu) * Vcb + site_cb; 9 planes (18 reals) or 6
//     planes (rows 0 and 1, "12-real" compression).
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <cstdint>
#include <cstdio>

namespace cpship {

#define CPSHIP_CHECK(cmd)                                                   \
  do {                                                                      \
    hipError_t e_ = (cmd);                                                  \
    if (e_ != hipSuccess) {                                                 \
      fprintf(stderr, "cps_hip: %s failed: %s (%s:%d)\n", #cmd,             \
              hipGetErrorString(e_), __FILE__, __LINE__);                   \
      return CPS_HIP_ERR_RUNTIME;                                           \
    }                                                                       \
  } while (0)

struct Geom {
  int L[4];     // local lattice extent x, y, z, t
  int Vcb;      // V/2
  int bc[4];    // 1 = antiperiodic in direction mu (sign folded into links)
};

// ---- scalar / complex helpers ------------------------------------------
template <typename R> struct cplx { R re, im; };

template <typename R>
__device__ __forceinline__ cplx<R> cmk(R a, R b) { cplx<R> c; c.re = a; c.im = b; return c; }

// a*b
template <typename R>
__device__ __forceinline__ cplx<R> cmul(cplx<R> a, cplx<R> b) {
  return cmk<R>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
// conj(a)*b
template <typename R>
__device__ __forceinline__ cplx<R> cmulc(cplx<R> a, cplx<R> b) {
  return cmk<R>(a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re);
}

// ---- spinor storage -------------------------------------------------------
// Storage tags.  load() returns the 12 complex components of one element in
// compute precision R; store() writes them.
struct SpinD {};   // double2 planes
struct SpinF {};   // float2 planes
struct SpinH {};   // __half2 planes + float norm

template <typename S> struct SpinTraits;
template <> struct SpinTraits<SpinD> { using elem = double2; using R = double; static constexpr bool has_norm = false; };
template <> struct SpinTraits<SpinF> { using elem = float2; using R = float; static constexpr bool has_norm = false; };
template <> struct SpinTraits<SpinH> { using elem = __half2; using R = float; static constexpr bool has_norm = true; };

struct FieldRef {
  void* data;       // base of plane 0
  float* norm;      // per-element norm (half storage) or nullptr
  int64_t plane;    // elements per plane
};

template <typename S, typename R>
__device__ __forceinline__ void spin_load(const FieldRef& f, int64_t e, cplx<R> v[12]) {
  using E = typename SpinTraits<S>::elem;
  const E* p = reinterpret_cast<const E*>(f.data) + e;
  if constexpr (SpinTraits<S>::has_norm) {
    float n = f.norm[e];
#pragma unroll
    for (int c = 0; c < 12; c++) {
      float2 h = __half22float2(p[c * f.plane]);
      v[c] = cmk<R>(R(h.x * n), R(h.y * n));
    }
  } else {
#pragma unroll
    for (int c = 0; c < 12; c++) {
      E h = p[c * f.plane];
      v[c] = cmk<R>(R(h.x), R(h.y));
    }
  }
}

// Load only components [c0, c0+6) (upper: 0, lower: 6 -- spins 0,1 or 2,3).
template <typename S, typename R>
__device__ __forceinline__ void spin_load_half(const FieldRef& f, int64_t e, int c0, cplx<R> v[6]) {
  using E = typename SpinTraits<S>::elem;
  const E* p = reinterpret_cast<const E*>(f.data) + e;
  if constexpr (SpinTraits<S>::has_norm) {
    float n = f.norm[e];
#pragma unroll
    for (int c = 0; c < 6; c++) {
      float2 h = __half22float2(p[(c0 + c) * f.plane]);
      v[c] = cmk<R>(R(h.x * n), R(h.y * n));
    }
  } else {
#pragma unroll
    for (int c = 0; c < 6; c++) {
      E h = p[(c0 + c) * f.plane];
      v[c] = cmk<R>(R(h.x), R(h.y));
    }
  }
}

template <typename S, typename R>
__device__ __forceinline__ void spin_store(const FieldRef& f, int64_t e, const cplx<R> v[12]) {
  using E = typename SpinTraits<S>::elem;
  E* p = reinterpret_cast<E*>(f.data) + e;
  if constexpr (SpinTraits<S>::has_norm) {
    float m = 0.f;
#pragma unroll
    for (int c = 0; c < 12; c++) m = fmaxf(m, fmaxf(fabsf(float(v[c].re)), fabsf(float(v[c].im))));
    float inv = m > 0.f ? 1.0f / m : 0.f;
    f.norm[e] = m;
#pragma unroll
    for (int c = 0; c < 12; c++)
      p[c * f.plane] = __floats2half2_rn(float(v[c].re) * inv, float(v[c].im) * inv);
  } else {
#pragma unroll
    for (int c = 0; c < 12; c++) {
      E h;
      h.x = v[c].re;
      h.y = v[c].im;
      p[c * f.plane] = h;
    }
  }
}

// ---- gauge storage ------------------------------------------------------------
struct LinkD18 {};  // double2 x 9
struct LinkF12 {};  // float2 x 6 (rows 0,1)
struct LinkH12 {};  // __half2 x 6
struct LinkB12 {};  // __hip_bfloat162 x 6

template <typename LK> struct LinkTraits;
template <> struct LinkTraits<LinkD18> { using elem = double2; static constexpr int nplanes = 9; };
template <> struct LinkTraits<LinkF12> { using elem = float2; static constexpr int nplanes = 6; };
template <> struct LinkTraits<LinkH12> { using elem = __half2; static constexpr int nplanes = 6; };
template <> struct LinkTraits<LinkB12> { using elem = __hip_bfloat162; static constexpr int nplanes = 6; };

struct GaugeRef {
  const void* data;
  int64_t plane;    // = 8 * Vcb
};

template <typename E>
__device__ __forceinline__ float2 link_elem_to_float2(E e);
template <> __device__ __forceinline__ float2 link_elem_to_float2<float2>(float2 e) { return e; }
template <> __device__ __forceinline__ float2 link_elem_to_float2<__half2>(__half2 e) { return __half22float2(e); }
template <> __device__ __forceinline__ float2 link_elem_to_float2<__hip_bfloat162>(__hip_bfloat162 e) {
  return make_float2(__bfloat162float(e.x), __bfloat162float(e.y));
}

// Load link (parity, mu, site_cb) as a full 3x3 matrix u[row*3+col].
// sign: boundary sign for the reconstructed third row of a 12-real link.
template <typename LK, typename R>
__device__ __forceinline__ void link_load(const GaugeRef& g, int parity, int mu, int site, int Vcb,
                                          R sign, cplx<R> u[9]) {
  using E = typename LinkTraits<LK>::elem;
  const E* p = reinterpret_cast<const E*>(g.data) + (int64_t)(parity * 4 + mu) * Vcb + site;
  if constexpr (LinkTraits<LK>::nplanes == 9) {
#pragma unroll
    for (int k = 0; k < 9; k++) {
      E h = p[k * g.plane];
      u[k] = cmk<R>(R(h.x), R(h.y));
    }
  } else {
#pragma unroll
    for (int k = 0; k < 6; k++) {
      float2 h = link_elem_to_float2<E>(p[k * g.plane]);
      u[k] = cmk<R>(R(h.x), R(h.y));
    }
    // row2 = sign * conj(row0 x row1)
    u[6] = cmk<R>(sign * (u[1].re * u[5].re - u[1].im * u[5].im - u[2].re * u[4].re + u[2].im * u[4].im),
                  -sign * (u[1].re * u[5].im + u[1].im * u[5].re - u[2].re * u[4].im - u[2].im * u[4].re));
    u[7] = cmk<R>(sign * (u[2].re * u[3].re - u[2].im * u[3].im - u[0].re * u[5].re + u[0].im * u[5].im),
                  -sign * (u[2].re * u[3].im + u[2].im * u[3].re - u[0].re * u[5].im - u[0].im * u[5].re));
    u[8] = cmk<R>(sign * (u[0].re * u[4].re - u[0].im * u[4].im - u[1].re * u[3].re + u[1].im * u[3].im),
                  -sign * (u[0].re * u[4].im + u[0].im * u[4].re - u[1].re * u[3].im - u[1].im * u[3].re));
  }
}

}  // namespace cpship
