// Internal definitions shared by the HIP kernels and the host driver.
//
// Device layout (structure of arrays of 16-byte chunks):
//   spinor field element e = (rhs * nsites + site) * nb + j
//     site : checkerboard site index (for fields that hold both 4-D parities,
//            site = parity * Vcb + site_cb); site_cb = x/2 + Lx/2*(y+Ly*(z+Lz*t))
//     j    : inner batch index (the Ls/2 fifth-dimension slots of a 5-D
//            checkerboarded DWF field, the Ls slices of a 4-D-preconditioned
//            Moebius field, 1 for a 4-D Wilson field)
//   The 12 complex components c = 3*spin + colour of an element are split
//   into 16-byte chunks; chunk k of every element forms plane k, so one
//   wavefront reads 64 consecutive 16-byte chunks per load instruction:
//     FP64 : 12 planes, chunk k = component k            (double2)
//     FP32 :  6 planes, chunk k = components 2k, 2k+1     (float4)
//     FP16 :  3 planes, chunk k = components 4k..4k+3     (8 halves)
//            + one float per element: the element's max |real|, by which
//              the stored halves are normalised
//   gauge: link (parity, mu, site_cb) at index (parity*4 + mu)*Vcb + site_cb
//     of every plane; complex element k = 3*row + col (canonical orientation)
//     FP64 18-real: 9 planes of double2
//     FP32 12-real: 3 planes of float4 (rows 0 and 1)
//     FP16 / BF16 12-real: one plane of 16 bytes (elements 0..3) followed by
//            one plane of 8 bytes (elements 4, 5)
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
  int bc[4];    // boundary condition of direction mu (CPS BndCndType: 0
                // periodic, 1 antiperiodic, 2 twisted), applied to the links
                // by the CANONICAL gauge import
  double bf[4][2];   // factor (re, im) of the links of the last mu slice that
                     // the CANONICAL import applies (1, -1 or the twist phase)
};

// ---- scalar / complex helpers ------------------------------------------
template <typename R> struct cplx { R re, im; };

template <typename R>
__device__ __forceinline__ cplx<R> cmk(R a, R b) { cplx<R> c; c.re = a; c.im = b; return c; }

// a*b
template <typename R>
__device__ __forceinline__ cplx<R> cadd(cplx<R> a, cplx<R> b) { return cmk<R>(a.re + b.re, a.im + b.im); }
template <typename R>
__device__ __forceinline__ cplx<R> csub(cplx<R> a, cplx<R> b) { return cmk<R>(a.re - b.re, a.im - b.im); }
template <typename R>
__device__ __forceinline__ cplx<R> cconj(cplx<R> a) { return cmk<R>(a.re, -a.im); }
template <typename R>
__device__ __forceinline__ cplx<R> cmul(cplx<R> a, cplx<R> b) {
  return cmk<R>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
// conj(a)*b
template <typename R>
__device__ __forceinline__ cplx<R> cmulc(cplx<R> a, cplx<R> b) {
  return cmk<R>(a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re);
}

__device__ __forceinline__ float2 h2f(uint32_t w) {
  __half2 h = *reinterpret_cast<__half2*>(&w);
  return __half22float2(h);
}
__device__ __forceinline__ uint32_t f2h(float a, float b) {
  __half2 h = __floats2half2_rn(a, b);
  return *reinterpret_cast<uint32_t*>(&h);
}
__device__ __forceinline__ float2 bf2f(uint32_t w) {
  // bf16 pair (low: real, high: imaginary) -> two floats
  return make_float2(__uint_as_float(w << 16), __uint_as_float(w & 0xffff0000u));
}

// ---- spinor storage -------------------------------------------------------
struct SpinD {};   // 12 planes of double2
struct SpinF {};   // 6 planes of float4
struct SpinH {};   // 3 planes of 8 halves + float norm

template <typename S> struct SpinTraits;
template <> struct SpinTraits<SpinD> { using R = double; static constexpr bool has_norm = false; static constexpr int nplanes = 12; };
template <> struct SpinTraits<SpinF> { using R = float; static constexpr bool has_norm = false; static constexpr int nplanes = 6; };
template <> struct SpinTraits<SpinH> { using R = float; static constexpr bool has_norm = true; static constexpr int nplanes = 3; };

struct FieldRef {
  void* data;       // base of plane 0
  float* norm;      // per-element norm (half storage) or nullptr
  int64_t plane;    // elements per plane
};

template <typename S, typename R>
__device__ __forceinline__ void spin_load(const FieldRef& f, int64_t e, cplx<R> v[12]) {
  if constexpr (std::is_same<S, SpinD>::value) {
    const double2* p = reinterpret_cast<const double2*>(f.data) + e;
#pragma unroll
    for (int c = 0; c < 12; c++) {
      const double2 h = p[c * f.plane];
      v[c] = cmk<R>(R(h.x), R(h.y));
    }
  } else if constexpr (std::is_same<S, SpinF>::value) {
    const float4* p = reinterpret_cast<const float4*>(f.data) + e;
#pragma unroll
    for (int k = 0; k < 6; k++) {
      const float4 h = p[k * f.plane];
      v[2 * k] = cmk<R>(R(h.x), R(h.y));
      v[2 * k + 1] = cmk<R>(R(h.z), R(h.w));
    }
  } else {
    const uint4* p = reinterpret_cast<const uint4*>(f.data) + e;
    const float n = f.norm[e];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const uint4 q = p[k * f.plane];
      const float2 a = h2f(q.x), b = h2f(q.y), c = h2f(q.z), d = h2f(q.w);
      v[4 * k + 0] = cmk<R>(R(a.x * n), R(a.y * n));
      v[4 * k + 1] = cmk<R>(R(b.x * n), R(b.y * n));
      v[4 * k + 2] = cmk<R>(R(c.x * n), R(c.y * n));
      v[4 * k + 3] = cmk<R>(R(d.x * n), R(d.y * n));
    }
  }
}

// Load only components [c0, c0+6) (c0 = 0: spins 0,1; c0 = 6: spins 2,3).
template <typename S, typename R>
__device__ __forceinline__ void spin_load_half(const FieldRef& f, int64_t e, int c0, cplx<R> v[6]) {
  if constexpr (std::is_same<S, SpinD>::value) {
    const double2* p = reinterpret_cast<const double2*>(f.data) + e;
#pragma unroll
    for (int c = 0; c < 6; c++) {
      const double2 h = p[(c0 + c) * f.plane];
      v[c] = cmk<R>(R(h.x), R(h.y));
    }
  } else if constexpr (std::is_same<S, SpinF>::value) {
    const float4* p = reinterpret_cast<const float4*>(f.data) + e;
    const int k0 = c0 >> 1;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const float4 h = p[(k0 + k) * f.plane];
      v[2 * k] = cmk<R>(R(h.x), R(h.y));
      v[2 * k + 1] = cmk<R>(R(h.z), R(h.w));
    }
  } else {
    const float n = f.norm[e];
    const uint4* p = reinterpret_cast<const uint4*>(f.data) + e;
    if (c0 == 0) {
      const uint4 q = p[0];
      const uint2 r = *reinterpret_cast<const uint2*>(p + f.plane);   // components 4, 5
      const uint32_t w[6] = {q.x, q.y, q.z, q.w, r.x, r.y};
#pragma unroll
      for (int c = 0; c < 6; c++) {
        const float2 a = h2f(w[c]);
        v[c] = cmk<R>(R(a.x * n), R(a.y * n));
      }
    } else {
      const uint2 r = *(reinterpret_cast<const uint2*>(p + f.plane) + 1);  // components 6, 7
      const uint4 q = p[2 * f.plane];                                     // 8..11
      const uint32_t w[6] = {r.x, r.y, q.x, q.y, q.z, q.w};
#pragma unroll
      for (int c = 0; c < 6; c++) {
        const float2 a = h2f(w[c]);
        v[c] = cmk<R>(R(a.x * n), R(a.y * n));
      }
    }
  }
}

template <typename T>
__device__ __forceinline__ void st_nt(T* p, const T& v) { __builtin_nontemporal_store(v, p); }

// NT = true: streaming store (the output is not re-read by this kernel)
template <typename S, typename R, bool NT = false>
__device__ __forceinline__ void spin_store(const FieldRef& f, int64_t e, const cplx<R> v[12]) {
  if constexpr (std::is_same<S, SpinD>::value) {
    double2* p = reinterpret_cast<double2*>(f.data) + e;
#pragma unroll
    for (int c = 0; c < 12; c++) {
      const double2 h = make_double2(double(v[c].re), double(v[c].im));
      if constexpr (NT) {
        st_nt(&p[c * f.plane].x, h.x);
        st_nt(&p[c * f.plane].y, h.y);
      } else {
        p[c * f.plane] = h;
      }
    }
  } else if constexpr (std::is_same<S, SpinF>::value) {
    float4* p = reinterpret_cast<float4*>(f.data) + e;
#pragma unroll
    for (int k = 0; k < 6; k++) {
      const float4 h = make_float4(float(v[2 * k].re), float(v[2 * k].im), float(v[2 * k + 1].re),
                                   float(v[2 * k + 1].im));
      p[k * f.plane] = h;
    }
  } else {
    float m = 0.f;
#pragma unroll
    for (int c = 0; c < 12; c++) m = fmaxf(m, fmaxf(fabsf(float(v[c].re)), fabsf(float(v[c].im))));
    // a site whose largest component is an FP32 subnormal (far from a point
    // source the residual decays below 1e-38) is stored as zero: 1/m would
    // overflow to inf and the stored halves would become inf/NaN
    const float inv = m >= 1.17549435e-38f ? 1.0f / m : 0.f;
    uint4* p = reinterpret_cast<uint4*>(f.data) + e;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      uint4 q;
      q.x = f2h(float(v[4 * k + 0].re) * inv, float(v[4 * k + 0].im) * inv);
      q.y = f2h(float(v[4 * k + 1].re) * inv, float(v[4 * k + 1].im) * inv);
      q.z = f2h(float(v[4 * k + 2].re) * inv, float(v[4 * k + 2].im) * inv);
      q.w = f2h(float(v[4 * k + 3].re) * inv, float(v[4 * k + 3].im) * inv);
      p[k * f.plane] = q;
    }
    f.norm[e] = m;
  }
}

template <typename R>
__device__ __forceinline__ R fma_r(R a, R b, R c) {
  if constexpr (std::is_same<R, float>::value) return __builtin_fmaf(a, b, c);
  else return __builtin_fma(a, b, c);
}

// The values spin_load<SpinH> returns after spin_store<SpinH> of v (the FP16
// quantisation with the element's norm, undone)
__device__ __forceinline__ void half_roundtrip(const cplx<float> v[12], cplx<float> r[12]) {
  float m = 0.f;
#pragma unroll
  for (int c = 0; c < 12; c++) m = fmaxf(m, fmaxf(fabsf(v[c].re), fabsf(v[c].im)));
  const float inv = m >= 1.17549435e-38f ? 1.0f / m : 0.f;
#pragma unroll
  for (int c = 0; c < 12; c++) {
    const float2 h = h2f(f2h(v[c].re * inv, v[c].im * inv));
    r[c] = cmk<float>(h.x * m, h.y * m);
  }
}

// ---- gauge storage ------------------------------------------------------------
struct LinkD18 {};  // 9 planes of double2
struct LinkF12 {};  // 3 planes of float4 (rows 0,1)
struct LinkH12 {};  // 16-byte plane + 8-byte plane of half pairs
struct LinkB12 {};  // same with bfloat16 pairs
struct LinkH18 {};  // two 16-byte planes + one 4-byte plane of half pairs (all 9 elements)
struct LinkB18 {};  // same with bfloat16 pairs

template <typename LK> struct LinkTraits;
template <> struct LinkTraits<LinkD18> { static constexpr int nreal = 18; static constexpr int bytes = 144; };
template <> struct LinkTraits<LinkF12> { static constexpr int nreal = 12; static constexpr int bytes = 48; };
template <> struct LinkTraits<LinkH12> { static constexpr int nreal = 12; static constexpr int bytes = 24; };
template <> struct LinkTraits<LinkB12> { static constexpr int nreal = 12; static constexpr int bytes = 24; };
template <> struct LinkTraits<LinkH18> { static constexpr int nreal = 18; static constexpr int bytes = 36; };
template <> struct LinkTraits<LinkB18> { static constexpr int nreal = 18; static constexpr int bytes = 36; };

struct GaugeRef {
  const void* data;
  int64_t plane;    // = 8 * Vcb links per plane
};

// Third row of a link stored by its first two rows: for any unitary U,
// row2 = det(U) conj(row0 x row1).  ph is det(U): 1 for SU(3) links, the
// boundary factor (-1 antiperiodic, e^{3 i theta} twisted by e^{i theta})
// for the links of the last slice (HopParams::bph).
template <typename R>
__device__ __forceinline__ void link_row2(cplx<R> u[9], float2 ph) {
  const R pr = R(ph.x), pi = R(ph.y);
  const R a0 = u[1].re * u[5].re - u[1].im * u[5].im - u[2].re * u[4].re + u[2].im * u[4].im;
  const R b0 = -(u[1].re * u[5].im + u[1].im * u[5].re - u[2].re * u[4].im - u[2].im * u[4].re);
  const R a1 = u[2].re * u[3].re - u[2].im * u[3].im - u[0].re * u[5].re + u[0].im * u[5].im;
  const R b1 = -(u[2].re * u[3].im + u[2].im * u[3].re - u[0].re * u[5].im - u[0].im * u[5].re);
  const R a2 = u[0].re * u[4].re - u[0].im * u[4].im - u[1].re * u[3].re + u[1].im * u[3].im;
  const R b2 = -(u[0].re * u[4].im + u[0].im * u[4].re - u[1].re * u[3].im - u[1].im * u[3].re);
  u[6] = cmk<R>(pr * a0 - pi * b0, pr * b0 + pi * a0);
  u[7] = cmk<R>(pr * a1 - pi * b1, pr * b1 + pi * a1);
  u[8] = cmk<R>(pr * a2 - pi * b2, pr * b2 + pi * a2);
}

// Load link (parity, mu, site_cb) as a full 3x3 matrix u[row*3+col].
// ph: factor of the reconstructed third row of a 12-real link (link_row2).
template <typename LK, typename R>
__device__ __forceinline__ void link_load(const GaugeRef& g, int parity, int mu, int site, int Vcb,
                                          float2 ph, cplx<R> u[9]) {
  const int64_t idx = (int64_t)(parity * 4 + mu) * Vcb + site;
  if constexpr (std::is_same<LK, LinkD18>::value) {
    const double2* p = reinterpret_cast<const double2*>(g.data) + idx;
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const double2 h = p[k * g.plane];
      u[k] = cmk<R>(R(h.x), R(h.y));
    }
    return;
  } else {
    if constexpr (std::is_same<LK, LinkF12>::value) {
      const float4* p = reinterpret_cast<const float4*>(g.data) + idx;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const float4 h = p[k * g.plane];
        u[2 * k] = cmk<R>(R(h.x), R(h.y));
        u[2 * k + 1] = cmk<R>(R(h.z), R(h.w));
      }
    } else {
      const uint4 a = reinterpret_cast<const uint4*>(g.data)[idx];
      const uint2 b = reinterpret_cast<const uint2*>(reinterpret_cast<const uint4*>(g.data) + g.plane)[idx];
      const uint32_t w[6] = {a.x, a.y, a.z, a.w, b.x, b.y};
#pragma unroll
      for (int k = 0; k < 6; k++) {
        const float2 h = std::is_same<LK, LinkH12>::value ? h2f(w[k]) : bf2f(w[k]);
        u[k] = cmk<R>(R(h.x), R(h.y));
      }
    }
    link_row2<R>(u, ph);
  }
}

// ---- raw (undecoded) loads ------------------------------------------------
// The hopping kernel issues the loads of several directions before decoding
// any of them, so each wavefront keeps many memory requests in flight.
template <typename S> struct SpinRaw;
template <> struct SpinRaw<SpinD> { double2 q[12]; };
template <> struct SpinRaw<SpinF> { float4 q[6]; };
template <> struct SpinRaw<SpinH> { uint4 q[3]; float n; };

// Load of type T at a uniform plane base plus a 32-bit per-lane byte offset
// (the plane base stays in scalar registers; no 64-bit address arithmetic
// per lane).  Every plane of every field is below 4 GiB (checked on the host).
template <typename T>
__device__ __forceinline__ T ld_off(const void* base, int64_t uniform_bytes, uint32_t off) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + uniform_bytes + off);
}

template <typename S>
__device__ __forceinline__ void raw_load(SpinRaw<S>& r, const FieldRef& f, int64_t e) {
  const uint32_t off = (uint32_t)e << 4;        // 16-byte chunks
  const int64_t pb = f.plane * 16;
  if constexpr (std::is_same<S, SpinD>::value) {
#pragma unroll
    for (int c = 0; c < 12; c++) r.q[c] = ld_off<double2>(f.data, c * pb, off);
  } else if constexpr (std::is_same<S, SpinF>::value) {
#pragma unroll
    for (int k = 0; k < 6; k++) r.q[k] = ld_off<float4>(f.data, k * pb, off);
  } else {
#pragma unroll
    for (int k = 0; k < 3; k++) r.q[k] = ld_off<uint4>(f.data, k * pb, off);
    r.n = ld_off<float>(f.norm, 0, (uint32_t)e << 2);
  }
}

template <typename S, typename R>
__device__ __forceinline__ void raw_decode(const SpinRaw<S>& r, cplx<R> v[12]) {
  if constexpr (std::is_same<S, SpinD>::value) {
#pragma unroll
    for (int c = 0; c < 12; c++) v[c] = cmk<R>(R(r.q[c].x), R(r.q[c].y));
  } else if constexpr (std::is_same<S, SpinF>::value) {
#pragma unroll
    for (int k = 0; k < 6; k++) {
      v[2 * k] = cmk<R>(R(r.q[k].x), R(r.q[k].y));
      v[2 * k + 1] = cmk<R>(R(r.q[k].z), R(r.q[k].w));
    }
  } else {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const float2 a = h2f(r.q[k].x), b = h2f(r.q[k].y), c = h2f(r.q[k].z), d = h2f(r.q[k].w);
      v[4 * k + 0] = cmk<R>(R(a.x * r.n), R(a.y * r.n));
      v[4 * k + 1] = cmk<R>(R(b.x * r.n), R(b.y * r.n));
      v[4 * k + 2] = cmk<R>(R(c.x * r.n), R(c.y * r.n));
      v[4 * k + 3] = cmk<R>(R(d.x * r.n), R(d.y * r.n));
    }
  }
}

template <typename LK> struct LinkRaw;
template <> struct LinkRaw<LinkD18> { double2 q[9]; };
template <> struct LinkRaw<LinkF12> { float4 q[3]; };
template <> struct LinkRaw<LinkH12> { uint4 a; uint2 b; };
template <> struct LinkRaw<LinkB12> { uint4 a; uint2 b; };
template <> struct LinkRaw<LinkH18> { uint4 a, b; uint32_t c; };
template <> struct LinkRaw<LinkB18> { uint4 a, b; uint32_t c; };

template <typename LK> struct LinkIs18Half { static constexpr bool value = false; };
template <> struct LinkIs18Half<LinkH18> { static constexpr bool value = true; };
template <> struct LinkIs18Half<LinkB18> { static constexpr bool value = true; };
template <typename LK> struct LinkIsBf16 { static constexpr bool value = false; };
template <> struct LinkIsBf16<LinkB12> { static constexpr bool value = true; };
template <> struct LinkIsBf16<LinkB18> { static constexpr bool value = true; };

template <typename LK>
__device__ __forceinline__ void link_raw_load(LinkRaw<LK>& r, const GaugeRef& g, int64_t idx) {
  const uint32_t i = (uint32_t)idx;
  const int64_t pb = g.plane * 16;
  if constexpr (std::is_same<LK, LinkD18>::value) {
#pragma unroll
    for (int k = 0; k < 9; k++) r.q[k] = ld_off<double2>(g.data, k * pb, i << 4);
  } else if constexpr (std::is_same<LK, LinkF12>::value) {
#pragma unroll
    for (int k = 0; k < 3; k++) r.q[k] = ld_off<float4>(g.data, k * pb, i << 4);
  } else if constexpr (LinkIs18Half<LK>::value) {
    r.a = ld_off<uint4>(g.data, 0, i << 4);
    r.b = ld_off<uint4>(g.data, pb, i << 4);
    r.c = ld_off<uint32_t>(g.data, 2 * pb, i << 2);
  } else {
    r.a = ld_off<uint4>(g.data, 0, i << 4);
    r.b = ld_off<uint2>(g.data, pb, i << 3);
  }
}

template <typename LK, typename R>
__device__ __forceinline__ void link_decode(const LinkRaw<LK>& r, float2 ph, cplx<R> u[9]) {
  if constexpr (std::is_same<LK, LinkD18>::value) {
#pragma unroll
    for (int k = 0; k < 9; k++) u[k] = cmk<R>(R(r.q[k].x), R(r.q[k].y));
    (void)ph;
  } else if constexpr (LinkIs18Half<LK>::value) {
    const uint32_t w[9] = {r.a.x, r.a.y, r.a.z, r.a.w, r.b.x, r.b.y, r.b.z, r.b.w, r.c};
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const float2 h = LinkIsBf16<LK>::value ? bf2f(w[k]) : h2f(w[k]);
      u[k] = cmk<R>(R(h.x), R(h.y));
    }
    (void)ph;
  } else {
    if constexpr (std::is_same<LK, LinkF12>::value) {
#pragma unroll
      for (int k = 0; k < 3; k++) {
        u[2 * k] = cmk<R>(R(r.q[k].x), R(r.q[k].y));
        u[2 * k + 1] = cmk<R>(R(r.q[k].z), R(r.q[k].w));
      }
    } else {
      const uint32_t w[6] = {r.a.x, r.a.y, r.a.z, r.a.w, r.b.x, r.b.y};
#pragma unroll
      for (int k = 0; k < 6; k++) {
        const float2 h = std::is_same<LK, LinkH12>::value ? h2f(w[k]) : bf2f(w[k]);
        u[k] = cmk<R>(R(h.x), R(h.y));
      }
    }
    link_row2<R>(u, ph);
  }
}

// how many hop directions the kernel loads before decoding them (1, 2, 4 or
// 8): more loads in flight per thread against more registers per thread
#ifndef CPSHIP_HOP_BATCH_D
#define CPSHIP_HOP_BATCH_D 2
#endif
#ifndef CPSHIP_HOP_BATCH_F
#define CPSHIP_HOP_BATCH_F 4
#endif
#ifndef CPSHIP_HOP_BATCH_H
#define CPSHIP_HOP_BATCH_H 4
#endif
template <typename S> struct HopBatch;
template <> struct HopBatch<SpinD> { static constexpr int value = CPSHIP_HOP_BATCH_D; };
template <> struct HopBatch<SpinF> { static constexpr int value = CPSHIP_HOP_BATCH_F; };
template <> struct HopBatch<SpinH> { static constexpr int value = CPSHIP_HOP_BATCH_H; };

// ---- clover field -----------------------------------------------------------
// 72 reals per site: two 6x6 Hermitian chiral blocks (block 0 acts on spins
// 0,1, block 1 on spins 2,3; components spin*3+colour), each packed in 36
// reals as CPS's clover_mat_mlt reads them (clover.C:127-172): the real
// diagonal (i,i) at i*i+2i, the lower element (i,j), i>j, at i*i+2j as
// (re,im).  On the device a field of one kappa and one checkerboard is a set
// of planes of 16-byte chunks, one chunk per site per plane: FP64 36 planes
// of double2, FP32 18 planes of float4, FP16 9 planes of 8 halves.  Sets are
// ordered [kappa][cb] (cb 0: A_ee^-1, cb 1: A_oo).
enum CloverStorage { CLOVER_D = 0, CLOVER_F = 1, CLOVER_H = 2, NCLOVER = 3 };
template <typename S> struct CloverOf;
template <> struct CloverOf<SpinD> { static constexpr int storage = CLOVER_D, nplanes = 36; };
template <> struct CloverOf<SpinF> { static constexpr int storage = CLOVER_F, nplanes = 18; };
template <> struct CloverOf<SpinH> { static constexpr int storage = CLOVER_H, nplanes = 9; };
inline int clover_nplanes(int storage) { return storage == CLOVER_D ? 36 : storage == CLOVER_F ? 18 : 9; }

struct CloverRef {
  const void* data;
  int64_t plane;    // chunks per plane (= Vcb)
};

// The 36 reals of chiral block B of the site whose chunk index (set base +
// site) is c0.
template <typename S, typename R, int B>
__device__ __forceinline__ void clover_load_block(const CloverRef& c, int64_t c0, R (&a)[36]) {
  if constexpr (std::is_same<S, SpinD>::value) {
    const double2* p = reinterpret_cast<const double2*>(c.data) + c0 + (int64_t)(18 * B) * c.plane;
#pragma unroll
    for (int q = 0; q < 18; q++) {
      const double2 h = p[q * c.plane];
      a[2 * q] = R(h.x);
      a[2 * q + 1] = R(h.y);
    }
  } else if constexpr (std::is_same<S, SpinF>::value) {
    const float4* p = reinterpret_cast<const float4*>(c.data) + c0 + (int64_t)(9 * B) * c.plane;
#pragma unroll
    for (int q = 0; q < 9; q++) {
      const float4 h = p[q * c.plane];
      a[4 * q] = R(h.x);
      a[4 * q + 1] = R(h.y);
      a[4 * q + 2] = R(h.z);
      a[4 * q + 3] = R(h.w);
    }
  } else {
    // 72 halves in 9 chunks: block 0 is chunks 0-3 and the first half of
    // chunk 4, block 1 the second half of chunk 4 and chunks 5-8
    const uint4* p = reinterpret_cast<const uint4*>(c.data) + c0;
    uint32_t w[18];
    const uint4 mid = p[4 * c.plane];
    if constexpr (B == 0) {
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const uint4 h = p[q * c.plane];
        w[4 * q] = h.x; w[4 * q + 1] = h.y; w[4 * q + 2] = h.z; w[4 * q + 3] = h.w;
      }
      w[16] = mid.x;
      w[17] = mid.y;
    } else {
      w[0] = mid.z;
      w[1] = mid.w;
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const uint4 h = p[(5 + q) * c.plane];
        w[2 + 4 * q] = h.x; w[3 + 4 * q] = h.y; w[4 + 4 * q] = h.z; w[5 + 4 * q] = h.w;
      }
    }
#pragma unroll
    for (int k = 0; k < 18; k++) {
      const float2 f = h2f(w[k]);
      a[2 * k] = R(f.x);
      a[2 * k + 1] = R(f.y);
    }
  }
}

// y = A x for one packed 6x6 Hermitian block (components O..O+5 of x, y).
template <typename R, int O>
__device__ __forceinline__ void clover_block_mul(const R (&a)[36], const cplx<R> (&x)[12], cplx<R> (&y)[12]) {
#pragma unroll
  for (int i = 0; i < 6; i++) {
    const R d = a[i * i + 2 * i];
    y[O + i] = cmk<R>(d * x[O + i].re, d * x[O + i].im);
  }
#pragma unroll
  for (int i = 1; i < 6; i++)
#pragma unroll
    for (int j = 0; j < 5; j++) {
      if (j < i) {
        const R lr = a[i * i + 2 * j], li = a[i * i + 2 * j + 1];
        y[O + i].re += lr * x[O + j].re - li * x[O + j].im;
        y[O + i].im += lr * x[O + j].im + li * x[O + j].re;
        y[O + j].re += lr * x[O + i].re + li * x[O + i].im;
        y[O + j].im += lr * x[O + i].im - li * x[O + i].re;
      }
    }
}

// y = A x for a full spinor (both chiral blocks) at chunk index c0.
template <typename S, typename R>
__device__ __forceinline__ void clover_mul(const CloverRef& c, int64_t c0, const cplx<R> (&x)[12], cplx<R> (&y)[12]) {
  {
    R a[36];
    clover_load_block<S, R, 0>(c, c0, a);
    clover_block_mul<R, 0>(a, x, y);
  }
  {
    R a[36];
    clover_load_block<S, R, 1>(c, c0, a);
    clover_block_mul<R, 6>(a, x, y);
  }
}


}  // namespace cpship
