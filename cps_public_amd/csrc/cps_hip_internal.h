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
    const float inv = m > 0.f ? 1.0f / m : 0.f;
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

// Load link (parity, mu, site_cb) as a full 3x3 matrix u[row*3+col].
// sign: boundary sign for the reconstructed third row of a 12-real link.
template <typename LK, typename R>
__device__ __forceinline__ void link_load(const GaugeRef& g, int parity, int mu, int site, int Vcb,
                                          R sign, cplx<R> u[9]) {
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
    // row2 = sign * conj(row0 x row1)
    u[6] = cmk<R>(sign * (u[1].re * u[5].re - u[1].im * u[5].im - u[2].re * u[4].re + u[2].im * u[4].im),
                  -sign * (u[1].re * u[5].im + u[1].im * u[5].re - u[2].re * u[4].im - u[2].im * u[4].re));
    u[7] = cmk<R>(sign * (u[2].re * u[3].re - u[2].im * u[3].im - u[0].re * u[5].re + u[0].im * u[5].im),
                  -sign * (u[2].re * u[3].im + u[2].im * u[3].re - u[0].re * u[5].im - u[0].im * u[5].re));
    u[8] = cmk<R>(sign * (u[0].re * u[4].re - u[0].im * u[4].im - u[1].re * u[3].re + u[1].im * u[3].im),
                  -sign * (u[0].re * u[4].im + u[0].im * u[4].re - u[1].re * u[3].im - u[1].im * u[3].re));
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
__device__ __forceinline__ void link_decode(const LinkRaw<LK>& r, R sign, cplx<R> u[9]) {
  if constexpr (std::is_same<LK, LinkD18>::value) {
#pragma unroll
    for (int k = 0; k < 9; k++) u[k] = cmk<R>(R(r.q[k].x), R(r.q[k].y));
    (void)sign;
  } else if constexpr (LinkIs18Half<LK>::value) {
    const uint32_t w[9] = {r.a.x, r.a.y, r.a.z, r.a.w, r.b.x, r.b.y, r.b.z, r.b.w, r.c};
#pragma unroll
    for (int k = 0; k < 9; k++) {
      const float2 h = LinkIsBf16<LK>::value ? bf2f(w[k]) : h2f(w[k]);
      u[k] = cmk<R>(R(h.x), R(h.y));
    }
    (void)sign;
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
    u[6] = cmk<R>(sign * (u[1].re * u[5].re - u[1].im * u[5].im - u[2].re * u[4].re + u[2].im * u[4].im),
                  -sign * (u[1].re * u[5].im + u[1].im * u[5].re - u[2].re * u[4].im - u[2].im * u[4].re));
    u[7] = cmk<R>(sign * (u[2].re * u[3].re - u[2].im * u[3].im - u[0].re * u[5].re + u[0].im * u[5].im),
                  -sign * (u[2].re * u[3].im + u[2].im * u[3].re - u[0].re * u[5].im - u[0].im * u[5].re));
    u[8] = cmk<R>(sign * (u[0].re * u[4].re - u[0].im * u[4].im - u[1].re * u[3].re + u[1].im * u[3].im),
                  -sign * (u[0].re * u[4].im + u[0].im * u[4].re - u[1].re * u[3].im - u[1].im * u[3].re));
  }
}

// how many hop directions the kernel loads before decoding (register budget)
template <typename S> struct HopBatch;
template <> struct HopBatch<SpinD> { static constexpr int value = 2; };
template <> struct HopBatch<SpinF> { static constexpr int value = 4; };
template <> struct HopBatch<SpinH> { static constexpr int value = 8; };

}  // namespace cpship
