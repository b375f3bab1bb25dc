// Layout transfer and precision conversion (see layout.hip).
#pragma once
#include "cps_hip_internal.h"

namespace cpship {

enum SpinStorage { STORE_D = 0, STORE_F = 1, STORE_H = 2 };
enum LinkStorage { LINK_D18 = 0, LINK_F12 = 1, LINK_H12 = 2, LINK_B12 = 3, LINK_H18 = 4, LINK_B18 = 5, NLINK = 6 };
enum SpinMap { MAP_4D = 0, MAP_DWF5 = 1, MAP_4DPREC = 2 };

hipError_t pack_spinor(int storage, const double* aos, const FieldRef& out, int map, int Vcb, int ls, int cb5,
                       int64_t n, hipStream_t st);
hipError_t unpack_spinor(int storage, const FieldRef& in, double* aos, int map, int Vcb, int ls, int cb5,
                         int64_t n, hipStream_t st);
hipError_t convert_spinor(int sin, const FieldRef& in, int sout, const FieldRef& out, int64_t n, hipStream_t st);
hipError_t pack_gauge(int link_storage, const double* aos, const GaugeRef& out, int Vcb, hipStream_t st);
// Determinants of the WILSON-order FP64 links aos: ref[mu] (2 doubles each)
// receives det U of the mu-link at x_mu = L_mu - 1, other coordinates 0;
// dev[2 mu + b] (as float bits, device memory zeroed by the caller) the
// largest |det U - e| over interior (b = 0, e = 1) and last-slice (b = 1,
// e = ref[mu]) links of direction mu.
// *flag (device) = 1 if the n 64-bit words of a and b differ anywhere, else 0
hipError_t words_differ(const uint64_t* a, const uint64_t* b, int64_t n, unsigned* flag, hipStream_t st);
hipError_t link_det_check(const double* aos, const Geom& g, double* ref, unsigned* dev, hipStream_t st);

inline size_t spin_elem_bytes(int storage) {
  return storage == STORE_D ? 16 : storage == STORE_F ? 8 : 4;
}
// bytes of one stored link (18 FP64 reals, or 12 reals of FP32 / FP16 / BF16)
inline size_t link_bytes(int ls_) {
  return ls_ == LINK_D18 ? 144 : ls_ == LINK_F12 ? 48 : (ls_ == LINK_H18 || ls_ == LINK_B18) ? 36 : 24;
}

}  // namespace cpship

namespace cpship {
// deterministic pseudo-random fill in [-1,1) (benchmark inputs)
hipError_t fill_spinor(int storage, const FieldRef& f, int64_t n, unsigned seed, hipStream_t st);
}
