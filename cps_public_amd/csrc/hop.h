// Launch parameters of the hopping kernels (host and device visible).
#pragma once
#include "cps_hip_internal.h"

namespace cpship {

enum HopKind { KIND_WILSON = 0, KIND_DWF = 1 };

// spinor storage / link storage / compute precision combinations
enum HopPrec {
  PREC_D_D18 = 0,   // FP64 spinors, FP64 18-real links, FP64 math
  PREC_F_F12 = 1,   // FP32 spinors, FP32 12-real links, FP32 math
  PREC_H_H12 = 2,   // FP16 spinors + norm, FP16 12-real links, FP32 math
  PREC_H_B12 = 3,   // FP16 spinors + norm, BF16 12-real links, FP32 math
};

struct HopParams {
  Geom geom;
  int kind;          // HopKind
  int prec;          // HopPrec
  int dag;
  int nb;            // inner batch (j) per site
  int nrhs;
  int nsites_out;    // Vcb (Wilson) or 2*Vcb (DWF)
  int out_parity;    // Wilson: parity of output sites
  int out_cb5;       // DWF: 5-D checkerboard of the output field
  int ls;
  int block;         // threads per block (divides nsites_out*nb)
  int link_bc_sign;  // 12-real links: apply the antiperiodic sign on reconstruction
  double two_over_a5, neg_mass_two_over_a5;
  double scale;      // out = xcoef*X + scale*hop
  double xcoef;
  FieldRef in, out, X;
  GaugeRef gauge;
  double* partial;   // per-block |out|^2, or nullptr
};

hipError_t launch_hop(const HopParams& P, hipStream_t st);

// block size dividing n (a power of two, at most 256)
inline int pick_block(long long n) {
  int b = 256;
  while (b > 1 && (n % b) != 0) b >>= 1;
  return b;
}

}  // namespace cpship
