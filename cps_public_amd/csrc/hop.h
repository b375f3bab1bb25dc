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
  PREC_H_H18 = 4,   // FP16 spinors + norm, FP16 18-real links
  PREC_H_B18 = 5,   // FP16 spinors + norm, BF16 18-real links
};

// launch variants chosen by the tuner
enum HopFlags {
  HOP_NT_STORE = 1,       // streaming (non-temporal) stores of the output
  HOP_NO_XCD_REMAP = 2,   // keep the hardware block order
  HOP_SPATIAL = 4,        // XCD x owns spatial block x of the lattice (all rhs),
                          // swept along t (Wilson kind, needs the XCD remap)
  HOP_SITEWAVE = 8,       // a wavefront computes one site for 64/nb rhs (Wilson
                          // kind, nb divides 64, no partials): the site's
                          // neighbour entries and links are wave-uniform
                          // (scalar loads), only the spinors are per lane
  HOP_LDS_LINKS = 16,     // half-precision spinors, nb >= 8: the block stages
                          // the 8 links of each of its sites in LDS once (f16
                          // words, third row rebuilt), and every thread keeps
                          // all 8 neighbour spinors in flight
  HOP_RHSGROUP = 32,      // Wilson kind, nb a power of two, no partials: a
                          // wavefront computes S sites x G rhs (x nb slots),
                          // G = 2 << ((flags >> 6) & 3), so the G lanes of a
                          // site share its link and clover loads
  HOP_LINK_ROW = 256,     // half-precision spinors, nb = 16, element order:
                          // the 16 lanes of a DPP row (one site's slots) load
                          // one 4-byte word of each link each and exchange
                          // them by row broadcasts
};
inline int hop_rhsgroup_G(int flags) { return (flags & HOP_RHSGROUP) ? (2 << ((flags >> 6) & 3)) : 1; }

struct HopParams {
  int flags;         // HopFlags
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
  // 12-real links: factor (re, im) of the rebuilt third row of a link that
  // crosses the lattice boundary in direction mu (the det of such links, set
  // from the gauge data at import: -1 antiperiodic, e^{3 i theta} for a link
  // twisted by e^{i theta}); interior links use 1
  float bph[4][2];
  double two_over_a5, neg_mass_two_over_a5;
  double scale;      // out = xcoef*X + scale*hop
  double xcoef;
  FieldRef in, out, X;
  GaugeRef gauge;
  double* partial;   // per-block |out|^2 (per-chunk with chunk_part), or nullptr
  // chunk_part (set by the launcher where hop_chunk_partials holds): one
  // partial per 64 consecutive elements of a rhs, summed as a pairwise tree
  // in element order (the same bits in every thread mapping that computes
  // it), at partial[rhs * (elements per rhs / 64) + chunk]
  int chunk_part;
  // clover epilogue (Wilson kind only): 0 none, CLOV_POST out = A (scale*hop),
  // CLOV_X out = xcoef*(A X) + kscale[set]*scale*hop, with A the clover set of
  // kappa index set = rhs / rhs_per_set at the output checkerboard and kscale
  // a device array (required for CLOV_X)
  // CLOV_S5X epilogue (Wilson kind with nb = Ls, the Moebius layout):
  // out = xcoef*X + x5coef*(P X) + scale*(hb5*hop + hc5*(P hop)), P the
  // fifth-dimension shift of the hop's dag with -x5mass at the walls; the
  // P hop term (hc5 != 0) takes the hop of slots s-1, s+1 from the
  // neighbouring lanes, so it needs nb a power of two dividing 64
  double x5coef, x5mass, hb5, hc5;
  // CLOV_S5X with FP16 fields and nb = 16: the slots s-1, s+1 of X and of the
  // hop come from the neighbouring lanes of the site's 16-lane row (DPP
  // rotations) instead of loads and lane permutes
  int s5_row;
  // rupd (with s5_row, or with the CLOV_X epilogue of FP16 hops): out is
  // not stored; instead the CG residual R (FP16)
  // of every rhs with ractive[rhs] becomes R - ralpha[rhs] * out, and the
  // partials are |R|^2 (0 for the other rhs)
  int rupd;
  FieldRef R;
  const double* ralpha;
  const int* ractive;
  const uint4* nbr;         // neighbour table (k_build_nbr)
  const uint32_t* nbr_neg;
  // HOP_SPATIAL: site_cb of logical site b*perm_S + q (the same table for
  // both parities), perm_S = Vcb / 8 sites per spatial block
  const uint32_t* perm;
  int perm_S;
  // launch geometry set by the launcher for the chosen block size: blocks
  // per rhs, blocks per (spatial block, rhs), log2(nb) or -1
  int bpr, bpr_b, nb_shift;
  // HOP_LDS_LINKS: sites per block the LDS area is sized for (set by the
  // launcher)
  int lds_sites;
  // HOP_RHSGROUP: log2 of the rhs per wavefront (set by the launcher)
  int rg_shift;
  // m5post (half-precision Wilson-kind hop with nb = Ls = 16, no epilogue
  // field): the Moebius fifth-dimension stage of the hop's dag applied to
  // out = scale*hop in registers, the 16 slots of a site being one 16-lane
  // row of the wavefront: dag 0 out = b5 M5^-1 h + c5 P M5^-1 h, dag 1
  // out = M5^-dag (b5 h + c5 P^dag h); M5 with ratio K = -m5a and the walls'
  // mass m5mass, m5inv_d = 1 / (1 + m5mass m5a^16)
  int m5post;
  double m5a, m5inv_d, m5mass, m5b, m5c;
  int clov;
  int rhs_per_set;
  int rhs0;          // clover set of local rhs r: (rhs0 + r) / rhs_per_set
  CloverRef clover;
  // CLOV_X: kscale[set] as above; CLOV_KS (a mass per rhs, Wilson kind):
  // the hop's scale of local rhs r is scale * kscale[rhs0 + r]
  const double* kscale;
  // CLOV_KS, DWF kind: the wall coefficient (-2 mf / a5) of local rhs r is
  // kwall[rhs0 + r] instead of neg_mass_two_over_a5
  const double* kwall;
};

// CLOV_S5XR: CLOV_S5X with the row exchange of s5_row; CLOV_S5XRR: the same
// with the CG residual update (rupd) instead of storing out; CLOV_XRR:
// CLOV_X with that residual update (kernels of their own, chosen by the
// launcher; not values of HopParams::clov)
// CLOV_KS: no clover term, a mass per rhs (kscale for the Wilson kind, kwall
// for the DWF kind), element order only; a value of HopParams::clov
// CLOV_NRR: no clover term, out = xcoef X + scale hop as for CLOV_NONE, and
// the CG residual update (rupd) instead of storing out (half-precision DWF
// hops; the kernel the launcher picks for CLOV_NONE with rupd, not a value
// of HopParams::clov)
enum { CLOV_NONE = 0, CLOV_POST = 1, CLOV_X = 2, CLOV_S5X = 3, CLOV_S5XR = 4, CLOV_S5XRR = 5, CLOV_XRR = 6,
       CLOV_KS = 7, CLOV_NRR = 8 };

hipError_t launch_hop(const HopParams& P, hipStream_t st);
hipError_t build_nbr(const Geom& g, uint4* nbr, uint32_t* nbr_neg, hipStream_t st);
hipError_t launch_hop_clover(const HopParams& P, hipStream_t st);

// link storage read by a hop precision
inline int link_storage_of(int hprec) {
  switch (hprec) {
    case PREC_D_D18: return 0;   // LINK_D18
    case PREC_F_F12: return 1;   // LINK_F12
    case PREC_H_H12: return 2;   // LINK_H12
    case PREC_H_B12: return 3;   // LINK_B12
    case PREC_H_H18: return 4;   // LINK_H18
    default: return 5;           // LINK_B18
  }
}

// HOP_LDS_LINKS layout: words 0-7 of the link of (block site sl, direction d)
// at sl*LDS_SITE_WORDS + 8d (16-byte aligned; consecutive sites start 4
// banks apart, so the few sites one wavefront reads do not conflict), word 8
// at lds_sites*LDS_SITE_WORDS + 8 sl + d.
constexpr int LDS_SITE_WORDS = 68;

// sites of a block of bs consecutive elements, nb elements per site (at most)
inline int lds_sites_per_block(int bs, int nb) { return bs % nb == 0 ? bs / nb : (bs - 1) / nb + 2; }
inline size_t lds_link_bytes(int sites) { return (size_t)sites * (LDS_SITE_WORDS + 8) * 4; }

// Shapes whose |out|^2 partials are per 64-element chunk rather than per
// block: Wilson kind with one element per site (the 4-D Wilson and clover
// operators) and a whole number of chunks per rhs
inline bool chunk_partials_shape(int kind, int nb, long long per_rhs) {
  return kind == KIND_WILSON && nb == 1 && per_rhs % 64 == 0;
}
inline bool hop_chunk_partials(const HopParams& P) {
  return P.partial && chunk_partials_shape(P.kind, P.nb, (long long)P.nsites_out * P.nb);
}

// HOP_RHSGROUP: whether the launch shape allows it at block size bs; sets
// the rhs-per-wavefront shift.  With partials only in the chunk form, and
// with G <= the wavefronts of a block, so that a chunk's G wavefronts (S =
// 64 / G sites each) lie in one block
inline bool rhsgroup_setup(HopParams& Q, int bs) {
  const int G = hop_rhsgroup_G(Q.flags);
  if (Q.partial && !(hop_chunk_partials(Q) && bs % 64 == 0 && G <= bs / 64)) return false;
  if (Q.kind != KIND_WILSON || (Q.flags & (HOP_SITEWAVE | HOP_LDS_LINKS | HOP_SPATIAL))) return false;
  if (Q.nb < 1 || (Q.nb & (Q.nb - 1)) || G * Q.nb > 64 || Q.nrhs % G || bs < 64 || bs % 64) return false;
  const int S = 64 / (G * Q.nb);
  if (Q.nsites_out % ((bs / 64) * S)) return false;
  Q.rg_shift = __builtin_ctz((unsigned)G);
  return true;
}

// HOP_LINK_ROW: half-precision spinors, Wilson kind with nb = 16 (a DPP row
// is one site's slots), element order (spatial allowed), blocks of whole
// rows, and gauge planes addressable with 32-bit offsets
inline bool link_row_setup(const HopParams& Q, int bs) {
  const bool half = Q.prec == PREC_H_H12 || Q.prec == PREC_H_B12 || Q.prec == PREC_H_H18 || Q.prec == PREC_H_B18;
  if (!half || Q.kind != KIND_WILSON || Q.nb != 16 || bs % 16) return false;
  if (Q.flags & (HOP_SITEWAVE | HOP_LDS_LINKS | HOP_RHSGROUP)) return false;
  return Q.gauge.plane * 16 * 3 < (int64_t(1) << 32);
}

// block size dividing n (a power of two, at most 256)
inline int pick_block(long long n) {
  int b = 256;
  while (b > 1 && (n % b) != 0) b >>= 1;
  return b;
}

}  // namespace cpship
