// Launchers of the Wilson hopping kernel with a fused epilogue: the clover
// multiply (out = A (scale hop) or out = xcoef A X + scale hop), the two
// halves of the clover preconditioned operator A_oo - kappa^2 D A_ee^-1 D
// (DiracOpClover::MatPcDagOrNot, d_op_clover.C:194-248); or the Moebius
// fifth-dimension term (out = X + r P X + scale hop, the last step of
// zmobius_m_orig, zmobius_m-orig.h:3-122).
#include "hop_kernel.h"

namespace cpship {

template <typename SIN, typename SOUT, typename LK, typename R>
static hipError_t launch_c(const HopParams& P, hipStream_t st) {
  if (P.kind != KIND_WILSON) return hipErrorInvalidValue;
  if ((P.clov == CLOV_POST || P.clov == CLOV_X) && (P.rhs_per_set < 1 || !P.clover.data)) return hipErrorInvalidValue;
  if (P.clov == CLOV_X && !P.kscale) return hipErrorInvalidValue;
  if (P.clov == CLOV_S5X && (!P.X.data || P.nb < 2)) return hipErrorInvalidValue;
  // the fused P hop term reads slots s-1, s+1 from lanes of one wavefront
  if (P.clov == CLOV_S5X && P.hc5 != 0.0 && ((P.nb & (P.nb - 1)) || 64 % P.nb || P.block % P.nb))
    return hipErrorInvalidValue;
  const int64_t nelem = (int64_t)P.nrhs * P.nsites_out * P.nb;
  const int bs = P.block;
  const int64_t nblk = (nelem + bs - 1) / bs;
  dim3 grid((unsigned)nblk), block(bs);
  // 32-bit block decomposition (hop_decode): per-rhs counts below 2^31
  const int64_t per_rhs = (int64_t)P.nsites_out * P.nb;
  if (per_rhs % bs || per_rhs >= (int64_t(1) << 31) || nblk >= (int64_t(1) << 31)) return hipErrorInvalidValue;
  HopParams Q = P;
  Q.bpr = (int)(per_rhs / bs);
  Q.bpr_b = (int)(((int64_t)P.perm_S * P.nb) / bs);
  Q.nb_shift = (P.nb & (P.nb - 1)) == 0 ? __builtin_ctz((unsigned)P.nb) : -1;
  // partials per 64-site chunk: wavefront-aligned blocks, and a mapping that
  // keeps a chunk's sites in one block (element order or the rhs groups)
  Q.chunk_part = hop_chunk_partials(P) ? 1 : 0;
  if (Q.chunk_part && (bs % 64 || (P.flags & (HOP_SPATIAL | HOP_SITEWAVE)))) return hipErrorInvalidValue;
  const bool nt = (P.flags & HOP_NT_STORE) != 0;
  // HOP_LDS_LINKS: half-precision spinors, one link per thread to stage
  const bool lds = (P.flags & HOP_LDS_LINKS) != 0;
  size_t shm = 0;
  if (lds) {
    Q.lds_sites = lds_sites_per_block(bs, P.nb);
    if (!std::is_same<SIN, SpinH>::value || (P.flags & HOP_SITEWAVE) || 8 * Q.lds_sites > bs)
      return hipErrorInvalidValue;
    shm = lds_link_bytes(Q.lds_sites);
  }
  if ((P.flags & HOP_RHSGROUP) && !rhsgroup_setup(Q, bs)) return hipErrorInvalidValue;
  if ((P.flags & HOP_LINK_ROW) && !link_row_setup(P, bs)) return hipErrorInvalidValue;
  if (P.m5post) return hipErrorInvalidValue;   // plain hops only
  if (P.s5_row && (P.clov != CLOV_S5X || P.nb != 16 || bs % 16)) return hipErrorInvalidValue;
  if (P.rupd && (!P.R.data || !P.ralpha || !P.ractive || (P.clov == CLOV_S5X ? !P.s5_row : P.clov != CLOV_X)))
    return hipErrorInvalidValue;
#define CPSHIP_HOP_LAUNCH2(DAG, CLOV, SW)                                                                        \
  if (nt) hipLaunchKernelGGL((k_hop<SIN, SOUT, LK, R, DAG, KIND_WILSON, true, CLOV, SW>), grid, block, shm, st, Q); \
  else hipLaunchKernelGGL((k_hop<SIN, SOUT, LK, R, DAG, KIND_WILSON, false, CLOV, SW>), grid, block, shm, st, Q);
#define CPSHIP_HOP_LAUNCH(DAG, CLOV)                                                              \
  if (P.flags & HOP_SITEWAVE) { CPSHIP_HOP_LAUNCH2(DAG, CLOV, 1) }                                \
  else if (P.flags & HOP_RHSGROUP) { CPSHIP_HOP_LAUNCH2(DAG, CLOV, 3) }                           \
  else if (P.flags & HOP_LINK_ROW) {                                                              \
    if constexpr (std::is_same<SIN, SpinH>::value && CLOV >= CLOV_S5X && CLOV <= CLOV_S5XRR) {    \
      CPSHIP_HOP_LAUNCH2(DAG, CLOV, 4)                                                            \
    }                                                                                             \
    else return hipErrorInvalidValue;                                                             \
  }                                                                                               \
  else if (lds) {                                                                                 \
    if constexpr (std::is_same<SIN, SpinH>::value && CLOV >= CLOV_S5X && CLOV <= CLOV_S5XRR) {                              \
      CPSHIP_HOP_LAUNCH2(DAG, CLOV, 2)                                                            \
    }                                                                                             \
    else return hipErrorInvalidValue;                                                             \
  } else { CPSHIP_HOP_LAUNCH2(DAG, CLOV, 0) }
  if (P.clov == CLOV_POST) {
    if (P.dag) { CPSHIP_HOP_LAUNCH(1, CLOV_POST) } else { CPSHIP_HOP_LAUNCH(0, CLOV_POST) }
  } else if (P.clov == CLOV_X && P.rupd) {
    if constexpr (std::is_same<SIN, SpinH>::value) {
      if (P.dag) { CPSHIP_HOP_LAUNCH(1, CLOV_XRR) } else { CPSHIP_HOP_LAUNCH(0, CLOV_XRR) }
    } else {
      return hipErrorInvalidValue;
    }
  } else if (P.clov == CLOV_X) {
    if (P.dag) { CPSHIP_HOP_LAUNCH(1, CLOV_X) } else { CPSHIP_HOP_LAUNCH(0, CLOV_X) }
  } else if (P.clov == CLOV_S5X && P.s5_row && P.rupd) {
    if constexpr (std::is_same<SIN, SpinH>::value) {
      if (P.dag) { CPSHIP_HOP_LAUNCH(1, CLOV_S5XRR) } else { CPSHIP_HOP_LAUNCH(0, CLOV_S5XRR) }
    } else {
      return hipErrorInvalidValue;
    }
  } else if (P.clov == CLOV_S5X && P.s5_row) {
    if constexpr (std::is_same<SIN, SpinH>::value) {
      if (P.dag) { CPSHIP_HOP_LAUNCH(1, CLOV_S5XR) } else { CPSHIP_HOP_LAUNCH(0, CLOV_S5XR) }
    } else {
      return hipErrorInvalidValue;
    }
  } else if (P.clov == CLOV_S5X) {
    if (P.dag) { CPSHIP_HOP_LAUNCH(1, CLOV_S5X) } else { CPSHIP_HOP_LAUNCH(0, CLOV_S5X) }
  } else {
    return hipErrorInvalidValue;
  }
#undef CPSHIP_HOP_LAUNCH
#undef CPSHIP_HOP_LAUNCH2
  return hipGetLastError();
}

hipError_t launch_hop_clover(const HopParams& P, hipStream_t st) {
  switch (P.prec) {
    case PREC_D_D18: return launch_c<SpinD, SpinD, LinkD18, double>(P, st);
    case PREC_F_F12: return launch_c<SpinF, SpinF, LinkF12, float>(P, st);
    case PREC_H_H12: return launch_c<SpinH, SpinH, LinkH12, float>(P, st);
    case PREC_H_B12: return launch_c<SpinH, SpinH, LinkB12, float>(P, st);
    case PREC_H_H18: return launch_c<SpinH, SpinH, LinkH18, float>(P, st);
    case PREC_H_B18: return launch_c<SpinH, SpinH, LinkB18, float>(P, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace cpship
