// Launchers of the Wilson / domain-wall hopping kernels (no clover term).
#include "hop_kernel.h"

namespace cpship {

template <typename SIN, typename SOUT, typename LK, typename R>
static hipError_t launch_t(const HopParams& P, hipStream_t st) {
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
  const bool row = (P.flags & HOP_LINK_ROW) != 0;
  if (row && !link_row_setup(P, bs)) return hipErrorInvalidValue;
  // the fused fifth-dimension stage needs the 16 slots of a site in one row
  if (P.m5post && (!std::is_same<SIN, SpinH>::value || P.kind != KIND_WILSON || P.nb != 16 || P.X.data || bs % 16))
    return hipErrorInvalidValue;
  if (P.rupd) {
    // the CG residual update instead of the store (CLOV_NRR): half-precision
    // DWF hops with X, element order, no variants
    if (P.clov != CLOV_NONE || P.kind != KIND_DWF || !P.X.data || !P.R.data || !P.ralpha || !P.ractive || P.m5post ||
        (P.flags & (HOP_SPATIAL | HOP_SITEWAVE | HOP_LDS_LINKS | HOP_RHSGROUP | HOP_LINK_ROW)))
      return hipErrorInvalidValue;
    if constexpr (std::is_same<SIN, SpinH>::value) {
      if (P.dag) {
        hipLaunchKernelGGL((k_hop<SIN, SOUT, LK, R, 1, KIND_DWF, false, CLOV_NRR, 0>), grid, block, 0, st, Q);
      } else {
        hipLaunchKernelGGL((k_hop<SIN, SOUT, LK, R, 0, KIND_DWF, false, CLOV_NRR, 0>), grid, block, 0, st, Q);
      }
      return hipGetLastError();
    } else {
      return hipErrorInvalidValue;
    }
  }
  if (P.clov == CLOV_KS) {
    // a mass per rhs: element order, no variants
    if (P.flags & (HOP_SPATIAL | HOP_SITEWAVE | HOP_LDS_LINKS | HOP_RHSGROUP | HOP_LINK_ROW) || P.m5post)
      return hipErrorInvalidValue;
    if (P.kind == KIND_DWF ? !P.kwall : !P.kscale) return hipErrorInvalidValue;
#define CPSHIP_HOP_KS(DAG, KIND) hipLaunchKernelGGL((k_hop<SIN, SOUT, LK, R, DAG, KIND, false, CLOV_KS, 0>), grid, block, 0, st, Q);
    if (P.kind == KIND_DWF) {
      if (P.dag) { CPSHIP_HOP_KS(1, KIND_DWF) } else { CPSHIP_HOP_KS(0, KIND_DWF) }
    } else {
      if (P.dag) { CPSHIP_HOP_KS(1, KIND_WILSON) } else { CPSHIP_HOP_KS(0, KIND_WILSON) }
    }
#undef CPSHIP_HOP_KS
    return hipGetLastError();
  }
#define CPSHIP_HOP_LAUNCH(DAG, KIND, SW)                                                                      \
  if (nt) hipLaunchKernelGGL((k_hop<SIN, SOUT, LK, R, DAG, KIND, true, CLOV_NONE, SW>), grid, block, shm, st, Q); \
  else hipLaunchKernelGGL((k_hop<SIN, SOUT, LK, R, DAG, KIND, false, CLOV_NONE, SW>), grid, block, shm, st, Q);
  if (row) {
    if constexpr (std::is_same<SIN, SpinH>::value) {
      if (P.dag) { CPSHIP_HOP_LAUNCH(1, KIND_WILSON, 4) } else { CPSHIP_HOP_LAUNCH(0, KIND_WILSON, 4) }
    } else {
      return hipErrorInvalidValue;
    }
  } else if (lds) {
    if constexpr (std::is_same<SIN, SpinH>::value) {
      if (P.kind == KIND_DWF) {
        if (P.dag) { CPSHIP_HOP_LAUNCH(1, KIND_DWF, 2) } else { CPSHIP_HOP_LAUNCH(0, KIND_DWF, 2) }
      } else {
        if (P.dag) { CPSHIP_HOP_LAUNCH(1, KIND_WILSON, 2) } else { CPSHIP_HOP_LAUNCH(0, KIND_WILSON, 2) }
      }
    }
  } else if (P.kind == KIND_DWF) {
    if (P.flags & HOP_SITEWAVE) return hipErrorInvalidValue;
    if (P.dag) { CPSHIP_HOP_LAUNCH(1, KIND_DWF, 0) } else { CPSHIP_HOP_LAUNCH(0, KIND_DWF, 0) }
  } else if (P.flags & HOP_SITEWAVE) {
    if (P.dag) { CPSHIP_HOP_LAUNCH(1, KIND_WILSON, 1) } else { CPSHIP_HOP_LAUNCH(0, KIND_WILSON, 1) }
  } else if (P.flags & HOP_RHSGROUP) {
    if (P.dag) { CPSHIP_HOP_LAUNCH(1, KIND_WILSON, 3) } else { CPSHIP_HOP_LAUNCH(0, KIND_WILSON, 3) }
  } else {
    if (P.dag) { CPSHIP_HOP_LAUNCH(1, KIND_WILSON, 0) } else { CPSHIP_HOP_LAUNCH(0, KIND_WILSON, 0) }
  }
#undef CPSHIP_HOP_LAUNCH
  return hipGetLastError();
}

hipError_t launch_hop(const HopParams& P, hipStream_t st) {
  if (P.clov != CLOV_NONE && P.clov != CLOV_KS) return launch_hop_clover(P, st);
  switch (P.prec) {
    case PREC_D_D18: return launch_t<SpinD, SpinD, LinkD18, double>(P, st);
    case PREC_F_F12: return launch_t<SpinF, SpinF, LinkF12, float>(P, st);
    case PREC_H_H12: return launch_t<SpinH, SpinH, LinkH12, float>(P, st);
    case PREC_H_B12: return launch_t<SpinH, SpinH, LinkB12, float>(P, st);
    case PREC_H_H18: return launch_t<SpinH, SpinH, LinkH18, float>(P, st);
    case PREC_H_B18: return launch_t<SpinH, SpinH, LinkB18, float>(P, st);
    default: return hipErrorInvalidValue;
  }
}

// Neighbour table: for every (parity p, site_cb s) at ns = p*Vcb + s, the
// checkerboard indices of the 8 neighbours (d = 0..3: +mu, 4..7: -mu) in
// nbr[ns] (d 0-3) and nbr[2 Vcb + ns] (d 4-7), and as bits of nbr_neg[ns]
// the directions whose link lies on the last slice of its direction (the
// links that carry the boundary factor; Nbr::neg).
__global__ void k_build_nbr(Geom g, uint4* __restrict__ nbr, uint32_t* __restrict__ nbr_neg) {
  const int64_t ns = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ns >= 2 * (int64_t)g.Vcb) return;
  const int p = (int)(ns / g.Vcb);
  const int s = (int)(ns - (int64_t)p * g.Vcb);
  const Nbr N = neighbours(g, site_coords(g, s, p));
  nbr[ns] = make_uint4(N.idx[0], N.idx[1], N.idx[2], N.idx[3]);
  nbr[2 * (int64_t)g.Vcb + ns] = make_uint4(N.idx[4], N.idx[5], N.idx[6], N.idx[7]);
  nbr_neg[ns] = N.neg;
}

hipError_t build_nbr(const Geom& g, uint4* nbr, uint32_t* nbr_neg, hipStream_t st) {
  const int64_t n = 2 * (int64_t)g.Vcb;
  hipLaunchKernelGGL(k_build_nbr, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, nbr, nbr_neg);
  return hipGetLastError();
}

}  // namespace cpship
