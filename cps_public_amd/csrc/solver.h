// Host-side objects behind the C ABI.
#pragma once
#include <map>
#include <vector>

#include "../../include/cps_hip.h"
#include "blas.h"
#include "cps_hip_internal.h"
#include "hop.h"
#include "layout.h"

namespace cpship {

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  ~DevBuf() { release(); }
  int ensure(size_t b);
  void release();
};

struct Field {
  int storage = -1;
  int64_t nelem = 0;
  DevBuf buf, nrm;
  int alloc(int st, int64_t n);
  FieldRef ref() const;
};

FieldRef null_field();

enum Slot {
  S_X = 0, S_SRC, S_R, S_P, S_AP, S_T1, S_T2, S_XS, S_RS, S_PS, S_APS, S_T1S, S_T2S, S_T3, S_T3S,
  S_IO_IN, S_IO_OUT, S_EVEN_IN, S_EVEN_OUT, S_TMP, S_BENCH_IN, S_BENCH_OUT, S_PS2, S_NSLOTS
};

struct OpLayout {
  int kind, map, nb, nsites;
  int64_t per_rhs;   // elements of one rhs
};

struct HopTuneKey {
  int kind, prec, dag, nb, nsites, nrhs, partial, hasX, clov;
  int L[4];   // geometry: the cache is shared by every handle of the process
  int mode;   // CPS_HIP_HOP_MODE restriction in force when tuned
  bool operator<(const HopTuneKey& o) const {
    const int a[14] = {kind, prec, dag, nb, nsites, nrhs, partial, hasX, clov, L[0], L[1], L[2], L[3], mode};
    const int b[14] = {o.kind, o.prec, o.dag, o.nb, o.nsites, o.nrhs, o.partial, o.hasX, o.clov,
                       o.L[0], o.L[1], o.L[2], o.L[3], o.mode};
    for (int k = 0; k < 14; k++)
      if (a[k] != b[k]) return a[k] < b[k];
    return false;
  }
};
struct HopTuneVal {
  int block, flags;
  double ms;
};

struct Lattice {
  Geom g;
  int ls = 1;
  hipStream_t stream = nullptr;
  DevBuf gauge_aos;
  DevBuf gauge[NLINK];
  bool gauge_valid[NLINK];
  bool has_gauge = false;
  long long gauge_generation = 0;
  // Determinant of the resident links, measured at import (link_det_check):
  // link_phase[mu] is det U of the links on the last mu slice (1 periodic,
  // -1 antiperiodic, e^{3 i theta} twisted by e^{i theta}); recon_ok says
  // every interior link has det 1 and every boundary link of direction mu
  // det link_phase[mu] (to 1e-5), the condition under which a link is
  // rebuilt from its first two rows (12-real storage).
  double link_phase[4][2];
  bool recon_ok = false;
  DevBuf det_scratch;
  DevBuf gauge_stage;   // a re-import lands here and is compared with gauge_aos
  int measure_link_phase();
  int set_hop_phase(HopParams& P, int hprec) const;
  std::vector<Field> scratch;
  DevBuf partial, scal, io, io2;
  DevBuf nbr, nbr_neg;   // neighbour table of the geometry (hop kernels)
  // spatial site order of the HOP_SPATIAL launches (built on first use):
  // 0 not built, 1 built, -1 the geometry has no 8-way (y, z) split
  DevBuf perm;
  int perm_state = 0;
  int ensure_perm();
  // clover sets [kappa][cb] (cb 0: A_ee^-1, cb 1: A_oo) per storage format;
  // FP64 is built from the gauge, FP32/FP16 are converted from it on demand
  DevBuf clover[NCLOVER];
  bool clover_valid[NCLOVER];
  DevBuf clover_ksq;   // kappa^2 per set (device)
  int n_clover = 0;
  // active-rhs window of a running CG: its operator calls see local rhs
  // 0 .. n-1 = global rhs win_lo .. win_lo+n-1 of win_total (0: no window)
  int win_lo = 0, win_total = 0;
  std::vector<double> clover_kappas;
  double clover_csw = 0;
  int last_rc = 0;

  Lattice();
  ~Lattice();
  int init(const int node_sites[4], int ls, const int bc[4]);
  int set_twist(const double angle[4]);   // twisted directions' phase (CANONICAL import)
  int import_gauge(const double* u_host);
  int import_gauge_device(const double* u_dev);   // same, from device memory
  int import_gauge_from(const double* u, hipMemcpyKind kind);
  int import_gauge_canonical(const double* u_canon);   // CANONICAL order + BondCond (canonical.hip)
  int ensure_gauge(int link_storage);
  GaugeRef gauge_ref(int link_storage) const;
  Field& field(int slot, int storage, int64_t nelem);
  // as field(), but an existing buffer of the same storage holding at least
  // nelem elements is kept (its plane stride is then its own size)
  Field& field_min(int slot, int storage, int64_t nelem);
  int ensure_partial(int64_t slots);
  long long device_bytes() const;
};

// One CG iteration's launch sequence replayed as a hipGraph.  The first
// iteration runs eagerly (the launch autotuner times its candidates there,
// which cannot happen inside a capture); the same sequence is then captured
// once and every later iteration is one graph launch.  Only kernel launches
// are captured (no memset nodes).  All kernel arguments are device pointers
// or values fixed while the operands stay the same, so one capture serves
// every iteration until reset() (the CG's active-rhs window shrinking).
// The process-wide controls (cps_hip_tuning): the field's value, or the
// environment variable CPS_HIP_<NAME> when that is set (read at every call,
// so a variable set between solves takes effect).
cps_hip_tuning tuning();
int knob(int cps_hip_tuning::*field, const char* env);

struct IterGraph {
  bool enabled;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // use_graph = 0 launches every iteration eagerly (e.g. under a profiler
  // that traces kernels but not graph replays)
  explicit IterGraph(bool on) : enabled(on && knob(&cps_hip_tuning::use_graph, "CPS_HIP_GRAPH") != 0) {}
  IterGraph(const IterGraph&) = delete;
  IterGraph& operator=(const IterGraph&) = delete;
  ~IterGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
  template <class F>
  int run(Lattice& L, int n, F&& body) {
    int rc = 0;
    if (n <= 0) return 0;
    if (!enabled) {
      for (int i = 0; i < n; i++)
        if ((rc = body())) return rc;
      return 0;
    }
    if (!exec) {
      if ((rc = body())) return rc;
      n--;
      CPSHIP_CHECK(hipStreamBeginCapture(L.stream, hipStreamCaptureModeThreadLocal));
      rc = body();
      hipGraph_t g = nullptr;
      const hipError_t ec = hipStreamEndCapture(L.stream, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (ec != hipSuccess) return CPS_HIP_ERR_RUNTIME;
      graph = g;
      CPSHIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
    for (int i = 0; i < n; i++) CPSHIP_CHECK(hipGraphLaunch(exec, L.stream));
    return 0;
  }
  // the launch sequence changed (new operands): capture again on the next run
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
};

struct CGState {
  int nrhs = 0;
  CGScalars s;
  double* rr_true = nullptr;    // FP64 |r|^2 per rhs (true residual evaluations)
  double* upd_store = nullptr;  // backing store of s.rr_upd
  double* alpha_prev_store = nullptr;  // backing store of s.alpha_prev
  int* mask = nullptr;          // per-rhs flags for masked BLAS launches
  int alloc(Lattice& L, int nrhs, int hist_len);
};

// rhs lo onwards of a batched field / of the per-rhs CG scalars (the same
// planes with the data, norm and scalar pointers advanced)
FieldRef rhs_view(const FieldRef& f, int64_t per_rhs, int lo);
CGScalars scalars_view(const CGScalars& s, int lo);

// After a reliable update of the mixed CG (rr_true = FP64 |r|^2 per rhs, rp
// = <r, p>): rhs meeting the stopping rule finish, the others continue from
// the FP64 residual with p re-orthogonalised by coef (p += coef r).
hipError_t mixed_after_update(const CGScalars& cg, const double* rr_true, const double* rp, double* coef, int nrhs,
                              hipStream_t st);

int op_layout(const cps_hip_op* op, int ls, OpLayout* o, const Geom& g);
double op_kappa(const cps_hip_op* op);
int hop(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& out, const FieldRef& in, int target, int dag,
        double scale, const FieldRef& X, double xcoef, double* partial, int nrhs);
// clover epilogue of a Wilson hop (HopParams::clov)
struct HopClover {
  int mode;
  CloverRef ref;
  const double* kscale;
  int rhs_per_set;
  int rhs0;   // global index of local rhs 0 (the CG's active window)
  // optional (CLOV_X, FP16): instead of storing out, the CG residual step
  // r -= ralpha[rhs] * out for the rhs with ractive[rhs], partials |r|^2
  const FieldRef* R = nullptr;
  const double* ralpha = nullptr;
  const int* ractive = nullptr;
  // mode CLOV_NONE: per-rhs masses (HopParams::kscale with rhs_per_set 1,
  // HopParams::kwall), device arrays indexed by rhs0 + local rhs
  const double* kwall = nullptr;
};
int hop_ex(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& out, const FieldRef& in, int target, int dag,
           double scale, const FieldRef& X, double xcoef, double* partial, int nrhs, const HopClover* hc);
// M and M^dag M of the Wilson or DWF kind with a mass per rhs (device
// arrays k2 = kappa^2, Wilson, or wall = -2 mf / a5, DWF, indexed by rhs0 +
// local rhs)
int apply_m_slots(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& out, const FieldRef& in,
                  const FieldRef& t, int dag, double* partial, int nrhs, const double* k2, const double* wall,
                  int rhs0);
int apply_mdagm_slots(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& out, const FieldRef& in,
                      const FieldRef& t1, const FieldRef& t2, double* partial, int nrhs, const double* k2,
                      const double* wall, int rhs0);
int hop_blocks_per_rhs(Lattice& L, const cps_hip_op* op);
int tuned_launch(Lattice& L, HopParams& P);
// While a NoAllocScope is alive on a thread, the hop launcher of that thread
// allocates nothing and times nothing (it builds no tables and uses a cached
// or the default launch choice): the no-allocation entry point
// cps_hip_matpc_dag_matpc_noalloc runs under it.
extern thread_local int tl_no_alloc;
struct NoAllocScope {
  NoAllocScope() { tl_no_alloc++; }
  ~NoAllocScope() { tl_no_alloc--; }
  NoAllocScope(const NoAllocScope&) = delete;
  NoAllocScope& operator=(const NoAllocScope&) = delete;
};
void precision_formats(int prec, const cps_hip_op* op, int ls, int nrhs, bool recon_ok, int* sp_store, int* hprec);
int apply_m(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& out, const FieldRef& in, const FieldRef& t,
            int dag, double* partial, int nrhs);
// t1_ready (Moebius): t1 already holds b in + c P in, the first step of M
int apply_mdagm(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& out, const FieldRef& in,
                const FieldRef& t1, const FieldRef& t2, double* partial, int nrhs, bool t1_ready = false);
int apply_matpc(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& out, const FieldRef& in,
                const FieldRef& t1, const FieldRef& t2, int dag, int nrhs);

// Moebius (mobius.hip) and clover (clover.hip)
int mobius_m(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& out, const FieldRef& in,
             const FieldRef& t1, const FieldRef& t2, int dag, double* partial, int nrhs, bool t1_ready = false);
int mobius_mdagm(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& out, const FieldRef& in,
                 const FieldRef& t1, const FieldRef& t2, double* partial, int nrhs, bool t1_ready = false);
int mobius_mdag_rupdate(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& r, const FieldRef& t3,
                        const FieldRef& t1, const FieldRef& t2, const double* alpha, const int* active,
                        double* partial, int nrhs);
int clover_matpc(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& out, const FieldRef& in,
                 const FieldRef& t1, int dag, double* partial, int nrhs);
int clover_mdagm(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& out, const FieldRef& in,
                 const FieldRef& t1, const FieldRef& t2, double* partial, int nrhs);
int clover_mdag_rupdate(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& r, const FieldRef& t2,
                        const FieldRef& t1, const double* alpha, const int* active, double* partial, int nrhs);
// DWF M^dag of t3 with the CG residual step in the epilogue of its second
// hop: r -= alpha[rhs] (M^dag t3) for the rhs with active[rhs], partial =
// |r|^2 (FP16 r and hops; M^dag M p is never stored)
int dwf_mdag_rupdate(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& r, const FieldRef& t3,
                     const FieldRef& t1, const double* alpha, const int* active, double* partial, int nrhs);
int clover_build(Lattice& L, const double* kappas, int n_kappa, double csw);
int clover_export(Lattice& L, int k, int cb, double* A);
// MatInv source preparation (src = MatPcDag of the prepared odd source) and
// even-half reconstruction for the Moebius and clover operators
int mobius_matinv_prep(Lattice& L, const cps_hip_op* op, const FieldRef& src, const FieldRef& in_o,
                       const FieldRef& in_e, const FieldRef& t1, const FieldRef& t2, const FieldRef& t3, int nrhs);
int mobius_matinv_reconstruct(Lattice& L, const cps_hip_op* op, const FieldRef& out_e, const FieldRef& out_o,
                              const FieldRef& in_e, const FieldRef& t1, const FieldRef& t2, const FieldRef& t3,
                              int nrhs);
int clover_matinv_prep(Lattice& L, const cps_hip_op* op, const FieldRef& src, const FieldRef& in_o,
                       const FieldRef& in_e, const FieldRef& t1, const FieldRef& t2, const FieldRef& t3, int nrhs);
int clover_matinv_reconstruct(Lattice& L, const cps_hip_op* op, const FieldRef& out_e, const FieldRef& out_o,
                              const FieldRef& in_e, const FieldRef& t1, const FieldRef& t2, const FieldRef& t3,
                              int nrhs);

// Where a 4-D site's spin-colour component lives in the device fields of
// an operator: the checkerboard (odd or even field) and the element index
// within one rhs.  For the 5-D operators s is the slice; DWF fields hold
// slice s = 2j + ((p + cb5) & 1) at element (p Vcb + site_cb) Ls/2 + j,
// Moebius fields Ls slots per site of one 4-D parity.
template <int KIND>
__host__ __device__ __forceinline__ void field_slot(int p, int site_cb, int s, int Vcb, int ls, bool* odd,
                                                    int64_t* e) {
  if constexpr (KIND == CPS_HIP_OP_DWF) {
    *odd = ((p + s) & 1) != 0;
    *e = ((int64_t)p * Vcb + site_cb) * (ls / 2) + (s >> 1);
  } else if constexpr (KIND == CPS_HIP_OP_MOBIUS) {
    *odd = p != 0;
    *e = (int64_t)site_cb * ls + s;
  } else {
    *odd = p != 0;
    *e = site_cb;
  }
}

// MatInv on device fields: in_o / in_e the source checkerboards (odd =
// the CG's checkerboard), x the initial guess and on return the odd
// solution, out_e the even solution; Moebius results are 2 kappa_b times the
// DiracOpMobius::MatInv ones (the caller scales)
int matinv_device(Lattice& L, const cps_hip_op* op, Field& in_o, Field& in_e, Field& x, Field& out_e, int nrhs,
                  const cps_hip_cg_arg* arg, cps_hip_cg_result* res);

// Moebius D_minus applied in place to the source checkerboards (FmatInv)
int mobius_dminus_fields(Lattice& L, const cps_hip_op* op, Field& f_odd, Field& f_even, int nrhs);
// CANONICAL <-> device fields (canonical.hip): rhs-major canonical vectors,
// [rhs][s][t][z][y][x][spin][colour][re,im]; scale multiplies the values
int pack_canonical(Lattice& L, int kind, const double* canon, Field& f_odd, Field& f_even, int nrhs, double scale);
int unpack_canonical(Lattice& L, int kind, const Field& f_odd, const Field& f_even, double* canon, int nrhs,
                     double scale);

// host <-> device field transfer with layout conversion; with on_device the
// double pointer is device memory in the CPS layout (no PCIe copy)
int upload(Lattice& L, const double* host, Field& f, const OpLayout& o, int cb5, int nrhs, int slot_io,
           bool on_device = false);
int download(Lattice& L, const Field& f, double* host, const OpLayout& o, int cb5, int nrhs, bool on_device = false);

int cg_double_device(Lattice& L, const cps_hip_op* op, Field& x, Field& src, int nrhs, const cps_hip_cg_arg* arg,
                     std::vector<int>& iters, std::vector<int>& status, std::vector<double>& true_res,
                     std::vector<double>& src_norm);
int cg_mixed_device(Lattice& L, const cps_hip_op* op, Field& x, Field& src, int nrhs, const cps_hip_cg_arg* arg,
                    std::vector<int>& iters, std::vector<int>& status, std::vector<double>& true_res,
                    std::vector<double>& src_norm, std::vector<int>& n_updates);

}  // namespace cpship

struct cps_hip_lattice {
  cpship::Lattice L;
};
