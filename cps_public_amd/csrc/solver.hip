// Host driver: lattice handle, device fields, operator application and the
// batched conjugate-gradient solvers behind the C ABI in include/cps_hip.h.
#include "solver.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

namespace cpship {

// ---------------------------------------------------------------------------
int DevBuf::ensure(size_t b) {
  if (b <= bytes) return 0;
  release();
  if (b == 0) return 0;
  if (hipMalloc(&p, b) != hipSuccess) {
    p = nullptr;
    bytes = 0;
    (void)hipGetLastError();
    return CPS_HIP_ERR_NOMEM;
  }
  bytes = b;
  return 0;
}
void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

int Field::alloc(int st, int64_t n) {
  storage = st;
  nelem = n;
  int rc = buf.ensure((size_t)n * 12 * spin_elem_bytes(st));
  if (rc) return rc;
  if (st == STORE_H) return nrm.ensure((size_t)n * sizeof(float));
  return 0;
}

FieldRef Field::ref() const {
  FieldRef r;
  r.data = buf.p;
  r.norm = storage == STORE_H ? (float*)nrm.p : nullptr;
  r.plane = nelem;
  return r;
}

FieldRef null_field() {
  FieldRef r;
  r.data = nullptr;
  r.norm = nullptr;
  r.plane = 0;
  return r;
}

// ---------------------------------------------------------------------------
Lattice::Lattice() {
  for (int k = 0; k < NLINK; k++) gauge_valid[k] = false;
  for (int k = 0; k < NCLOVER; k++) clover_valid[k] = false;
  // every slot exists up front: references handed out by field() stay valid
  scratch.resize(S_NSLOTS);
}

Lattice::~Lattice() {
  if (stream) (void)hipStreamDestroy(stream);
}

int Lattice::init(const int node_sites[4], int ls_, const int bc[4]) {
  for (int mu = 0; mu < 4; mu++) {
    if (node_sites[mu] <= 0 || (node_sites[mu] & 1)) return CPS_HIP_ERR_ARG;
    g.L[mu] = node_sites[mu];
    g.bc[mu] = bc ? bc[mu] : 0;
    // BND_CND_PRD, BND_CND_APRD, BND_CND_TWISTED (enum.x:271-277); the
    // G-parity kinds need the second flavour's field and are not supported
    if (g.bc[mu] < 0 || g.bc[mu] > 2) return CPS_HIP_ERR_ARG;
    g.bf[mu][0] = g.bc[mu] == 1 ? -1.0 : 1.0;   // a twist starts at angle 0
    g.bf[mu][1] = 0.0;
  }
  if (ls_ < 1) return CPS_HIP_ERR_ARG;
  ls = ls_;
  g.Vcb = g.L[0] * g.L[1] * g.L[2] * g.L[3] / 2;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return CPS_HIP_ERR_NODEV;
  }
  if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) return CPS_HIP_ERR_RUNTIME;
  int rc = nbr.ensure((size_t)4 * g.Vcb * sizeof(uint4));
  if (rc) return rc;
  if ((rc = nbr_neg.ensure((size_t)2 * g.Vcb * sizeof(uint32_t)))) return rc;
  CPSHIP_CHECK(build_nbr(g, (uint4*)nbr.p, (uint32_t*)nbr_neg.p, stream));
  CPSHIP_CHECK(hipStreamSynchronize(stream));
  return 0;
}

// GJP::TwistPhase (gjp.C:475-493): e^{-i pi theta}, theta in units of pi
int Lattice::set_twist(const double angle[4]) {
  for (int mu = 0; mu < 4; mu++) {
    if (g.bc[mu] != 2) continue;
    const double th = angle ? angle[mu] : 0.0;
    g.bf[mu][0] = std::cos(M_PI * th);
    g.bf[mu][1] = -std::sin(M_PI * th);
  }
  return 0;
}

long long Lattice::device_bytes() const {
  long long b = (long long)gauge_aos.bytes + partial.bytes + scal.bytes + io.bytes + io2.bytes + nbr.bytes +
                nbr_neg.bytes + det_scratch.bytes + gauge_stage.bytes;
  for (int k = 0; k < NLINK; k++) b += gauge[k].bytes;
  for (auto& f : scratch) b += f.buf.bytes + f.nrm.bytes;
  for (int k = 0; k < NCLOVER; k++) b += clover[k].bytes;
  return b;
}

int Lattice::import_gauge(const double* u_host) { return import_gauge_from(u_host, hipMemcpyHostToDevice); }

int Lattice::import_gauge_device(const double* u_dev) { return import_gauge_from(u_dev, hipMemcpyDeviceToDevice); }

// A field equal (bit for bit) to the resident one keeps every copy derived
// from it (low-precision links, clover sets): callers that pass the gauge
// with every operator call, as CPS's library functions take it, pay one
// transfer and a device comparison instead of a repack, and a clover field
// built earlier stays valid.
int Lattice::import_gauge_from(const double* u, hipMemcpyKind kind) {
  const size_t n = (size_t)72 * 2 * g.Vcb;
  int rc;
  if (has_gauge && gauge_aos.bytes >= n * sizeof(double)) {
    if ((rc = gauge_stage.ensure(n * sizeof(double)))) return rc;
    if ((rc = det_scratch.ensure(8 * sizeof(double) + 9 * sizeof(unsigned)))) return rc;
    unsigned* diff = (unsigned*)((double*)det_scratch.p + 8) + 8;
    CPSHIP_CHECK(hipMemcpyAsync(gauge_stage.p, u, n * sizeof(double), kind, stream));
    CPSHIP_CHECK(words_differ((const uint64_t*)gauge_stage.p, (const uint64_t*)gauge_aos.p, (int64_t)n, diff, stream));
    unsigned h_diff = 1;
    CPSHIP_CHECK(hipMemcpyAsync(&h_diff, diff, sizeof(unsigned), hipMemcpyDeviceToHost, stream));
    CPSHIP_CHECK(hipStreamSynchronize(stream));
    if (!h_diff) return 0;
    std::swap(gauge_stage.p, gauge_aos.p);
    std::swap(gauge_stage.bytes, gauge_aos.bytes);
  } else {
    if ((rc = gauge_aos.ensure(n * sizeof(double)))) return rc;
    CPSHIP_CHECK(hipMemcpyAsync(gauge_aos.p, u, n * sizeof(double), kind, stream));
  }
  for (int k = 0; k < NLINK; k++) gauge_valid[k] = false;
  // a clover field built from the previous gauge field is stale
  n_clover = 0;
  for (int k = 0; k < NCLOVER; k++) clover_valid[k] = false;
  rc = ensure_gauge(LINK_D18);
  if (rc) return rc;
  if ((rc = measure_link_phase())) return rc;   // synchronises the stream
  has_gauge = true;
  gauge_generation++;
  return 0;
}

// Boundary factor of the resident links, from the data itself: the FP64
// links carry whatever BondCond applied (CPS's DiracOp constructor, or the
// CANONICAL import), so the 12-real formats never depend on the handle's bc.
int Lattice::measure_link_phase() {
  int rc = det_scratch.ensure(8 * sizeof(double) + 9 * sizeof(unsigned));
  if (rc) return rc;
  double* ref = (double*)det_scratch.p;
  unsigned* dev = (unsigned*)(ref + 8);
  CPSHIP_CHECK(link_det_check((const double*)gauge_aos.p, g, ref, dev, stream));
  double h_ref[8];
  unsigned h_dev[8];
  CPSHIP_CHECK(hipMemcpyAsync(h_ref, ref, sizeof(h_ref), hipMemcpyDeviceToHost, stream));
  CPSHIP_CHECK(hipMemcpyAsync(h_dev, dev, sizeof(h_dev), hipMemcpyDeviceToHost, stream));
  CPSHIP_CHECK(hipStreamSynchronize(stream));
  const double tol = 1e-5;
  recon_ok = true;
  for (int mu = 0; mu < 4; mu++) {
    float e0, e1;
    std::memcpy(&e0, &h_dev[2 * mu], 4);
    std::memcpy(&e1, &h_dev[2 * mu + 1], 4);
    double pr = h_ref[2 * mu], pi = h_ref[2 * mu + 1];
    const double m = std::sqrt(pr * pr + pi * pi);
    if (!(e0 <= tol && e1 <= tol && std::fabs(m - 1.0) <= tol)) recon_ok = false;
    // the exact value for the periodic and antiperiodic cases, else the
    // measured phase on the unit circle
    if (std::fabs(pr - 1.0) <= tol && std::fabs(pi) <= tol) pr = 1.0, pi = 0.0;
    else if (std::fabs(pr + 1.0) <= tol && std::fabs(pi) <= tol) pr = -1.0, pi = 0.0;
    else if (m > 0) pr /= m, pi /= m;
    link_phase[mu][0] = pr;
    link_phase[mu][1] = pi;
  }
  return 0;
}

// The boundary factors of a hop reading 12-real links; such links need the
// resident field to be SU(3) times the boundary factor (recon_ok).
int Lattice::set_hop_phase(HopParams& P, int hprec) const {
  const bool twelve = hprec == PREC_F_F12 || hprec == PREC_H_H12 || hprec == PREC_H_B12;
  if (twelve && !recon_ok) return CPS_HIP_ERR_ARG;
  for (int mu = 0; mu < 4; mu++) {
    P.bph[mu][0] = (float)link_phase[mu][0];
    P.bph[mu][1] = (float)link_phase[mu][1];
  }
  return 0;
}

int Lattice::ensure_gauge(int link_storage) {
  if (gauge_valid[link_storage]) return 0;
  const int64_t plane = (int64_t)8 * g.Vcb;
  int rc = gauge[link_storage].ensure((size_t)plane * link_bytes(link_storage));
  if (rc) return rc;
  GaugeRef gr;
  gr.data = gauge[link_storage].p;
  gr.plane = plane;
  CPSHIP_CHECK(pack_gauge(link_storage, (const double*)gauge_aos.p, gr, g.Vcb, stream));
  gauge_valid[link_storage] = true;
  return 0;
}

GaugeRef Lattice::gauge_ref(int link_storage) const {
  GaugeRef gr;
  gr.data = gauge[link_storage].p;
  gr.plane = (int64_t)8 * g.Vcb;
  return gr;
}

// The (y, z) plane is split into 8 blocks (by x bz = 8, the split with the
// smallest halo), one per XCD; inside a block the order is t, z, y, x, so an
// XCD sweeps its block along t and the t-neighbour reuse stays within the
// window of a few block slices its L2 holds.
int Lattice::ensure_perm() {
  if (perm_state) return perm_state > 0 ? 0 : CPS_HIP_ERR_ARG;
  int best_by = 0, best_bz = 0;
  double best = 1e30;
  for (int by = 1; by <= 8; by *= 2) {
    const int bz = 8 / by;
    if (g.L[1] % by || g.L[2] % bz) continue;
    const double ny = g.L[1] / by, nz = g.L[2] / bz;
    const double halo = (ny + 2) * (nz + 2) / (ny * nz);
    if (halo < best) {
      best = halo;
      best_by = by;
      best_bz = bz;
    }
  }
  if (!best_by) {
    perm_state = -1;
    return CPS_HIP_ERR_ARG;
  }
  const int lh = g.L[0] / 2, ny = g.L[1] / best_by, nz = g.L[2] / best_bz;
  std::vector<uint32_t> h((size_t)g.Vcb);
  size_t q = 0;
  for (int bz = 0; bz < best_bz; bz++)
    for (int by = 0; by < best_by; by++)
      for (int t = 0; t < g.L[3]; t++)
        for (int z = bz * nz; z < (bz + 1) * nz; z++)
          for (int y = by * ny; y < (by + 1) * ny; y++)
            for (int xh = 0; xh < lh; xh++) h[q++] = (uint32_t)(xh + lh * (y + g.L[1] * (z + g.L[2] * t)));
  int rc = perm.ensure(h.size() * sizeof(uint32_t));
  if (rc) return rc;
  CPSHIP_CHECK(hipMemcpy(perm.p, h.data(), h.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
  perm_state = 1;
  return 0;
}

Field& Lattice::field(int slot, int storage, int64_t nelem) {
  Field& f = scratch[slot];   // slot < S_NSLOTS by construction
  if (f.storage != storage || f.nelem != nelem || !f.buf.p) {
    if (f.storage != storage) {
      f.buf.release();
      f.nrm.release();
    }
    last_rc = f.alloc(storage, nelem);
  }
  return f;
}

Field& Lattice::field_min(int slot, int storage, int64_t nelem) {
  Field& f = scratch[slot];
  if (f.buf.p && f.storage == storage && f.nelem >= nelem) return f;
  return field(slot, storage, nelem);
}

int Lattice::ensure_partial(int64_t slots) {
  return partial.ensure((size_t)std::max<int64_t>(slots, 1) * sizeof(double));
}

// ---------------------------------------------------------------------------
// Operator description and application

int op_layout(const cps_hip_op* op, int ls, OpLayout* o, const Geom& g) {
  if (!op) return CPS_HIP_ERR_ARG;
  o->kind = op->kind;
  switch (op->kind) {
    case CPS_HIP_OP_WILSON:
    case CPS_HIP_OP_CLOVER:
      o->map = MAP_4D;
      o->nb = 1;
      o->nsites = g.Vcb;
      o->per_rhs = g.Vcb;
      break;
    case CPS_HIP_OP_DWF:
      if (ls < 2 || (ls & 1)) return CPS_HIP_ERR_ARG;
      o->map = MAP_DWF5;
      o->nb = ls / 2;
      o->nsites = 2 * g.Vcb;
      o->per_rhs = (int64_t)g.Vcb * ls;
      break;
    case CPS_HIP_OP_MOBIUS:
      if (ls < 2) return CPS_HIP_ERR_ARG;
      o->map = MAP_4DPREC;
      o->nb = ls;
      o->nsites = g.Vcb;
      o->per_rhs = (int64_t)g.Vcb * ls;
      break;
    default:
      return CPS_HIP_ERR_ARG;
  }
  return 0;
}

double op_kappa(const cps_hip_op* op) {
  if (op->kind == CPS_HIP_OP_DWF) return 1.0 / (2 * (4 + op->a5_inv - op->m5));  // dwf_init.C:97-98
  return 1.0 / (2.0 * (op->mass + 4.0));                                         // d_op_wilson.C:136
}

// out = xcoef*X + scale*D in, D the 4-D (Wilson) or 5-D (DWF) hop whose
// output lives on `target` (Wilson: 4-D parity; DWF: 5-D checkerboard).
int hop(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& out, const FieldRef& in, int target, int dag,
        double scale, const FieldRef& X, double xcoef, double* partial, int nrhs) {
  return hop_ex(L, op, hprec, out, in, target, dag, scale, X, xcoef, partial, nrhs, nullptr);
}

int hop_ex(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& out, const FieldRef& in, int target, int dag,
           double scale, const FieldRef& X, double xcoef, double* partial, int nrhs, const HopClover* hc) {
  OpLayout o;
  int rc = op_layout(op, L.ls, &o, L.g);
  if (rc) return rc;
  HopParams P;
  std::memset(&P, 0, sizeof(P));
  P.geom = L.g;
  P.kind = (op->kind == CPS_HIP_OP_DWF) ? KIND_DWF : KIND_WILSON;
  P.prec = hprec;
  P.dag = dag;
  P.nb = o.nb;
  P.nrhs = nrhs;
  P.nsites_out = o.nsites;
  P.out_parity = target;
  P.out_cb5 = target;
  P.ls = L.ls;
  P.block = pick_block((long long)o.nsites * o.nb);
  if ((rc = L.set_hop_phase(P, hprec))) return rc;
  P.two_over_a5 = 2.0 * op->a5_inv;
  P.neg_mass_two_over_a5 = -2.0 * op->mass * op->a5_inv;
  P.scale = scale;
  P.xcoef = xcoef;
  P.in = in;
  P.out = out;
  P.X = X;
  int lstore = link_storage_of(hprec);
  rc = L.ensure_gauge(lstore);
  if (rc) return rc;
  P.gauge = L.gauge_ref(lstore);
  P.partial = partial;
  if (hc) {
    P.clov = hc->mode;
    P.clover = hc->ref;
    P.kscale = hc->kscale;
    P.rhs_per_set = hc->rhs_per_set;
    P.rhs0 = hc->rhs0;
    P.kwall = hc->kwall;
    if (hc->R) {
      P.rupd = 1;
      P.R = *hc->R;
      P.ralpha = hc->ralpha;
      P.ractive = hc->ractive;
    }
  }
  rc = tuned_launch(L, P);
  return rc;
}

// ---------------------------------------------------------------------------
// Autotuner: the first launch of each hop shape (kind, precision, dag, batch,
// volume, rhs count, epilogue) times every candidate block size and launch
// variant on the real operands and keeps the fastest for the lattice
// handle's lifetime.  Launches that write per-block reduction partials keep
// the default block size, because the reduction reads one partial per block;
// per-chunk partials (4-D operators) are summed in an order that does not
// depend on the block or mapping.  The candidates compute every element with
// the same source arithmetic: FP64 results are the same bits whichever wins;
// the half-precision kernels of different mappings are separate compilations
// and may round a few FP32 results differently (FMA contraction).
// CPS_HIP_TUNE=0 switches tuning off (default block, no variants).

static std::mutex tuning_mu;
static cps_hip_tuning tuning_fields = {1, 1, 0, 0, 1, 1, 1, 1, 1, 0, 1, 1, 1, 0, 0, 0.0};

static int env_int(const char* name, int v) {
  const char* e = getenv(name);
  return (e && *e) ? atoi(e) : v;
}

cps_hip_tuning tuning() {
  cps_hip_tuning t;
  {
    std::lock_guard<std::mutex> lk(tuning_mu);
    t = tuning_fields;
  }
  t.use_graph = env_int("CPS_HIP_GRAPH", t.use_graph);
  t.tune = env_int("CPS_HIP_TUNE", t.tune);
  t.tune_report = env_int("CPS_HIP_TUNE_REPORT", t.tune_report);
  if (const char* m = getenv("CPS_HIP_HOP_MODE"))
    t.hop_mode = !strcmp(m, "sitewave") ? 2 : !strcmp(m, "element") ? 1 : !strcmp(m, "lds") ? 3
               : !strcmp(m, "rhsgroup") ? 4 : !strcmp(m, "row") ? 5 : 0;
  t.rhs_window = env_int("CPS_HIP_RHS_WINDOW", t.rhs_window);
  t.lagged_x = env_int("CPS_HIP_LAGGED_X", t.lagged_x);
  t.paired_x = env_int("CPS_HIP_PAIRED_X", t.paired_x);
  t.s5_xpay = env_int("CPS_HIP_S5_XPAY", t.s5_xpay);
  t.r_in_hop = env_int("CPS_HIP_R_IN_HOP", t.r_in_hop);
  t.clover_r_in_hop = env_int("CPS_HIP_CLOVER_R_IN_HOP", t.clover_r_in_hop);
  t.m5_fused = env_int("CPS_HIP_M5_FUSED", t.m5_fused);
  t.s5_fused = env_int("CPS_HIP_S5_FUSED", t.s5_fused);
  t.s5x_row = env_int("CPS_HIP_S5X_ROW", t.s5x_row);
  if (const char* e = getenv("CPS_HIP_SLOPPY_VECTORS"))
    t.sloppy_vectors = !strcmp(e, "single") ? 1 : !strcmp(e, "xhalf") ? 2 : !strcmp(e, "half") ? 3 : 0;
  t.verbose = env_int("CPS_HIP_VERBOSE", t.verbose);
  if (const char* e = getenv("CPS_HIP_PROGRESS")) t.progress_seconds = atof(e);
  return t;
}

// one integer control: the field, or its environment override
int knob(int cps_hip_tuning::*field, const char* env) {
  int v;
  {
    std::lock_guard<std::mutex> lk(tuning_mu);
    v = tuning_fields.*field;
  }
  return env_int(env, v);
}

}  // namespace cpship

extern "C" int cps_hip_get_tuning(cps_hip_tuning* t) {
  if (!t) return CPS_HIP_ERR_ARG;
  std::lock_guard<std::mutex> lk(cpship::tuning_mu);
  *t = cpship::tuning_fields;
  return 0;
}

extern "C" int cps_hip_set_tuning(const cps_hip_tuning* t) {
  if (!t || t->hop_mode < 0 || t->hop_mode > 5 || t->sloppy_vectors < 0 || t->sloppy_vectors > 3)
    return CPS_HIP_ERR_ARG;
  std::lock_guard<std::mutex> lk(cpship::tuning_mu);
  cpship::tuning_fields = *t;
  return 0;
}

namespace cpship {

static bool tuning_enabled() { return knob(&cps_hip_tuning::tune, "CPS_HIP_TUNE") != 0; }

// whether a launch may use HOP_SPATIAL at block size b
static bool spatial_ok(Lattice& L, const HopParams& P, int b) {
  if (P.kind != KIND_WILSON || L.g.Vcb % 8) return false;
  const long long per_b = (long long)(L.g.Vcb / 8) * P.nb;
  if (per_b % b) return false;
  return L.ensure_perm() == 0;
}

// whether a launch may use HOP_SITEWAVE at block size b (one site per
// wavefront for 64/nb rhs; launches with partials keep the element order)
static bool sitewave_ok(Lattice& L, const HopParams& P, int b, bool spatial) {
  if (P.kind != KIND_WILSON || P.partial || b < 64 || 64 % P.nb) return false;
  const int W = b / 64, g = 64 / P.nb;
  if (P.nrhs % g || P.nsites_out % W) return false;
  if (spatial && (L.g.Vcb / 8) % W) return false;
  return true;
}

// whether a launch may use HOP_RHSGROUP with these flags at block size b
static bool rhsgroup_ok(const HopParams& P, int b, int flags) {
  HopParams Q = P;
  Q.flags = flags;
  return rhsgroup_setup(Q, b);
}

// whether a launch may use HOP_LDS_LINKS at block size b (half-precision
// spinors, at most one link per thread to stage)
static bool lds_links_ok(const HopParams& P, int b) {
  const bool half = P.prec == PREC_H_H12 || P.prec == PREC_H_B12 || P.prec == PREC_H_H18 || P.prec == PREC_H_B18;
  if (!half || (P.clov != CLOV_NONE && P.clov != CLOV_S5X)) return false;
  return 8 * lds_sites_per_block(b, P.nb) <= b;
}

// CPS_HIP_TUNE_REPORT=1 prints each tuning decision to stderr, =2 also the
// time of every candidate
static int tune_report() { return knob(&cps_hip_tuning::tune_report, "CPS_HIP_TUNE_REPORT"); }

// the tuner's winners are kept per process (every handle, engine and thread
// of the same geometry reuses them); a launch shape is timed once
static std::mutex tune_mu;
static std::map<HopTuneKey, HopTuneVal> tune_cache;

thread_local int tl_no_alloc = 0;

int tuned_launch(Lattice& L, HopParams& P) {
  P.nbr = (const uint4*)L.nbr.p;
  P.nbr_neg = (const uint32_t*)L.nbr_neg.p;
  P.perm = nullptr;
  P.perm_S = L.g.Vcb / 8;
  const long long per_rhs = (long long)P.nsites_out * P.nb;
  const int def_block = pick_block(per_rhs);
  if (tl_no_alloc) {
    // no allocation and no timing (cps_hip_matpc_dag_matpc_noalloc): the
    // choice cached for this launch shape if it needs no table the handle
    // lacks, else the default launch (results do not depend on the choice)
    if (P.kind == KIND_WILSON && L.perm_state > 0) P.perm = (const uint32_t*)L.perm.p;
    HopTuneVal chosen{def_block, 0, 0.0};
    if (P.clov != CLOV_KS) {
      const HopTuneKey key{P.kind, P.prec, P.dag, P.nb, P.nsites_out, P.nrhs, P.partial != nullptr,
                           P.X.data != nullptr, P.clov + 16 * P.m5post + 32 * P.rupd,
                           {L.g.L[0], L.g.L[1], L.g.L[2], L.g.L[3]}, 0};
      std::lock_guard<std::mutex> lk(tune_mu);
      auto f = tune_cache.find(key);
      if (f != tune_cache.end()) {
        const HopTuneVal c = f->second;
        const bool ok = (!(c.flags & HOP_SPATIAL) || P.perm) &&
                        (!(c.flags & HOP_SITEWAVE) || sitewave_ok(L, P, c.block, (c.flags & HOP_SPATIAL) != 0)) &&
                        (!(c.flags & HOP_RHSGROUP) || rhsgroup_ok(P, c.block, c.flags));
        if (ok) chosen = c;
      }
    }
    P.block = chosen.block;
    P.flags = chosen.flags;
    CPSHIP_CHECK(launch_hop(P, L.stream));
    return 0;
  }
  if (P.kind == KIND_WILSON && L.perm_state >= 0 && L.g.Vcb % 8 == 0 && L.ensure_perm() == 0)
    P.perm = (const uint32_t*)L.perm.p;
  if (P.clov == CLOV_KS) {
    // a mass per rhs: one kernel in element order (no variants to time)
    P.block = def_block;
    P.flags = 0;
    CPSHIP_CHECK(launch_hop(P, L.stream));
    return 0;
  }
  // hop_mode: the field, or CPS_HIP_HOP_MODE=element|sitewave|lds|rhsgroup|row
  int mode;
  {
    std::lock_guard<std::mutex> lk(tuning_mu);
    mode = tuning_fields.hop_mode;
  }
  if (const char* m = getenv("CPS_HIP_HOP_MODE"))
    mode = !strcmp(m, "sitewave") ? 2 : !strcmp(m, "element") ? 1 : !strcmp(m, "lds") ? 3 : !strcmp(m, "rhsgroup") ? 4
         : !strcmp(m, "row") ? 5 : 0;
  HopTuneKey key{P.kind, P.prec, P.dag, P.nb, P.nsites_out, P.nrhs, P.partial != nullptr, P.X.data != nullptr,
                 P.clov + 16 * P.m5post + 32 * P.rupd, {L.g.L[0], L.g.L[1], L.g.L[2], L.g.L[3]}, mode};
  HopTuneVal chosen;
  bool have = false;
  {
    std::lock_guard<std::mutex> lk(tune_mu);
    auto f = tune_cache.find(key);
    if (f != tune_cache.end()) {
      chosen = f->second;
      have = true;
    } else {
      // the same shape tuned for another rhs count (a CG's active-rhs window
      // shrinking): reuse that choice where it is valid for this count (for
      // launches with per-block partials it is the fixed-order choice, which
      // does not depend on the count; chunk partials are summed in the same
      // order in every mapping, so a rhs-group choice that the count no
      // longer allows falls back to element order)
      for (const auto& kv : tune_cache) {
        HopTuneKey k2 = kv.first;
        k2.nrhs = key.nrhs;
        if (k2 < key || key < k2) continue;
        chosen = kv.second;
        if ((chosen.flags & HOP_SITEWAVE) && !sitewave_ok(L, P, chosen.block, (chosen.flags & HOP_SPATIAL) != 0))
          chosen.flags &= ~HOP_SITEWAVE;
        if ((chosen.flags & HOP_RHSGROUP) && !rhsgroup_ok(P, chosen.block, chosen.flags))
          chosen.flags &= ~(HOP_RHSGROUP | (3 << 6));
        have = true;
        tune_cache.emplace(key, chosen);
        break;
      }
    }
  }
  if (!have) {
    // launches that write per-block reduction partials use a fixed order
    // (results do not depend on tuning): spatial for the 5-D layouts (nb >= 8
    // slots per site) whenever the geometry allows it, element order
    // otherwise (with one site per lane the spatial sweep is the slower
    // order: 6.2 against 4.6 ms for the 32^3x64 clover hop with 48 rhs); with
    // per-chunk partials the element order is the default and the rhs-group
    // mapping a candidate
    HopTuneVal best{def_block, (P.nb >= 8 && spatial_ok(L, P, def_block)) ? HOP_SPATIAL : 0, 0.0};
    // a launch that updates a field in place (rupd) is never timed on the
    // live data: it keeps the fixed order of launches with partials
    if (tuning_enabled() && !P.rupd) {
      std::vector<int> blocks;
      if (P.partial && !hop_chunk_partials(P)) blocks.push_back(def_block);
      else
        for (int b : {64, 128, 256})
          if (b <= def_block && per_rhs % b == 0) blocks.push_back(b);
      hipEvent_t e0, e1;
      CPSHIP_CHECK(hipEventCreate(&e0));
      CPSHIP_CHECK(hipEventCreate(&e1));
      float best_ms = 1e30f;
      // candidate flags: every combination of the first five variants, then
      // the rhs-group mapping for G = 2..16 with and without streaming
      // stores, then the row-shared links with the order variants (tried
      // with CPS_HIP_HOP_MODE=row only)
      std::vector<int> cand;
      for (int f = 0; f < 32; f++) cand.push_back(f);
      for (int g = 0; g < 4; g++)
        for (int nt = 0; nt < 2; nt++) cand.push_back(HOP_RHSGROUP | (g << 6) | nt);
      for (int f = 0; f < 8; f++) cand.push_back(HOP_LINK_ROW | f);
      for (int b : blocks)
        for (int f : cand) {
          {
            // element order with the links shared by each row: only with
            // CPS_HIP_HOP_MODE=row (then the only candidate where possible);
            // measured no faster than the element order on the 24^3x64x16
            // Moebius operator with 12 rhs (the M5^-1 hop 4.47 against 3.63
            // ms, the s-term hop a tie), so the default search leaves it out
            HopParams Q = P;
            Q.flags = HOP_LINK_ROW;
            const bool row_ok = link_row_setup(Q, b);
            if (f & HOP_LINK_ROW) {
              if (mode != 5 || !row_ok) continue;
            } else if (mode == 5 && row_ok) {
              continue;
            }
          }
          if (f & HOP_RHSGROUP) {
            if (!rhsgroup_ok(P, b, f) || mode == 1 || mode == 2 || mode == 3) continue;
          } else if (mode == 4 && (rhsgroup_ok(P, b, HOP_RHSGROUP) || rhsgroup_ok(P, def_block, HOP_RHSGROUP))) {
            continue;   // CPS_HIP_HOP_MODE=rhsgroup
          }
          if (f & HOP_SPATIAL) {
            if ((f & HOP_NO_XCD_REMAP) || !spatial_ok(L, P, b)) continue;
          }
          if ((f & HOP_SITEWAVE) && !sitewave_ok(L, P, b, (f & HOP_SPATIAL) != 0)) continue;
          if ((f & HOP_LDS_LINKS) && ((f & HOP_SITEWAVE) || !lds_links_ok(P, b))) continue;
          // CPS_HIP_HOP_MODE=sitewave / lds / element restricts the
          // candidates (where that mapping is possible) to that thread mapping
          if ((mode == 1 || mode == 2) && sitewave_ok(L, P, b, false) && ((f & HOP_SITEWAVE) != 0) != (mode == 2))
            continue;
          if ((mode == 1 || mode == 3) && lds_links_ok(P, b) && ((f & HOP_LDS_LINKS) != 0) != (mode == 3)) continue;
          if (mode == 3 && (f & HOP_SITEWAVE)) continue;
          if (P.partial && (f & HOP_SPATIAL) != (best.flags & HOP_SPATIAL)) continue;
          P.block = b;
          P.flags = f;
          CPSHIP_CHECK(launch_hop(P, L.stream));   // warm-up
          CPSHIP_CHECK(hipEventRecord(e0, L.stream));
          for (int r = 0; r < 3; r++) CPSHIP_CHECK(launch_hop(P, L.stream));
          CPSHIP_CHECK(hipEventRecord(e1, L.stream));
          CPSHIP_CHECK(hipEventSynchronize(e1));
          float ms = 0;
          CPSHIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
          if (tune_report() >= 2)
            fprintf(stderr, "cps_hip tune:   block %d flags %d: %.3f ms\n", b, f, ms / 3);
          if (ms < best_ms) {
            best_ms = ms;
            best.block = b;
            best.flags = f;
          }
        }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      best.ms = best_ms / 3;
      if (tune_report() >= 1)
        fprintf(stderr, "cps_hip tune: kind %d prec %d dag %d nb %d nsites %d nrhs %d clov %d -> block %d flags %d (%.3f ms)\n",
                P.kind, P.prec, P.dag, P.nb, P.nsites_out, P.nrhs, P.clov, best.block, best.flags, best.ms);
    }
    chosen = best;
    std::lock_guard<std::mutex> lk(tune_mu);
    tune_cache.emplace(key, best);
  }
  P.block = chosen.block;
  P.flags = chosen.flags;
  CPSHIP_CHECK(launch_hop(P, L.stream));
  return 0;
}

int hop_blocks_per_rhs(Lattice& L, const cps_hip_op* op) {
  OpLayout o;
  if (op_layout(op, L.ls, &o, L.g)) return 1;
  const long long n = (long long)o.nsites * o.nb;
  if (chunk_partials_shape(op->kind == CPS_HIP_OP_DWF ? KIND_DWF : KIND_WILSON, o.nb, n)) return (int)(n / 64);
  return (int)(n / pick_block(n));
}

// out = M in (dag 0) or M^dagger in (dag 1), M = 1 - kappa^2 D D
// (wilson_m.C:73-91 / dwf_m.C:31-62); partial receives |out|^2 per block.
int apply_m(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& out, const FieldRef& in, const FieldRef& t,
            int dag, double* partial, int nrhs) {
  const double k = op_kappa(op);
  int rc = hop(L, op, hprec, t, in, 0, dag, 1.0, null_field(), 0.0, nullptr, nrhs);
  if (rc) return rc;
  return hop(L, op, hprec, out, t, 1, dag, -k * k, in, 1.0, partial, nrhs);
}

// M^dag t3 = t3 - kappa^2 D^dag D^dag t3 (dwf_mdag.C:32-62) with r -= alpha
// M^dag t3 and |r|^2 in the epilogue of the second hop (CLOV_NRR)
int dwf_mdag_rupdate(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& r, const FieldRef& t3,
                     const FieldRef& t1, const double* alpha, const int* active, double* partial, int nrhs) {
  if (op->kind != CPS_HIP_OP_DWF || hprec == PREC_D_D18 || hprec == PREC_F_F12) return CPS_HIP_ERR_ARG;
  const double k = op_kappa(op);
  int rc = hop(L, op, hprec, t1, t3, 0, 1, 1.0, null_field(), 0.0, nullptr, nrhs);
  if (rc) return rc;
  HopClover h{};
  h.mode = CLOV_NONE;
  h.R = &r;
  h.ralpha = alpha;
  h.ractive = active;
  return hop_ex(L, op, hprec, r, t1, 1, 1, -k * k, t3, 1.0, partial, nrhs, &h);
}

// apply_m with a mass per rhs (Wilson and DWF kinds): rhs r of the launch
// uses kappa^2 = k2[rhs0 + r] (Wilson: the second hop's scale) or the wall
// coefficient wall[rhs0 + r] = -2 mf / a5 (DWF: both hops; kappa_5 does not
// depend on the mass), device arrays.
int apply_m_slots(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& out, const FieldRef& in,
                  const FieldRef& t, int dag, double* partial, int nrhs, const double* k2, const double* wall,
                  int rhs0) {
  HopClover h{};
  h.mode = CLOV_KS;
  h.rhs_per_set = 1;
  h.rhs0 = rhs0;
  if (op->kind == CPS_HIP_OP_WILSON) {
    int rc = hop(L, op, hprec, t, in, 0, dag, 1.0, null_field(), 0.0, nullptr, nrhs);
    if (rc) return rc;
    h.kscale = k2;
    return hop_ex(L, op, hprec, out, t, 1, dag, -1.0, in, 1.0, partial, nrhs, &h);
  }
  if (op->kind != CPS_HIP_OP_DWF) return CPS_HIP_ERR_ARG;
  const double k = op_kappa(op);
  h.kwall = wall;
  int rc = hop_ex(L, op, hprec, t, in, 0, dag, 1.0, null_field(), 0.0, nullptr, nrhs, &h);
  if (rc) return rc;
  return hop_ex(L, op, hprec, out, t, 1, dag, -k * k, in, 1.0, partial, nrhs, &h);
}

int apply_mdagm_slots(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& out, const FieldRef& in,
                      const FieldRef& t1, const FieldRef& t2, double* partial, int nrhs, const double* k2,
                      const double* wall, int rhs0) {
  int rc = apply_m_slots(L, op, hprec, t2, in, t1, 0, partial, nrhs, k2, wall, rhs0);
  if (rc) return rc;
  return apply_m_slots(L, op, hprec, out, t2, t1, 1, nullptr, nrhs, k2, wall, rhs0);
}

// MatPcDagMatPc: out = M^dag M in; partial (optional) = |M in|^2 per block
int apply_mdagm(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& out, const FieldRef& in,
                const FieldRef& t1, const FieldRef& t2, double* partial, int nrhs, bool t1_ready) {
  if (op->kind == CPS_HIP_OP_MOBIUS) return mobius_mdagm(L, op, hprec, out, in, t1, t2, partial, nrhs, t1_ready);
  if (op->kind == CPS_HIP_OP_CLOVER) return clover_mdagm(L, op, hprec, out, in, t1, t2, partial, nrhs);
  int rc = apply_m(L, op, hprec, t2, in, t1, 0, partial, nrhs);
  if (rc) return rc;
  return apply_m(L, op, hprec, out, t2, t1, 1, nullptr, nrhs);
}

int apply_matpc(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& out, const FieldRef& in,
                const FieldRef& t1, const FieldRef& t2, int dag, int nrhs) {
  if (op->kind == CPS_HIP_OP_MOBIUS) return mobius_m(L, op, hprec, out, in, t1, t2, dag, nullptr, nrhs);
  if (op->kind == CPS_HIP_OP_CLOVER) return clover_matpc(L, op, hprec, out, in, t1, dag, nullptr, nrhs);
  return apply_m(L, op, hprec, out, in, t1, dag, nullptr, nrhs);
}

// ---------------------------------------------------------------------------
// host <-> device vector transfer in CPS order

int upload(Lattice& L, const double* host, Field& f, const OpLayout& o, int cb5, int nrhs, int slot_io,
           bool on_device) {
  const int64_t n = o.per_rhs * nrhs;
  const double* src = host;
  if (!on_device) {
    DevBuf& io = slot_io ? L.io2 : L.io;
    int rc = io.ensure((size_t)n * 24 * sizeof(double));
    if (rc) return rc;
    CPSHIP_CHECK(hipMemcpyAsync(io.p, host, (size_t)n * 24 * sizeof(double), hipMemcpyHostToDevice, L.stream));
    src = (const double*)io.p;
  }
  CPSHIP_CHECK(pack_spinor(f.storage, src, f.ref(), o.map, L.g.Vcb, L.ls, cb5, n, L.stream));
  return 0;
}

int download(Lattice& L, const Field& f, double* host, const OpLayout& o, int cb5, int nrhs, bool on_device) {
  const int64_t n = o.per_rhs * nrhs;
  if (on_device) {
    CPSHIP_CHECK(unpack_spinor(f.storage, f.ref(), host, o.map, L.g.Vcb, L.ls, cb5, n, L.stream));
    CPSHIP_CHECK(hipStreamSynchronize(L.stream));
    return 0;
  }
  int rc = L.io.ensure((size_t)n * 24 * sizeof(double));
  if (rc) return rc;
  CPSHIP_CHECK(unpack_spinor(f.storage, f.ref(), (double*)L.io.p, o.map, L.g.Vcb, L.ls, cb5, n, L.stream));
  CPSHIP_CHECK(hipMemcpyAsync(host, L.io.p, (size_t)n * 24 * sizeof(double), hipMemcpyDeviceToHost, L.stream));
  CPSHIP_CHECK(hipStreamSynchronize(L.stream));
  return 0;
}

// ---------------------------------------------------------------------------
// CG scalar block

int CGState::alloc(Lattice& L, int nrhs_, int hist_len_) {
  nrhs = nrhs_;
  const size_t nd = 11, ni = 4;
  const size_t hist = (size_t)nrhs * hist_len_;
  int rc = L.scal.ensure((nd * nrhs + hist) * sizeof(double) + ni * nrhs * sizeof(int) + 256);
  if (rc) return rc;
  double* d = (double*)L.scal.p;
  s.rr = d + 0 * nrhs;
  s.rr_prev = d + 1 * nrhs;
  s.d = d + 2 * nrhs;
  s.alpha = d + 3 * nrhs;
  s.beta = d + 4 * nrhs;
  s.stp = d + 5 * nrhs;
  s.src_norm = d + 6 * nrhs;
  s.rsd = d + 7 * nrhs;
  rr_true = d + 8 * nrhs;
  s.rr_upd = nullptr;          // set by the mixed solver
  upd_store = d + 9 * nrhs;
  alpha_prev_store = d + 10 * nrhs;
  s.alpha_prev = nullptr;      // set by the mixed solver
  s.delta2 = 0.0;
  s.hist = hist_len_ ? d + nd * nrhs : nullptr;
  s.hist_len = hist_len_ ? hist_len_ : 1;
  int* ip = (int*)(d + nd * nrhs + hist);
  s.active = ip;
  s.iters = ip + nrhs;
  s.status = ip + 2 * nrhs;
  mask = ip + 3 * nrhs;
  return 0;
}

// ---------------------------------------------------------------------------
// FP64 CG: DiracOp::InvCgShift without shift, batched over nrhs.


// CPS_HIP_VERBOSE=1 prints the residual history to stderr (the VRB.Flow
// lines of inv_cg.C:352-353)
static bool verbose() { return knob(&cps_hip_tuning::verbose, "CPS_HIP_VERBOSE") != 0; }

// CPS_HIP_PROGRESS=<seconds> prints one line to stderr at most every that
// many seconds while a solve runs (iteration, rhs still active, elapsed).
struct Progress {
  double every = 0.0;
  std::chrono::steady_clock::time_point t0, last;
  const char* what;
  explicit Progress(const char* w) : what(w) {
    every = tuning().progress_seconds;
    t0 = last = std::chrono::steady_clock::now();
  }
  // rr: device |r|^2 per rhs (the largest is printed)
  void tick(Lattice& L, int iter, const std::vector<int>& act, const double* rr) {
    if (every <= 0.0) return;
    const auto now = std::chrono::steady_clock::now();
    if (std::chrono::duration<double>(now - last).count() < every) return;
    last = now;
    int n = 0;
    for (int a : act) n += a != 0;
    std::vector<double> h(act.size(), 0.0);
    double mx = 0.0;
    if (hipMemcpyAsync(h.data(), rr, h.size() * sizeof(double), hipMemcpyDeviceToHost, L.stream) == hipSuccess &&
        hipStreamSynchronize(L.stream) == hipSuccess)
      for (double v : h) mx = std::max(mx, v);
    fprintf(stderr, "cps_hip %s CG: iteration %d, %d of %zu rhs active, max |r|^2 %.3e, %.1f s\n", what, iter, n,
            act.size(), mx, std::chrono::duration<double>(now - t0).count());
    fflush(stderr);
  }
};

static int sync_read(Lattice& L, void* dst, const void* src, size_t bytes) {
  CPSHIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, L.stream));
  CPSHIP_CHECK(hipStreamSynchronize(L.stream));
  return 0;
}

static int auto_check_every(const OpLayout& o, int nrhs) {
  const double elems = (double)o.per_rhs * nrhs;
  if (elems < 1e5) return 32;
  if (elems < 1e6) return 16;
  return 8;
}

// ---------------------------------------------------------------------------
// Active-rhs window.  Once the rhs at either end of a batch have converged,
// every kernel of an iteration runs on the contiguous range [lo, hi) of rhs
// still iterating: the same planes, the data and norm pointers advanced by
// lo rhs, the per-rhs scalars by lo entries.  Each rhs's arithmetic (and its
// reduction order) is the same inside or outside the window, so the results
// do not change; converged rhs in the middle of the range keep their masked
// updates.  CPS_HIP_RHS_WINDOW=0 keeps the whole batch.

static bool rhs_window_enabled() { return knob(&cps_hip_tuning::rhs_window, "CPS_HIP_RHS_WINDOW") != 0; }

FieldRef rhs_view(const FieldRef& f, int64_t per_rhs, int lo) {
  FieldRef v = f;
  if (v.data) v.data = static_cast<char*>(v.data) + (int64_t)lo * per_rhs * 16;   // 16-byte chunks
  if (v.norm) v.norm += (int64_t)lo * per_rhs;
  return v;
}

CGScalars scalars_view(const CGScalars& s, int lo) {
  CGScalars v = s;
  v.rr += lo;
  v.rr_prev += lo;
  v.d += lo;
  v.alpha += lo;
  v.beta += lo;
  v.stp += lo;
  v.src_norm += lo;
  v.rsd += lo;
  v.active += lo;
  v.iters += lo;
  v.status += lo;
  if (v.rr_upd) v.rr_upd += lo;
  if (v.alpha_prev) v.alpha_prev += lo;
  if (v.hist) v.hist += (int64_t)lo * v.hist_len;
  return v;
}

struct RhsWindow {
  Lattice& L;
  bool on;
  int lo, hi;
  RhsWindow(Lattice& l, int nrhs) : L(l), on(rhs_window_enabled()), lo(0), hi(nrhs) {
    L.win_lo = 0;
    L.win_total = nrhs;
  }
  ~RhsWindow() {
    L.win_lo = 0;
    L.win_total = 0;
  }
  // shrink to the range of flags still set; true if the range changed
  bool update(const std::vector<int>& live) {
    if (!on) return false;
    int a = 0, b = (int)live.size();
    while (a < b && !live[a]) a++;
    while (b > a && !live[b - 1]) b--;
    if (a >= b || (a == lo && b == hi)) return false;
    lo = a;
    hi = b;
    L.win_lo = lo;
    return true;
  }
  int n() const { return hi - lo; }
};

int cg_double_device(Lattice& L, const cps_hip_op* op, Field& x, Field& src, int nrhs, const cps_hip_cg_arg* arg,
                     std::vector<int>& iters, std::vector<int>& status, std::vector<double>& true_res,
                     std::vector<double>& src_norm) {
  OpLayout o;
  int rc = op_layout(op, L.ls, &o, L.g);
  if (rc) return rc;
  const int64_t n = o.per_rhs * nrhs;
  Field& r = L.field(S_R, STORE_D, n);
  Field& p = L.field(S_P, STORE_D, n);
  Field& ap = L.field(S_AP, STORE_D, n);
  Field& t1 = L.field(S_T1, STORE_D, n);
  Field& t2 = L.field(S_T2, STORE_D, n);
  if (L.last_rc) return L.last_rc;
  const int bsb = pick_block(o.per_rhs);
  const int bpr_blas = (int)(o.per_rhs / bsb);
  const int bpr_hop = hop_blocks_per_rhs(L, op);
  rc = L.ensure_partial((int64_t)std::max(bpr_blas, bpr_hop) * nrhs);
  if (rc) return rc;
  double* part = (double*)L.partial.p;
  CGState cg;
  rc = cg.alloc(L, nrhs, 0);
  if (rc) return rc;
  cg.s.max_itr = arg->max_num_iter - 1;
  {
    std::vector<double> rsd(nrhs, arg->stop_rsd);
    CPSHIP_CHECK(hipMemcpyAsync(cg.s.rsd, rsd.data(), nrhs * sizeof(double), hipMemcpyHostToDevice, L.stream));
  }
  // |src|^2 and stp_cnd (inv_cg.C:166-174)
  CPSHIP_CHECK(blas_norm(STORE_D, src.ref(), o.per_rhs, nrhs, bsb, part, L.stream));
  CPSHIP_CHECK(reduce_rhs(part, bpr_blas, nrhs, nullptr, RED_SRC_NORM, cg.s, L.stream));
  // r = src - MdagM x ; p = r ; |r|^2 (inv_cg.C:248-269)
  rc = apply_mdagm(L, op, PREC_D_D18, ap.ref(), x.ref(), t1.ref(), t2.ref(), nullptr, nrhs);
  if (rc) return rc;
  CPSHIP_CHECK(blas_sub_norm_d(src.ref(), ap.ref(), r.ref(), o.per_rhs, nrhs, bsb, part, L.stream));
  CPSHIP_CHECK(reduce_rhs(part, bpr_blas, nrhs, nullptr, RED_CG_INIT, cg.s, L.stream));
  CPSHIP_CHECK(convert_spinor(STORE_D, r.ref(), STORE_D, p.ref(), n, L.stream));

  const int chunk = arg->check_every > 0 ? arg->check_every : auto_check_every(o, nrhs);
  std::vector<int> act(nrhs);
  int done_iters = 0;
  const int max_itr = std::max(arg->max_num_iter - 1, 0);
  rc = sync_read(L, act.data(), cg.s.active, nrhs * sizeof(int));
  if (rc) return rc;
  auto any_active = [&]() {
    for (int k = 0; k < nrhs; k++)
      if (act[k]) return true;
    return false;
  };
  Progress prog("fp64");
  IterGraph iter_graph(arg->use_graph != 0);
  RhsWindow win(L, nrhs);
  while (any_active() && done_iters < max_itr) {
    prog.tick(L, done_iters, act, cg.s.rr);
    if (win.update(act)) iter_graph.reset();
    const int nw = win.n();
    const FieldRef xw = rhs_view(x.ref(), o.per_rhs, win.lo), pw = rhs_view(p.ref(), o.per_rhs, win.lo);
    const FieldRef rw = rhs_view(r.ref(), o.per_rhs, win.lo), apw = rhs_view(ap.ref(), o.per_rhs, win.lo);
    const CGScalars sw = scalars_view(cg.s, win.lo);
    const int todo = std::min(chunk, max_itr - done_iters);
    rc = iter_graph.run(L, todo, [&]() -> int {
      // mmp = MdagM p, d = |M p|^2 (inv_cg.C:302)
      int e = apply_mdagm(L, op, PREC_D_D18, apw, pw, t1.ref(), t2.ref(), part, nw);
      if (e) return e;
      CPSHIP_CHECK(reduce_rhs(part, bpr_hop, nw, nullptr, RED_CG_ALPHA, sw, L.stream));
      CPSHIP_CHECK(blas_cg_xr_d(xw, pw, rw, apw, sw, o.per_rhs, nw, bsb, part, L.stream));
      CPSHIP_CHECK(reduce_rhs(part, bpr_blas, nw, nullptr, RED_CG_BETA, sw, L.stream));
      CPSHIP_CHECK(blas_xpay(STORE_D, pw, rw, sw.beta, sw.active, o.per_rhs, nw, bsb, L.stream));
      return 0;
    });
    if (rc) return rc;
    done_iters += todo;
    rc = sync_read(L, act.data(), cg.s.active, nrhs * sizeof(int));
    if (rc) return rc;
  }
  win.lo = 0;
  win.hi = nrhs;
  L.win_lo = 0;
  // true residual (inv_cg.C:403-417)
  rc = apply_mdagm(L, op, PREC_D_D18, ap.ref(), x.ref(), t1.ref(), t2.ref(), nullptr, nrhs);
  if (rc) return rc;
  CPSHIP_CHECK(blas_sub_norm_d(src.ref(), ap.ref(), r.ref(), o.per_rhs, nrhs, bsb, part, L.stream));
  double* trn = cg.s.d;  // reuse as output slot
  CPSHIP_CHECK(reduce_rhs(part, bpr_blas, nrhs, trn, RED_PLAIN, cg.s, L.stream));
  std::vector<double> tr(nrhs), sn(nrhs);
  iters.resize(nrhs);
  status.resize(nrhs);
  rc = sync_read(L, tr.data(), trn, nrhs * sizeof(double));
  if (rc) return rc;
  rc = sync_read(L, sn.data(), cg.s.src_norm, nrhs * sizeof(double));
  if (rc) return rc;
  rc = sync_read(L, iters.data(), cg.s.iters, nrhs * sizeof(int));
  if (rc) return rc;
  rc = sync_read(L, status.data(), cg.s.status, nrhs * sizeof(int));
  if (rc) return rc;
  true_res.resize(nrhs);
  src_norm = sn;
  for (int k = 0; k < nrhs; k++) {
    true_res[k] = std::sqrt(tr[k] / sn[k]);
    if (status[k] == CG_STATUS_RUNNING) status[k] = CG_STATUS_MAXITER;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Mixed-precision CG with FP64 reliable updates.
//
// Sloppy iteration: FP32 arithmetic, p / Ap / operator temporaries stored in
// the sloppy spinor format, the sloppy residual rs in FP32, the partial
// solution xs in FP32.  Every `reliable_every` iterations, and whenever the
// sloppy residual says a rhs has converged, the FP64-critical steps run:
//   x += xs (FP64 accumulate), r = src - MdagM x (FP64 operator and links),
//   rs = r, xs = 0;
// a rhs stops only when that FP64 residual meets |r|^2 <= stop_rsd^2 |src|^2.

// Storage / arithmetic of a solver precision mode for an operator.  Links are
// stored 18-real when a site carries at least 4 spinors (DWF slots, Moebius
// slices, right-hand sides): the links are then a small share of the traffic
// and skipping the third-row reconstruction saves arithmetic; otherwise the
// 12-real form saves bandwidth.  The 12-real form needs links that are SU(3)
// times the boundary factor (recon_ok, measured at import); without it the
// half-precision modes keep 18 reals, and MIXED_SINGLE (FP32 links exist only
// 12-real) is refused by the hop (CPS_HIP_ERR_ARG).
void precision_formats(int prec, const cps_hip_op* op, int ls, int nrhs, bool recon_ok, int* sp_store, int* hprec) {
  int per_site = nrhs;
  if (op->kind == CPS_HIP_OP_DWF) per_site *= ls / 2;
  if (op->kind == CPS_HIP_OP_MOBIUS) per_site *= ls;
  const bool full = per_site >= 4 || !recon_ok;
  switch (prec) {
    case CPS_HIP_PREC_DOUBLE: *sp_store = STORE_D; *hprec = PREC_D_D18; break;
    case CPS_HIP_PREC_MIXED_SINGLE: *sp_store = STORE_F; *hprec = PREC_F_F12; break;
    case CPS_HIP_PREC_MIXED_BF16: *sp_store = STORE_H; *hprec = full ? PREC_H_B18 : PREC_H_B12; break;
    default: *sp_store = STORE_H; *hprec = full ? PREC_H_H18 : PREC_H_H12; break;
  }
}

// Storage of the sloppy solution xs and residual rs with half-precision
// sloppy spinors: xs FP32 and rs FP16 with per-site norms by default (the
// same iteration counts as an FP32 rs on the 16^3x32x16 DWF, 24^3x64x16
// Moebius and 32^3x64 clover solves, and 44 fewer bytes per element each
// time rs is read or written); CPS_HIP_SLOPPY_VECTORS (read
// per solve) "single" keeps both FP32, "xhalf" stores xs in FP16 and rs in
// FP32, "half" both in FP16 (an FP16 xs triples the iteration count).
static void sloppy_vector_storage(int sps, int* xst, int* rst) {
  *xst = *rst = STORE_F;
  if (sps != STORE_H) return;
  *rst = STORE_H;
  switch (tuning().sloppy_vectors) {
    case 1: *rst = STORE_F; break;
    case 2: *xst = STORE_H; *rst = STORE_F; break;
    case 3: *xst = *rst = STORE_H; break;
    default: break;
  }
}

static void mixed_formats(const Lattice& L, int prec, const cps_hip_op* op, int nrhs, int* sp_store, int* hprec) {
  precision_formats(prec, op, L.ls, nrhs, L.recon_ok, sp_store, hprec);
}

// After a reliable update (rr_true = FP64 |src - MdagM x|^2 per rhs, the
// sloppy residual already replaced by the FP64 one, rp = <r, p> with the
// direction p of the last step, whose own update p = r + beta p is still to
// come at the start of the next iteration):
//  - a rhs whose FP64 residual meets the stopping rule is finished;
//  - every other unfinished rhs continues from the FP64 residual: |r|^2 is
//    the true one, beta = |r|^2 / |r_prev|^2, and p is re-orthogonalised
//    against the new r, p -= (<r,p>/|r|^2) r (coef[k]; 0 for the others), as
//    reliable-update CG does so that the sloppy drift of r does not break
//    the CG relations of the next step.
__global__ void k_mixed_after_update(CGScalars cg, const double* rr_true, const double* rp, double* coef, int nrhs) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nrhs) return;
  // rp aliases alpha: read it, then leave no solution step pending
  const double rpk = rp[k];
  cg.alpha[k] = 0.0;
  if (cg.alpha_prev) cg.alpha_prev[k] = 0.0;
  coef[k] = 0.0;
  const int st = cg.status[k];
  if (st == CG_STATUS_D_ZERO || st == CG_STATUS_MAXITER) return;
  if (st == CG_STATUS_CONVERGED && !cg.active[k] && cg.iters[k] == 0 && rr_true[k] <= cg.stp[k]) return;
  const double v = rr_true[k];
  cg.rr_upd[k] = v;
  if (v <= cg.stp[k]) {
    cg.rr[k] = v;
    cg.active[k] = 0;
    cg.beta[k] = 0.0;
    cg.status[k] = CG_STATUS_CONVERGED;
    return;
  }
  cg.beta[k] = v / cg.rr_prev[k];
  coef[k] = v > 0.0 ? -rpk / v : 0.0;
  cg.rr[k] = v;
  cg.active[k] = 1;
  cg.status[k] = CG_STATUS_RUNNING;
}

hipError_t mixed_after_update(const CGScalars& cg, const double* rr_true, const double* rp, double* coef, int nrhs,
                              hipStream_t st) {
  hipLaunchKernelGGL(k_mixed_after_update, dim3((nrhs + 63) / 64), dim3(64), 0, st, cg, rr_true, rp, coef, nrhs);
  return hipGetLastError();
}

int cg_mixed_device(Lattice& L, const cps_hip_op* op, Field& x, Field& src, int nrhs, const cps_hip_cg_arg* arg,
                    std::vector<int>& iters, std::vector<int>& status, std::vector<double>& true_res,
                    std::vector<double>& src_norm, std::vector<int>& n_updates) {
  OpLayout o;
  int rc = op_layout(op, L.ls, &o, L.g);
  if (rc) return rc;
  int sps, hprec;
  mixed_formats(L, arg->precision, op, nrhs, &sps, &hprec);
  const int64_t n = o.per_rhs * nrhs;
  Field& r = L.field(S_R, STORE_D, n);
  Field& ap = L.field(S_AP, STORE_D, n);
  Field& t1 = L.field(S_T1, STORE_D, n);
  Field& t2 = L.field(S_T2, STORE_D, n);
  // sloppy solution and residual: FP32 by default; with half-precision
  // sloppy storage and CPS_HIP_SLOPPY_VECTORS=half they are FP16 too (less
  // BLAS traffic, more reliance on the reliable updates)
  int xvec, svec;
  sloppy_vector_storage(sps, &xvec, &svec);
  Field& xs = L.field(S_XS, xvec, n);
  Field& rs = L.field(S_RS, svec, n);
  Field& ps = L.field(S_PS, sps, n);
  Field& aps = L.field(S_APS, sps, n);
  Field& t1s = L.field(S_T1S, sps, n);
  Field& t2s = L.field(S_T2S, sps, n);
  if (L.last_rc) return L.last_rc;
  const int bsb = pick_block(o.per_rhs);
  const int bpr_blas = (int)(o.per_rhs / bsb);
  const int bpr_hop = hop_blocks_per_rhs(L, op);
  rc = L.ensure_partial((int64_t)std::max(bpr_blas, bpr_hop) * nrhs);
  if (rc) return rc;
  double* part = (double*)L.partial.p;
  CGState cg;
  rc = cg.alloc(L, nrhs, 0);
  if (rc) return rc;
  const int max_itr = std::max(arg->max_num_iter - 1, 0);
  cg.s.max_itr = max_itr;
  {
    std::vector<double> rsd(nrhs, arg->stop_rsd);
    CPSHIP_CHECK(hipMemcpyAsync(cg.s.rsd, rsd.data(), nrhs * sizeof(double), hipMemcpyHostToDevice, L.stream));
  }
  double* rr_true = cg.rr_true;
  cg.s.rr_upd = cg.upd_store;
  {
    const double delta = arg->reliable_delta > 0 ? arg->reliable_delta : 0.1;
    cg.s.delta2 = delta * delta;
  }

  CPSHIP_CHECK(blas_norm(STORE_D, src.ref(), o.per_rhs, nrhs, bsb, part, L.stream));
  CPSHIP_CHECK(reduce_rhs(part, bpr_blas, nrhs, nullptr, RED_SRC_NORM, cg.s, L.stream));
  // FP64 initial residual
  rc = apply_mdagm(L, op, PREC_D_D18, ap.ref(), x.ref(), t1.ref(), t2.ref(), nullptr, nrhs);
  if (rc) return rc;
  CPSHIP_CHECK(blas_sub_norm_d(src.ref(), ap.ref(), r.ref(), o.per_rhs, nrhs, bsb, part, L.stream));
  CPSHIP_CHECK(reduce_rhs(part, bpr_blas, nrhs, cg.s.rr_upd, RED_CG_INIT, cg.s, L.stream));
  CPSHIP_CHECK(convert_spinor(STORE_D, r.ref(), svec, rs.ref(), n, L.stream));
  CPSHIP_CHECK(convert_spinor(STORE_D, r.ref(), sps, ps.ref(), n, L.stream));
  CPSHIP_CHECK(blas_zero(xvec, xs.ref(), o.per_rhs, nrhs, bsb, L.stream));
  {
    // the first iteration's direction update p = r + 0 p, and no pending
    // solution step
    std::vector<double> zeros(nrhs, 0.0);
    CPSHIP_CHECK(hipMemcpyAsync(cg.s.beta, zeros.data(), nrhs * sizeof(double), hipMemcpyHostToDevice, L.stream));
    CPSHIP_CHECK(hipMemcpyAsync(cg.s.alpha, zeros.data(), nrhs * sizeof(double), hipMemcpyHostToDevice, L.stream));
  }
  // the solution step x += alpha p is applied one iteration late, by the
  // kernel that reads p next (the direction update), which saves one read of
  // p per iteration; a reliable update applies the pending step first
  // (CPS_HIP_LAGGED_X=0: in the same iteration, with r)
  const cps_hip_tuning tn = tuning();
  const bool lagged = tn.lagged_x != 0;
  // Moebius, FP16 directions, Ls 16: the direction update also writes the
  // first step of M, b p + c P p, into the operator's first scratch field
  // (CPS_HIP_S5_XPAY=0: the operator computes it)
  const bool s5_in_xpay = lagged && tn.s5_xpay != 0 && op->kind == CPS_HIP_OP_MOBIUS && sps == STORE_H &&
                          xvec == STORE_F && L.ls == 16;
  // with it, the residual step r -= alpha A p runs in the epilogue of the
  // last hop of M^dag (alpha is known after the first half of M^dag M), so A p
  // is never stored (CPS_HIP_R_IN_HOP=0: a separate pass)
  const bool r_in_hop = s5_in_xpay && svec == STORE_H && tn.r_in_hop != 0;
  // clover, FP16 directions and residual, on request
  // (CPS_HIP_CLOVER_R_IN_HOP=1): likewise the residual step in the epilogue
  // of the last hop of MatPcDag, after the alpha of the first half.  Off by
  // default: that epilogue (clover multiply, X and r in registers) spills
  // at the hop's register budget and the 32^3x64 clover solve with 48 rhs
  // takes 3.85 instead of 2.56 s.
  const bool clover_rih = lagged && op->kind == CPS_HIP_OP_CLOVER && sps == STORE_H && svec == STORE_H &&
                          tn.clover_r_in_hop != 0;
  // DWF, FP16 directions and residual: the residual step in the epilogue of
  // the second hop of M^dag (CPS_HIP_R_IN_HOP=0: a separate pass)
  const bool dwf_rih = lagged && op->kind == CPS_HIP_OP_DWF && sps == STORE_H && svec == STORE_H &&
                       tn.r_in_hop != 0;
  Field& t3s = L.field_min(S_T3S, sps, n);
  if (L.last_rc) return L.last_rc;
  // FP16 directions and an FP32 solution: iterations run in pairs with the
  // direction in two buffers, and the solution is updated every second
  // iteration with both pending steps, in order (bit-identical to one step
  // per iteration, one pass over x less per pair; CPS_HIP_PAIRED_X=0 turns
  // it off)
  const bool paired = lagged && sps == STORE_H && xvec == STORE_F && tn.paired_x != 0;
  Field* ps2 = nullptr;
  if (paired) {
    ps2 = &L.field(S_PS2, sps, n);
    if (L.last_rc) return L.last_rc;
    cg.s.alpha_prev = cg.alpha_prev_store;
    std::vector<double> zeros(nrhs, 0.0);
    CPSHIP_CHECK(hipMemcpyAsync(cg.s.alpha_prev, zeros.data(), nrhs * sizeof(double), hipMemcpyHostToDevice,
                                L.stream));
  }
  DevBuf coef_buf;
  if ((rc = coef_buf.ensure(nrhs * sizeof(double)))) return rc;
  double* coef = (double*)coef_buf.p;

  const int chunk = arg->check_every > 0 ? arg->check_every : auto_check_every(o, nrhs);
  const int every = arg->reliable_every > 0 ? arg->reliable_every : 100;
  std::vector<int> act(nrhs), st(nrhs);
  n_updates.assign(nrhs, 0);
  rc = sync_read(L, act.data(), cg.s.active, nrhs * sizeof(int));
  if (rc) return rc;
  std::vector<int> live = act;  // rhs not yet converged on the FP64 residual
  int done = 0, since = 0;
  auto any = [&](const std::vector<int>& v) {
    for (int k = 0; k < nrhs; k++)
      if (v[k]) return true;
    return false;
  };
  Progress prog("mixed");
  IterGraph iter_graph(arg->use_graph != 0);
  RhsWindow win(L, nrhs);   // over the rhs not yet converged on the FP64 residual
  while (any(live) && done < max_itr) {
    prog.tick(L, done, live, cg.s.rr);
    if (win.update(live)) iter_graph.reset();
    const int lo = win.lo, nw = win.n();
    const FieldRef xw = rhs_view(x.ref(), o.per_rhs, lo), srcw = rhs_view(src.ref(), o.per_rhs, lo);
    const FieldRef rw = rhs_view(r.ref(), o.per_rhs, lo), apw = rhs_view(ap.ref(), o.per_rhs, lo);
    const FieldRef xsw = rhs_view(xs.ref(), o.per_rhs, lo), rsw = rhs_view(rs.ref(), o.per_rhs, lo);
    const FieldRef psw = rhs_view(ps.ref(), o.per_rhs, lo), apsw = rhs_view(aps.ref(), o.per_rhs, lo);
    const CGScalars sw = scalars_view(cg.s, lo);
    int todo = std::min(std::min(chunk, max_itr - done), every - since);
    // the operator on this iteration's direction pw, alpha, the residual step
    // (in the last hop of M^dag, or its own pass) and beta
    auto opstep = [&](const FieldRef& pw) -> int {
      if (clover_rih) {
        // M p with |M p|^2, alpha, then r -= alpha M^dag (M p) with |r|^2
        int e = clover_matpc(L, op, hprec, t3s.ref(), pw, t1s.ref(), 0, part, nw);
        if (e) return e;
        CPSHIP_CHECK(reduce_rhs(part, bpr_hop, nw, nullptr, RED_CG_ALPHA, sw, L.stream));
        e = clover_mdag_rupdate(L, op, hprec, rsw, t3s.ref(), t1s.ref(), sw.alpha, sw.active, part, nw);
        if (e) return e;
        CPSHIP_CHECK(reduce_rhs(part, bpr_hop, nw, nullptr, RED_CG_BETA, sw, L.stream));
        return 0;
      }
      if (r_in_hop) {
        // M p with |M p|^2, alpha, then r -= alpha M^dag (M p) with |r|^2
        int e = mobius_m(L, op, hprec, t3s.ref(), pw, t1s.ref(), t2s.ref(), 0, part, nw, true);
        if (e) return e;
        CPSHIP_CHECK(reduce_rhs(part, bpr_hop, nw, nullptr, RED_CG_ALPHA, sw, L.stream));
        e = mobius_mdag_rupdate(L, op, hprec, rsw, t3s.ref(), t1s.ref(), t2s.ref(), sw.alpha, sw.active, part, nw);
        if (e) return e;
        CPSHIP_CHECK(reduce_rhs(part, bpr_hop, nw, nullptr, RED_CG_BETA, sw, L.stream));
        return 0;
      }
      if (dwf_rih) {
        // M p with |M p|^2, alpha, then r -= alpha M^dag (M p) with |r|^2
        int e = apply_m(L, op, hprec, t3s.ref(), pw, t1s.ref(), 0, part, nw);
        if (e) return e;
        CPSHIP_CHECK(reduce_rhs(part, bpr_hop, nw, nullptr, RED_CG_ALPHA, sw, L.stream));
        e = dwf_mdag_rupdate(L, op, hprec, rsw, t3s.ref(), t1s.ref(), sw.alpha, sw.active, part, nw);
        if (e) return e;
        CPSHIP_CHECK(reduce_rhs(part, bpr_hop, nw, nullptr, RED_CG_BETA, sw, L.stream));
        return 0;
      }
      int e = apply_mdagm(L, op, hprec, apsw, pw, t1s.ref(), t2s.ref(), part, nw, s5_in_xpay);
      if (e) return e;
      CPSHIP_CHECK(reduce_rhs(part, bpr_hop, nw, nullptr, RED_CG_ALPHA, sw, L.stream));
      CPSHIP_CHECK(blas_cg_r_mixed(svec, rsw, sps, apsw, sw, o.per_rhs, nw, bsb, part, L.stream));
      CPSHIP_CHECK(reduce_rhs(part, bpr_blas, nw, nullptr, RED_CG_BETA, sw, L.stream));
      return 0;
    };
    if (paired) {
      // iteration a: p_B = r + beta p_A ; iteration b: x += alpha_prev p_A +
      // alpha p_B, p_A = r + beta p_B (chunks are whole pairs, so the latest
      // direction is in A at every chunk boundary)
      const FieldRef pbw = rhs_view(ps2->ref(), o.per_rhs, lo);
      const FieldRef t1ref = t1s.ref();
      const FieldRef* t5 = s5_in_xpay ? &t1ref : nullptr;
      const int pairs = (todo + 1) / 2;
      todo = 2 * pairs;
      rc = iter_graph.run(L, pairs, [&]() -> int {
        CPSHIP_CHECK(blas_xpay_pair(psw, pbw, svec, rsw, xsw, t5, false, nullptr, nullptr, sw.beta, sw.active,
                                    o.per_rhs, nw, bsb, op->mobius_b, op->mobius_c, op->mass, L.stream));
        int e = opstep(pbw);
        if (e) return e;
        CPSHIP_CHECK(blas_xpay_pair(pbw, psw, svec, rsw, xsw, t5, true, sw.alpha_prev, sw.alpha, sw.beta,
                                    sw.active, o.per_rhs, nw, bsb, op->mobius_b, op->mobius_c, op->mass, L.stream));
        return opstep(psw);
      });
    } else
    // one iteration: the previous step's direction update p = r + beta p
    // first, then Ap, alpha, x and r, |r|^2 and beta -- so that a reliable
    // update at a chunk boundary falls between r's update and p's
    rc = iter_graph.run(L, todo, [&]() -> int {
      if (s5_in_xpay)
        CPSHIP_CHECK(blas_xpay_x_s5(psw, svec, rsw, xsw, t1s.ref(), sw.alpha, sw.beta, sw.active, o.per_rhs, nw, bsb,
                                    op->mobius_b, op->mobius_c, op->mass, L.stream));
      else if (lagged)
        CPSHIP_CHECK(blas_xpay_x_mixed(sps, psw, svec, rsw, xvec, xsw, sw.alpha, sw.beta, sw.active, o.per_rhs, nw,
                                       bsb, L.stream));
      else
        CPSHIP_CHECK(blas_xpay_mixed(sps, psw, svec, rsw, sw.beta, sw.active, o.per_rhs, nw, bsb, L.stream));
      if (clover_rih || dwf_rih) return opstep(psw);
      if (r_in_hop) {
        // M p with |M p|^2, alpha, then r -= alpha M^dag (M p) with |r|^2
        int e = mobius_m(L, op, hprec, t3s.ref(), psw, t1s.ref(), t2s.ref(), 0, part, nw, true);
        if (e) return e;
        CPSHIP_CHECK(reduce_rhs(part, bpr_hop, nw, nullptr, RED_CG_ALPHA, sw, L.stream));
        e = mobius_mdag_rupdate(L, op, hprec, rsw, t3s.ref(), t1s.ref(), t2s.ref(), sw.alpha, sw.active, part, nw);
        if (e) return e;
        CPSHIP_CHECK(reduce_rhs(part, bpr_hop, nw, nullptr, RED_CG_BETA, sw, L.stream));
        return 0;
      }
      int e = apply_mdagm(L, op, hprec, apsw, psw, t1s.ref(), t2s.ref(), part, nw, s5_in_xpay);
      if (e) return e;
      CPSHIP_CHECK(reduce_rhs(part, bpr_hop, nw, nullptr, RED_CG_ALPHA, sw, L.stream));
      if (lagged)
        CPSHIP_CHECK(blas_cg_r_mixed(svec, rsw, sps, apsw, sw, o.per_rhs, nw, bsb, part, L.stream));
      else
        CPSHIP_CHECK(blas_cg_xr_mixed(xvec, xsw, sps, psw, svec, rsw, sps, apsw, sw, o.per_rhs, nw, bsb, part,
                                      L.stream));
      CPSHIP_CHECK(reduce_rhs(part, bpr_blas, nw, nullptr, RED_CG_BETA, sw, L.stream));
      return 0;
    });
    if (rc) return rc;
    done += todo;
    since += todo;
    rc = sync_read(L, act.data(), cg.s.active, nrhs * sizeof(int));
    if (rc) return rc;
    rc = sync_read(L, st.data(), cg.s.status, nrhs * sizeof(int));
    if (rc) return rc;
    bool paused = false;   // a rhs stopped on the sloppy residual or the delta trigger
    for (int k = 0; k < nrhs; k++)
      if (live[k] && !act[k]) paused = true;
    if (verbose()) {
      std::vector<double> rr(nrhs), al(nrhs), be(nrhs), dd(nrhs);
      sync_read(L, rr.data(), cg.s.rr, nrhs * sizeof(double));
      sync_read(L, al.data(), cg.s.alpha, nrhs * sizeof(double));
      sync_read(L, be.data(), cg.s.beta, nrhs * sizeof(double));
      sync_read(L, dd.data(), cg.s.d, nrhs * sizeof(double));
      fprintf(stderr, "cps_hip mixed CG iter %d window [%d,%d):", done, lo, lo + nw);
      for (int k = 0; k < nrhs; k++)
        fprintf(stderr, " rhs %d |r_s|^2 %a alpha %a beta %a d %a active %d;", k, rr[k], al[k], be[k], dd[k], act[k]);
      fprintf(stderr, "\n");
    }
    if (since >= every || paused || done >= max_itr) {
      // ---- FP64-critical steps: accumulate the solution, recompute r ----
      // (on the window: the rhs outside it have converged and keep x)
      if (lagged)   // the pending step x += alpha p of the last iteration
        CPSHIP_CHECK(blas_axpy(sps, psw, xvec, xsw, sw.alpha, 0.0, o.per_rhs, nw, bsb, L.stream));
      CPSHIP_CHECK(blas_axpy(xvec, xsw, STORE_D, xw, nullptr, 1.0, o.per_rhs, nw, bsb, L.stream));
      CPSHIP_CHECK(blas_zero(xvec, xsw, o.per_rhs, nw, bsb, L.stream));
      rc = apply_mdagm(L, op, PREC_D_D18, apw, xw, t1.ref(), t2.ref(), nullptr, nw);
      if (rc) return rc;
      CPSHIP_CHECK(blas_sub_norm_d(srcw, apw, rw, o.per_rhs, nw, bsb, part, L.stream));
      CPSHIP_CHECK(reduce_rhs(part, bpr_blas, nw, rr_true + lo, RED_PLAIN, sw, L.stream));
      CPSHIP_CHECK(convert_spinor(STORE_D, rw, svec, rsw, o.per_rhs * nw, L.stream));
      CPSHIP_CHECK(blas_redot_mixed(svec, rsw, sps, psw, o.per_rhs, nw, bsb, part, L.stream));
      // <r, p> into alpha (rewritten by the next step)
      CPSHIP_CHECK(reduce_rhs(part, bpr_blas, nw, sw.alpha, RED_PLAIN, sw, L.stream));
      hipLaunchKernelGGL(k_mixed_after_update, dim3((nw + 63) / 64), dim3(64), 0, L.stream, sw, rr_true + lo,
                         (const double*)sw.alpha, coef + lo, nw);
      CPSHIP_CHECK(hipGetLastError());
      // p -= (<r,p>/|r|^2) r for the continuing rhs (coef 0 for the others)
      CPSHIP_CHECK(blas_axpy(svec, rsw, sps, psw, coef + lo, 0.0, o.per_rhs, nw, bsb, L.stream));
      for (int k = 0; k < nrhs; k++)
        if (live[k]) n_updates[k]++;
      since = 0;
      rc = sync_read(L, act.data(), cg.s.active, nrhs * sizeof(int));
      if (rc) return rc;
      rc = sync_read(L, st.data(), cg.s.status, nrhs * sizeof(int));
      if (rc) return rc;
      for (int k = 0; k < nrhs; k++)
        if (live[k] && st[k] == CG_STATUS_CONVERGED) live[k] = 0;
      if (verbose()) {
        std::vector<double> rt(nrhs);
        sync_read(L, rt.data(), rr_true, nrhs * sizeof(double));
        fprintf(stderr, "cps_hip mixed CG reliable update at %d: true |r|^2", done);
        for (int k = 0; k < nrhs; k++) fprintf(stderr, " %a", rt[k]);
        fprintf(stderr, "\n");
      }
    }
  }
  win.lo = 0;
  win.hi = nrhs;
  L.win_lo = 0;
  // final FP64 true residual of MdagM (inv_cg.C:403-417)
  rc = apply_mdagm(L, op, PREC_D_D18, ap.ref(), x.ref(), t1.ref(), t2.ref(), nullptr, nrhs);
  if (rc) return rc;
  CPSHIP_CHECK(blas_sub_norm_d(src.ref(), ap.ref(), r.ref(), o.per_rhs, nrhs, bsb, part, L.stream));
  CPSHIP_CHECK(reduce_rhs(part, bpr_blas, nrhs, rr_true, RED_PLAIN, cg.s, L.stream));
  std::vector<double> tr(nrhs), sn(nrhs);
  iters.resize(nrhs);
  status.resize(nrhs);
  rc = sync_read(L, tr.data(), rr_true, nrhs * sizeof(double));
  if (rc) return rc;
  rc = sync_read(L, sn.data(), cg.s.src_norm, nrhs * sizeof(double));
  if (rc) return rc;
  rc = sync_read(L, iters.data(), cg.s.iters, nrhs * sizeof(int));
  if (rc) return rc;
  rc = sync_read(L, status.data(), cg.s.status, nrhs * sizeof(int));
  if (rc) return rc;
  true_res.resize(nrhs);
  src_norm = sn;
  for (int k = 0; k < nrhs; k++) {
    true_res[k] = std::sqrt(tr[k] / sn[k]);
    if (status[k] == CG_STATUS_RUNNING) status[k] = CG_STATUS_MAXITER;
  }
  return 0;
}

}  // namespace cpship
