// Resident solve service: gauge fields stay on the device, and incoming
// propagator solves (any registered lattice, operator and mass, 1..n
// right-hand sides) are served by continuous batching.
//
// An engine is one batched CG (FP64: the iteration of DiracOp::InvCg,
// inv_cg.C:293-364; mixed: the sloppy iteration with FP64 reliable updates)
// over S right-hand-side slots for one lattice, operator and precision, on
// its own HIP stream; Wilson and DWF engines carry a mass per slot.  Every
// slot carries its own CG scalars, so a slot that converges is harvested
// (FP64 true residual, solution copied out) and refilled with the next
// waiting right-hand side while the other slots keep iterating: requests
// that arrive while others are on the GPU join as slots free up.  Engines of
// different lattices or operators run concurrently on their own streams; the
// service thread only schedules them, and the copies between the callers'
// host memory and the device (sources in, solutions out) run on a few host
// worker threads.  A new engine is created only while the HBM budget allows
// (admission by footprint), and engines idle for 30 s are released (sooner
// when the budget needs their memory).
//
// A right-hand side's arithmetic never depends on the other slots: the hop
// and BLAS kernels are element-wise, and every reduction sums one slot's
// block (or 64-site chunk) partials in a fixed order.  A solve therefore gives the same bits
// alone and in any mix.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "solver.h"

using clk = std::chrono::steady_clock;

namespace cpship {
namespace {

struct Request {
  long long id = 0;
  int gauge_id = 0;
  cps_hip_op op;
  const double* src = nullptr;   // caller's host memory, nrhs vectors
  double* sol = nullptr;
  int nrhs = 0;
  double stop_rsd = 1e-10;
  int max_iter = 10000;
  int precision = CPS_HIP_PREC_DOUBLE;
  int reliable_every = 100;
  double reliable_delta = 0.1;
  int next_rhs = 0;              // next rhs to place in a slot
  int done_rhs = 0;
  int rc = 0;
  std::vector<cps_hip_cg_result> res;
  clk::time_point t_submit, t_done;
  bool finished = false;
};

struct GaugeEntry {
  int L[4];
  int ls;
  int bc[4];
  DevBuf u;   // WILSON order FP64, resident
};

// Wilson and DWF engines take a mass per slot (the mass enters only the
// second hop's kappa^2, or the DWF wall coefficient), so requests of any
// mass on one lattice share an engine; Moebius and clover engines are per
// mass (the mass enters M5^-1 and the clover term).
bool per_slot_mass(const cps_hip_op& op) { return op.kind == CPS_HIP_OP_WILSON || op.kind == CPS_HIP_OP_DWF; }

bool same_engine_op(const cps_hip_op& a, const cps_hip_op& b) {
  return a.kind == b.kind && (per_slot_mass(a) || a.mass == b.mass) && a.m5 == b.m5 && a.a5_inv == b.a5_inv &&
         a.mobius_b == b.mobius_b && a.mobius_c == b.mobius_c && a.csw == b.csw;
}

// An engine: one batched CG over S right-hand-side slots for one lattice,
// operator and solver precision, on the lattice's own stream.
//  - FP64: the iteration of DiracOp::InvCg (inv_cg.C:293-364) per slot; a
//    slot that stopped gets its FP64 true residual (inv_cg.C:403-417) alone.
//  - mixed (FP32 arithmetic on the sloppy fields of the precision mode):
//    the sloppy CG per slot with FP64 reliable updates of all slots at
//    least every `every` iterations and whenever a slot's sloppy residual
//    stops it (below delta times its last FP64 residual, or the stopping
//    rule): x += xs and r = b - M^dag M x in FP64, the sloppy CG continuing
//    from that r; a slot finishes only on its FP64 residual.
struct Engine {
  int gauge_id = -1;
  cps_hip_op op;
  int max_iter = 0;
  int prec = CPS_HIP_PREC_DOUBLE;
  int every = 100;
  double delta = 0.1;
  int S = 0;
  bool mixed = false, slot_mass = false;
  int hprec = PREC_D_D18, sps = STORE_D, svec = STORE_D, xvec = STORE_D;
  Lattice L;
  OpLayout o;
  int bsb = 0, bpr_blas = 0, bpr_hop = 0;
  double* part = nullptr;
  CGState cg;
  Field *x = nullptr, *b = nullptr, *r = nullptr, *p = nullptr, *ap = nullptr, *t1 = nullptr, *t2 = nullptr,
        *t3 = nullptr;
  Field *xs = nullptr, *rs = nullptr, *ps = nullptr, *aps = nullptr, *t1s = nullptr, *t2s = nullptr;
  DevBuf mass_buf;   // per slot: kappa^2 [0, S), wall coefficient [S, 2S)
  DevBuf coef_buf;
  std::vector<Request*> owner;
  std::vector<int> rhs_of, updates;
  int since = 0;     // mixed: iterations since the last reliable update
  // mixed: a slot stopped and the FP64 steps must run before it is harvested
  bool need_update = false;
  // mixed: slot s has had no sloppy iteration since the last reliable update
  // (its solution and FP64 residual are current)
  std::vector<char> fresh;
  long long bytes = 0;
  clk::time_point idle_since = clk::now();
  // the iteration graph of each slot window [0, nw) (captured on first use)
  std::vector<std::unique_ptr<IterGraph>> graphs;
  // the engine's last chunk of iterations is on the GPU until `done` fires
  hipEvent_t done = nullptr;
  bool in_flight = false;
  // a host worker is uploading sources into the engine or copying its
  // results out (the service thread leaves the engine alone meanwhile)
  bool busy_host = false;

  ~Engine() {
    if (done) (void)hipEventDestroy(done);
  }

  static bool is_mixed(int prec) { return prec != CPS_HIP_PREC_DOUBLE; }

  static long long footprint(const GaugeEntry& ge, const cps_hip_op& op, int S, int prec) {
    const long long V = (long long)ge.L[0] * ge.L[1] * ge.L[2] * ge.L[3];
    const long long per_rhs = (op.kind == CPS_HIP_OP_DWF || op.kind == CPS_HIP_OP_MOBIUS) ? V / 2 * ge.ls : V / 2;
    const long long vec = per_rhs * 192;
    // FP64 fields x, b, r, Ap, t1, t2, t3 (+ p for FP64 CG), the CPS-order
    // staging vector, links and neighbour table
    long long b = (is_mixed(prec) ? 7 : 8) * vec * S + vec + 2 * V * 72 * 8 + V * 48 + 64 * 1024 * 1024;
    if (is_mixed(prec)) b += S * per_rhs * (96 + 5 * 96) + 2 * V * 72 * 4;   // sloppy fields (at most FP32), links
    return b;
  }

  const double* k2() const { return slot_mass ? (const double*)mass_buf.p : nullptr; }
  const double* wall() const { return slot_mass ? (const double*)mass_buf.p + S : nullptr; }

  // M^dag M on slots [lo, lo + n) of the engine's fields
  int mdagm(int hp, const FieldRef& out, const FieldRef& in, const FieldRef& a, const FieldRef& c, double* pt, int n,
            int lo) {
    if (slot_mass) return apply_mdagm_slots(L, &op, hp, out, in, a, c, pt, n, k2(), wall(), lo);
    return apply_mdagm(L, &op, hp, out, in, a, c, pt, n);
  }

  int setup(const GaugeEntry& ge, int gid, const Request& rq, int S_) {
    gauge_id = gid;
    op = rq.op;
    S = S_;
    max_iter = rq.max_iter;
    prec = rq.precision;
    every = rq.reliable_every > 0 ? rq.reliable_every : 100;
    delta = rq.reliable_delta > 0 ? rq.reliable_delta : 0.1;
    mixed = is_mixed(prec);
    slot_mass = per_slot_mass(op);
    int rc = L.init(ge.L, ge.ls, ge.bc);
    if (rc) return rc;
    if ((rc = L.import_gauge_device((const double*)ge.u.p))) return rc;
    if ((rc = op_layout(&op, L.ls, &o, L.g))) return rc;
    if (op.kind == CPS_HIP_OP_CLOVER) {
      const double kappa = op_kappa(&op);
      if ((rc = clover_build(L, &kappa, 1, op.csw))) return rc;
    }
    const int64_t n = o.per_rhs * S;
    x = &L.field(S_X, STORE_D, n);
    b = &L.field(S_SRC, STORE_D, n);
    r = &L.field(S_R, STORE_D, n);
    if (!is_mixed(prec)) p = &L.field(S_P, STORE_D, n);   // mixed: the direction is sloppy (ps)
    ap = &L.field(S_AP, STORE_D, n);
    t1 = &L.field(S_T1, STORE_D, n);
    t2 = &L.field(S_T2, STORE_D, n);
    t3 = &L.field(S_T3, STORE_D, n);
    if (mixed) {
      // the batch solver's formats for this mode with S rhs; the sloppy
      // residual in the direction's storage, the sloppy solution FP32
      precision_formats(prec, &op, L.ls, S, L.recon_ok, &sps, &hprec);
      svec = sps;
      xvec = STORE_F;
      xs = &L.field(S_XS, xvec, n);
      rs = &L.field(S_RS, svec, n);
      ps = &L.field(S_PS, sps, n);
      aps = &L.field(S_APS, sps, n);
      t1s = &L.field(S_T1S, sps, n);
      t2s = &L.field(S_T2S, sps, n);
    }
    if (L.last_rc) return L.last_rc;
    bsb = pick_block(o.per_rhs);
    bpr_blas = (int)(o.per_rhs / bsb);
    bpr_hop = hop_blocks_per_rhs(L, &op);
    if ((rc = L.ensure_partial((int64_t)std::max(bpr_blas, bpr_hop) * S))) return rc;
    part = (double*)L.partial.p;
    if ((rc = cg.alloc(L, S, 0))) return rc;
    cg.s.max_itr = max_iter - 1;
    if (mixed) {
      cg.s.rr_upd = cg.upd_store;
      cg.s.delta2 = delta * delta;
    }
    if ((rc = mass_buf.ensure(2 * S * sizeof(double)))) return rc;
    if ((rc = coef_buf.ensure(S * sizeof(double)))) return rc;
    if ((rc = L.io.ensure((size_t)o.per_rhs * 24 * sizeof(double)))) return rc;
    // empty slots hold zeros and are inactive (and converged, so that a
    // reliable update leaves them alone)
    for (Field* f : {x, b, r, p, ap, t1, t2, t3})
      if (f) CPSHIP_CHECK(blas_zero(STORE_D, f->ref(), o.per_rhs, S, bsb, L.stream));
    if (mixed)
      for (Field* f : {xs, rs, ps, aps, t1s, t2s})
        CPSHIP_CHECK(blas_zero(f->storage, f->ref(), o.per_rhs, S, bsb, L.stream));
    std::vector<int> zeros(S, 0), conv(S, CG_STATUS_CONVERGED);
    std::vector<double> dz(2 * S, 0.0);
    CPSHIP_CHECK(hipMemcpyAsync(cg.s.active, zeros.data(), S * sizeof(int), hipMemcpyHostToDevice, L.stream));
    CPSHIP_CHECK(hipMemcpyAsync(cg.s.iters, zeros.data(), S * sizeof(int), hipMemcpyHostToDevice, L.stream));
    CPSHIP_CHECK(hipMemcpyAsync(cg.s.status, conv.data(), S * sizeof(int), hipMemcpyHostToDevice, L.stream));
    for (double* d : {cg.s.stp, cg.s.src_norm, cg.s.alpha, cg.s.beta, cg.s.rr, cg.s.rr_prev})
      CPSHIP_CHECK(hipMemcpyAsync(d, dz.data(), S * sizeof(double), hipMemcpyHostToDevice, L.stream));
    CPSHIP_CHECK(hipMemcpyAsync(mass_buf.p, dz.data(), 2 * S * sizeof(double), hipMemcpyHostToDevice, L.stream));
    CPSHIP_CHECK(hipStreamSynchronize(L.stream));
    owner.assign(S, nullptr);
    rhs_of.assign(S, -1);
    updates.assign(S, 0);
    fresh.assign(S, 0);
    bytes = L.device_bytes();
    graphs.clear();
    graphs.resize(S + 1);
    CPSHIP_CHECK(hipEventCreateWithFlags(&done, hipEventDisableTiming));
    return 0;
  }

  int free_slots() const {
    int n = 0;
    for (auto* q : owner) n += q == nullptr;
    return n;
  }
  bool busy() const { return free_slots() < S; }

  // slot s assigned to rhs k of request rq (service thread; the device
  // work is upload())
  void claim(int s, Request* rq, int k) {
    owner[s] = rq;
    rhs_of[s] = k;
    updates[s] = 0;
    fresh[s] = 0;
  }

  // rhs k of request rq into slot s: x = 0, r = p = src, |src|^2, stp, |r|^2
  // (mixed: the sloppy residual = src, xs = 0, p = 0 with beta = 0, and the
  // slot's last FP64 residual = |src|^2); the slot's mass.  Copies from the
  // caller's memory, so it runs on a host worker.
  int upload(int s, Request* rq, int k) {
    const double* h = rq->src + (size_t)k * o.per_rhs * 24;
    CPSHIP_CHECK(hipMemcpyAsync(L.io.p, h, (size_t)o.per_rhs * 24 * sizeof(double), hipMemcpyHostToDevice, L.stream));
    const FieldRef bs = rhs_view(b->ref(), o.per_rhs, s), xs0 = rhs_view(x->ref(), o.per_rhs, s);
    const FieldRef rs0 = rhs_view(r->ref(), o.per_rhs, s);
    CPSHIP_CHECK(pack_spinor(STORE_D, (const double*)L.io.p, bs, o.map, L.g.Vcb, L.ls, 1, o.per_rhs, L.stream));
    CPSHIP_CHECK(blas_zero(STORE_D, xs0, o.per_rhs, 1, bsb, L.stream));
    CPSHIP_CHECK(convert_spinor(STORE_D, bs, STORE_D, rs0, o.per_rhs, L.stream));
    if (p) CPSHIP_CHECK(convert_spinor(STORE_D, bs, STORE_D, rhs_view(p->ref(), o.per_rhs, s), o.per_rhs, L.stream));
    if (slot_mass) {
      const double kap = 1.0 / (2.0 * (rq->op.mass + 4.0));
      const double k2v = kap * kap, wv = -2.0 * rq->op.mass * op.a5_inv;
      CPSHIP_CHECK(hipMemcpyAsync((double*)mass_buf.p + s, &k2v, sizeof(double), hipMemcpyHostToDevice, L.stream));
      CPSHIP_CHECK(hipMemcpyAsync((double*)mass_buf.p + S + s, &wv, sizeof(double), hipMemcpyHostToDevice, L.stream));
    }
    CPSHIP_CHECK(hipMemcpyAsync(cg.s.rsd + s, &rq->stop_rsd, sizeof(double), hipMemcpyHostToDevice, L.stream));
    const CGScalars v = scalars_view(cg.s, s);
    CPSHIP_CHECK(blas_norm(STORE_D, bs, o.per_rhs, 1, bsb, part, L.stream));
    CPSHIP_CHECK(reduce_rhs(part, bpr_blas, 1, nullptr, RED_SRC_NORM, v, L.stream));
    CPSHIP_CHECK(blas_norm(STORE_D, rs0, o.per_rhs, 1, bsb, part, L.stream));
    CPSHIP_CHECK(reduce_rhs(part, bpr_blas, 1, mixed ? v.rr_upd : nullptr, RED_CG_INIT, v, L.stream));
    if (mixed) {
      // a reliable update may come before the slot's first iteration (one
      // requested for the other slots): its beta = |r|^2 / rr_prev must then
      // be finite (p is zero, so the value itself does not matter)
      CPSHIP_CHECK(hipMemcpyAsync(cg.s.rr_prev + s, cg.s.rr + s, sizeof(double), hipMemcpyDeviceToDevice, L.stream));
      CPSHIP_CHECK(convert_spinor(STORE_D, bs, svec, rhs_view(rs->ref(), o.per_rhs, s), o.per_rhs, L.stream));
      CPSHIP_CHECK(blas_zero(xvec, rhs_view(xs->ref(), o.per_rhs, s), o.per_rhs, 1, bsb, L.stream));
      CPSHIP_CHECK(blas_zero(sps, rhs_view(ps->ref(), o.per_rhs, s), o.per_rhs, 1, bsb, L.stream));
      const double z = 0.0;
      CPSHIP_CHECK(hipMemcpyAsync(cg.s.beta + s, &z, sizeof(double), hipMemcpyHostToDevice, L.stream));
      CPSHIP_CHECK(hipMemcpyAsync(cg.s.alpha + s, &z, sizeof(double), hipMemcpyHostToDevice, L.stream));
    }
    return 0;
  }

  // the slot window [0, nw): every slot in use lies below nw (slots are
  // filled lowest first), so the kernels leave out the empty slots above
  int window() const {
    int nw = 0;
    for (int s = 0; s < S; s++)
      if (owner[s]) nw = s + 1;
    return nw;
  }

  IterGraph& graph_for(int nw) {
    if (!graphs[nw]) graphs[nw].reset(new IterGraph(true));
    return *graphs[nw];
  }

  // up to `iters` CG iterations over the slot window (inactive slots are
  // frozen by the per-slot scalars), enqueued on the engine's stream;
  // returns the number run (mixed: up to the next reliable update).  Mixed:
  // when a slot has stopped or `every` iterations have passed, the step is
  // the reliable update alone, so the service thread never waits for it.
  int step(int iters, int* ran) {
    const int nw = window();
    *ran = 0;
    if (nw == 0) return 0;
    if (!mixed) {
      *ran = iters;
      return graph_for(nw).run(L, iters, [&]() -> int {
        int e = mdagm(PREC_D_D18, ap->ref(), p->ref(), t1->ref(), t2->ref(), part, nw, 0);
        if (e) return e;
        CPSHIP_CHECK(reduce_rhs(part, bpr_hop, nw, nullptr, RED_CG_ALPHA, cg.s, L.stream));
        CPSHIP_CHECK(blas_cg_xr_d(x->ref(), p->ref(), r->ref(), ap->ref(), cg.s, o.per_rhs, nw, bsb, part, L.stream));
        CPSHIP_CHECK(reduce_rhs(part, bpr_blas, nw, nullptr, RED_CG_BETA, cg.s, L.stream));
        CPSHIP_CHECK(blas_xpay(STORE_D, p->ref(), r->ref(), cg.s.beta, cg.s.active, o.per_rhs, nw, bsb, L.stream));
        return 0;
      });
    }
    if (need_update || since >= every) {
      need_update = false;
      for (int s = 0; s < S; s++) fresh[s] = owner[s] != nullptr;
      return reliable_update();
    }
    const int n = std::max(1, std::min(iters, every - since));
    *ran = n;
    since += n;
    std::fill(fresh.begin(), fresh.end(), 0);
    // one sloppy iteration: p = r + beta p, A p, alpha, xs += alpha p and
    // r -= alpha A p with |r|^2, beta (the batch solver's unfused order)
    return graph_for(nw).run(L, n, [&]() -> int {
      CPSHIP_CHECK(blas_xpay_mixed(sps, ps->ref(), svec, rs->ref(), cg.s.beta, cg.s.active, o.per_rhs, nw, bsb,
                                   L.stream));
      int e = mdagm(hprec, aps->ref(), ps->ref(), t1s->ref(), t2s->ref(), part, nw, 0);
      if (e) return e;
      CPSHIP_CHECK(reduce_rhs(part, bpr_hop, nw, nullptr, RED_CG_ALPHA, cg.s, L.stream));
      CPSHIP_CHECK(blas_cg_xr_mixed(xvec, xs->ref(), sps, ps->ref(), svec, rs->ref(), sps, aps->ref(), cg.s, o.per_rhs,
                                    nw, bsb, part, L.stream));
      CPSHIP_CHECK(reduce_rhs(part, bpr_blas, nw, nullptr, RED_CG_BETA, cg.s, L.stream));
      return 0;
    });
  }

  // the FP64-critical steps for every slot: x += xs, r = b - M^dag M x,
  // |r|^2 into rr_true, the sloppy residual restarted from r, p
  // re-orthogonalised; slots whose FP64 residual meets the stopping rule
  // finish (mixed_after_update)
  int reliable_update() {
    const int nw = window();
    if (nw == 0) return 0;
    CPSHIP_CHECK(blas_axpy(xvec, xs->ref(), STORE_D, x->ref(), nullptr, 1.0, o.per_rhs, nw, bsb, L.stream));
    CPSHIP_CHECK(blas_zero(xvec, xs->ref(), o.per_rhs, nw, bsb, L.stream));
    int e = mdagm(PREC_D_D18, ap->ref(), x->ref(), t1->ref(), t2->ref(), nullptr, nw, 0);
    if (e) return e;
    CPSHIP_CHECK(blas_sub_norm_d(b->ref(), ap->ref(), r->ref(), o.per_rhs, nw, bsb, part, L.stream));
    CPSHIP_CHECK(reduce_rhs(part, bpr_blas, nw, cg.rr_true, RED_PLAIN, cg.s, L.stream));
    CPSHIP_CHECK(convert_spinor(STORE_D, r->ref(), svec, rs->ref(), o.per_rhs * nw, L.stream));
    CPSHIP_CHECK(blas_redot_mixed(svec, rs->ref(), sps, ps->ref(), o.per_rhs, nw, bsb, part, L.stream));
    CPSHIP_CHECK(reduce_rhs(part, bpr_blas, nw, cg.s.alpha, RED_PLAIN, cg.s, L.stream));
    double* coef = (double*)coef_buf.p;
    CPSHIP_CHECK(mixed_after_update(cg.s, cg.rr_true, cg.s.alpha, coef, nw, L.stream));
    CPSHIP_CHECK(blas_axpy(svec, rs->ref(), sps, ps->ref(), coef, 0.0, o.per_rhs, nw, bsb, L.stream));
    for (int s = 0; s < S; s++)
      if (owner[s]) updates[s]++;
    since = 0;
    return 0;
  }

  // Slots whose CG stopped, in three parts.  scan() (service thread, when the
  // engine's last step has finished and its stream is idle) reads the slots'
  // flags with one synchronisation and lists the finished slots.  Mixed: a
  // stopped slot is finished only once a reliable update has folded its
  // sloppy solution into x and decided on its FP64 residual (which restarts
  // it if the residual is not small enough); until then it asks for that
  // update, which the next step runs.
  int scan(std::vector<int>& done, std::vector<int>& st_done) {
    std::vector<int> act(S), st(S);
    CPSHIP_CHECK(hipMemcpyAsync(act.data(), cg.s.active, S * sizeof(int), hipMemcpyDeviceToHost, L.stream));
    CPSHIP_CHECK(hipMemcpyAsync(st.data(), cg.s.status, S * sizeof(int), hipMemcpyDeviceToHost, L.stream));
    CPSHIP_CHECK(hipStreamSynchronize(L.stream));
    done.clear();
    st_done.clear();
    for (int s = 0; s < S; s++) {
      if (!owner[s] || act[s]) continue;
      if (mixed && (!fresh[s] || st[s] == CG_STATUS_NEED_UPDATE || st[s] == CG_STATUS_RUNNING)) {
        need_update = true;
        continue;
      }
      if (st[s] != CG_STATUS_NEED_UPDATE && st[s] != CG_STATUS_RUNNING) {
        done.push_back(s);
        st_done.push_back(st[s]);
      }
    }
    return 0;
  }

  // collect() (host worker): the FP64 true residual of the finished slots
  // (inv_cg.C:403-417; mixed engines have it from the reliable update), their
  // counters, and their solutions copied to the callers' memory, all enqueued
  // together and read back with one synchronisation.
  int collect(const std::vector<int>& done, std::vector<double>& tr, std::vector<double>& sn, std::vector<int>& it) {
    double* trn = cg.rr_true;
    if (!mixed) {
      for (int s : done) {
        const FieldRef xw = rhs_view(x->ref(), o.per_rhs, s), aw = rhs_view(ap->ref(), o.per_rhs, s);
        int e = mdagm(PREC_D_D18, aw, xw, t1->ref(), t2->ref(), nullptr, 1, s);
        if (e) return e;
        CPSHIP_CHECK(blas_sub_norm_d(rhs_view(b->ref(), o.per_rhs, s), aw, rhs_view(t3->ref(), o.per_rhs, s), o.per_rhs,
                                     1, bsb, part, L.stream));
        CPSHIP_CHECK(reduce_rhs(part, bpr_blas, 1, trn + s, RED_PLAIN, scalars_view(cg.s, s), L.stream));
      }
    }
    tr.assign(S, 0.0);
    sn.assign(S, 0.0);
    it.assign(S, 0);
    CPSHIP_CHECK(hipMemcpyAsync(tr.data(), trn, S * sizeof(double), hipMemcpyDeviceToHost, L.stream));
    CPSHIP_CHECK(hipMemcpyAsync(sn.data(), cg.s.src_norm, S * sizeof(double), hipMemcpyDeviceToHost, L.stream));
    CPSHIP_CHECK(hipMemcpyAsync(it.data(), cg.s.iters, S * sizeof(int), hipMemcpyDeviceToHost, L.stream));
    // solutions in CPS order: slot s unpacked into the t3 field's memory at
    // slot s (FP64, per_rhs * 24 doubles each), then copied out
    double* stage = static_cast<double*>(t3->ref().data);
    for (int s : done) {
      double* dst = stage + (size_t)s * o.per_rhs * 24;
      CPSHIP_CHECK(unpack_spinor(STORE_D, rhs_view(x->ref(), o.per_rhs, s), dst, o.map, L.g.Vcb, L.ls, 1, o.per_rhs,
                                 L.stream));
      Request* rq = owner[s];
      CPSHIP_CHECK(hipMemcpyAsync(rq->sol + (size_t)rhs_of[s] * o.per_rhs * 24, dst,
                                  (size_t)o.per_rhs * 24 * sizeof(double), hipMemcpyDeviceToHost, L.stream));
    }
    CPSHIP_CHECK(hipStreamSynchronize(L.stream));
    return 0;
  }

  // a host job failed with rc: its slots are freed and their requests fail
  void fail(const std::vector<int>& slots, int rc, std::vector<Request*>& touched) {
    for (int s : slots) {
      Request* rq = owner[s];
      if (!rq) continue;
      rq->rc = rc;
      rq->done_rhs++;
      owner[s] = nullptr;
      rhs_of[s] = -1;
      touched.push_back(rq);
    }
    if (!busy()) idle_since = clk::now();
  }

  // release() (service thread): the results into the requests, slots freed
  void release(const std::vector<int>& done, const std::vector<int>& st, const std::vector<double>& tr,
               const std::vector<double>& sn, const std::vector<int>& it, std::vector<Request*>& touched) {
    for (size_t i = 0; i < done.size(); i++) {
      const int s = done[i];
      Request* rq = owner[s];
      const int k = rhs_of[s];
      cps_hip_cg_result& res = rq->res[k];
      res.iterations = it[s] + 1;
      res.status = st[i];
      res.src_norm_sq = sn[s];
      res.true_res = sn[s] > 0 ? std::sqrt(tr[s] / sn[s]) : 0.0;
      res.reliable_updates = mixed ? updates[s] : 0;
      if (st[i] == CG_STATUS_D_ZERO) rq->rc = CPS_HIP_ERR_D_ZERO;
      rq->done_rhs++;
      owner[s] = nullptr;
      rhs_of[s] = -1;
      touched.push_back(rq);
    }
    if (!busy()) idle_since = clk::now();
  }
};

// Host work of one engine done by a worker thread: sources uploaded into
// claimed slots, or the results of finished slots copied to the callers
struct HostJob {
  Engine* e = nullptr;
  bool harvest = false;
  std::vector<int> slots;            // admit: the claimed slots; harvest: the finished ones
  std::vector<Request*> rqs;         // admit: request and rhs of each slot
  std::vector<int> ks;
  std::vector<int> st;               // harvest: CG status of each finished slot
  std::vector<double> tr, sn;        // harvest: per slot of the engine
  std::vector<int> it;
  int rc = 0;
};

}  // namespace

// an engine with no slot in use for this long is released
constexpr double kIdleReleaseSeconds = 30.0;
// host worker threads: copies between the callers' memory and the device run
// there, so a large lattice's transfers never hold up the other engines
constexpr int kHostWorkers = 4;

struct Service {
  long long budget = 0;
  int max_slots = 12;
  int chunk = 8;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> incoming;
  std::map<long long, std::unique_ptr<Request>> requests;
  std::vector<std::unique_ptr<GaugeEntry>> gauges;
  std::vector<std::unique_ptr<Engine>> engines;
  std::deque<Request*> pending;   // service thread only
  long long next_id = 1;
  bool stop = false;
  int fatal = 0;
  std::thread th;
  cps_hip_service_stats stats;
  int device = 0;   // the caller's device at creation; every service thread uses it
  // host jobs: queued by the service thread, run by the workers, their
  // results applied by the service thread
  std::mutex job_mu;
  std::condition_variable job_cv;
  std::deque<std::unique_ptr<HostJob>> jobs, jobs_done;
  std::vector<std::thread> workers;
  bool workers_stop = false;
  int jobs_out = 0;   // service thread only

  void worker() {
    (void)hipSetDevice(device);
    for (;;) {
      std::unique_ptr<HostJob> j;
      {
        std::unique_lock<std::mutex> lk(job_mu);
        job_cv.wait(lk, [&] { return workers_stop || !jobs.empty(); });
        if (jobs.empty()) return;
        j = std::move(jobs.front());
        jobs.pop_front();
      }
      Engine* e = j->e;
      if (j->harvest) {
        j->rc = e->collect(j->slots, j->tr, j->sn, j->it);
      } else {
        for (size_t i = 0; i < j->slots.size() && !j->rc; i++) j->rc = e->upload(j->slots[i], j->rqs[i], j->ks[i]);
        if (!j->rc && hipStreamSynchronize(e->L.stream) != hipSuccess) j->rc = CPS_HIP_ERR_RUNTIME;
      }
      {
        std::lock_guard<std::mutex> lk(job_mu);
        jobs_done.push_back(std::move(j));
      }
    }
  }

  void dispatch(std::unique_ptr<HostJob> j) {
    j->e->busy_host = true;
    jobs_out++;
    {
      std::lock_guard<std::mutex> lk(job_mu);
      jobs.push_back(std::move(j));
    }
    job_cv.notify_one();
  }

  long long used_bytes() const {
    long long b = 0;
    for (auto& g : gauges) b += (long long)g->u.bytes;
    for (auto& e : engines) b += e->bytes;
    return b;
  }

  Engine* find_engine(const Request* rq) {
    for (auto& e : engines)
      if (e->gauge_id == rq->gauge_id && e->max_iter == rq->max_iter && e->prec == rq->precision &&
          (rq->precision == CPS_HIP_PREC_DOUBLE ||
           (e->every == rq->reliable_every && e->delta == rq->reliable_delta)) &&
          same_engine_op(e->op, rq->op))
        return e.get();
    return nullptr;
  }

  // creates an engine for rq if the HBM budget allows it, releasing idle
  // engines first when needed
  int make_engine(Request* rq, Engine** out) {
    *out = nullptr;
    const GaugeEntry& ge = *gauges[rq->gauge_id];
    const int slots = max_slots;
    const long long need = Engine::footprint(ge, rq->op, slots, rq->precision);
    while (used_bytes() + need > budget) {
      auto idle = std::find_if(engines.begin(), engines.end(), [](const std::unique_ptr<Engine>& e) {
        return !e->busy() && !e->in_flight && !e->busy_host;
      });
      if (idle == engines.end()) return 0;   // wait: no room now
      engines.erase(idle);
    }
    std::unique_ptr<Engine> e(new Engine);
    int rc = e->setup(ge, rq->gauge_id, *rq, slots);
    if (rc) return rc;
    *out = e.get();
    engines.push_back(std::move(e));
    std::lock_guard<std::mutex> lk(mu);
    stats.engines_created++;
    return 0;
  }

  // rq complete: wait() may free it from here on, so this is the service
  // thread's last access to it
  void finish(Request* rq) {
    std::lock_guard<std::mutex> lk(mu);
    rq->finished = true;
    rq->t_done = clk::now();
    stats.requests_done++;
    stats.rhs_done += rq->nrhs;
    cv_done.notify_all();
  }

  // the requests whose slots just finished (a request appears once per
  // slot): each complete one is finished once, and none is read after that
  void finish_touched(std::vector<Request*>& touched) {
    std::sort(touched.begin(), touched.end());
    touched.erase(std::unique(touched.begin(), touched.end()), touched.end());
    for (Request* rq : touched)
      if (rq->done_rhs == rq->nrhs && !rq->finished) finish(rq);
  }

  void loop() {
    (void)hipSetDevice(device);
    for (int w = 0; w < kHostWorkers; w++) workers.emplace_back([this] { worker(); });
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu);
        const bool idle = pending.empty() && jobs_out == 0 &&
                          std::none_of(engines.begin(), engines.end(),
                                       [](const std::unique_ptr<Engine>& e) { return e->busy(); });
        if (idle) cv_work.wait(lk, [&] { return stop || !incoming.empty(); });
        if (stop && incoming.empty() && idle) break;
        while (!incoming.empty()) {
          pending.push_back(incoming.front());
          incoming.pop_front();
        }
      }
      bool progress = false;
      // finished host jobs: uploads done (the engine may step), results
      // copied out (slots freed, requests completed)
      std::deque<std::unique_ptr<HostJob>> fin;
      {
        std::lock_guard<std::mutex> lk(job_mu);
        fin.swap(jobs_done);
      }
      for (auto& j : fin) {
        jobs_out--;
        progress = true;
        j->e->busy_host = false;
        std::vector<Request*> touched;
        if (j->rc) {
          fatal = j->rc;
          j->e->fail(j->slots, j->rc, touched);
        } else if (j->harvest) {
          j->e->release(j->slots, j->st, j->tr, j->sn, j->it, touched);
        }
        finish_touched(touched);
      }
      // Engines run independently: an engine whose last chunk finished (its
      // event fired) is scanned for finished slots (their results go to a
      // host worker), takes waiting rhs into its free slots (uploaded by a
      // worker) and gets its next chunk; an engine on the GPU or with a host
      // job is left alone, so a small lattice's engine cycles at its own pace
      // beside a large one and beside any engine's transfers.
      for (auto& e : engines) {
        if (!e->in_flight) continue;
        const hipError_t q = hipEventQuery(e->done);
        if (q == hipErrorNotReady) continue;
        (void)hipGetLastError();
        e->in_flight = false;
        progress = true;
        std::unique_ptr<HostJob> j(new HostJob);
        j->e = e.get();
        j->harvest = true;
        if (int rc = e->scan(j->slots, j->st)) fatal = rc;
        if (!j->slots.empty()) dispatch(std::move(j));
      }
      // admission: FIFO over waiting requests, one rhs per free slot of an
      // engine that is neither on the GPU nor in a host job; the slots are
      // claimed here and their sources uploaded by one job per engine
      std::map<Engine*, std::unique_ptr<HostJob>> uploads;
      for (auto it = pending.begin(); it != pending.end();) {
        Request* rq = *it;
        Engine* e = find_engine(rq);
        if (e && (e->in_flight || e->busy_host)) {
          ++it;
          continue;
        }
        int rc = e ? 0 : make_engine(rq, &e);
        if (!rc && !e && engines.empty()) rc = CPS_HIP_ERR_NOMEM;   // can never fit the budget
        while (!rc && e && rq->next_rhs < rq->nrhs && e->free_slots() > 0) {
          int slot = 0;
          while (e->owner[slot]) slot++;
          std::unique_ptr<HostJob>& j = uploads[e];
          if (!j) {
            j.reset(new HostJob);
            j->e = e;
          }
          e->claim(slot, rq, rq->next_rhs);
          j->slots.push_back(slot);
          j->rqs.push_back(rq);
          j->ks.push_back(rq->next_rhs);
          rq->next_rhs++;
        }
        if (rc) {
          // the request fails with rc: its remaining rhs are not solved, the
          // ones already in slots finish normally
          rq->rc = rc;
          rq->done_rhs += rq->nrhs - rq->next_rhs;
          rq->next_rhs = rq->nrhs;
        }
        if (rq->next_rhs >= rq->nrhs) {
          it = pending.erase(it);
          // last access: once finished, wait() may free the request
          if (rc && rq->done_rhs == rq->nrhs) finish(rq);
        } else {
          ++it;
        }
      }
      for (auto& kv : uploads) {
        progress = true;
        dispatch(std::move(kv.second));
      }
      // every busy engine that is neither on the GPU nor in a host job gets
      // its next chunk
      long long stepped = 0;
      for (auto& e : engines)
        if (e->busy() && !e->in_flight && !e->busy_host) {
          int ran = 0;
          if (int rc = e->step(chunk, &ran)) fatal = rc;
          if (hipEventRecord(e->done, e->L.stream) != hipSuccess) fatal = CPS_HIP_ERR_RUNTIME;
          e->in_flight = true;
          stepped += ran;
          progress = true;
        }
      if (!progress) std::this_thread::sleep_for(std::chrono::microseconds(50));
      // release engines idle for a while (freeing device memory waits for
      // the whole device, so an engine that may soon be needed again is kept;
      // make_engine also releases idle engines when the budget requires it)
      const auto now = clk::now();
      engines.erase(std::remove_if(engines.begin(), engines.end(),
                                   [&](const std::unique_ptr<Engine>& e) {
                                     return !e->busy() && !e->in_flight && !e->busy_host &&
                                            std::chrono::duration<double>(now - e->idle_since).count() >
                                                kIdleReleaseSeconds;
                                   }),
                    engines.end());
      {
        std::lock_guard<std::mutex> lk(mu);
        stats.iterations += stepped;
        stats.engines_live = (int)engines.size();
        stats.peak_bytes = std::max(stats.peak_bytes, used_bytes());
        if (fatal && !stats.last_error) stats.last_error = fatal;
      }
    }
    {
      std::lock_guard<std::mutex> lk(job_mu);
      workers_stop = true;
    }
    job_cv.notify_all();
    for (auto& w : workers) w.join();
    workers.clear();
  }
};

}  // namespace cpship

struct cps_hip_service {
  cpship::Service S;
};

extern "C" {

int cps_hip_service_create(cps_hip_service** s, long long hbm_budget_bytes, int max_slots) {
  if (!s || hbm_budget_bytes <= 0 || max_slots < 1) return CPS_HIP_ERR_ARG;
  *s = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return CPS_HIP_ERR_NODEV;
  }
  cps_hip_service* h = new (std::nothrow) cps_hip_service;
  if (!h) return CPS_HIP_ERR_NOMEM;
  h->S.budget = hbm_budget_bytes;
  h->S.max_slots = max_slots;
  if (hipGetDevice(&h->S.device) != hipSuccess) {
    (void)hipGetLastError();
    delete h;
    return CPS_HIP_ERR_RUNTIME;
  }
  std::memset(&h->S.stats, 0, sizeof(h->S.stats));
  h->S.th = std::thread([h] { h->S.loop(); });
  *s = h;
  return 0;
}

int cps_hip_service_destroy(cps_hip_service* s) {
  if (!s) return CPS_HIP_ERR_ARG;
  {
    std::lock_guard<std::mutex> lk(s->S.mu);
    s->S.stop = true;
  }
  s->S.cv_work.notify_all();
  s->S.th.join();
  delete s;
  return 0;
}

int cps_hip_service_add_gauge(cps_hip_service* s, const int node_sites[4], int ls, const int bc[4], const double* u,
                              int* gauge_id) {
  if (!s || !node_sites || !u || !gauge_id || ls < 1) return CPS_HIP_ERR_ARG;
  std::unique_ptr<cpship::GaugeEntry> g(new cpship::GaugeEntry);
  long long V = 1;
  for (int mu = 0; mu < 4; mu++) {
    if (node_sites[mu] <= 0 || (node_sites[mu] & 1)) return CPS_HIP_ERR_ARG;
    g->L[mu] = node_sites[mu];
    g->bc[mu] = bc ? bc[mu] : 0;
    V *= node_sites[mu];
  }
  g->ls = ls;
  const size_t bytes = (size_t)V * 72 * sizeof(double);
  if (int rc = g->u.ensure(bytes)) return rc;
  if (hipMemcpy(g->u.p, u, bytes, hipMemcpyHostToDevice) != hipSuccess) return CPS_HIP_ERR_RUNTIME;
  std::lock_guard<std::mutex> lk(s->S.mu);
  *gauge_id = (int)s->S.gauges.size();
  s->S.gauges.push_back(std::move(g));
  return 0;
}

int cps_hip_service_submit(cps_hip_service* s, int gauge_id, const cps_hip_op* op, const double* src, double* sol,
                           int nrhs, double stop_rsd, int max_num_iter, long long* request_id) {
  cps_hip_cg_arg arg;
  std::memset(&arg, 0, sizeof(arg));
  arg.max_num_iter = max_num_iter;
  arg.stop_rsd = stop_rsd;
  arg.precision = CPS_HIP_PREC_DOUBLE;
  return cps_hip_service_submit_arg(s, gauge_id, op, src, sol, nrhs, &arg, request_id);
}

int cps_hip_service_submit_arg(cps_hip_service* s, int gauge_id, const cps_hip_op* op, const double* src, double* sol,
                               int nrhs, const cps_hip_cg_arg* arg, long long* request_id) {
  if (!s || !op || !src || !sol || nrhs < 1 || !request_id || !arg || arg->max_num_iter < 1) return CPS_HIP_ERR_ARG;
  if (arg->precision < CPS_HIP_PREC_DOUBLE || arg->precision > CPS_HIP_PREC_MIXED_SINGLE) return CPS_HIP_ERR_ARG;
  std::unique_ptr<cpship::Request> rq(new cpship::Request);
  rq->gauge_id = gauge_id;
  rq->op = *op;
  rq->src = src;
  rq->sol = sol;
  rq->nrhs = nrhs;
  rq->stop_rsd = arg->stop_rsd;
  rq->max_iter = arg->max_num_iter;
  rq->precision = arg->precision;
  rq->reliable_every = arg->reliable_every > 0 ? arg->reliable_every : 100;
  rq->reliable_delta = arg->reliable_delta > 0 ? arg->reliable_delta : 0.1;
  rq->res.resize(nrhs);
  rq->t_submit = clk::now();
  {
    std::lock_guard<std::mutex> lk(s->S.mu);
    if (gauge_id < 0 || gauge_id >= (int)s->S.gauges.size()) return CPS_HIP_ERR_ARG;
    cpship::OpLayout o;
    if (int rc = cpship::op_layout(op, s->S.gauges[gauge_id]->ls, &o, cpship::Geom{})) return rc;
    rq->id = s->S.next_id++;
    *request_id = rq->id;
    s->S.incoming.push_back(rq.get());
    s->S.requests[rq->id] = std::move(rq);
  }
  s->S.cv_work.notify_all();
  return 0;
}

int cps_hip_service_wait(cps_hip_service* s, long long request_id, cps_hip_cg_result* res, double* latency_s) {
  if (!s) return CPS_HIP_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->S.mu);
  auto it = s->S.requests.find(request_id);
  if (it == s->S.requests.end()) return CPS_HIP_ERR_ARG;
  cpship::Request* rq = it->second.get();
  s->S.cv_done.wait(lk, [&] { return rq->finished; });
  if (res)
    for (int k = 0; k < rq->nrhs; k++) res[k] = rq->res[k];
  if (latency_s) *latency_s = std::chrono::duration<double>(rq->t_done - rq->t_submit).count();
  const int rc = rq->rc;
  s->S.requests.erase(it);
  return rc;
}

int cps_hip_service_get_stats(cps_hip_service* s, cps_hip_service_stats* st) {
  if (!s || !st) return CPS_HIP_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->S.mu);
  *st = s->S.stats;
  return 0;
}

}  // extern "C"
