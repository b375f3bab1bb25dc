// Resident solve service: gauge fields stay on the device, and incoming
// propagator solves (any registered lattice, operator and mass, 1..n
// right-hand sides) are served by continuous batching.
//
// An engine is one batched FP64 CG (the iteration of DiracOp::InvCg,
// inv_cg.C:293-364) over S right-hand-side slots for one (lattice, operator)
// pair, on its own HIP stream.  Every slot carries its own CG scalars, so a
// slot that converges is harvested (FP64 true residual, solution copied out)
// and refilled with the next waiting right-hand side while the other slots
// keep iterating: requests that arrive while others are on the GPU join as
// slots free up.  Engines of different lattices or masses run concurrently
// on their own streams; a new engine is created only while the HBM budget
// allows (admission by footprint), and idle engines are released.
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

bool same_op(const cps_hip_op& a, const cps_hip_op& b) {
  return a.kind == b.kind && a.mass == b.mass && a.m5 == b.m5 && a.a5_inv == b.a5_inv && a.mobius_b == b.mobius_b &&
         a.mobius_c == b.mobius_c && a.csw == b.csw;
}

FieldRef slot_ref(const Field& f, int64_t per_rhs, int k) {
  FieldRef r = f.ref();
  r.data = static_cast<char*>(r.data) + (size_t)k * per_rhs * 16;   // FP64: 16-byte chunks per plane
  return r;
}

CGScalars slot_scalars(const CGScalars& s, int k) {
  CGScalars v = s;
  v.rr += k; v.rr_prev += k; v.d += k; v.alpha += k; v.beta += k; v.stp += k; v.src_norm += k; v.rsd += k;
  v.active += k; v.iters += k; v.status += k;
  return v;
}

struct Engine {
  int gauge_id = -1;
  cps_hip_op op;
  int max_iter = 0;
  int S = 0;
  Lattice L;
  OpLayout o;
  int bsb = 0, bpr_blas = 0, bpr_hop = 0;
  double* part = nullptr;
  CGState cg;
  Field *x = nullptr, *b = nullptr, *r = nullptr, *p = nullptr, *ap = nullptr, *t1 = nullptr, *t2 = nullptr,
        *t3 = nullptr;
  std::vector<Request*> owner;
  std::vector<int> rhs_of;
  long long bytes = 0;
  clk::time_point idle_since = clk::now();
  std::unique_ptr<IterGraph> graph;

  static long long footprint(const GaugeEntry& ge, const cps_hip_op& op, int S) {
    const long long V = (long long)ge.L[0] * ge.L[1] * ge.L[2] * ge.L[3];
    const long long per_rhs = (op.kind == CPS_HIP_OP_DWF || op.kind == CPS_HIP_OP_MOBIUS) ? V / 2 * ge.ls : V / 2;
    const long long vec = per_rhs * 192;
    return 8 * vec * S + vec + 2 * V * 72 * 8 + V * 48 + 64 * 1024 * 1024;
  }

  int setup(const GaugeEntry& ge, int gid, const cps_hip_op& op_, int S_, int max_iter_) {
    gauge_id = gid;
    op = op_;
    S = S_;
    max_iter = max_iter_;
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
    p = &L.field(S_P, STORE_D, n);
    ap = &L.field(S_AP, STORE_D, n);
    t1 = &L.field(S_T1, STORE_D, n);
    t2 = &L.field(S_T2, STORE_D, n);
    t3 = &L.field(S_T3, STORE_D, n);
    if (L.last_rc) return L.last_rc;
    bsb = pick_block(o.per_rhs);
    bpr_blas = (int)(o.per_rhs / bsb);
    bpr_hop = hop_blocks_per_rhs(L, &op);
    if ((rc = L.ensure_partial((int64_t)std::max(bpr_blas, bpr_hop) * S))) return rc;
    part = (double*)L.partial.p;
    if ((rc = cg.alloc(L, S, 0))) return rc;
    cg.s.max_itr = max_iter - 1;
    if ((rc = L.io.ensure((size_t)o.per_rhs * 24 * sizeof(double)))) return rc;
    // empty slots hold zeros and are inactive
    for (Field* f : {x, b, r, p, ap, t1, t2, t3}) CPSHIP_CHECK(blas_zero(STORE_D, f->ref(), o.per_rhs, S, bsb, L.stream));
    std::vector<int> zeros(S, 0);
    CPSHIP_CHECK(hipMemcpyAsync(cg.s.active, zeros.data(), S * sizeof(int), hipMemcpyHostToDevice, L.stream));
    CPSHIP_CHECK(hipStreamSynchronize(L.stream));
    owner.assign(S, nullptr);
    rhs_of.assign(S, -1);
    bytes = L.device_bytes();
    graph.reset(new IterGraph(true));
    return 0;
  }

  int free_slots() const {
    int n = 0;
    for (auto* q : owner) n += q == nullptr;
    return n;
  }
  bool busy() const { return free_slots() < S; }

  // rhs k of request rq into slot s: x = 0, r = p = src, |src|^2, stp, |r|^2
  int admit(int s, Request* rq, int k) {
    const double* h = rq->src + (size_t)k * o.per_rhs * 24;
    CPSHIP_CHECK(hipMemcpyAsync(L.io.p, h, (size_t)o.per_rhs * 24 * sizeof(double), hipMemcpyHostToDevice, L.stream));
    const FieldRef bs = slot_ref(*b, o.per_rhs, s), xs = slot_ref(*x, o.per_rhs, s);
    const FieldRef rs = slot_ref(*r, o.per_rhs, s), ps = slot_ref(*p, o.per_rhs, s);
    CPSHIP_CHECK(pack_spinor(STORE_D, (const double*)L.io.p, bs, o.map, L.g.Vcb, L.ls, 1, o.per_rhs, L.stream));
    CPSHIP_CHECK(blas_zero(STORE_D, xs, o.per_rhs, 1, bsb, L.stream));
    CPSHIP_CHECK(convert_spinor(STORE_D, bs, STORE_D, rs, o.per_rhs, L.stream));
    CPSHIP_CHECK(convert_spinor(STORE_D, bs, STORE_D, ps, o.per_rhs, L.stream));
    CPSHIP_CHECK(hipMemcpyAsync(cg.s.rsd + s, &rq->stop_rsd, sizeof(double), hipMemcpyHostToDevice, L.stream));
    const CGScalars v = slot_scalars(cg.s, s);
    CPSHIP_CHECK(blas_norm(STORE_D, bs, o.per_rhs, 1, bsb, part, L.stream));
    CPSHIP_CHECK(reduce_rhs(part, bpr_blas, 1, nullptr, RED_SRC_NORM, v, L.stream));
    CPSHIP_CHECK(blas_norm(STORE_D, rs, o.per_rhs, 1, bsb, part, L.stream));
    CPSHIP_CHECK(reduce_rhs(part, bpr_blas, 1, nullptr, RED_CG_INIT, v, L.stream));
    owner[s] = rq;
    rhs_of[s] = k;
    return 0;
  }

  // `iters` CG iterations over all slots (inactive slots are frozen by the
  // per-slot scalars), enqueued on the engine's stream
  int step(int iters) {
    return graph->run(L, iters, [&]() -> int {
      int e = apply_mdagm(L, &op, PREC_D_D18, ap->ref(), p->ref(), t1->ref(), t2->ref(), part, S);
      if (e) return e;
      CPSHIP_CHECK(reduce_rhs(part, bpr_hop, S, nullptr, RED_CG_ALPHA, cg.s, L.stream));
      CPSHIP_CHECK(blas_cg_xr_d(x->ref(), p->ref(), r->ref(), ap->ref(), cg.s, o.per_rhs, S, bsb, part, L.stream));
      CPSHIP_CHECK(reduce_rhs(part, bpr_blas, S, nullptr, RED_CG_BETA, cg.s, L.stream));
      CPSHIP_CHECK(blas_xpay(STORE_D, p->ref(), r->ref(), cg.s.beta, cg.s.active, o.per_rhs, S, bsb, L.stream));
      return 0;
    });
  }

  // Slots whose CG stopped: FP64 true residual (inv_cg.C:403-417), solution
  // copied to the caller, slot freed.  Returns the requests touched.
  int harvest(std::vector<Request*>& touched) {
    std::vector<int> act(S), st(S);
    CPSHIP_CHECK(hipMemcpyAsync(act.data(), cg.s.active, S * sizeof(int), hipMemcpyDeviceToHost, L.stream));
    CPSHIP_CHECK(hipMemcpyAsync(st.data(), cg.s.status, S * sizeof(int), hipMemcpyDeviceToHost, L.stream));
    CPSHIP_CHECK(hipStreamSynchronize(L.stream));
    bool any = false;
    for (int s = 0; s < S; s++)
      if (owner[s] && !act[s]) any = true;
    if (!any) return 0;
    int e = apply_mdagm(L, &op, PREC_D_D18, ap->ref(), x->ref(), t1->ref(), t2->ref(), nullptr, S);
    if (e) return e;
    CPSHIP_CHECK(blas_sub_norm_d(b->ref(), ap->ref(), t3->ref(), o.per_rhs, S, bsb, part, L.stream));
    double* trn = cg.s.d;   // free between iterations (the next one rewrites it first)
    CPSHIP_CHECK(reduce_rhs(part, bpr_blas, S, trn, RED_PLAIN, cg.s, L.stream));
    std::vector<double> tr(S), sn(S);
    std::vector<int> it(S);
    CPSHIP_CHECK(hipMemcpyAsync(tr.data(), trn, S * sizeof(double), hipMemcpyDeviceToHost, L.stream));
    CPSHIP_CHECK(hipMemcpyAsync(sn.data(), cg.s.src_norm, S * sizeof(double), hipMemcpyDeviceToHost, L.stream));
    CPSHIP_CHECK(hipMemcpyAsync(it.data(), cg.s.iters, S * sizeof(int), hipMemcpyDeviceToHost, L.stream));
    CPSHIP_CHECK(hipStreamSynchronize(L.stream));
    for (int s = 0; s < S; s++) {
      Request* rq = owner[s];
      if (!rq || act[s]) continue;
      const int k = rhs_of[s];
      CPSHIP_CHECK(unpack_spinor(STORE_D, slot_ref(*x, o.per_rhs, s), (double*)L.io.p, o.map, L.g.Vcb, L.ls, 1,
                                 o.per_rhs, L.stream));
      CPSHIP_CHECK(hipMemcpyAsync(rq->sol + (size_t)k * o.per_rhs * 24, L.io.p, (size_t)o.per_rhs * 24 * sizeof(double),
                                  hipMemcpyDeviceToHost, L.stream));
      CPSHIP_CHECK(hipStreamSynchronize(L.stream));
      cps_hip_cg_result& res = rq->res[k];
      res.iterations = it[s] + 1;
      res.status = st[s];
      res.src_norm_sq = sn[s];
      res.true_res = sn[s] > 0 ? std::sqrt(tr[s] / sn[s]) : 0.0;
      res.reliable_updates = 0;
      if (st[s] == CG_STATUS_D_ZERO) rq->rc = CPS_HIP_ERR_D_ZERO;
      rq->done_rhs++;
      owner[s] = nullptr;
      rhs_of[s] = -1;
      touched.push_back(rq);
    }
    if (!busy()) idle_since = clk::now();
    return 0;
  }
};

}  // namespace

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

  long long used_bytes() const {
    long long b = 0;
    for (auto& g : gauges) b += (long long)g->u.bytes;
    for (auto& e : engines) b += e->bytes;
    return b;
  }

  Engine* find_engine(const Request* rq) {
    for (auto& e : engines)
      if (e->gauge_id == rq->gauge_id && e->max_iter == rq->max_iter && same_op(e->op, rq->op)) return e.get();
    return nullptr;
  }

  // creates an engine for rq if the HBM budget allows it, releasing idle
  // engines first when needed
  int make_engine(Request* rq, Engine** out) {
    *out = nullptr;
    const GaugeEntry& ge = *gauges[rq->gauge_id];
    const long long need = Engine::footprint(ge, rq->op, max_slots);
    while (used_bytes() + need > budget) {
      auto idle = std::find_if(engines.begin(), engines.end(), [](const std::unique_ptr<Engine>& e) { return !e->busy(); });
      if (idle == engines.end()) return 0;   // wait: no room now
      engines.erase(idle);
    }
    std::unique_ptr<Engine> e(new Engine);
    int rc = e->setup(ge, rq->gauge_id, rq->op, max_slots, rq->max_iter);
    if (rc) return rc;
    *out = e.get();
    engines.push_back(std::move(e));
    std::lock_guard<std::mutex> lk(mu);
    stats.engines_created++;
    return 0;
  }

  void finish(Request* rq) {
    std::lock_guard<std::mutex> lk(mu);
    rq->finished = true;
    rq->t_done = clk::now();
    stats.requests_done++;
    stats.rhs_done += rq->nrhs;
    cv_done.notify_all();
  }

  void loop() {
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu);
        const bool idle = pending.empty() && std::none_of(engines.begin(), engines.end(),
                                                         [](const std::unique_ptr<Engine>& e) { return e->busy(); });
        if (idle) cv_work.wait(lk, [&] { return stop || !incoming.empty(); });
        if (stop && incoming.empty() && idle) return;
        while (!incoming.empty()) {
          pending.push_back(incoming.front());
          incoming.pop_front();
        }
      }
      // admission: FIFO over waiting requests, one rhs per free slot
      for (auto it = pending.begin(); it != pending.end();) {
        Request* rq = *it;
        Engine* e = find_engine(rq);
        int rc = e ? 0 : make_engine(rq, &e);
        if (!rc && !e && engines.empty()) rc = CPS_HIP_ERR_NOMEM;   // can never fit the budget
        while (!rc && e && rq->next_rhs < rq->nrhs && e->free_slots() > 0) {
          int slot = 0;
          while (e->owner[slot]) slot++;
          rc = e->admit(slot, rq, rq->next_rhs);
          if (!rc) rq->next_rhs++;
        }
        if (rc) {
          // the request fails with rc: its remaining rhs are not solved, the
          // ones already in slots finish normally
          rq->rc = rc;
          rq->done_rhs += rq->nrhs - rq->next_rhs;
          rq->next_rhs = rq->nrhs;
          if (rq->done_rhs == rq->nrhs) finish(rq);
        }
        if (rq->next_rhs >= rq->nrhs) it = pending.erase(it);
        else ++it;
      }
      // every busy engine advances one chunk; the launches of all engines are
      // queued before any stream is synchronised, so their work overlaps
      long long stepped = 0;
      for (auto& e : engines)
        if (e->busy()) {
          if (int rc = e->step(chunk)) fatal = rc;
          stepped += chunk;
        }
      for (auto& e : engines) {
        if (!e->busy()) continue;
        std::vector<Request*> touched;
        if (int rc = e->harvest(touched)) fatal = rc;
        for (Request* rq : touched)
          if (rq->done_rhs == rq->nrhs && !rq->finished) finish(rq);
      }
      // release engines idle for a while
      const auto now = clk::now();
      engines.erase(std::remove_if(engines.begin(), engines.end(),
                                   [&](const std::unique_ptr<Engine>& e) {
                                     return !e->busy() &&
                                            std::chrono::duration<double>(now - e->idle_since).count() > 2.0;
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
  if (!s || !op || !src || !sol || nrhs < 1 || !request_id || max_num_iter < 1) return CPS_HIP_ERR_ARG;
  std::unique_ptr<cpship::Request> rq(new cpship::Request);
  rq->gauge_id = gauge_id;
  rq->op = *op;
  rq->src = src;
  rq->sol = sol;
  rq->nrhs = nrhs;
  rq->stop_rsd = stop_rsd;
  rq->max_iter = max_num_iter;
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
