#!/usr/bin/env python3
"""Benchmark: domain-wall CG time-to-solution per right-hand side at true
residual 1e-10, mixed precision (+ batched right-hand sides) against plain
FP64 CG, and the HBM bandwidth of the dominant hopping kernel -- on one
MI355X through libcps_hip.so.

A "step" is one complete CG solve of the workload's right-hand sides, from
sources resident in HBM (FP64, CPS layout) to solutions resident in HBM with
their FP64 true residual computed; the gauge field is resident beforehand.

Workloads (--config):
  c3h (default) Moebius DWF 24^3x64, Ls=16, b=1.5, c=0.5, M5=1.8, mf=0.005,
                links exp(i 0.2 H), 12 point sources (spin x colour) in one
                batched solve; mixed = FP32 math, FP16 spinors and links, FP64
                reliable updates at least every 100 iterations.
  c3            c3h with BF16 instead of FP16 link storage.
  c2            Shamir DWF 16^3x32, Ls=16, M5=1.8, mf=0.01, links exp(i 0.2 H),
                one Gaussian source, mixed as c3h.
  c4            clover 32^3x64, c_sw 1.0, links exp(i 0.15 H), 12 point sources
                x 4 kappas (0.1360-0.1375) = 48 rhs in one batched solve, FP16
                spinors and links.
  c1            Wilson 8^4, kappa 0.12, links exp(i 0.3 H), FP64 only.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
# the solver reports progress on stderr every 20 s during long solves
os.environ.setdefault("CPS_HIP_PROGRESS", "20")
sys.path.insert(0, ROOT)

import cps_public_amd as C  # noqa: E402
from cps_public_amd import fields as F  # noqa: E402
from cps_public_amd import layout as LY  # noqa: E402

HBM_PEAK_GBS = 8000.0

CONFIGS = {
    "c1": dict(name="wilson_8^4_kappa0.12", L=[8, 8, 8, 8], ls=1, kind=C.OP_WILSON, eps=0.3,
               mass=1.0 / (2 * 0.12) - 4.0, nrhs=1, src="gauss", mixed=None, seed=101),
    # reliable-update trigger 0.03: 172 iterations and 7 FP64 updates
    # (0.1: 171 and 10, 0.01: 244 and 6)
    "c2": dict(name="shamir_dwf_16^3x32_ls16_mf0.01", L=[16, 16, 16, 32], ls=16, kind=C.OP_DWF, eps=0.2,
               mass=0.01, m5=1.8, nrhs=1, src="gauss", mixed=C.PREC_MIXED_HALF, seed=202, reliable_delta=0.03),
    "c3": dict(name="mobius_dwf_24^3x64_ls16_b1.5_c0.5_mf0.005_12rhs", L=[24, 24, 24, 64], ls=16,
               kind=C.OP_MOBIUS, eps=0.2, mass=0.005, m5=1.8, b=1.5, c=0.5, nrhs=12, src="point",
               mixed=C.PREC_MIXED_BF16, seed=303),
    # the same workload with FP16 instead of BF16 link storage (same bytes;
    # BF16's 8-bit mantissa costs 2.5x the CG iterations at this mass)
    "c3h": dict(name="mobius_dwf_24^3x64_ls16_b1.5_c0.5_mf0.005_12rhs_fp16links", L=[24, 24, 24, 64], ls=16,
                kind=C.OP_MOBIUS, eps=0.2, mass=0.005, m5=1.8, b=1.5, c=0.5, nrhs=12, src="point",
                mixed=C.PREC_MIXED_HALF, seed=303, reliable_delta=0.01),
    # reliable-update trigger 0.03: 86-101 iterations, 7-8 FP64 updates per
    # rhs (0.01: 90-109 and 5-6; 0.1: 86-100 and 10-11)
    "c4": dict(name="clover_32^3x64_csw1.0_kappa0.1360-0.1375_48rhs", L=[32, 32, 32, 64], ls=1,
               kind=C.OP_CLOVER, eps=0.15, kappas=[0.1360, 0.1365, 0.1370, 0.1375], csw=1.0, nrhs=48,
               src="point", mixed=C.PREC_MIXED_HALF, seed=404, reliable_delta=0.03),
}


def make_inputs(cfg):
    L, ls = cfg["L"], cfg["ls"]
    u = F.gauge_field(L, cfg["eps"], cfg["seed"])
    u_w = LY.gauge_canon_to_wilson(LY.bond_cond(u, L, [0, 0, 0, 1]), L)
    vcb = int(np.prod(L)) // 2
    per = 24 * vcb * ls
    if cfg["src"] == "gauss":
        src = np.concatenate([F.gaussian_fermion(vcb * ls, cfg["seed"] + 1 + k) for k in range(cfg["nrhs"])])
    else:
        # point sources delta(x - (1,0,0,0)) delta_{s,0} e_{spin,colour}: (1,0,0,0)
        # is an odd site, so the source lies on the checkerboard CG works on
        src = np.zeros(per * cfg["nrhs"])
        site_cb = 0  # x/2 + Lx/2*(y + ...) for x = 1
        for k in range(cfg["nrhs"]):
            spin, colour = divmod(k % 12, 3)
            src[k * per + 24 * site_cb + 2 * (colour + 3 * spin)] = 1.0
    return u_w, src


def op_for(cfg):
    if cfg["kind"] == C.OP_CLOVER:
        # rhs r uses kappas[r // (nrhs / len(kappas))]; the op names the first
        return C.make_op(C.OP_CLOVER, mass=1.0 / (2 * cfg["kappas"][0]) - 4.0, csw=cfg["csw"])
    return C.make_op(cfg["kind"], mass=cfg["mass"], m5=cfg.get("m5", 1.8), b=cfg.get("b", 1.0),
                     c=cfg.get("c", 0.0))


def time_solves(lat, op, src_dev, arg, nrhs, warmup, steps, keep=False):
    """Each step is one complete solve of all nrhs sources, which are resident
    in HBM (CPS layout, FP64) before the clock starts; the solutions are
    resident in HBM and their FP64 true residual computed when it stops
    (cps_hip_invcg_device returns after a device synchronisation).  With
    keep, the last solutions are returned (host copy, FP64) as well."""
    sol = C.DeviceBuffer(src_dev.nbytes)
    zeros = np.zeros(src_dev.nbytes // 8)
    res = None
    for _ in range(warmup):
        sol.upload(zeros)
        res = lat.invcg_device(op, sol, src_dev, arg, nrhs=nrhs)
    t = []
    for _ in range(steps):
        sol.upload(zeros)
        t0 = time.perf_counter()
        res = lat.invcg_device(op, sol, src_dev, arg, nrhs=nrhs)
        t.append(time.perf_counter() - t0)
        print(f"# solve: {t[-1]:.4f} s, iterations {[r.iterations for r in res]}, "
              f"true_res {max(r.true_res for r in res):.3e}", file=sys.stderr, flush=True)
    x = sol.download() if keep else None
    sol.free()
    return (t, res, x) if keep else (t, res)


def rel_err_per_rhs(x, ref, nrhs):
    """|x - ref| / |ref| for each rhs (FP64 on the host)."""
    n = x.size // nrhs
    return [float(np.linalg.norm(x[k * n:(k + 1) * n] - ref[k * n:(k + 1) * n])
                  / np.linalg.norm(ref[k * n:(k + 1) * n])) for k in range(nrhs)]


def host_info():
    """The host the CPU leg ran on: CPU model, logical CPUs this process may
    use, OpenMP threads, memory."""
    model, mem_kb = "unknown", 0
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
        with open("/proc/meminfo") as f:
            for line in f:
                if line.startswith("MemTotal"):
                    mem_kb = int(line.split()[1])
                    break
    except OSError:
        pass
    ncpu = len(os.sched_getaffinity(0))
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return {"cpu": model, "cpus_available": ncpu, "omp_num_threads": omp or None,
            "mem_gb": round(mem_kb / 2 ** 20, 1)}


def cpu_threads():
    """Threads of the multi-threaded CPU leg: OMP_NUM_THREADS if set (the
    machine's CPU share), else the CPUs this process may run on."""
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return omp if omp > 0 else len(os.sched_getaffinity(0))


def _oracle_cg(O, cfg, u_w, s0, stop_rsd, max_iter, clov=None):
    L, ls = cfg["L"], cfg["ls"]
    if cfg["kind"] == C.OP_DWF:
        return O.dwf_invcg(u_w, s0, cfg["mass"], stop_rsd, max_iter, O.dwf_struct(L, ls, cfg["m5"]))
    if cfg["kind"] == C.OP_MOBIUS:
        m = O.mobius_struct(L, ls, cfg["m5"], cfg["b"], cfg["c"])
        return O.mobius_invcg(u_w, s0, cfg["mass"], stop_rsd, max_iter, m)
    if cfg["kind"] == C.OP_CLOVER:
        kappa, A_oo, A_ee_inv = clov
        return O.clover_invcg(u_w, A_oo, A_ee_inv, s0, kappa, stop_rsd, max_iter, L)
    return O.wilson_invcg(u_w, s0, cfg["mass"], stop_rsd, max_iter, L)


def cpu_baseline(cfg, u_w, src, gpu_iters, full=False, legs=None):
    """The CPU restatement (oracle/) of the CPS noarch FP64 CG on the same
    gauge field and rhs 0, on this machine's host cores: with OpenMP over the
    sites of the hopping term at the machine's thread count (the reported
    value) and with one thread (CPS noarch: wilson_dslash.C:280-477 has no
    OpenMP).  full: to convergence at 1e-10 (C1, C2); otherwise a bounded
    sample: a few CG iterations at full size, per-iteration time times the
    GPU FP64 iteration count (a projection, labelled as such)."""
    from oracle import oracle as O
    O.build()
    O.use_timing_build(True)
    try:
        return _cpu_baseline(O, cfg, u_w, src, gpu_iters, full, legs)
    finally:
        O.use_timing_build(False)


def _cpu_baseline(O, cfg, u_w, src, gpu_iters, full, legs):
    L, ls = cfg["L"], cfg["ls"]
    per = 24 * (int(np.prod(L)) // 2) * ls
    s0 = np.ascontiguousarray(src[:per])
    nthr = cpu_threads()
    clov = None
    if cfg["kind"] == C.OP_CLOVER:
        # the clover matrices are set-up work (CloverMatChkb), outside the clock
        kappa = cfg["kappas"][0]
        O.set_threads(nthr)
        clov = (kappa, O.clover_mat_chkb(u_w, 1, 0, kappa, cfg["csw"], L),
                O.clover_mat_chkb(u_w, 0, 1, kappa, cfg["csw"], L))
    out = {"unit": "s/RHS", "kind": "port", "host": host_info(), "build": O.TIMING_BUILD_FLAGS}
    if full:
        todo = legs if legs else ([nthr, 1] if nthr > 1 else [1])
        nthr = todo[0]
        legs = {}
        O.set_progress(10)   # the solve reports its own progress on stderr
        for thr in todo:
            O.set_threads(thr)
            t0 = time.perf_counter()
            _, it, tr = _oracle_cg(O, cfg, u_w, s0, 1e-10, 100000, clov)
            el = time.perf_counter() - t0
            legs[thr] = {"value": el, "cores": thr, "iterations": it, "true_res": tr}
            print(f"# cpu leg: {thr} threads, {it} iterations, {el:.1f} s, true_res {tr:.2e}",
                  file=sys.stderr, flush=True)
        out.update(value=legs[nthr]["value"], cores=nthr,
                   sample=f"rhs 0 solved to convergence (stop_rsd 1e-10) by the oracle FP64 CG with {nthr} "
                          f"threads: {legs[nthr]['iterations']} iterations, true residual "
                          f"{legs[nthr]['true_res']:.2e}")
        for thr in todo[1:]:
            out.setdefault("other_legs", []).append(
                {"value": legs[thr]["value"], "cores": thr,
                 "sample": f"the same solve with {thr} thread(s): {legs[thr]['iterations']} iterations"})
        return out
    # bounded sample at full size with nthr threads: n_it CG iterations plus
    # the two extra M^dag M of InvCg (initial residual, true residual)
    vol5 = int(np.prod(L)) * ls
    n_it = 2 if vol5 > 4_000_000 else (5 if vol5 > 100_000 else 50)
    O.set_threads(nthr)
    t0 = time.perf_counter()
    _, it, _ = _oracle_cg(O, cfg, u_w, s0, 1e-30, n_it + 1, clov)
    el = time.perf_counter() - t0
    per_iter = el / (n_it + 2)
    out.update(value=per_iter * gpu_iters, cores=nthr,
               sample=f"{n_it} CG iterations (+2 M^dag M) of the oracle FP64 CG on rhs 0 at full size with "
                      f"{nthr} threads, {el:.1f} s; per-iteration time x {gpu_iters} iterations (the GPU FP64 "
                      f"count): a projection, not a solve to convergence")
    # one thread, as CPS noarch runs the hopping term: one M^dag M on a slab
    # of Lt/8 time slices (same Ls and parameters; cost linear in volume)
    if vol5 > 4_000_000 and cfg["kind"] in (C.OP_MOBIUS, C.OP_DWF):
        Lt_sub = max(2, L[3] // 8)
        Ls_ = [L[0], L[1], L[2], Lt_sub]
        u_sub = LY.gauge_canon_to_wilson(LY.bond_cond(F.gauge_field(Ls_, cfg["eps"], cfg["seed"]), Ls_, [0, 0, 0, 1]),
                                         Ls_)
        s_sub = F.gaussian_fermion(int(np.prod(Ls_)) // 2 * ls, cfg["seed"] + 1)
        O.set_threads(1)
        t0 = time.perf_counter()
        if cfg["kind"] == C.OP_MOBIUS:
            O.mobius_mdagm(u_sub, s_sub, cfg["mass"], O.mobius_struct(Ls_, ls, cfg["m5"], cfg["b"], cfg["c"]))
        else:
            O.dwf_mdagm(u_sub, s_sub, cfg["mass"], O.dwf_struct(Ls_, ls, cfg["m5"]))
        el1 = (time.perf_counter() - t0) * L[3] / Lt_sub
        out["one_thread"] = {"value": el1 * gpu_iters, "cores": 1,
                             "sample": f"one M^dag M of the oracle with one thread on a {L[0]}x{L[1]}x{L[2]}x"
                                       f"{Lt_sub} slab, x {L[3] // Lt_sub} (volume) = {el1:.1f} s per M^dag M, x "
                                       f"{gpu_iters} iterations (projection; CG vector updates not counted)"}
    O.set_threads(nthr)
    return out


SERVE_LATTICES = [([8, 8, 8, 8], 8), ([16, 16, 16, 32], 16), ([24, 24, 24, 48], 16)]
SERVE_MASSES = [0.005, 0.01, 0.02, 0.05]


def serve_stream(a):
    """Solve-serving stream: requests with Poisson arrivals at a fixed rate,
    each a Wilson or DWF propagator solve (Ls of the lattice's gauge entry) on
    one of three resident lattices, a mass from SERVE_MASSES and 1-12 point
    sources (spin-colour e_k at a random site on the checkerboard CG works
    on), solved to FP64 true residual 1e-10 by the resident service
    (continuous batching over CG slots; mixed precision with FP64 reliable
    updates unless --serve-fp64).  The latency of a request runs from its
    arrival (submit) to its last solution being harvested."""
    import threading
    rng = np.random.default_rng(505)
    svc = C.Service(240 << 30, max_slots=12)
    gids = []
    lattices = SERVE_LATTICES if a.full else SERVE_LATTICES[:2]
    for L, ls in lattices:
        u = F.gauge_field(L, 0.2, 500 + len(gids))
        gids.append(svc.add_gauge(L, ls, LY.gauge_canon_to_wilson(LY.bond_cond(u, L, [0, 0, 0, 1]), L)))
    print("# service: gauges resident", file=sys.stderr, flush=True)
    # request mix over the lattices: 50 / 40 / 10 % with the 24^3x48 lattice
    # (--full), 60 / 40 % without it
    probs = [0.5, 0.4, 0.1] if a.full else [0.6, 0.4]
    prec = C.PREC_DOUBLE if a.serve_fp64 else C.PREC_MIXED_HALF
    plan = []
    t = 0.0
    for k in range(a.requests):
        t += rng.exponential(1.0 / a.rate)
        li = int(rng.choice(len(probs), p=probs))
        kind = C.OP_WILSON if rng.random() < 0.5 else C.OP_DWF
        mass = float(rng.choice(SERVE_MASSES))
        nrhs = int(rng.integers(1, 13))
        site = int(rng.integers(0, int(np.prod(lattices[li][0])) // 2))
        plan.append((t, li, kind, mass, nrhs, site))
    lat_s, rhs_total, results = [], 0, []
    lock = threading.Lock()
    waiters = []

    def waiter(rid, meta):
        sol, res, lat = svc.wait(rid)
        del sol
        with lock:
            lat_s.append(lat)
            results.append((meta, max(r.true_res for r in res), [r.iterations for r in res], lat))

    t0 = time.perf_counter()
    for (ta, li, kind, mass, nrhs, site) in plan:
        now = time.perf_counter() - t0
        if ta > now:
            time.sleep(ta - now)
        L, ls = lattices[li]
        vcb = int(np.prod(L)) // 2
        per = 24 * vcb * (1 if kind == C.OP_WILSON else ls)
        # point sources e_{k mod 12} at checkerboard site `site` (slice s = 0);
        # np.zeros pages stay unbacked until written
        src = np.zeros(per * nrhs)
        for r in range(nrhs):
            spin, colour = divmod(r % 12, 3)
            src[r * per + 24 * site + 2 * (colour + 3 * spin)] = 1.0
        op = C.make_op(kind, mass=mass, m5=1.8)
        rid = svc.submit(gids[li], op, src, nrhs, precision=prec)
        rhs_total += nrhs
        th = threading.Thread(target=waiter, args=(rid, (li, kind, mass, nrhs)))
        th.start()
        waiters.append(th)
        if len(waiters) % 50 == 0:
            print(f"# service: {len(waiters)} requests submitted, {len(lat_s)} done", file=sys.stderr, flush=True)
    for th in waiters:
        th.join()
    wall = time.perf_counter() - t0
    st = svc.stats()
    svc.close()
    lat = np.array(lat_s)
    worst = max(r[1] for r in results)
    per_lat = {}
    for li, (L, ls) in enumerate(lattices):
        ls_ = [r[3] for r in results if r[0][0] == li]
        if ls_:
            per_lat["x".join(str(v) for v in L)] = {"requests": len(ls_), "p50_s": float(np.percentile(ls_, 50)),
                                                    "p99_s": float(np.percentile(ls_, 99))}
    names = ", ".join(f"{'x'.join(map(str, L))} (Ls {ls})" for L, ls in lattices)
    line = {
        "metric": "solve-serving stream: p50/p99 latency and solves/s (resident gauge, continuous batching)",
        "value": len(lat) / wall, "unit": "solves/s", "n_gpus": 1, "higher_is_better": True,
        "requests": len(lat), "rhs": rhs_total, "wall_s": wall, "rate_per_s": a.rate,
        "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
        "rhs_per_s": rhs_total / wall, "true_res_max": worst, "per_lattice": per_lat,
        "dtype": "f64" if a.serve_fp64 else "f32",
        "data": "synthetic (seeded SU(3) links exp(i 0.2 H), point sources at random sites)",
        "config": {"workload": f"serve: Wilson/DWF on {names}; lattice mix " + "/".join(str(p) for p in probs)
                               + "; masses " + ",".join(str(m) for m in SERVE_MASSES)
                               + "; 1-12 rhs; stop_rsd 1e-10; "
                               + ("FP64 CG" if a.serve_fp64 else "mixed FP16/FP32 CG with FP64 reliable updates")},
        "service": st,
    }
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3h", choices=sorted(CONFIGS) + ["c5"])
    ap.add_argument("--requests", type=int, default=48, help="c5: number of requests in the stream")
    ap.add_argument("--rate", type=float, default=1.0, help="c5: Poisson arrival rate (requests/s)")
    ap.add_argument("--full", action="store_true",
                    help="c5: include the 24^3x48 lattice in the mix (off by default to keep the default run "
                         "within minutes)")
    ap.add_argument("--serve-fp64", action="store_true", help="c5: FP64 CG engines instead of mixed precision")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-full", action="store_true",
                    help="CPU leg: solve rhs 0 to convergence with the machine's thread count and with one thread "
                         "(instead of the bounded sample); meant for c1 and c2")
    ap.add_argument("--cpu-only", action="store_true", help="run only the CPU leg (no GPU work)")
    ap.add_argument("--cpu-threads", default="",
                    help="with --cpu-full: comma-separated thread counts of the legs (default: the machine's, then 1)")
    ap.add_argument("--no-bf16", action="store_true",
                    help="c3h: skip the extra solve of the same workload with BF16 link storage")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison solve")
    ap.add_argument("--no-dirac-res", action="store_true",
                    help="skip the untimed solve that reports |xi - M_pc x| / |xi| for rhs 0")
    ap.add_argument("--hop-reps", type=int, default=50)
    ap.add_argument("--max-iter", type=int, default=20000,
                    help="CG iteration cap (a small cap gives a bounded run for kernel traces)")
    ap.add_argument("--reliable-every", type=int, default=100,
                    help="mixed CG: FP64 reliable update at least every this many iterations")
    ap.add_argument("--reliable-delta", type=float, default=None,
                    help="mixed CG: reliable update when |r_sloppy| < delta |r| of the last update (default: the "
                         "workload's value, 0.01 for c3h and c4 and 0.1 otherwise)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only launch the hopping kernels (mixed and FP64) and the reference copy kernel, "
                         "for hardware-counter runs")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (the library has no multi-GPU path); got --gpus {a.gpus}",
              file=sys.stderr)
        sys.exit(2)
    if a.config == "c5":
        serve_stream(a)
        return
    cfg = CONFIGS[a.config]
    print(f"# {cfg['name']}: building inputs", file=sys.stderr, flush=True)
    u_w, src = make_inputs(cfg)
    print("# inputs ready", file=sys.stderr, flush=True)
    if a.cpu_only:
        print(json.dumps({"config": {"workload": cfg["name"]},
                          "cpu_baseline": cpu_baseline(cfg, u_w, src, 0, full=True,
                                                       legs=[int(v) for v in a.cpu_threads.split(",") if v])}))
        return
    lat = C.Lattice(cfg["L"], cfg["ls"], (0, 0, 0, 1))
    lat.import_gauge(u_w)
    if cfg["kind"] == C.OP_CLOVER:
        lat.clover_build(cfg["kappas"], cfg["csw"])
    op = op_for(cfg)
    nrhs = cfg["nrhs"]

    mixed = cfg["mixed"]
    prec_main = mixed if mixed is not None else C.PREC_DOUBLE
    if a.kernels_only:
        sec, byt = lat.bench_hop(cfg["kind"], prec_main, nrhs, a.hop_reps)
        sec64, byt64 = lat.bench_hop(cfg["kind"], C.PREC_DOUBLE, nrhs, a.hop_reps)
        ncopy = (int(np.prod(cfg["L"])) // 2) * cfg["ls"] * nrhs
        secc = lat.bench_copy(ncopy, a.hop_reps)
        print(json.dumps({"hop_mixed": {"s": sec, "bytes": byt}, "hop_fp64": {"s": sec64, "bytes": byt64},
                          "copy": {"s": secc, "bytes_read": 192 * ncopy, "bytes_written": 192 * ncopy}}))
        lat.close()
        return
    src_dev = C.DeviceBuffer.from_array(src)
    delta = a.reliable_delta if a.reliable_delta is not None else cfg.get("reliable_delta", 0.1)
    arg_main = C.make_cg_arg(1e-10, a.max_iter, precision=prec_main, reliable_every=a.reliable_every,
                             reliable_delta=delta)
    t_main, res_main, x_main = time_solves(lat, op, src_dev, arg_main, nrhs, a.warmup, a.steps, keep=True)
    ms = 1e3 * float(np.mean(t_main))
    s_per_rhs = float(np.mean(t_main)) / nrhs

    fp64 = None
    x64 = None
    if mixed is not None and not a.no_fp64:
        # a few-iteration FP64 solve first, so that kernel autotuning is not timed
        time_solves(lat, op, src_dev, C.make_cg_arg(1e-10, 3, precision=C.PREC_DOUBLE), nrhs, 1, 0)
        arg64 = C.make_cg_arg(1e-10, a.max_iter, precision=C.PREC_DOUBLE)
        t64, res64, x64 = time_solves(lat, op, src_dev, arg64, nrhs, 0, 1, keep=True)
        fp64 = {"s_per_rhs": float(np.mean(t64)) / nrhs, "iterations": [r.iterations for r in res64],
                "true_res_max": max(r.true_res for r in res64)}
    gpu_fp64_iters = (fp64["iterations"][0] if fp64 else res_main[0].iterations)
    # per-solve error of the mixed solution against the FP64 GPU CG solution
    # of the same system (both to true residual 1e-10)
    err64 = rel_err_per_rhs(x_main, x64, nrhs) if x64 is not None else None
    if err64:
        print(f"# |x_mixed - x_fp64| / |x_fp64| per rhs: max {max(err64):.2e}", file=sys.stderr, flush=True)

    # the same workload with BF16 link storage (the link format of the
    # Moebius workload as specified), one timed solve after a warm-up
    bf16 = None
    if a.config == "c3h" and not a.no_bf16 and fp64 is not None:
        cb = CONFIGS["c3"]
        arg_b = C.make_cg_arg(1e-10, a.max_iter, precision=C.PREC_MIXED_BF16, reliable_every=100,
                              reliable_delta=cb.get("reliable_delta", 0.1))
        time_solves(lat, op, src_dev, C.make_cg_arg(1e-10, 3, precision=C.PREC_MIXED_BF16), nrhs, 1, 0)
        tb, resb, xb = time_solves(lat, op, src_dev, arg_b, nrhs, 0, 1, keep=True)
        sb = float(np.mean(tb)) / nrhs
        bf16 = {"workload": cb["name"], "s_per_rhs": sb, "speedup_vs_fp64": fp64["s_per_rhs"] / sb,
                "iterations": [r.iterations for r in resb], "true_res_max": max(r.true_res for r in resb),
                "reliable_delta": cb.get("reliable_delta", 0.1),
                "err_vs_fp64_max": max(rel_err_per_rhs(xb, x64, nrhs))}
        del xb

    # residual of the Dirac equation M_pc x = xi itself (not of the normal
    # equations the CG works on) for rhs 0, untimed: the same solver on the
    # source M_pc^dag xi, then |xi - M_pc x| / |xi| in FP64 (clover with
    # several kappas: the first rhs of each kappa, solved together)
    dirac_res = None
    if not a.no_dirac_res:
        per = src.size // nrhs
        nsets = len(cfg["kappas"]) if cfg["kind"] == C.OP_CLOVER else 1
        idx = [k * (nrhs // nsets) for k in range(nsets)]
        xi = np.concatenate([src[i * per:(i + 1) * per] for i in idx])
        x1, r1 = lat.invcg(op, lat.matpc(op, xi, dag=1, nrhs=nsets), arg_main, nrhs=nsets)
        rd = xi - lat.matpc(op, x1, dag=0, nrhs=nsets)
        res_d = [float(np.linalg.norm(rd[k * per:(k + 1) * per]) / np.linalg.norm(xi[k * per:(k + 1) * per]))
                 for k in range(nsets)]
        dirac_res = {"rhs": idx[0] if nsets == 1 else idx, "res": max(res_d),
                     "iterations": r1[0].iterations if nsets == 1 else [r.iterations for r in r1],
                     "normal_eq_true_res": max(r.true_res for r in r1)}
        print(f"# Dirac-equation residual (rhs {dirac_res['rhs']}): {dirac_res['res']:.3e}", file=sys.stderr,
              flush=True)

    # dominant kernel: the hopping kernel of the headline path, timed with HIP
    # events on its launch stream
    hk = cfg["kind"] if cfg["kind"] != C.OP_MOBIUS else C.OP_MOBIUS
    sec, byt = lat.bench_hop(hk, prec_main, nrhs, a.hop_reps)
    achieved = byt / sec / 1e9
    sec64, byt64 = lat.bench_hop(hk, C.PREC_DOUBLE, nrhs, a.hop_reps)

    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{a.config}.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(cfg, u_w, src, gpu_fp64_iters, full=a.cpu_full)

    line = {
        "metric": "DWF CG time-to-solution/RHS @1e-10 & dslash HBM GB/s: mixed+batched vs FP64",
        "value": s_per_rhs,
        "unit": "s/RHS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if mixed is not None else "f64",
        "data": "synthetic (seeded SU(3) links exp(i eps H), Philox)",
        "config": {"workload": cfg["name"], "nrhs": nrhs, "stop_rsd": 1e-10,
                   "reliable_delta": delta if mixed is not None else None, "reliable_every": a.reliable_every,
                   "precision": {0: "fp64", 1: "mixed_fp16", 2: "mixed_bf16links", 3: "mixed_fp32"}[prec_main],
                   "storage": {0: "f64 spinors and links",
                               1: "f16 spinors (per-site norms) + f16 links, f32 CG vectors, f64 reliable updates",
                               2: "f16 spinors (per-site norms) + bf16 links, f32 CG vectors, f64 reliable updates",
                               3: "f32 spinors + f32 links, f64 reliable updates"}[prec_main]},
        "iterations": [r.iterations for r in res_main],
        "true_res_max": max(r.true_res for r in res_main),
        "dirac_eq_res": dirac_res,
        "reliable_updates": [r.reliable_updates for r in res_main],
        "fp64": fp64,
        "speedup_vs_fp64": (fp64["s_per_rhs"] / s_per_rhs) if fp64 else None,
        "err_vs_fp64": err64,
        "bf16_links": bf16,
        "dslash": {"kernel": "k_hop (5-D DWF)" if cfg["kind"] == C.OP_DWF else "k_hop (4-D, batched)",
                   "seconds_per_launch": sec, "algorithmic_bytes": byt, "GBps": achieved,
                   "fp64_GBps": byt64 / sec64 / 1e9, "fp64_seconds_per_launch": sec64},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    lat.close()


if __name__ == "__main__":
    main()
