"""MI355X-native CPS Wilson / domain-wall even-odd preconditioned operators and CG.

The compute path is libcps_hip.so (HIP kernels for gfx950 behind the C ABI in
include/cps_hip.h).  This module loads it with ctypes and offers a thin
Python mirror of the CPS entry points for host-side callers and tests.
Importing works without a GPU; any call that needs the device raises
CpsHipError when the library or the device is missing -- there is no CPU
fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# CPS_HIP_LIB selects another build of the library (same ABI)
LIB_PATH = os.environ.get("CPS_HIP_LIB") or os.path.join(_HERE, "libcps_hip.so")
HEADER_PATH = os.path.join(_HERE, "..", "include", "cps_hip.h")

OP_WILSON, OP_DWF, OP_MOBIUS, OP_CLOVER = 0, 1, 2, 3
PREC_DOUBLE, PREC_MIXED_HALF, PREC_MIXED_BF16, PREC_MIXED_SINGLE = 0, 1, 2, 3
ERR_NAMES = {-1: "bad argument", -2: "HIP runtime failure", -3: "no gauge field resident",
             -4: "CG d == 0 (inv_cg.C:326-329)", -5: "device allocation failed", -6: "no HIP device",
             -7: "file missing, truncated or malformed", -8: "NERSC checksum mismatch",
             -9: "NERSC plaquette mismatch", -10: "request exceeds the no-allocation scratch array"}
ERR_TOO_LARGE = -10


class CpsHipError(RuntimeError):
    def __init__(self, code, what=""):
        super().__init__(f"{what}: cps_hip error {code} ({ERR_NAMES.get(code, 'unknown')})")
        self.code = code


class Op(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("mass", ctypes.c_double), ("m5", ctypes.c_double),
                ("a5_inv", ctypes.c_double), ("mobius_b", ctypes.c_double), ("mobius_c", ctypes.c_double),
                ("csw", ctypes.c_double)]


class CgArg(ctypes.Structure):
    _fields_ = [("max_num_iter", ctypes.c_int), ("stop_rsd", ctypes.c_double), ("precision", ctypes.c_int),
                ("reliable_every", ctypes.c_int), ("reliable_delta", ctypes.c_double),
                ("check_every", ctypes.c_int), ("use_graph", ctypes.c_int)]


class ServiceStats(ctypes.Structure):
    _fields_ = [("requests_done", ctypes.c_longlong), ("rhs_done", ctypes.c_longlong),
                ("iterations", ctypes.c_longlong), ("peak_bytes", ctypes.c_longlong),
                ("engines_created", ctypes.c_int), ("engines_live", ctypes.c_int), ("last_error", ctypes.c_int)]


class NerscInfo(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int * 4), ("recon_row_3", ctypes.c_int), ("bytes_per_real", ctypes.c_int),
                ("big_endian", ctypes.c_int), ("checksum_header", ctypes.c_uint), ("checksum_data", ctypes.c_uint),
                ("plaquette_header", ctypes.c_double), ("plaquette_data", ctypes.c_double),
                ("link_trace_header", ctypes.c_double), ("link_trace_data", ctypes.c_double)]


class Tuning(ctypes.Structure):
    """cps_hip_tuning: process-wide solver and launch controls (an environment
    variable CPS_HIP_<NAME>, when set, overrides the field)."""
    _fields_ = [("use_graph", ctypes.c_int), ("tune", ctypes.c_int), ("tune_report", ctypes.c_int),
                ("hop_mode", ctypes.c_int), ("rhs_window", ctypes.c_int), ("lagged_x", ctypes.c_int),
                ("paired_x", ctypes.c_int), ("s5_xpay", ctypes.c_int), ("r_in_hop", ctypes.c_int),
                ("clover_r_in_hop", ctypes.c_int), ("m5_fused", ctypes.c_int), ("s5_fused", ctypes.c_int),
                ("s5x_row", ctypes.c_int), ("sloppy_vectors", ctypes.c_int), ("verbose", ctypes.c_int),
                ("progress_seconds", ctypes.c_double)]


class CgResult(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("status", ctypes.c_int), ("true_res", ctypes.c_double),
                ("src_norm_sq", ctypes.c_double), ("reliable_updates", ctypes.c_int), ("seconds", ctypes.c_double)]


_lib = None


def lib():
    """Load libcps_hip.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CpsHipError(-2, f"{LIB_PATH} is not built (run __graft_entry__.build())")
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(l):
    vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    P = ctypes.POINTER
    sig = {
        "cps_hip_lattice_create": [P(vp), vp, i, vp],
        "cps_hip_lattice_destroy": [vp],
        "cps_hip_gauge_import": [vp, vp],
        "cps_hip_lattice_set_twist": [vp, vp],
        "cps_hip_gauge_boundary_phase": [vp, vp, P(i)],
        "cps_hip_wilson_dslash": [vp, vp, vp, vp, i, i],
        "cps_hip_wilson_m": [vp, vp, vp, vp, d],
        "cps_hip_wilson_mdag": [vp, vp, vp, vp, d],
        "cps_hip_wilson_mdagm": [vp, vp, vp, vp, vp, d],
        "cps_hip_dwf_dslash": [vp, vp, vp, vp, d, i, i, d, d],
        "cps_hip_dwf_m": [vp, vp, vp, vp, d, d, d],
        "cps_hip_dwf_mdag": [vp, vp, vp, vp, d, d, d],
        "cps_hip_dwf_mdagm": [vp, vp, vp, vp, vp, d, d, d],
        "cps_hip_matpc_dag_matpc": [vp, P(Op), vp, vp, vp, i],
        "cps_hip_matpc": [vp, P(Op), vp, vp, i, i],
        "cps_hip_matpc_dag_matpc_noalloc": [vp, P(Op), vp, vp, vp, i],
        "cps_hip_matpc_dag_matpc_prec": [vp, P(Op), i, vp, vp, vp, i],
        "cps_hip_clover_build": [vp, vp, i, d],
        "cps_hip_clover_export": [vp, i, i, vp],
        "cps_hip_invcg": [vp, P(Op), vp, vp, i, P(CgArg), P(CgResult)],
        "cps_hip_invcg_device": [vp, P(Op), vp, vp, i, P(CgArg), P(CgResult)],
        "cps_hip_matinv": [vp, P(Op), vp, vp, i, P(CgArg), P(CgResult)],
        "cps_hip_qpropw_point": [vp, P(Op), P(ctypes.c_int), P(CgArg), vp, P(CgResult)],
        "cps_hip_gauge_import_canonical": [vp, vp],
        "cps_hip_nersc_read_header": [ctypes.c_char_p, P(NerscInfo)],
        "cps_hip_nersc_read": [ctypes.c_char_p, vp, P(NerscInfo), ctypes.c_double],
        "cps_hip_nersc_write": [ctypes.c_char_p, P(ctypes.c_int), vp, i, i, ctypes.c_char_p, i],
        "cps_hip_fmatinv": [vp, P(Op), vp, vp, i, P(CgArg), P(CgResult)],
        "cps_hip_bench_hop": [vp, i, i, i, i, P(d), P(d)],
        "cps_hip_bench_copy": [vp, ctypes.c_longlong, i, P(d)],
        "cps_hip_device_count": [],
        "cps_hip_device_alloc": [P(vp), ctypes.c_size_t],
        "cps_hip_service_create": [P(vp), ctypes.c_longlong, i],
        "cps_hip_service_destroy": [vp],
        "cps_hip_service_add_gauge": [vp, vp, i, vp, vp, P(i)],
        "cps_hip_service_submit": [vp, i, P(Op), vp, vp, i, d, i, P(ctypes.c_longlong)],
        "cps_hip_service_submit_arg": [vp, i, P(Op), vp, vp, i, P(CgArg), P(ctypes.c_longlong)],
        "cps_hip_get_tuning": [P(Tuning)],
        "cps_hip_set_tuning": [P(Tuning)],
        "cps_hip_service_wait": [vp, ctypes.c_longlong, P(CgResult), P(d)],
        "cps_hip_service_get_stats": [vp, P(ServiceStats)],
        "cps_hip_device_free": [vp],
        "cps_hip_memcpy": [vp, vp, ctypes.c_size_t, i],
    }
    for name, args in sig.items():
        f = getattr(l, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    l.cps_hip_version.restype = ctypes.c_char_p
    l.cps_hip_noalloc_scratch_bytes.argtypes = []
    l.cps_hip_noalloc_scratch_bytes.restype = ctypes.c_size_t
    l.cps_hip_lattice_device_bytes.argtypes = [vp]
    l.cps_hip_lattice_device_bytes.restype = ctypes.c_longlong


def _check(rc, what):
    if rc != 0:
        raise CpsHipError(rc, what)


def _ptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.c_void_p)


def noalloc_scratch_bytes():
    """Size of the static device array of cps_hip_matpc_dag_matpc_noalloc."""
    return int(lib().cps_hip_noalloc_scratch_bytes())


def get_tuning():
    t = Tuning()
    _check(lib().cps_hip_get_tuning(ctypes.byref(t)), "cps_hip_get_tuning")
    return t


def set_tuning(t):
    _check(lib().cps_hip_set_tuning(ctypes.byref(t)), "cps_hip_set_tuning")


def make_op(kind, mass=0.0, m5=1.8, a5_inv=1.0, b=1.0, c=0.0, csw=0.0):
    return Op(kind, mass, m5, a5_inv, b, c, csw)


def make_cg_arg(stop_rsd=1e-10, max_num_iter=10000, precision=PREC_DOUBLE, reliable_every=100,
                reliable_delta=0.1, check_every=0, use_graph=1):
    return CgArg(max_num_iter, stop_rsd, precision, reliable_every, reliable_delta, check_every, use_graph)


def nersc_read(path, chkprec=0.01):
    """ReadLatticeParallel: the CANONICAL FP64 gauge field of a NERSC file and
    its header/check information (raises CpsHipError on a checksum or
    plaquette mismatch)."""
    info = NerscInfo()
    _check(lib().cps_hip_nersc_read_header(path.encode(), ctypes.byref(info)), "cps_hip_nersc_read_header")
    u = np.zeros(int(np.prod(list(info.dims))) * 72)
    rc = lib().cps_hip_nersc_read(path.encode(), _ptr(u), ctypes.byref(info), float(chkprec))
    _check(rc, "cps_hip_nersc_read")
    return u, info


def nersc_write(path, dims, u_canon, recon_row_3=1, bytes_per_real=8, ensemble_id="", sequence_number=0):
    """WriteLatticeParallel: a CANONICAL FP64 gauge field to a NERSC file
    (IEEE32BIG or IEEE64BIG)."""
    d = (ctypes.c_int * 4)(*[int(v) for v in dims])
    u = np.ascontiguousarray(u_canon, dtype=np.float64)
    _check(lib().cps_hip_nersc_write(path.encode(), d, _ptr(u), int(recon_row_3), int(bytes_per_real),
                                     ensemble_id.encode(), int(sequence_number)), "cps_hip_nersc_write")


class Lattice:
    """Device handle: geometry, resident gauge field and scratch (wilson_init /
    dwf_init equivalent).  node_sites = (Lx, Ly, Lz, Lt)."""

    def __init__(self, node_sites, ls=1, bc=(0, 0, 0, 1), twist=None):
        """bc[mu]: CPS BndCndType, 0 periodic, 1 antiperiodic, 2 twisted
        (twist[mu] = GJP.TwistAngle in units of pi); used by the CANONICAL
        import only."""
        self.L = [int(v) for v in node_sites]
        self.ls = int(ls)
        self.bc = [int(v) for v in bc]
        h = ctypes.c_void_p()
        dims = (ctypes.c_int * 4)(*self.L)
        bcs = (ctypes.c_int * 4)(*self.bc)
        _check(lib().cps_hip_lattice_create(ctypes.byref(h), dims, self.ls, bcs), "cps_hip_lattice_create")
        self.h = h
        if twist is not None:
            tw = np.ascontiguousarray(np.asarray(twist, dtype=np.float64))
            _check(lib().cps_hip_lattice_set_twist(self.h, _ptr(tw)), "cps_hip_lattice_set_twist")

    def boundary_phase(self):
        """det U of the last-slice links per direction (complex array of 4)
        and whether the field is SU(3) times those factors (recon_ok)."""
        ph = np.zeros(8)
        ok = ctypes.c_int()
        _check(lib().cps_hip_gauge_boundary_phase(self.h, _ptr(ph), ctypes.byref(ok)), "cps_hip_gauge_boundary_phase")
        return ph[0::2] + 1j * ph[1::2], bool(ok.value)

    @property
    def vcb(self):
        return self.L[0] * self.L[1] * self.L[2] * self.L[3] // 2

    def close(self):
        if getattr(self, "h", None):
            lib().cps_hip_lattice_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def import_gauge(self, u_wilson):
        _check(lib().cps_hip_gauge_import(self.h, _ptr(u_wilson)), "cps_hip_gauge_import")

    def device_bytes(self):
        return int(lib().cps_hip_lattice_device_bytes(self.h))

    # ---- Wilson library ----
    def wilson_dslash(self, psi, cb, dag, u=None):
        chi = np.zeros(24 * self.vcb)
        _check(lib().cps_hip_wilson_dslash(self.h, _ptr(chi), None if u is None else _ptr(u), _ptr(psi), cb, dag),
               "cps_hip_wilson_dslash")
        return chi

    def wilson_m(self, psi, kappa, dag=0, u=None):
        chi = np.zeros(24 * self.vcb)
        f = lib().cps_hip_wilson_mdag if dag else lib().cps_hip_wilson_m
        _check(f(self.h, _ptr(chi), None if u is None else _ptr(u), _ptr(psi), kappa), "cps_hip_wilson_m")
        return chi

    def wilson_mdagm(self, psi, kappa):
        chi = np.zeros(24 * self.vcb)
        mp = np.zeros(1)
        _check(lib().cps_hip_wilson_mdagm(self.h, _ptr(chi), None, _ptr(psi), _ptr(mp), kappa), "cps_hip_wilson_mdagm")
        return chi, float(mp[0])

    # ---- DWF library ----
    def dwf_dslash(self, inp, mass, cb, dag, m5=1.8, a5_inv=1.0):
        out = np.zeros(24 * self.vcb * self.ls)
        _check(lib().cps_hip_dwf_dslash(self.h, _ptr(out), None, _ptr(inp), mass, cb, dag, m5, a5_inv),
               "cps_hip_dwf_dslash")
        return out

    def dwf_mdagm(self, inp, mass, m5=1.8, a5_inv=1.0):
        out = np.zeros(24 * self.vcb * self.ls)
        dp = np.zeros(1)
        _check(lib().cps_hip_dwf_mdagm(self.h, _ptr(out), None, _ptr(inp), _ptr(dp), mass, m5, a5_inv),
               "cps_hip_dwf_mdagm")
        return out, float(dp[0])

    # ---- generic operator entry points ----
    def _nreal(self, op):
        return 24 * self.vcb * (1 if op.kind in (OP_WILSON, OP_CLOVER) else self.ls)

    def mdagm(self, op, inp, nrhs=1):
        out = np.zeros(self._nreal(op) * nrhs)
        dp = np.zeros(nrhs)
        _check(lib().cps_hip_matpc_dag_matpc(self.h, ctypes.byref(op), _ptr(out), _ptr(inp), _ptr(dp), nrhs),
               "cps_hip_matpc_dag_matpc")
        return out, dp

    def mdagm_noalloc(self, op, inp, nrhs=1):
        """MatPcDagMatPc (Wilson or DWF) through the entry point that
        allocates no device memory (cps_hip_matpc_dag_matpc_noalloc)."""
        inp = np.ascontiguousarray(inp, dtype=np.float64)
        out = np.zeros(self._nreal(op) * nrhs)
        dp = np.zeros(nrhs)
        _check(lib().cps_hip_matpc_dag_matpc_noalloc(self.h, ctypes.byref(op), _ptr(out), _ptr(inp), _ptr(dp), nrhs),
               "cps_hip_matpc_dag_matpc_noalloc")
        return out, dp

    def mdagm_prec(self, op, inp, precision, nrhs=1):
        out = np.zeros(self._nreal(op) * nrhs)
        dp = np.zeros(nrhs)
        _check(lib().cps_hip_matpc_dag_matpc_prec(self.h, ctypes.byref(op), precision, _ptr(out), _ptr(inp),
                                                  _ptr(dp), nrhs), "cps_hip_matpc_dag_matpc_prec")
        return out, dp

    def matpc(self, op, inp, dag=0, nrhs=1):
        out = np.zeros(self._nreal(op) * nrhs)
        _check(lib().cps_hip_matpc(self.h, ctypes.byref(op), _ptr(out), _ptr(inp), dag, nrhs), "cps_hip_matpc")
        return out

    def clover_build(self, kappas, csw):
        k = np.ascontiguousarray(np.asarray(kappas, dtype=np.float64))
        _check(lib().cps_hip_clover_build(self.h, _ptr(k), len(k), csw), "cps_hip_clover_build")

    def clover_export(self, k, cb):
        A = np.zeros(72 * self.vcb)
        _check(lib().cps_hip_clover_export(self.h, k, cb, _ptr(A)), "cps_hip_clover_export")
        return A

    def invcg(self, op, src, arg, guess=None, nrhs=1):
        sol = np.zeros(self._nreal(op) * nrhs) if guess is None else np.ascontiguousarray(guess, dtype=np.float64).copy()
        res = (CgResult * nrhs)()
        rc = lib().cps_hip_invcg(self.h, ctypes.byref(op), _ptr(sol), _ptr(src), nrhs, ctypes.byref(arg), res)
        _check(rc, "cps_hip_invcg")
        return sol, [res[k] for k in range(nrhs)]

    def invcg_device(self, op, sol, src, arg, nrhs=1):
        """InvCg on device-resident FP64 vectors in CPS layout: sol and src are
        DeviceBuffer objects of nrhs * f_size_cb reals; sol holds the initial
        guess and receives the solution in place."""
        n = self._nreal(op) * nrhs * 8
        for b, name in ((sol, "sol"), (src, "src")):
            if not isinstance(b, DeviceBuffer) or b.nbytes != n:
                raise ValueError(f"invcg_device: {name} must be a DeviceBuffer of {n} bytes")
        res = (CgResult * nrhs)()
        rc = lib().cps_hip_invcg_device(self.h, ctypes.byref(op), sol.ptr, src.ptr, nrhs, ctypes.byref(arg), res)
        _check(rc, "cps_hip_invcg_device")
        return [res[k] for k in range(nrhs)]

    def matinv(self, op, b, arg, guess=None, nrhs=1):
        out = np.zeros(2 * self._nreal(op) * nrhs) if guess is None else np.ascontiguousarray(guess, dtype=np.float64).copy()
        res = (CgResult * nrhs)()
        _check(lib().cps_hip_matinv(self.h, ctypes.byref(op), _ptr(out), _ptr(b), nrhs, ctypes.byref(arg), res),
               "cps_hip_matinv")
        return out, [res[k] for k in range(nrhs)]

    def import_gauge_canonical(self, u_canon):
        """CANONICAL gauge without boundary signs; the handle's bc signs are
        applied on the device (Lattice::Convert + BondCond)."""
        a = np.ascontiguousarray(u_canon, dtype=np.float64)
        _check(lib().cps_hip_gauge_import_canonical(self.h, _ptr(a)), "cps_hip_gauge_import_canonical")

    def fmatinv(self, op, b, arg, guess=None, nrhs=1):
        """Lattice::FmatInv on CANONICAL full vectors (5-D ones s-major)."""
        b = np.ascontiguousarray(b, dtype=np.float64)
        out = np.zeros_like(b) if guess is None else np.ascontiguousarray(guess, dtype=np.float64).copy()
        res = (CgResult * nrhs)()
        _check(lib().cps_hip_fmatinv(self.h, ctypes.byref(op), _ptr(out), _ptr(b), nrhs, ctypes.byref(arg), res),
               "cps_hip_fmatinv")
        return out, [res[k] for k in range(nrhs)]

    def qpropw_point(self, op, x0, arg):
        """QPropW::Run for a point source at x0: the 12 spin-colour solves as
        one batched MatInv; returns the propagator as an array
        [t][z][y][x][s1][c1][spin][colour] of complex (sink s1, c1; source
        spin, colour) and the 12 solve results (rhs = 3 spin + colour)."""
        V = int(np.prod(self.L))
        prop = np.zeros(V * 288)
        res = (CgResult * 12)()
        xs = (ctypes.c_int * 4)(*[int(v) for v in x0])
        _check(lib().cps_hip_qpropw_point(self.h, ctypes.byref(op), xs, ctypes.byref(arg), _ptr(prop), res),
               "cps_hip_qpropw_point")
        p = prop.reshape(V, 4, 3, 4, 3, 2)
        z = (p[..., 0] + 1j * p[..., 1]).reshape(self.L[3], self.L[2], self.L[1], self.L[0], 4, 3, 4, 3)
        return z, [res[k] for k in range(12)]

    def bench_hop(self, kind, precision, nrhs=1, reps=20):
        s, b = ctypes.c_double(), ctypes.c_double()
        _check(lib().cps_hip_bench_hop(self.h, kind, precision, nrhs, reps, ctypes.byref(s), ctypes.byref(b)),
               "cps_hip_bench_hop")
        return s.value, b.value


    def bench_copy(self, n_elem, reps=20):
        s = ctypes.c_double()
        _check(lib().cps_hip_bench_copy(self.h, n_elem, reps, ctypes.byref(s)), "cps_hip_bench_copy")
        return s.value


class Service:
    """Resident solve service (cps_hip_service_*): gauge fields stay on the
    device; submit() queues a solve of nrhs sources and returns a request id
    at once, wait() blocks for its solutions, per-rhs results and latency."""

    def __init__(self, hbm_budget_bytes, max_slots=12):
        h = ctypes.c_void_p()
        _check(lib().cps_hip_service_create(ctypes.byref(h), int(hbm_budget_bytes), int(max_slots)),
               "cps_hip_service_create")
        self.h = h
        self._gauge_ls = {}
        self._live = {}

    def add_gauge(self, node_sites, ls, u_wilson, bc=(0, 0, 0, 1)):
        gid = ctypes.c_int()
        dims = (ctypes.c_int * 4)(*[int(v) for v in node_sites])
        bcs = (ctypes.c_int * 4)(*[int(v) for v in bc])
        u = np.ascontiguousarray(u_wilson, dtype=np.float64)
        _check(lib().cps_hip_service_add_gauge(self.h, dims, int(ls), bcs, _ptr(u), ctypes.byref(gid)),
               "cps_hip_service_add_gauge")
        self._gauge_ls[gid.value] = (list(node_sites), int(ls))
        return gid.value

    def submit(self, gauge_id, op, src, nrhs=1, stop_rsd=1e-10, max_num_iter=10000, precision=PREC_DOUBLE,
               reliable_every=100, reliable_delta=0.1):
        """precision PREC_DOUBLE: FP64 CG; a mixed mode: the sloppy CG of that
        mode with FP64 reliable updates (cps_hip_service_submit_arg)."""
        src = np.ascontiguousarray(src, dtype=np.float64)
        sol = np.zeros_like(src)
        rid = ctypes.c_longlong()
        if precision == PREC_DOUBLE:
            _check(lib().cps_hip_service_submit(self.h, int(gauge_id), ctypes.byref(op), _ptr(src), _ptr(sol),
                                                int(nrhs), float(stop_rsd), int(max_num_iter), ctypes.byref(rid)),
                   "cps_hip_service_submit")
        else:
            arg = make_cg_arg(stop_rsd, max_num_iter, precision, reliable_every, reliable_delta)
            _check(lib().cps_hip_service_submit_arg(self.h, int(gauge_id), ctypes.byref(op), _ptr(src), _ptr(sol),
                                                    int(nrhs), ctypes.byref(arg), ctypes.byref(rid)),
                   "cps_hip_service_submit_arg")
        self._live[rid.value] = (src, sol, nrhs)   # buffers stay alive until wait()
        return rid.value

    def wait(self, request_id):
        src, sol, nrhs = self._live[request_id]
        res = (CgResult * nrhs)()
        lat = ctypes.c_double()
        rc = lib().cps_hip_service_wait(self.h, ctypes.c_longlong(request_id), res, ctypes.byref(lat))
        del self._live[request_id]
        _check(rc, "cps_hip_service_wait")
        return sol, [res[k] for k in range(nrhs)], lat.value

    def stats(self):
        st = ServiceStats()
        _check(lib().cps_hip_service_get_stats(self.h, ctypes.byref(st)), "cps_hip_service_get_stats")
        return {f: getattr(st, f) for f, _ in ServiceStats._fields_}

    def close(self):
        if getattr(self, "h", None):
            lib().cps_hip_service_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    """A device allocation made by the library's own HIP runtime (so that no
    second runtime is loaded into the process)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _check(lib().cps_hip_device_alloc(ctypes.byref(p), self.nbytes), "cps_hip_device_alloc")
        self.ptr = p

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a):
        assert a.dtype == np.float64 and a.flags.c_contiguous and a.nbytes == self.nbytes
        _check(lib().cps_hip_memcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), self.nbytes, 0), "cps_hip_memcpy")

    def download(self):
        a = np.empty(self.nbytes // 8)
        _check(lib().cps_hip_memcpy(a.ctypes.data_as(ctypes.c_void_p), self.ptr, self.nbytes, 1), "cps_hip_memcpy")
        return a

    def copy_from(self, other):
        assert other.nbytes == self.nbytes
        _check(lib().cps_hip_memcpy(self.ptr, other.ptr, self.nbytes, 2), "cps_hip_memcpy")

    def free(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            lib().cps_hip_device_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_count():
    return int(lib().cps_hip_device_count())
