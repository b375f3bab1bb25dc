/*
 * cps_hip.h -- C ABI of the MI355X (gfx950) implementation of the CPS
 * even-odd preconditioned Wilson / domain-wall / Moebius / clover operators
 * and their conjugate-gradient inverter.
 *
 * The entry points mirror the CPS fermion "library" functions that the
 * DiracOp classes call (include/util/wilson.h:187-284, include/util/dwf.h:
 * 97-248 of the CPS source tree cps_pp/) and DiracOp::InvCg
 * (src/util/dirac_op/d_op_base/comsrc/inv_cg.C:84-454).  Vectors passed as
 * `double *` are host arrays in the CPS WILSON order the reference functions
 * take:
 *   4-D checkerboard spinor : [site_cb][spin 4][colour 3][re,im]
 *   full 4-D spinor         : [odd checkerboard][even checkerboard]
 *   5-D DWF checkerboard    : Ls slices of V/2 sites, slice s of checkerboard
 *                             cb at 4-D parity (s+cb)%2
 *   5-D Moebius (4-D prec.) : Ls slices of V/2 sites of one 4-D parity
 *   gauge                   : [parity][site_cb][mu][col][row][re,im]
 *                             (boundary factors already applied, as the
 *                             DiracOp constructor leaves it)
 * Multi-RHS variants take nrhs such vectors back to back.
 *
 * Every function returns 0 on success or a negative CPS_HIP_ERR_* code; no
 * function aborts the process (the C++ DiracOp front end maps codes to the
 * CPS ERR.* behaviour).  A handle is not thread safe; distinct handles may be
 * used from distinct threads.
 */
#ifndef CPS_HIP_H
#define CPS_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CPS_HIP_OK 0
#define CPS_HIP_ERR_ARG (-1)       /* bad argument (cb/dag not 0/1, sizes) */
#define CPS_HIP_ERR_RUNTIME (-2)   /* HIP runtime failure */
#define CPS_HIP_ERR_NOGAUGE (-3)   /* no gauge field resident */
#define CPS_HIP_ERR_D_ZERO (-4)    /* CG: <p,Ap> = 0 (inv_cg.C:326-329) */
#define CPS_HIP_ERR_NOMEM (-5)     /* device allocation failed / over budget */
#define CPS_HIP_ERR_NODEV (-6)     /* no HIP device */
#define CPS_HIP_ERR_IO (-7)        /* file missing, truncated or malformed */
#define CPS_HIP_ERR_CHECKSUM (-8)  /* NERSC checksum mismatch (ReadLatticePar.C:161-176) */
#define CPS_HIP_ERR_PLAQUETTE (-9) /* NERSC plaquette mismatch (ReadLatticePar.C:283-286) */
#define CPS_HIP_ERR_TOO_LARGE (-10) /* request exceeds the no-allocation scratch array */

/* operator kinds */
#define CPS_HIP_OP_WILSON 0
#define CPS_HIP_OP_DWF 1
#define CPS_HIP_OP_MOBIUS 2
#define CPS_HIP_OP_CLOVER 3

/* solver precision modes */
#define CPS_HIP_PREC_DOUBLE 0        /* plain FP64 CG (the CPS algorithm) */
/* Link storage of the sloppy operator: FP16/BF16 links are kept 18-real when
 * a site carries at least 4 spinors (DWF slots x rhs, Moebius slices x rhs,
 * rhs) and 12-real (third row rebuilt) otherwise; FP32 links are always
 * 12-real.  12-real links need a field that is SU(3) times the boundary
 * factor (checked at import, cps_hip_gauge_boundary_phase). */
#define CPS_HIP_PREC_MIXED_HALF 1    /* FP32 math, FP16 spinors + FP16 links (18- or 12-real) */
#define CPS_HIP_PREC_MIXED_BF16 2    /* FP32 math, FP16 spinors + BF16 links (18- or 12-real) */
#define CPS_HIP_PREC_MIXED_SINGLE 3  /* FP32 math, FP32 spinors + FP32 12-real links */

typedef struct cps_hip_lattice cps_hip_lattice;

/* Operator parameters: the union of what DiracOpWilson (d_op_wilson.C:134),
 * DiracOpDwf (dwf_init.C:97), DiracOpMobius (d_op_mobius.C:106-122) and
 * DiracOpClover (d_op_clover.C:107-122) read from CgArg and GJP. */
typedef struct {
  int kind;          /* CPS_HIP_OP_* */
  double mass;       /* CgArg.mass */
  double m5;         /* GJP.DwfHeight() */
  double a5_inv;     /* GJP.DwfA5Inv() */
  double mobius_b;   /* GJP.Mobius_b() */
  double mobius_c;   /* GJP.Mobius_c() */
  double csw;        /* GJP.CloverCoeff() */
} cps_hip_op;

/* CgArg subset used by InvCg (include/input/cg_arg.x:2-26) plus the
 * mixed-precision controls. */
typedef struct {
  int max_num_iter;      /* CgArg.max_num_iter */
  double stop_rsd;       /* CgArg.stop_rsd */
  int precision;         /* CPS_HIP_PREC_* */
  int reliable_every;    /* mixed: FP64 reliable update at least every N iterations */
  double reliable_delta; /* mixed: also when |r_sloppy| < delta * |r| at last update */
  int check_every;       /* host convergence polling interval (iterations), 0 = auto */
  int use_graph;         /* 1: replay the CG iteration as a captured hipGraph, 0: eager launches */
} cps_hip_cg_arg;

/* Per-rhs result of a solve. */
typedef struct {
  int iterations;        /* InvCg return value: iterations + 1 (inv_cg.C:446) */
  int status;            /* 1 converged, 2 max iterations, 3 d == 0 */
  double true_res;       /* |src - MdagM sol| / |src| in FP64 (inv_cg.C:403-417) */
  double src_norm_sq;
  int reliable_updates;  /* mixed precision only */
  double seconds;        /* wall time of the batched solve (same for every rhs) */
} cps_hip_cg_result;

/* ---- solver and launch controls ------------------------------------------
 * Process-wide switches of the solver's fused steps and of the hop launcher.
 * The defaults are the measured fastest path; each switch exists to compare
 * a fusion against its unfused form (the results of the two are
 * bit-identical except where noted) or to restrict the autotuner.  An
 * environment variable CPS_HIP_<NAME> (e.g. CPS_HIP_LAGGED_X=0), when set,
 * overrides the field of the same name.  No counterpart in CPS. */
typedef struct {
  int use_graph;        /* 1: CG iterations replayed as a captured hipGraph (ANDed with cps_hip_cg_arg.use_graph) */
  int tune;             /* 1: the hop autotuner times its candidates; 0: default launch */
  int tune_report;      /* 1: print the tuner's choices, 2: also every candidate's time */
  int hop_mode;         /* restrict the hop's thread mapping: 0 any, 1 element, 2 sitewave, 3 lds, 4 rhsgroup, 5 row-shared links (only reachable this way) */
  int rhs_window;       /* 1: kernels run on the contiguous range of rhs still iterating */
  int lagged_x;         /* mixed CG: solution step applied by the next direction update */
  int paired_x;         /* mixed CG: two solution steps per pass over x (FP16 directions) */
  int s5_xpay;          /* mixed Moebius Ls 16: first fifth-dimension step of M in the direction update */
  int r_in_hop;         /* mixed Moebius Ls 16 and DWF: residual step in the last hop of M^dag (rounds differently) */
  int clover_r_in_hop;  /* mixed clover: residual step in the last hop of MatPcDag (default 0: it spills) */
  int m5_fused;         /* half precision Ls 16: M5^-1 in the first hop's epilogue */
  int s5_fused;         /* Moebius M^dag: last fifth-dimension step in the hop epilogue */
  int s5x_row;          /* FP16 Ls 16: slots s+-1 of the s-term epilogue by DPP row rotations */
  int sloppy_vectors;   /* half precision: 0 x FP32 + r FP16, 1 both FP32, 2 x FP16 + r FP32, 3 both FP16 */
  int verbose;          /* 1: per-chunk CG diagnostics on stderr */
  double progress_seconds; /* > 0: a progress line on stderr at most this often during long solves */
} cps_hip_tuning;
int cps_hip_get_tuning(cps_hip_tuning *t);
int cps_hip_set_tuning(const cps_hip_tuning *t);

/* ---- lattice handle ------------------------------------------------------
 * Replaces wilson_init/dwf_init (wilson_init.C:52-112, dwf_init.C:41-99): it
 * holds the local geometry, the device scratch fields and the resident gauge.
 * bc[mu] is GJP.Bc(mu) as CPS's BndCndType (enum.x:271-277): 0 periodic,
 * 1 antiperiodic, 2 twisted (the G-parity kinds are refused).  bc is used
 * only by cps_hip_gauge_import_canonical, which applies the boundary factor
 * the way Lattice::BondCond does; a WILSON-order field passed to
 * cps_hip_gauge_import or to the library calls carries its factors already
 * (as the DiracOp constructor leaves it), so a handle for such fields may
 * pass bc = {0,0,0,0}.  ls may be 1 for 4-D operators. */
int cps_hip_lattice_create(cps_hip_lattice **lat, const int node_sites[4], int ls, const int bc[4]);
int cps_hip_lattice_destroy(cps_hip_lattice *lat);
/* Twist angles theta[mu] (GJP.TwistAngle, in units of pi) of the directions
 * created with bc = 2: the CANONICAL import multiplies their last slice by
 * GJP.TwistPhase = e^{-i pi theta} (gjp.C:475-493, lattice_base.C:5327-5349).
 * Takes effect at the next cps_hip_gauge_import_canonical. */
int cps_hip_lattice_set_twist(cps_hip_lattice *lat, const double twist_angle[4]);
/* Upload the WILSON-order FP64 gauge field (72*V/2*2 reals) and build the
 * device copies (FP64 18-real, and the low-precision copies on first use).
 * The boundary factor of every direction is measured from the field itself
 * (det U of its last-slice links): the 12-real formats rebuild the third row
 * as det(U) conj(row0 x row1), so they never depend on the handle's bc.  A
 * field bit-identical to the resident one is compared on the device and
 * changes nothing (derived copies and clover sets stay valid). */
int cps_hip_gauge_import(cps_hip_lattice *lat, const double *u_wilson);
/* The boundary factors measured at the last import: phase[2 mu], phase[2 mu
 * + 1] = det U (re, im) of the last-slice mu-links; *recon_ok (may be NULL)
 * = 1 if every interior link has det 1 and every last-slice link of each
 * direction the same det, to 1e-5 (required by the 12-real formats). */
int cps_hip_gauge_boundary_phase(const cps_hip_lattice *lat, double phase[8], int *recon_ok);
/* Device bytes currently held by the handle. */
long long cps_hip_lattice_device_bytes(const cps_hip_lattice *lat);

/* ---- Wilson library (wilson.h:187-284) ---------------------------------
 * u: the WILSON-order gauge; if non-NULL it is imported first (equivalent to
 * cps_hip_gauge_import: an unchanged field costs one transfer and a device
 * comparison), if NULL the resident gauge is used. */
/* wilson_dslash (wilson_dslash.C:173-515): chi on parity 1-cb from psi on cb */
int cps_hip_wilson_dslash(cps_hip_lattice *lat, double *chi, const double *u, const double *psi, int cb, int dag);
/* wilson_m / wilson_mdag (wilson_m.C:49-93, wilson_mdag.C:50-95) */
int cps_hip_wilson_m(cps_hip_lattice *lat, double *chi, const double *u, const double *psi, double kappa);
int cps_hip_wilson_mdag(cps_hip_lattice *lat, double *chi, const double *u, const double *psi, double kappa);
/* wilson_mdagm (wilson_mdagm.C:83-164); mp_sq (may be NULL) = |M psi|^2 */
int cps_hip_wilson_mdagm(cps_hip_lattice *lat, double *chi, const double *u, const double *psi, double *mp_sq,
                         double kappa);

/* ---- domain-wall library (dwf.h:97-248) --------------------------------- */
/* dwf_dslash (dwf_dslash.C:32-50): 4-D hop on every slice + 5th-dim hop */
int cps_hip_dwf_dslash(cps_hip_lattice *lat, double *out, const double *u, const double *in, double mass, int cb,
                       int dag, double m5, double a5_inv);
/* dwf_m / dwf_mdag / dwf_mdagm (dwf_m.C:31-62, dwf_mdag.C:32-62, dwf_mdagm.C:32-61) */
int cps_hip_dwf_m(cps_hip_lattice *lat, double *out, const double *u, const double *in, double mass, double m5,
                  double a5_inv);
int cps_hip_dwf_mdag(cps_hip_lattice *lat, double *out, const double *u, const double *in, double mass, double m5,
                     double a5_inv);
int cps_hip_dwf_mdagm(cps_hip_lattice *lat, double *out, const double *u, const double *in, double *dot_prd,
                      double mass, double m5, double a5_inv);

/* ---- generic operator entry points (DiracOp*::MatPcDagMatPc / MatPc) -----
 * For every kind: MatPcDagMatPc (dirac_op.h:262), with dot_prd = |MatPc in|^2
 * per rhs (may be NULL).  Clover requires cps_hip_clover_build first. */
int cps_hip_matpc_dag_matpc(cps_hip_lattice *lat, const cps_hip_op *op, double *out, const double *in,
                            double *dot_prd, int nrhs);
int cps_hip_matpc(cps_hip_lattice *lat, const cps_hip_op *op, double *out, const double *in, int dag, int nrhs);
/* cps_hip_matpc_dag_matpc for the Wilson and DWF operators without any
 * device allocation.  No counterpart in CPS: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The call's device buffers (staging of in
 * and out, the operator's fields, reduction partials) live in one static
 * device array of cps_hip_noalloc_scratch_bytes() (256 MiB) per GPU, held
 * until the process ends; a request that needs more (five FP64 fields of
 * nrhs vectors plus partials) returns CPS_HIP_ERR_TOO_LARGE.  The gauge field
 * must be resident (an import beforehand).  Every call on a GPU shares that
 * GPU's array: calls that overlap on one GPU (several threads or handles at
 * once) race.  Same results as cps_hip_matpc_dag_matpc. */
int cps_hip_matpc_dag_matpc_noalloc(cps_hip_lattice *lat, const cps_hip_op *op, double *out, const double *in,
                                    double *dot_prd, int nrhs);
size_t cps_hip_noalloc_scratch_bytes(void);
/* MatPcDagMatPc evaluated in the storage and arithmetic of a solver
 * precision mode (CPS_HIP_PREC_*): the FP64 host input is converted to that
 * format on the device, the operator runs there, and the result comes back
 * as FP64.  This is the sloppy operator the mixed-precision CG iterates with;
 * it has no counterpart in CPS (whose operators are FP64 only). */
int cps_hip_matpc_dag_matpc_prec(cps_hip_lattice *lat, const cps_hip_op *op, int precision, double *out,
                                 const double *in, double *dot_prd, int nrhs);

/* Clover: build A_oo and A_ee^{-1} on the device from the resident gauge for
 * the given kappa list (CloverMatChkb, d_op_clover_supp.C:492-540).  Solves
 * with nrhs = n_kappa * k use clover field rhs / k. */
int cps_hip_clover_build(cps_hip_lattice *lat, const double *kappas, int n_kappa, double csw);
/* Export the packed clover matrices of kappa index k (72 reals per site,
 * CPS clover_mat_mlt layout) of one checkerboard: cb 1 = A_oo, cb 0 = A_ee^{-1}. */
int cps_hip_clover_export(cps_hip_lattice *lat, int k, int cb, double *A);

/* ---- DiracOp::InvCg (inv_cg.C:84-454) -------------------------------------
 * Solves MdagM sol = src for nrhs right-hand sides at once.  sol holds the
 * initial guesses on entry.  With precision CPS_HIP_PREC_DOUBLE each rhs
 * follows the reference algorithm (same recurrences, stopping rule
 * |r|^2 <= stop_rsd^2 |src|^2 and FP64 true residual).  Mixed modes run the
 * sloppy CG with FP64 reliable updates and stop on the FP64 residual. */
int cps_hip_invcg(cps_hip_lattice *lat, const cps_hip_op *op, double *sol, const double *src, int nrhs,
                  const cps_hip_cg_arg *arg, cps_hip_cg_result *res);
/* The same solve with sol and src in device memory (CPS layout, FP64, on the
 * current device): no PCIe transfer; sources and solutions stay in HBM.
 * Same inputs, semantics and results as cps_hip_invcg. */
int cps_hip_invcg_device(cps_hip_lattice *lat, const cps_hip_op *op, double *sol, const double *src, int nrhs,
                         const cps_hip_cg_arg *arg, cps_hip_cg_result *res);

/* ---- MatInv (d_op_wilson.C:257-334, d_op_dwf.C:316-509,
 * d_op_mobius.C:389-640, d_op_clover.C:280-352) ----------------------------
 * Full-vector solve M out = in by even-odd preconditioning + InvCg; in and
 * out are full vectors in the DiracOp's order: [odd][even] (WILSON, DWF),
 * [even][odd] for Moebius (DWF_4D_EOPREC_EE, d_op_mobius.C:410-415).  out
 * holds the initial guess on entry.  in is preserved.  Clover needs exactly
 * one kappa set built (cps_hip_clover_build). */
int cps_hip_matinv(cps_hip_lattice *lat, const cps_hip_op *op, double *out, const double *in, int nrhs,
                   const cps_hip_cg_arg *arg, cps_hip_cg_result *res);

/* ---- CANONICAL order (Lattice::Convert / Fconvert on the device) ----------
 * cps_hip_gauge_import_canonical: u is the CANONICAL gauge field
 * [t][z][y][x][mu][row][col][re,im] (lattice_base.C:5350-5351) WITHOUT
 * boundary factors; the import transposes it to the WILSON order
 * (convert_func.C:318-415) and multiplies the last slice of each direction by
 * the factor of the handle's bc (-1 antiperiodic, e^{-i pi theta} twisted;
 * Lattice::BondCond, lattice_base.C:5313-5369), replacing what the DiracOp
 * constructor does to the Lattice's field (dirac_op_base.C:164,
 * d_op_wilson.C:118-121). */
int cps_hip_gauge_import_canonical(cps_hip_lattice *lat, const double *u_canon);
/* cps_hip_fmatinv: Lattice::FmatInv with CNV_FRM_YES (Fwilson::FmatInv
 * f_wilson.C:269-285, FdwfBase / Fmobius::FmatInv f_mobius.C:45-110,
 * Fclover::FmatInv): in and out are CANONICAL full vectors, 5-D ones s-major
 * ([s][t][z][y][x][spin][colour][re,im]), nrhs of them back to back; out
 * holds the initial guess.  The conversion to and from the DiracOp order
 * runs on the device.  Moebius multiplies the source by D_minus first, as
 * Fmobius::FmatInv does by default (dminus = 1, lattice.h:1124). */
int cps_hip_fmatinv(cps_hip_lattice *lat, const cps_hip_op *op, double *out, const double *in, int nrhs,
                    const cps_hip_cg_arg *arg, cps_hip_cg_result *res);

/* ---- NERSC gauge files (ReadLatticeParallel, ReadLatticePar.C:60-286;
 * LatticeHeader::read/write, latheader.C:210-312) -------------------------
 * Host-side reader and writer of the NERSC format: DATATYPE 4D_SU3_GAUGE
 * (two rows stored, third rebuilt) or 4D_SU3_GAUGE_3x3, FLOATING_POINT
 * IEEE32/64 big or little endian.  The field is the CANONICAL gauge field
 * [t][z][y][x][mu][row][col][re,im] in FP64 (V*72 doubles), e.g. for
 * cps_hip_gauge_import_canonical.  The reader checks the CHECKSUM (32-bit
 * sum of the data words, fpconv.C:320-361) and the PLAQUETTE to relative
 * precision chkprec (the reference's default is 0.01); the link trace is
 * reported, not enforced (ReadLatticePar.C:278-281).  The writer stores
 * IEEE32BIG or IEEE64BIG. */
typedef struct {
  int dims[4];
  int recon_row_3;      /* 1: two rows per link stored (4D_SU3_GAUGE) */
  int bytes_per_real;   /* 4 or 8 */
  int big_endian;
  unsigned int checksum_header, checksum_data;
  double plaquette_header, plaquette_data;
  double link_trace_header, link_trace_data;
} cps_hip_nersc_info;

int cps_hip_nersc_read_header(const char *path, cps_hip_nersc_info *info);
int cps_hip_nersc_read(const char *path, double *u_canon, cps_hip_nersc_info *info, double chkprec);
int cps_hip_nersc_write(const char *path, const int dims[4], const double *u_canon, int recon_row_3, int bytes_per_real,
                        const char *ensemble_id, int sequence_number);

/* ---- point-source propagator (QPropW::Run, QPropW.C:566-613) -------------
 * The 12 spin-colour point sources at site x0 (FermionVectorTp::
 * SetPointSource, FermionVector.C:1567-1644; for DWF/Moebius placed at
 * s = 0 / Ls-1 by Ffour2five, f_dwf_base.C:961-1050), solved as ONE batched
 * MatInv (QPropW::CG, QPropW.C:1077-1267; Moebius multiplies the source by
 * D_minus first, f_mobius.C:68-78), the 4-D solution taken by Ffive2four
 * (f_dwf_base.C:1068-1168) and stored as WilsonMatrix columns (LoadRow,
 * QPropW.C:1343-1364): prop[site][s1][c1][spin][colour][re,im], 288 doubles
 * per CANONICAL site x + Lx(y + Ly(z + Lz t)).  res receives the 12 solve
 * results (rhs k = 3 spin + colour).  Gauge fixing of the sink (FixSol) and
 * the midpoint propagator are not part of it. */
int cps_hip_qpropw_point(cps_hip_lattice *lat, const cps_hip_op *op, const int x0[4], const cps_hip_cg_arg *arg,
                         double *prop, cps_hip_cg_result *res);

/* ---- measurement helpers ------------------------------------------------ */
/* Time `reps` applications of the single-parity hop kernel of an operator
 * kind on resident random data, in the given precision; returns the mean
 * kernel time in seconds (HIP events on the launch stream) and the
 * algorithmic bytes moved per launch. */
int cps_hip_bench_hop(cps_hip_lattice *lat, int kind, int precision, int nrhs, int reps, double *seconds,
                      double *bytes);

/* Time `reps` launches of a plain device copy kernel (FP64 field to FP64
 * field) of n_elem spinors (192 bytes read + 192 bytes written each): the
 * known-size reference that scales hardware traffic counters. */
int cps_hip_bench_copy(cps_hip_lattice *lat, long long n_elem, int reps, double *seconds);

/* ---- resident solve service -----------------------------------------------
 * Keeps gauge fields on the device and serves propagator solves by
 * continuous batching: every (lattice, operator, precision) in use gets an
 * engine, a batched CG (FP64: inv_cg.C:293-364 per slot; or mixed precision
 * with FP64 reliable updates) over max_slots right-hand side slots on its own
 * stream; a slot whose rhs converged is harvested and refilled with the next
 * waiting rhs while the others iterate.  Wilson and DWF engines take a mass
 * per slot.  Engines run concurrently; an engine is created only while
 * the estimated device footprint fits hbm_budget_bytes (idle engines are
 * released first; an engine idle for 30 s is released anyway).  submit()
 * returns at once; src and sol (host, CPS layout,
 * nrhs vectors) must stay valid until wait() returns.  No counterpart in CPS. */
typedef struct cps_hip_service cps_hip_service;
typedef struct {
  long long requests_done;
  long long rhs_done;
  long long iterations;    /* CG iterations summed over engines */
  long long peak_bytes;    /* peak device footprint of gauges + engines */
  int engines_created;
  int engines_live;
  int last_error;
} cps_hip_service_stats;
int cps_hip_service_create(cps_hip_service **s, long long hbm_budget_bytes, int max_slots);
int cps_hip_service_destroy(cps_hip_service *s);
/* Registers a gauge field (WILSON order, boundary signs applied) for lattices
 * of the given extent and Ls; copies it to the device. */
int cps_hip_service_add_gauge(cps_hip_service *s, const int node_sites[4], int ls, const int bc[4],
                              const double *u, int *gauge_id);
int cps_hip_service_submit(cps_hip_service *s, int gauge_id, const cps_hip_op *op, const double *src, double *sol,
                           int nrhs, double stop_rsd, int max_num_iter, long long *request_id);
/* The same with the solver controls of cps_hip_cg_arg: precision (FP64 or a
 * mixed mode: the sloppy CG of that mode with FP64 reliable updates at least
 * every reliable_every iterations and when a slot's sloppy residual falls
 * below reliable_delta times its last FP64 residual; a slot finishes only on
 * its FP64 residual), stop_rsd, max_num_iter.  Requests of one lattice and
 * operator kind share an engine when precision and reliable-update controls
 * agree; Wilson and DWF requests of different masses share it too. */
int cps_hip_service_submit_arg(cps_hip_service *s, int gauge_id, const cps_hip_op *op, const double *src,
                               double *sol, int nrhs, const cps_hip_cg_arg *arg, long long *request_id);
/* Blocks until the request is done; res has nrhs entries; latency_s is the
 * time from submit to the last rhs harvested. */
int cps_hip_service_wait(cps_hip_service *s, long long request_id, cps_hip_cg_result *res, double *latency_s);
int cps_hip_service_get_stats(cps_hip_service *s, cps_hip_service_stats *st);

/* ---- device memory for device-resident callers ---------------------------
 * Plain device buffers on the library's own HIP runtime, for callers that
 * keep sources and solutions in HBM (cps_hip_invcg_device).  No counterpart
 * in CPS, whose vectors are host memory (smalloc). */
enum { CPS_HIP_H2D = 0, CPS_HIP_D2H = 1, CPS_HIP_D2D = 2 };
int cps_hip_device_alloc(void **p, size_t bytes);
int cps_hip_device_free(void *p);
/* Synchronous copy; kind is CPS_HIP_H2D, CPS_HIP_D2H or CPS_HIP_D2D. */
int cps_hip_memcpy(void *dst, const void *src, size_t bytes, int kind);

int cps_hip_device_count(void);
const char *cps_hip_version(void);

#ifdef __cplusplus
}
#endif
#endif
