// MatPcDagMatPc without device allocation: the variant of
// cps_hip_matpc_dag_matpc for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  Every device buffer of the call beyond the resident gauge
// field and neighbour table of the handle -- the CPS-order staging of in and
// out, the operator's input, output and two intermediate fields, the
// reduction partials and the |M in|^2 values -- lives in one file-scope
// device array of NOALLOC_SCRATCH_BYTES per GPU, held until the process ends.
#include <cstring>
#include <mutex>

#include "solver.h"

namespace cpship {

constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(256) << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

// the device address of `scratch` on each GPU (each holds its own copy of
// the array): taken once per device with that device current; host code
// uses these addresses, never the array's name
static std::mutex scratch_mu;
static void* scratch_dev[NOALLOC_MAX_DEVICES];

static int scratch_ptr(void** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
    (void)hipGetLastError();
    return CPS_HIP_ERR_RUNTIME;
  }
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (!scratch_dev[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return CPS_HIP_ERR_RUNTIME;
    }
    scratch_dev[dev] = p;
  }
  *out = scratch_dev[dev];
  return 0;
}

static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace cpship

using namespace cpship;

extern "C" {

size_t cps_hip_noalloc_scratch_bytes(void) { return NOALLOC_SCRATCH_BYTES; }

int cps_hip_matpc_dag_matpc_noalloc(cps_hip_lattice* lat, const cps_hip_op* op, double* out, const double* in,
                                    double* dot_prd, int nrhs) {
  if (!lat || !op || !out || !in || nrhs < 1) return CPS_HIP_ERR_ARG;
  Lattice& L = lat->L;
  if (!L.has_gauge || !L.gauge_valid[0]) return CPS_HIP_ERR_NOGAUGE;
  // the Wilson and DWF operators apply as four hops through two
  // intermediate fields; Moebius and clover need further library fields
  if (op->kind != CPS_HIP_OP_WILSON && op->kind != CPS_HIP_OP_DWF) return CPS_HIP_ERR_ARG;
  OpLayout o;
  int rc = op_layout(op, L.ls, &o, L.g);
  if (rc) return rc;
  const int64_t n = o.per_rhs * nrhs;
  const int bpr = hop_blocks_per_rhs(L, op);
  // carve: CPS-order staging, operator input, output, t1, t2 (FP64 fields of
  // n spinors each), partials, |M in|^2
  const size_t field = align256((size_t)n * 24 * sizeof(double));
  const size_t parts = align256((size_t)bpr * nrhs * sizeof(double));
  const size_t dots = align256((size_t)nrhs * sizeof(double));
  const size_t need = 5 * field + parts + dots;
  if (need > NOALLOC_SCRATCH_BYTES) return CPS_HIP_ERR_TOO_LARGE;
  void* base = nullptr;
  if ((rc = scratch_ptr(&base))) return rc;
  char* b = static_cast<char*>(base);
  double* stage = reinterpret_cast<double*>(b);
  auto fref = [&](int k) {
    FieldRef f;
    f.data = b + (size_t)k * field;
    f.norm = nullptr;
    f.plane = n;   // FP64: 12 planes of n 16-byte chunks
    return f;
  };
  const FieldRef fin = fref(1), fout = fref(2), t1 = fref(3), t2 = fref(4);
  double* part = reinterpret_cast<double*>(b + 5 * field);
  double* dv = reinterpret_cast<double*>(b + 5 * field + parts);
  NoAllocScope scope;   // the hop launcher neither tunes nor builds tables
  CPSHIP_CHECK(hipMemcpyAsync(stage, in, (size_t)n * 24 * sizeof(double), hipMemcpyHostToDevice, L.stream));
  CPSHIP_CHECK(pack_spinor(STORE_D, stage, fin, o.map, L.g.Vcb, L.ls, 1, n, L.stream));
  if ((rc = apply_mdagm(L, op, PREC_D_D18, fout, fin, t1, t2, part, nrhs))) return rc;
  CPSHIP_CHECK(unpack_spinor(STORE_D, fout, stage, o.map, L.g.Vcb, L.ls, 1, n, L.stream));
  CPSHIP_CHECK(hipMemcpyAsync(out, stage, (size_t)n * 24 * sizeof(double), hipMemcpyDeviceToHost, L.stream));
  if (dot_prd) {
    CGScalars dummy;
    std::memset(&dummy, 0, sizeof(dummy));
    CPSHIP_CHECK(reduce_rhs(part, bpr, nrhs, dv, RED_PLAIN, dummy, L.stream));
    CPSHIP_CHECK(hipMemcpyAsync(dot_prd, dv, nrhs * sizeof(double), hipMemcpyDeviceToHost, L.stream));
  }
  CPSHIP_CHECK(hipStreamSynchronize(L.stream));
  return 0;
}

}  // extern "C"
