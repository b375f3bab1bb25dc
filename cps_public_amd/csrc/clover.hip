// Clover-improved Wilson operator (filled in by the clover kernels).
#include "solver.h"

namespace cpship {

int clover_build(Lattice& L, const double* kappas, int n_kappa, double csw) {
  (void)L; (void)kappas; (void)n_kappa; (void)csw;
  return CPS_HIP_ERR_ARG;
}

int clover_matpc(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& out, const FieldRef& in,
                 const FieldRef& t1, int dag, double* partial, int nrhs) {
  (void)L; (void)op; (void)hprec; (void)out; (void)in; (void)t1; (void)dag; (void)partial; (void)nrhs;
  return CPS_HIP_ERR_ARG;
}

int clover_mdagm(Lattice& L, const cps_hip_op* op, int hprec, const FieldRef& out, const FieldRef& in,
                 const FieldRef& t1, const FieldRef& t2, double* partial, int nrhs) {
  (void)L; (void)op; (void)hprec; (void)out; (void)in; (void)t1; (void)t2; (void)partial; (void)nrhs;
  return CPS_HIP_ERR_ARG;
}

}  // namespace cpship
