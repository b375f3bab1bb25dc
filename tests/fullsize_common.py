"""Shared inputs and tolerances of the full-size GPU tests
(tests/test_gpu_full_c*.py), which run the benchmark workloads at their real
sizes: the operators against the oracle, a fixed number of FP64 CG
iterations against the oracle's CG vector for vector, and the complete
mixed-precision solves to the FP64 stopping rule -- the reference's own
solver tests self-check at the size they run (tests/f_wilson_inv/main.C:
183-246).  One module per workload, so that each workload's device fields are
released before the next one starts.

Tolerances: FP64 operators 1e-13 relative to the oracle (summation order
only); |M x|^2 (dot_prd) 1e-11 relative: the oracle sums the 10^8-10^9
squares of a full-size field one after another, so its own rounding error
grows like sqrt(N) eps ~ 1e-12, while the device sums pairwise; FP64 CG after
a fixed iteration count 1e-11 relative (the rounding of the two summation
orders, grown over the iterations); mixed solves: FP64 true residual <= 1e-10
by the CPS definition (inv_cg.C:403-417), the oracle's M^dag M applied to the
solution reproduces the source to 1e-10 relative, and the solution agrees
with the FP64 GPU CG's to 1e-6 relative (two solutions of one system with
normal-equation residuals <= 1e-10 differ by up to the condition number times
that residual; the measured values are printed)."""
import numpy as np

SOL_TOL = 1e-6
DOT_TOL = 1e-11
# the solver stops at a true residual <= 1e-10; the oracle's recomputation of
# |b - M^dag M x| / |b| differs from the device's by rounding (~1e-15)
ORA_RES_TOL = 1.01e-10


def point_sources(vcb, ls, nrhs, site_cb=0):
    """delta(x - (1,0,0,0)) delta_{s,0} e_{spin,colour}, rhs k = 3 spin +
    colour (mod 12): the odd site (1,0,0,0) is site_cb 0 of the checkerboard
    the CG works on."""
    per = 24 * vcb * ls
    src = np.zeros(per * nrhs)
    for k in range(nrhs):
        spin, colour = divmod(k % 12, 3)
        src[k * per + 24 * site_cb + 2 * (colour + 3 * spin)] = 1.0
    return src
