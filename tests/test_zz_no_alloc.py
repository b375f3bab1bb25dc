"""GPU: MatPcDagMatPc without device allocation
(cps_hip_matpc_dag_matpc_noalloc) against the allocating entry point
cps_hip_matpc_dag_matpc, bit for bit (the same kernels on the same inputs),
on two calls in a row, the second with a smaller input, for the Wilson and
DWF operators; a request beyond the scratch array is refused."""
import numpy as np
import pytest

import cps_public_amd as C
from cps_public_amd import fields as F
from tests.helpers import wilson_gauge

pytestmark = pytest.mark.gpu


def _case(kind):
    if kind == "wilson":
        L, ls = [8, 8, 8, 8], 1
        op = C.make_op(C.OP_WILSON, mass=0.1)
    else:
        L, ls = [4, 4, 4, 8], 8
        op = C.make_op(C.OP_DWF, mass=0.02, m5=1.8)
    lat = C.Lattice(L, ls, (0, 0, 0, 1))
    lat.import_gauge(wilson_gauge(L, 0.2, 610))
    return lat, op


@pytest.mark.parametrize("kind", ["wilson", "dwf"])
def test_noalloc_matches_allocating_entry_point(kind):
    lat, op = _case(kind)
    n = lat._nreal(op)
    print("noalloc case", kind, flush=True)
    # first call 3 rhs, second call 1 rhs (smaller): both equal the
    # allocating entry point's output and |M in|^2
    for call, nrhs in enumerate((3, 1)):
        src = np.concatenate([F.gaussian_fermion(n // 24, 620 + 10 * call + k) for k in range(nrhs)])
        ref, dref = lat.mdagm(op, src, nrhs=nrhs)
        got, dgot = lat.mdagm_noalloc(op, src, nrhs=nrhs)
        print("noalloc call", call, nrhs, float(np.max(np.abs(got - ref))), flush=True)
        assert np.array_equal(got, ref)
        assert np.array_equal(dgot, dref)
    lat.close()


def test_noalloc_refuses_beyond_scratch():
    lat, op = _case("dwf")
    n = lat._nreal(op)
    limit = C.noalloc_scratch_bytes()
    nrhs = limit // (5 * n * 8) + 1   # five FP64 fields of nrhs vectors exceed the array
    src = np.zeros(n * nrhs)
    with pytest.raises(C.CpsHipError) as e:
        lat.mdagm_noalloc(op, src, nrhs=nrhs)
    assert e.value.code == C.ERR_TOO_LARGE
    lat.close()
