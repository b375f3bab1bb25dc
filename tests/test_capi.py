"""CPU checks of the C ABI library: it is built, it loads without a GPU, and
it exports every entry point include/cps_hip.h declares."""
import ctypes
import os
import re

import pytest

import cps_public_amd as C


def declared_functions():
    text = open(C.HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cps_hip_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ["cps_hip_wilson_dslash", "cps_hip_wilson_mdagm", "cps_hip_dwf_dslash", "cps_hip_dwf_mdagm",
                 "cps_hip_invcg", "cps_hip_matinv", "cps_hip_matpc_dag_matpc", "cps_hip_lattice_create"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(C.LIB_PATH), "libcps_hip.so not built"
    lib = ctypes.CDLL(C.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_version_string():
    assert b"gfx950" in C.lib().cps_hip_version()


def test_null_handle_is_an_argument_error():
    # every entry point validates its handle before touching the device
    l = C.lib()
    assert l.cps_hip_lattice_destroy(None) == -1
    assert l.cps_hip_gauge_import(None, None) == -1
    assert l.cps_hip_wilson_dslash(None, None, None, None, 0, 0) == -1


def test_tuning_controls_round_trip():
    """cps_hip_set_tuning / cps_hip_get_tuning (no device needed): the fields
    read back as set, out-of-range values are refused, the defaults come
    back."""
    base = C.get_tuning()
    assert (base.use_graph, base.tune, base.lagged_x, base.paired_x, base.clover_r_in_hop) == (1, 1, 1, 1, 0)
    t = C.get_tuning()
    t.paired_x, t.hop_mode, t.progress_seconds = 0, 2, 5.0
    C.set_tuning(t)
    got = C.get_tuning()
    assert (got.paired_x, got.hop_mode, got.progress_seconds) == (0, 2, 5.0)
    bad = C.get_tuning()
    bad.hop_mode = 9
    with pytest.raises(C.CpsHipError):
        C.set_tuning(bad)
    C.set_tuning(base)
    assert C.get_tuning().paired_x == 1
