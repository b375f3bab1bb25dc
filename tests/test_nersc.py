"""NERSC gauge files (cps_hip_nersc_read / _write, host code, no GPU).

The reader is pinned by files written here independently in numpy (header
text, big- and little-endian data, two-row and 3x3 links), by a checksum
computed from the file bytes (32-bit big-endian words summed mod 2^32,
fpconv.C:320-361), and by a plaquette computed with numpy matrix products;
the writer by reading its files back with the same independent code."""
import os

import numpy as np
import pytest

import cps_public_amd as C
from cps_public_amd import fields as F

L = [4, 4, 4, 6]


def plaquette_numpy(u, L):
    """sum_x sum_{mu<nu} ReTr U_mu(x) U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+ / (18 V)
    and sum ReTr U / (12 V) for a CANONICAL field."""
    U = u.reshape(L[3], L[2], L[1], L[0], 4, 3, 3, 2)
    U = U[..., 0] + 1j * U[..., 1]                   # [t][z][y][x][mu][r][c]
    V = int(np.prod(L))
    axis = {0: 3, 1: 2, 2: 1, 3: 0}                  # direction -> array axis
    s = 0.0
    for mu in range(4):
        for nu in range(mu + 1, 4):
            Um, Un = U[..., mu, :, :], U[..., nu, :, :]
            Un_xm = np.roll(Un, -1, axis=axis[mu])
            Um_xn = np.roll(Um, -1, axis=axis[nu])
            P = Um @ Un_xm @ np.conj(np.swapaxes(Um_xn, -1, -2)) @ np.conj(np.swapaxes(Un, -1, -2))
            s += np.trace(P, axis1=-2, axis2=-1).real.sum()
    lt = np.trace(U, axis1=-2, axis2=-1).real.sum() / (12.0 * V)
    return s / (18.0 * V), lt


def write_numpy(path, u, L, rows=2, dtype=">f8", checksum=None, plaq=None, fp="IEEE64BIG"):
    """An independent NERSC writer: header lines, then per site the 4 links'
    first `rows` rows (complex, row-major) in `dtype`."""
    V = int(np.prod(L))
    links = u.reshape(V, 4, 3, 3, 2)[:, :, :rows, :, :]
    data = np.ascontiguousarray(links.astype(dtype)).tobytes()
    words = np.frombuffer(data, dtype=">u4" if dtype.startswith(">") else "<u4")
    cs = int(words.astype(np.uint64).sum() % (1 << 32)) if checksum is None else checksum
    p, lt = plaquette_numpy(u, L)
    if plaq is None:
        plaq = p
    hdr = ["BEGIN_HEADER", "HDR_VERSION = 1.0",
           "DATATYPE = " + ("4D_SU3_GAUGE" if rows == 2 else "4D_SU3_GAUGE_3x3"), "STORAGE_FORMAT = 1.0"]
    hdr += [f"DIMENSION_{d + 1} = {L[d]}" for d in range(4)]
    hdr += [f"LINK_TRACE = {lt:.10f}", f"PLAQUETTE  = {plaq:.10f}"]
    hdr += [f"BOUNDARY_{d + 1} = PERIODIC" for d in range(4)]
    hdr += [f"CHECKSUM = {cs:x}", "ENSEMBLE_ID = test", "SEQUENCE_NUMBER = 7", f"FLOATING_POINT = {fp}",
            "END_HEADER"]
    with open(path, "wb") as f:
        f.write(("\n".join(hdr) + "\n").encode())
        f.write(data)
    return cs, p, lt


@pytest.fixture(scope="module")
def gauge():
    return F.gauge_field(L, 0.3, 240)


@pytest.mark.parametrize("rows,dtype,fp,tol", [(3, ">f8", "IEEE64BIG", 0.0), (2, ">f8", "IEEE64BIG", 1e-14),
                                               (2, "<f8", "IEEE64LITTLE", 1e-14), (3, ">f4", "IEEE32BIG", 1e-6),
                                               (2, ">f4", "IEEE32", 2e-6)])
def test_read_independent_file(tmp_path, gauge, rows, dtype, fp, tol):
    path = str(tmp_path / "cfg.nersc")
    cs, p, lt = write_numpy(path, gauge, L, rows, dtype, fp=fp)
    u, info = C.nersc_read(path)
    assert list(info.dims) == L
    assert info.recon_row_3 == (1 if rows == 2 else 0)
    assert info.checksum_data == cs == info.checksum_header
    assert np.abs(u - gauge).max() <= tol
    assert abs(info.plaquette_data - p) < 1e-12 + tol
    assert abs(info.link_trace_data - lt) < 1e-12 + tol


def test_unit_gauge_plaquette_one(tmp_path):
    u = F.unit_gauge(L)
    path = str(tmp_path / "unit.nersc")
    C.nersc_write(path, L, u, recon_row_3=1, bytes_per_real=8)
    v, info = C.nersc_read(path)
    assert np.array_equal(v, u)
    assert abs(info.plaquette_data - 1.0) < 1e-15 and abs(info.link_trace_data - 1.0) < 1e-15


@pytest.mark.parametrize("recon,bpr", [(0, 8), (1, 8), (1, 4)])
def test_writer_round_trip(tmp_path, gauge, recon, bpr):
    """The writer's file read back by the reader, and its header and data
    checked independently (checksum of the bytes, numpy plaquette)."""
    path = str(tmp_path / "w.nersc")
    C.nersc_write(path, L, gauge, recon_row_3=recon, bytes_per_real=bpr, ensemble_id="e1", sequence_number=3)
    u, info = C.nersc_read(path)
    tol = 0.0 if (recon == 0 and bpr == 8) else (1e-14 if bpr == 8 else 2e-6)
    assert np.abs(u - gauge).max() <= tol
    raw = open(path, "rb").read()
    data = raw[raw.index(b"END_HEADER\n") + len(b"END_HEADER\n"):]
    V = int(np.prod(L))
    assert len(data) == V * 4 * (12 if recon else 18) * bpr
    cs = int(np.frombuffer(data, dtype=">u4").astype(np.uint64).sum() % (1 << 32))
    assert info.checksum_header == cs
    p, _ = plaquette_numpy(gauge, L)
    assert abs(info.plaquette_header - p) < 1e-9


def test_checksum_and_plaquette_errors(tmp_path, gauge):
    path = str(tmp_path / "bad.nersc")
    cs, p, _ = write_numpy(path, gauge, L, 3, ">f8", checksum=None)
    write_numpy(path, gauge, L, 3, ">f8", checksum=(cs + 1) % (1 << 32))
    with pytest.raises(C.CpsHipError) as e:
        C.nersc_read(path)
    assert e.value.code == -8
    write_numpy(path, gauge, L, 3, ">f8", plaq=p * 1.05)
    with pytest.raises(C.CpsHipError) as e:
        C.nersc_read(path)
    assert e.value.code == -9
    C.nersc_read(path, chkprec=0.1)   # within a looser precision


def test_truncated_and_missing(tmp_path, gauge):
    path = str(tmp_path / "t.nersc")
    write_numpy(path, gauge, L, 2, ">f8")
    raw = open(path, "rb").read()
    open(path, "wb").write(raw[:-100])
    with pytest.raises(C.CpsHipError) as e:
        C.nersc_read(path)
    assert e.value.code == -7
    with pytest.raises(C.CpsHipError):
        C.nersc_read(str(tmp_path / "none.nersc"))
