"""Host-side CPS layout conversion (numpy, vectorised).

CANONICAL <-> WILSON order for the gauge field (convert_func.C:318-415: site
(x - x%2 + Lx*(y+Ly*(z+Lz*t)) + V*parity)/2 with even sites first, link
element (row,col) moved to 2*(row+3*col)), for 4-D fermions
(convert_func.C:429-494: odd sites first) and for 5-D domain-wall fermions
(convert.C:635-760), plus the antiperiodic boundary toggle of
Lattice::BondCond (src/util/lattice/lattice_base/lattice_base.C:5313-5369).
"""
import numpy as np


def _coords(L):
    Lx, Ly, Lz, Lt = L
    t, z, y, x = np.meshgrid(np.arange(Lt), np.arange(Lz), np.arange(Ly), np.arange(Lx), indexing="ij")
    return x.ravel(), y.ravel(), z.ravel(), t.ravel()


def _cb_index(L, x, y, z, t):
    return x // 2 + (L[0] // 2) * (y + L[1] * (z + L[2] * t))


def gauge_site_perm(L):
    """WILSON site index of every CANONICAL site (even parity block first)."""
    x, y, z, t = _coords(L)
    vcb = int(np.prod(L)) // 2
    par = (x + y + z + t) % 2
    return _cb_index(L, x, y, z, t) + vcb * par


def ferm_site_perm(L):
    """WILSON (odd first) site index of every CANONICAL site."""
    x, y, z, t = _coords(L)
    vcb = int(np.prod(L)) // 2
    par = (x + y + z + t + 1) % 2
    return _cb_index(L, x, y, z, t) + vcb * par


def boundary_factor(bc, twist=None):
    """Factor Lattice::BondCond applies to the last slice of each direction
    (lattice_base.C:5313-5369): 1 periodic, -1 antiperiodic, GJP::TwistPhase
    e^{-i pi theta} twisted (gjp.C:475-493; theta = twist[mu] in units of pi)."""
    f = []
    for mu in range(4):
        if bc[mu] == 1:
            f.append(-1.0 + 0j)
        elif bc[mu] == 2:
            th = 0.0 if twist is None else float(twist[mu])
            f.append(complex(np.cos(np.pi * th), -np.sin(np.pi * th)))
        else:
            f.append(1.0 + 0j)
    return f


def bond_cond(u_canon, L, bc, twist=None):
    """Multiply the mu-links on the last slice of every non-periodic
    direction by its boundary factor (CANONICAL order; real fields stay real
    for periodic / antiperiodic conditions)."""
    fac = boundary_factor(bc, twist)
    u = u_canon.reshape(L[3], L[2], L[1], L[0], 4, 9, 2).copy()
    sl = {0: (slice(None), slice(None), slice(None), L[0] - 1),
          1: (slice(None), slice(None), L[1] - 1, slice(None)),
          2: (slice(None), L[2] - 1, slice(None), slice(None)),
          3: (L[3] - 1, slice(None), slice(None), slice(None))}
    for mu in range(4):
        f = fac[mu]
        if f == 1:
            continue
        blk = u[sl[mu] + (mu,)]
        re, im = blk[..., 0].copy(), blk[..., 1].copy()
        blk[..., 0] = f.real * re - f.imag * im
        blk[..., 1] = f.real * im + f.imag * re
        u[sl[mu] + (mu,)] = blk
    return u.reshape(-1)


def gauge_canon_to_wilson(u_canon, L):
    v = int(np.prod(L))
    src = u_canon.reshape(v, 4, 3, 3, 2)              # [site][mu][row][col][ri]
    dst = np.empty((v, 4, 3, 3, 2))                    # [site][mu][col][row][ri]
    dst[gauge_site_perm(L)] = np.transpose(src, (0, 1, 3, 2, 4))
    return dst.reshape(-1)


def gauge_wilson_to_canon(u_w, L):
    v = int(np.prod(L))
    src = u_w.reshape(v, 4, 3, 3, 2)[gauge_site_perm(L)]
    return np.ascontiguousarray(np.transpose(src, (0, 1, 3, 2, 4))).reshape(-1)


def ferm_canon_to_wilson(f, L):
    v = int(np.prod(L))
    dst = np.empty((v, 24))
    dst[ferm_site_perm(L)] = f.reshape(v, 24)
    return dst.reshape(-1)


def ferm_wilson_to_canon(f, L):
    v = int(np.prod(L))
    return f.reshape(v, 24)[ferm_site_perm(L)].reshape(-1)


def ferm5_canon_to_wilson(f, L, ls):
    """5-D DWF order: odd 5-D checkerboard (slice s at 4-D parity (s+1)%2)
    then even 5-D checkerboard (slice s at 4-D parity s%2)."""
    v = int(np.prod(L))
    vcb = v // 2
    w = np.stack([ferm_canon_to_wilson(f[s * 24 * v:(s + 1) * 24 * v], L).reshape(2, vcb * 24) for s in range(ls)])
    # w[s, 0] = odd 4-D half, w[s, 1] = even 4-D half
    odd5 = [w[s, 0] if (s + 1) % 2 == 1 else w[s, 1] for s in range(ls)]
    even5 = [w[s, 0] if s % 2 == 1 else w[s, 1] for s in range(ls)]
    return np.concatenate(odd5 + even5)


def ferm5_wilson_to_canon(f, L, ls):
    v = int(np.prod(L))
    vcb = v // 2
    h = f.reshape(2, ls, vcb * 24)
    out = []
    for s in range(ls):
        odd4 = h[0, s] if (s + 1) % 2 == 1 else h[1, s]
        even4 = h[1, s] if (s + 1) % 2 == 1 else h[0, s]
        out.append(ferm_wilson_to_canon(np.concatenate([odd4, even4]), L))
    return np.concatenate(out)


def ferm5_canon_to_4dprec(f, L, ls):
    """DWF_4D_EOPREC order: the odd 4-D half of every slice, then the even."""
    v = int(np.prod(L))
    vcb = v // 2
    w = np.stack([ferm_canon_to_wilson(f[s * 24 * v:(s + 1) * 24 * v], L).reshape(2, vcb * 24) for s in range(ls)])
    return np.concatenate([w[:, 0].reshape(-1), w[:, 1].reshape(-1)])


def ferm5_4dprec_to_canon(f, L, ls):
    v = int(np.prod(L))
    vcb = v // 2
    h = f.reshape(2, ls, vcb * 24)
    return np.concatenate([ferm_wilson_to_canon(np.concatenate([h[0, s], h[1, s]]), L) for s in range(ls)])
