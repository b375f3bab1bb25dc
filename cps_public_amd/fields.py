"""Seeded synthetic lattice fields (host side, numpy).

The gauge field follows the construction every benchmark configuration names:
per link U = exp(i * eps * H), with H a traceless Hermitian 3x3 matrix whose
entries are Gaussian, from a counter-based generator (numpy's Philox) keyed by
a 64-bit seed.  The exponential uses scaling and squaring of a Taylor series,
and each link is then re-unitarised (Gram-Schmidt on rows 0 and 1, row 2 the
conjugated cross product) so that it is SU(3) to rounding, which the 12-real
link compression relies on.

Fields are produced in the CPS CANONICAL layouts:
  gauge  : [t][z][y][x][mu][row][col][re,im]   (float64, 72 reals per site)
  fermion: [t][z][y][x][spin][colour][re,im]   (float64, 24 reals per site)
  5-D    : Ls canonical 4-D fields, s-major.
"""
import numpy as np


def _rng(seed, stream=0):
    return np.random.Generator(np.random.Philox(key=[int(seed) & (2**64 - 1), int(stream)]))


def _mul3(a, b):
    """Product of two stacks of 3x3 matrices held as [row][col] lists of
    arrays (structure of arrays: one vector per matrix element)."""
    return [[a[i][0] * b[0][j] + a[i][1] * b[1][j] + a[i][2] * b[2][j] for j in range(3)] for i in range(3)]


def _expm_i_eps_h(h, eps, order=10, squarings=5, chunk=1 << 16):
    """exp(i*eps*h) for a stack of 3x3 Hermitian matrices: Taylor series of
    order `order` after `squarings` halvings, then repeated squaring
    (|eps h| / 2^5 < 0.1 for the spreads used here, so the truncation error
    is below 1e-16).  Chunks of the stack are processed element-wise."""
    n = h.shape[0]
    out = np.empty((n, 3, 3), dtype=np.complex128)
    f = 1j * eps / (2 ** squarings)
    for c0 in range(0, n, chunk):
        hc = h[c0:c0 + chunk]
        a = [[f * hc[:, i, j] for j in range(3)] for i in range(3)]
        one = np.ones(hc.shape[0], dtype=np.complex128)
        zero = np.zeros(hc.shape[0], dtype=np.complex128)
        term = [[one if i == j else zero for j in range(3)] for i in range(3)]
        res = [[term[i][j].copy() for j in range(3)] for i in range(3)]
        for k in range(1, order + 1):
            term = _mul3(term, a)
            term = [[term[i][j] / k for j in range(3)] for i in range(3)]
            res = [[res[i][j] + term[i][j] for j in range(3)] for i in range(3)]
        for _ in range(squarings):
            res = _mul3(res, res)
        for i in range(3):
            for j in range(3):
                out[c0:c0 + chunk, i, j] = res[i][j]
    return out


def reunitarize(u):
    """Project a stack of 3x3 complex matrices onto SU(3) (rows orthonormal,
    det = +1 through the cross-product third row)."""
    r0 = u[:, 0, :]
    r0 = r0 / np.linalg.norm(r0, axis=1, keepdims=True)
    r1 = u[:, 1, :]
    r1 = r1 - np.sum(np.conj(r0) * r1, axis=1, keepdims=True) * r0
    r1 = r1 / np.linalg.norm(r1, axis=1, keepdims=True)
    r2 = np.conj(np.cross(r0, r1))
    return np.stack([r0, r1, r2], axis=1)


def random_su3(n, eps, seed, stream=0, chunk=1 << 16):
    """n SU(3) matrices exp(i eps H), H traceless Hermitian with N(0,1) entries
    (real parts of all n matrices drawn first, then the imaginary parts);
    processed in chunks so the temporaries stay cache-sized."""
    g = _rng(seed, stream)
    re = g.standard_normal((n, 3, 3))
    im = g.standard_normal((n, 3, 3))
    out = np.empty((n, 3, 3), dtype=np.complex128)
    eye = np.eye(3)[None]
    for c0 in range(0, n, chunk):
        a = re[c0:c0 + chunk] + 1j * im[c0:c0 + chunk]
        h = 0.5 * (a + np.conj(np.transpose(a, (0, 2, 1))))
        tr = np.trace(h, axis1=1, axis2=2) / 3.0
        h = h - tr[:, None, None] * eye
        out[c0:c0 + chunk] = reunitarize(_expm_i_eps_h(h, eps))
    return out


def gauge_field(L, eps, seed):
    """CANONICAL-order gauge field (float64, 72*V reals)."""
    v = int(np.prod(L))
    u = random_su3(4 * v, eps, seed, stream=1)
    out = np.empty((4 * v, 3, 3, 2))
    out[..., 0] = u.real
    out[..., 1] = u.imag
    return out.reshape(-1)


def unit_gauge(L):
    v = int(np.prod(L))
    u = np.zeros((4 * v, 3, 3, 2))
    for i in range(3):
        u[:, i, i, 0] = 1.0
    return u.reshape(-1)


def gaussian_fermion(nsites, seed, stream=2):
    """N(0,1) for every real of nsites spinors (24 reals each)."""
    return _rng(seed, stream).standard_normal(24 * int(nsites))


def point_source(L, ls, spin, colour, site=(0, 0, 0, 0), s=0):
    """CANONICAL point source delta(x-site) delta_{s,s0} e_{spin,colour}."""
    v = int(np.prod(L))
    f = np.zeros(24 * v * ls)
    x, y, z, t = site
    idx = x + L[0] * (y + L[1] * (z + L[2] * t))
    f[24 * (s * v + idx) + 2 * (colour + 3 * spin)] = 1.0
    return f
