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


def _expm_i_eps_h(h, eps, order=14, squarings=5):
    """exp(i*eps*h) for a stack of 3x3 Hermitian matrices."""
    a = (1j * eps / (2 ** squarings)) * h
    n = h.shape[0]
    eye = np.broadcast_to(np.eye(3, dtype=np.complex128), (n, 3, 3))
    result = eye.copy()
    term = eye.copy()
    for k in range(1, order + 1):
        term = np.matmul(term, a) / k
        result = result + term
    for _ in range(squarings):
        result = np.matmul(result, result)
    return result


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


def random_su3(n, eps, seed, stream=0):
    """n SU(3) matrices exp(i eps H), H traceless Hermitian with N(0,1) entries."""
    g = _rng(seed, stream)
    re = g.standard_normal((n, 3, 3))
    im = g.standard_normal((n, 3, 3))
    a = re + 1j * im
    h = 0.5 * (a + np.conj(np.transpose(a, (0, 2, 1))))
    tr = np.trace(h, axis1=1, axis2=2) / 3.0
    h = h - tr[:, None, None] * np.eye(3)[None]
    return reunitarize(_expm_i_eps_h(h, eps))


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
