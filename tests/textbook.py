"""Textbook lattice Dirac operators, built as sparse matrices in the CANONICAL
site order, used to pin the oracle.

Nothing here calls the oracle or the product.  The definitions are the ones
the reference documents:
  - gamma matrices of the chiral basis, include/util/dirac_op.h:955-978
    (gamma_5 = gamma_x gamma_y gamma_z gamma_t = diag(1,1,-1,-1));
  - the Wilson hopping matrix
      H_xy = sum_mu (1 - gamma_mu) U_mu(x) delta_{y,x+mu}
                  + (1 + gamma_mu) U_mu(x-mu)^dag delta_{y,x-mu},
    M = A - kappa H with A = 1 (Wilson) or the clover matrix
      A(x) = 1 - 1/2 kappa c_sw sum_{mu<nu} [gamma_mu, gamma_nu] F_munu(x),
      F_munu = 1/8 (P_munu - P_munu^dag), P_munu the four clover leaves,
    include/util/dirac_op.h:915-951;
  - the Shamir domain-wall operator (5-D hopping 2/a5 P_+ psi(s-1) +
    2/a5 P_- psi(s+1), walls -m_f, kappa_5 = 1/(2(4 + 1/a5 - M5)),
    src/util/dirac_op/d_op_dwf/noarch/dwf_dslash_5_plus.C:111-340,
    dwf_init.C:97-98);
  - the Moebius operator (b D_W + 1) + (c D_W - 1) P with D_W the Wilson
    operator at mass -M5 (d_op_mobius.C:106-109 for kappa_b, kappa_c).

Vectors are complex arrays of 12 components (spin*3 + colour) per site; the
CANONICAL site index is n = x + Lx (y + Ly (z + Lz t)), and 5-D vectors are
[s][site].
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

GX = np.array([[0, 0, 0, 1j], [0, 0, 1j, 0], [0, -1j, 0, 0], [-1j, 0, 0, 0]])
GY = np.array([[0, 0, 0, -1], [0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0]], dtype=complex)
GZ = np.array([[0, 0, 1j, 0], [0, 0, 0, -1j], [-1j, 0, 0, 0], [0, 1j, 0, 0]])
GT = np.array([[0, 0, 1, 0], [0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0]], dtype=complex)
GAMMA = [GX, GY, GZ, GT]
G5 = np.diag([1, 1, -1, -1]).astype(complex)
PPLUS = np.diag([1, 1, 0, 0]).astype(complex)
PMINUS = np.diag([0, 0, 1, 1]).astype(complex)
I4 = np.eye(4, dtype=complex)


def coords(L):
    """(V, 4) integer coordinates of the CANONICAL sites."""
    V = int(np.prod(L))
    n = np.arange(V)
    x = n % L[0]
    y = (n // L[0]) % L[1]
    z = (n // (L[0] * L[1])) % L[2]
    t = n // (L[0] * L[1] * L[2])
    return np.stack([x, y, z, t], axis=1)


def index(L, c):
    c = np.asarray(c) % np.asarray(L)
    return c[..., 0] + L[0] * (c[..., 1] + L[1] * (c[..., 2] + L[2] * c[..., 3]))


def parity(L):
    return coords(L).sum(axis=1) % 2


def links(u_canon, L):
    """CANONICAL gauge reals [site][mu][row][col][re,im] -> (V, 4, 3, 3) complex."""
    a = np.asarray(u_canon).reshape(int(np.prod(L)), 4, 3, 3, 2)
    return a[..., 0] + 1j * a[..., 1]


def _block_sparse(rows, cols, blocks, n):
    """Sparse matrix from per-site 12x12 blocks placed at (rows[k], cols[k])."""
    k = np.arange(12)
    r = (12 * rows[:, None, None] + k[None, :, None]) + 0 * k[None, None, :]
    c = (12 * cols[:, None, None] + k[None, None, :]) + 0 * k[None, :, None]
    return sp.coo_matrix((blocks.reshape(-1), (r.reshape(-1), c.reshape(-1))), shape=(n, n)).tocsr()


def hopping(U, L, antiperiodic=(0, 0, 0, 0)):
    """Sparse 12V x 12V Wilson hopping matrix H (dag 0).  U: (V,4,3,3) links
    WITHOUT boundary signs; antiperiodic[mu] = 1 multiplies every hop across
    the mu boundary by -1."""
    V = int(np.prod(L))
    c = coords(L)
    n = np.arange(V)
    H = sp.csr_matrix((12 * V, 12 * V), dtype=complex)
    for mu in range(4):
        e = np.zeros(4, dtype=int)
        e[mu] = 1
        fwd = index(L, c + e)
        bwd = index(L, c - e)
        sf = np.where((c[:, mu] == L[mu] - 1) & bool(antiperiodic[mu]), -1.0, 1.0)
        sb = np.where((c[:, mu] == 0) & bool(antiperiodic[mu]), -1.0, 1.0)
        Uf = U[:, mu] * sf[:, None, None]
        Ub = np.conj(np.transpose(U[bwd, mu], (0, 2, 1))) * sb[:, None, None]
        bf = np.einsum("ab,nij->naibj", I4 - GAMMA[mu], Uf).reshape(V, 12, 12)
        bb = np.einsum("ab,nij->naibj", I4 + GAMMA[mu], Ub).reshape(V, 12, 12)
        H = H + _block_sparse(n, fwd, bf, 12 * V) + _block_sparse(n, bwd, bb, 12 * V)
    return H


def clover_blocks(U, L, kappa, csw):
    """(V, 12, 12) clover matrices A(x) of the documented formula (links
    U must already carry any boundary signs; closed loops are insensitive
    to them)."""
    V = int(np.prod(L))
    c = coords(L)
    dag = lambda m: np.conj(np.transpose(m, (0, 2, 1)))
    mm = lambda a, b: np.einsum("nij,njk->nik", a, b)

    def link(mu, shift):
        return U[index(L, c + shift), mu]

    A = np.zeros((V, 12, 12), dtype=complex)
    A[:] = np.eye(12)
    for mu in range(4):
        for nu in range(mu + 1, 4):
            em = np.zeros(4, dtype=int)
            en = np.zeros(4, dtype=int)
            em[mu] = 1
            en[nu] = 1
            z = np.zeros(4, dtype=int)
            p1 = mm(mm(link(mu, z), link(nu, em)), mm(dag(link(mu, en)), dag(link(nu, z))))
            p2 = mm(mm(link(nu, z), dag(link(mu, -em + en))), mm(dag(link(nu, -em)), link(mu, -em)))
            p3 = mm(mm(dag(link(mu, -em)), dag(link(nu, -em - en))), mm(link(mu, -em - en), link(nu, -en)))
            p4 = mm(mm(dag(link(nu, -en)), link(mu, -en)), mm(link(nu, em - en), dag(link(mu, z))))
            P = p1 + p2 + p3 + p4
            F = (P - dag(P)) / 8.0
            comm = GAMMA[mu] @ GAMMA[nu] - GAMMA[nu] @ GAMMA[mu]
            A -= 0.5 * kappa * csw * np.einsum("ab,nij->naibj", comm, F).reshape(V, 12, 12)
    return A


def block_diag_sparse(blocks):
    V = blocks.shape[0]
    n = np.arange(V)
    return _block_sparse(n, n, blocks, 12 * V)


def s_shift(ls, mass):
    """Ls x Ls pieces of the domain-wall 5-D hop T = P_+ psi(s-1) + P_- psi(s+1)
    with -mass at the walls, as (S_minus, S_plus): T = S_minus (x) P_+ +
    S_plus (x) P_-."""
    Sm = np.zeros((ls, ls))
    Sp = np.zeros((ls, ls))
    for s in range(ls):
        if s > 0:
            Sm[s, s - 1] = 1.0
        else:
            Sm[0, ls - 1] = -mass
        if s < ls - 1:
            Sp[s, s + 1] = 1.0
        else:
            Sp[ls - 1, 0] = -mass
    return Sm, Sp


def five_d_T(ls, mass, V):
    """Sparse T on a 5-D vector [s][site][12]."""
    Sm, Sp = s_shift(ls, mass)
    spin = lambda P: np.kron(P, np.eye(3))
    return (sp.kron(sp.csr_matrix(Sm), sp.kron(sp.identity(V), spin(PPLUS)))
            + sp.kron(sp.csr_matrix(Sp), sp.kron(sp.identity(V), spin(PMINUS)))).tocsr()


def dwf_operator(H, ls, mass, m5, a5_inv=1.0):
    """Normalised Shamir operator D = 1 - kappa_5 (H (x) 1_s + 2/a5 T)."""
    V = H.shape[0] // 12
    k5 = 1.0 / (2 * (4 + a5_inv - m5))
    Dh = sp.kron(sp.identity(ls), H) + 2 * a5_inv * five_d_T(ls, mass, V)
    return (sp.identity(12 * V * ls) - k5 * Dh).tocsr(), Dh, k5


def mobius_kappas(m5, b, c, a5_inv=1.0):
    return 1.0 / (2 * (b * (4 - m5) + a5_inv)), 1.0 / (2 * (c * (4 - m5) - a5_inv))


def mobius_operator(H, ls, mass, m5, b, c, a5_inv=1.0):
    """2 kappa_b x [(b D_W + 1/a5) + (c D_W - 1/a5) T] with D_W = (4 - M5) - H/2,
    D_W acting after the s-hop: = M5 - kappa_b H (b + c T),
    M5 = 1 + (kappa_b/kappa_c) T.  Returns (D, M5)."""
    V = H.shape[0] // 12
    kb, kc = mobius_kappas(m5, b, c, a5_inv)
    T = five_d_T(ls, mass, V)
    I = sp.identity(12 * V * ls)
    M5 = (I + (kb / kc) * T).tocsr()
    Hs = sp.kron(sp.identity(ls), H)
    D = (M5 - kb * Hs @ (b * I + c * T)).tocsr()
    return D, M5


def schur(D, odd_mask):
    """Even-odd Schur complement S = D_oo - D_oe D_ee^-1 D_eo as a function of
    a vector on the odd set (and its adjoint)."""
    o = np.nonzero(odd_mask)[0]
    e = np.nonzero(~odd_mask)[0]
    Doo = D[o][:, o]
    Doe = D[o][:, e]
    Deo = D[e][:, o]
    Dee = D[e][:, e].tocsc()
    lu = spla.splu(Dee)
    luh = spla.splu(Dee.conj().T.tocsc())

    def apply(v):
        return Doo @ v - Doe @ lu.solve(Deo @ v)

    def apply_dag(v):
        return Doo.conj().T @ v - Deo.conj().T @ luh.solve(Doe.conj().T @ v)

    return apply, apply_dag


def dof_mask(site_mask):
    """Expand a per-site boolean mask to the 12 components of each site."""
    return np.repeat(site_mask, 12)
