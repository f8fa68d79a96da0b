"""Exact Hamiltonians built from Kronecker products of local matrices,
independently of any MPO: the exact-diagonalisation oracle that the reference's
own tests use against DMRG (test/runtests.jl:423-451, :773-778).

Site ordering: site 1 is the slowest index of the C-order state vector (the
same ordering as oracle.dmrg.to_dense).  Spin matrices follow
utilities/spin_siteops.jl:4-11 (sigma^+ = [0 0; 1 0], sigma^z = diag(-1, 1)).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

SP = np.array([[0.0, 0.0], [1.0, 0.0]])
SM = SP.T.copy()
SZ = np.diag([-1.0, 1.0])
SX = SP + SM


def _site(op, i, L, d):
    left = sp.identity(d ** i, format="csr")
    right = sp.identity(d ** (L - i - 1), format="csr")
    return sp.kron(sp.kron(left, sp.csr_matrix(op)), right, format="csr")


def xxz_hamiltonian(L, J=1.0, Jzz=1.0, hz=0.0):
    """models.jl:27-42 (Pauli convention)."""
    H = sp.csr_matrix((2 ** L, 2 ** L))
    for i in range(L):
        if hz:
            H = H + hz * _site(SZ, i, L, 2)
    for i in range(L - 1):
        H = H + 2 * J * (_site(SP, i, L, 2) @ _site(SM, i + 1, L, 2)
                         + _site(SM, i, L, 2) @ _site(SP, i + 1, L, 2))
        H = H + Jzz * (_site(SZ, i, L, 2) @ _site(SZ, i + 1, L, 2))
    return H


def tfim_hamiltonian(L, J=1.0, h=1.0):
    H = sp.csr_matrix((2 ** L, 2 ** L))
    for i in range(L):
        H = H - h * _site(SZ, i, L, 2)
    for i in range(L - 1):
        H = H - J * (_site(SX, i, L, 2) @ _site(SX, i + 1, L, 2))
    return H


def j1j2_hamiltonian(L, J1=1.0, J2=0.0):
    """J1 sum S_i.S_{i+1} + J2 sum S_i.S_{i+2}, S = sigma/2."""
    def ss(i, j):
        return (0.5 * (_site(SP, i, L, 2) @ _site(SM, j, L, 2) + _site(SM, i, L, 2) @ _site(SP, j, L, 2))
                + 0.25 * _site(SZ, i, L, 2) @ _site(SZ, j, L, 2))
    H = sp.csr_matrix((2 ** L, 2 ** L))
    for i in range(L - 1):
        H = H + J1 * ss(i, i + 1)
    for i in range(L - 2):
        H = H + J2 * ss(i, i + 2)
    return H


def hubbard_hamiltonian(L, t=1.0, U=4.0, mu=0.0):
    """Spin-1/2 Fermi-Hubbard chain on 2L Jordan-Wigner modes, mode 2i (up),
    2i+1 (down); occupied = basis index 1 (n = (sigma^z + 1)/2)."""
    M = 2 * L
    par = np.diag([1.0, -1.0])  # (-1)^n of one mode

    def cdag(m):
        op = sp.identity(1, format="csr")
        for k in range(M):
            loc = par if k < m else (SP if k == m else np.eye(2))
            op = sp.kron(op, sp.csr_matrix(loc), format="csr")
        return op

    cd = [cdag(m) for m in range(M)]
    c = [x.T.tocsr() for x in cd]
    n = [cd[m] @ c[m] for m in range(M)]
    H = sp.csr_matrix((2 ** M, 2 ** M))
    for i in range(L - 1):
        for s in range(2):
            a, b = 2 * i + s, 2 * (i + 1) + s
            H = H - t * (cd[a] @ c[b] + cd[b] @ c[a])
    for i in range(L):
        H = H + U * (n[2 * i] @ n[2 * i + 1])
        if mu:
            H = H - mu * (n[2 * i] + n[2 * i + 1])
    return H


def hubbard_sector_energy(L, nup, ndn, t=1.0, U=4.0):
    """Lowest energy of the Hubbard chain with nup up and ndn down electrons:
    the Kronecker Hamiltonian restricted to the basis states of that sector
    (mode 2i counts up, 2i+1 down electrons; occupied = bit 1, mode 0 the
    most significant bit)."""
    H = hubbard_hamiltonian(L, t, U)
    M = 2 * L
    idx = np.arange(2 ** M)
    bits = (idx[:, None] >> (M - 1 - np.arange(M))[None, :]) & 1
    sel = np.nonzero((bits[:, 0::2].sum(1) == nup) & (bits[:, 1::2].sum(1) == ndn))[0]
    Hs = H[sel][:, sel]
    return ground_energy(Hs)[0]


def ground_energy(H, k=1):
    if H.shape[0] <= 512:
        return np.linalg.eigvalsh(H.toarray())[:k]
    vals = spla.eigsh(H, k=k, which="SA", tol=1e-13, maxiter=100000)[0]
    return np.sort(vals)
