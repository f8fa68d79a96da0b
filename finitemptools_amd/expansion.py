"""Optimal subspace expansion (algorithms/expansion/optimalexpand.jl:4-69):
enlarge the bond between two sites by the dominant directions of the
two-site H_eff residual projected onto the complements of both site spaces.
The two-site H_eff application runs on the engine; the null spaces and the
small SVDs are FP64 dense linear algebra.
"""
import torch

from .derivatives import ac2_prime
from .linalg import as_matrix, from_matrix, gemm, two_site
from .truncation import MPSTruncation


def leftnull(A):
    """Orthonormal complement of the column space of A[(a,s), b]:
    N[(a,s), k] with N^dag A = 0 (full QR, columns past the rank)."""
    M = as_matrix(A, 2)
    m, n = M.shape
    if m <= n:
        return M.new_zeros((m, 0))
    q, _ = torch.linalg.qr(M, mode="complete")
    return q[:, n:]


def rightnull(M):
    """Orthonormal complement of the row space of a matrix M (r x c):
    N (c - r x c) with M N^dag = 0."""
    r, c = M.shape
    if c <= r:
        return M.new_zeros((0, c))
    q, _ = torch.linalg.qr(M.conj().t(), mode="complete")
    return q[:, r:].conj().t()


def _tsvd_mat(M, trunc, kmax=None):
    U, S, Vh = torch.linalg.svd(M, full_matrices=False)
    k, _ = trunc.truncate(S.detach().cpu().numpy())
    if kmax is not None:
        k = min(k, kmax)
    k = max(k, 1)
    return U[:, :k], S[:k], Vh[:k]


def _room(trunc, Dm):
    """Expansion vectors allowed on a bond of dimension Dm.  The second SVD of
    the expansion acts on a matrix with orthonormal rows (columns), whose
    singular values are all 1; cutting such a spectrum to D would keep an
    arbitrary subspace (the order of degenerate singular vectors depends on
    the LAPACK implementation), so at most D - Dm vectors are added and that
    SVD never truncates.  When D does not bind this is the reference's rule."""
    return trunc.D - Dm


def _project(NL, AC2, NR):
    """NL^dag . AC2[(a,s1),(s2,b)] . NR^dag on the engine's GEMM."""
    t = gemm(NL.t(), as_matrix(AC2, 2), conjA=True)
    return gemm(t, NR.conj().t().contiguous())


class OptimalExpansion:
    """optimalexpand.jl:4-6 (the reference's default expansion)."""


def right_expansion(m, site, trunc, prec="f64"):
    """optimalexpand.jl:42-69: expand bond (site, site+1) while sweeping right."""
    mpo, mps, hs = m.mpo, m.mps, m.hstorage
    A, ACi = mps[site], mps[site + 1]
    Dl, d, Dm = A.shape
    _, d2, Dr = ACi.shape
    AC2 = ac2_prime(two_site(A, ACi), mpo[site], mpo[site + 1], hs[site], hs[site + 2], prec=prec)
    NL = leftnull(A)                        # (Dl d) x k_l
    NR = rightnull(as_matrix(ACi, 1))       # k_r x (d2 Dr)
    if NL.shape[1] == 0 or NR.shape[0] == 0 or _room(trunc, Dm) <= 0:
        m.updateright(site + 1)
        return
    inter = _project(NL, AC2, NR)
    U1, S1, V1 = _tsvd_mat(inter, trunc, kmax=_room(trunc, Dm))
    k1 = V1.shape[0]
    a = torch.cat([as_matrix(A, 2), A.new_zeros((Dl * d, k1))], dim=1)       # (Dl d) x (Dm + k1)
    b = torch.cat([as_matrix(ACi, 1), V1 @ NR], dim=0)                          # (Dm + k1) x (d2 Dr)
    u, s, v = _tsvd_mat(b, trunc)
    mps[site + 1] = from_matrix(v, (v.shape[0], d2, Dr))
    mps[site] = from_matrix(a @ (u * s.to(u.dtype).reshape(1, -1)), (Dl, d, u.shape[1]))
    m.updateright(site + 1)


def left_expansion(m, site, trunc, prec="f64"):
    """optimalexpand.jl:12-40: expand bond (site-1, site) while sweeping left."""
    mpo, mps, hs = m.mpo, m.mps, m.hstorage
    ACi, B = mps[site - 1], mps[site]
    Dl, d, Dm = ACi.shape
    _, d2, Dr = B.shape
    AC2 = ac2_prime(two_site(ACi, B), mpo[site - 1], mpo[site], hs[site - 1], hs[site + 1], prec=prec)
    NL = leftnull(ACi)
    NR = rightnull(as_matrix(B, 1))
    if NL.shape[1] == 0 or NR.shape[0] == 0 or _room(trunc, Dm) <= 0:
        m.updateleft(site - 1)
        return
    inter = _project(NL, AC2, NR)
    U1, S1, V1 = _tsvd_mat(inter, trunc, kmax=_room(trunc, Dm))
    k1 = U1.shape[1]
    a = torch.cat([as_matrix(ACi, 2), NL @ U1], dim=1)                          # (Dl d) x (Dm + k1)
    b = torch.cat([as_matrix(B, 1), B.new_zeros((k1, d2 * Dr))], dim=0)        # (Dm + k1) x (d2 Dr)
    u, s, v = _tsvd_mat(a, trunc)
    mps[site - 1] = from_matrix(u, (Dl, d, u.shape[1]))
    mps[site] = from_matrix((v * s.to(v.dtype).reshape(-1, 1)) @ b, (u.shape[1], d2, Dr))
    m.updateleft(site - 1)
