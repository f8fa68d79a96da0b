"""Subspace expansion of the single-site algorithms (DMRG1S, TDVP1S):

- optimal expansion (algorithms/expansion/optimalexpand.jl:4-69, the
  reference's default): enlarge the bond between two sites by the dominant
  directions of the two-site H_eff residual projected onto the complements of
  both site spaces;
- CH expansion (expansion/chexpand.jl:1-63, Hubig et al., PRB 91, 155115):
  append alpha * L.A.W (the site's H_eff with its far bond left open) to the
  bond and truncate by an SVD.

The H_eff contractions run on the engine; the null spaces and the small SVDs
are FP64 dense linear algebra.
"""
import torch

from .derivatives import ac2_prime, ac_prime_open
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


class CHExpansion:
    """chexpand.jl:5-17: expansion term alpha * L.A.W; ``momentum`` is kept
    (and checked) as in the reference, which does not use it further."""

    def __init__(self, eps=0.01, momentum=0.9):
        if not (0.0 < momentum <= 1.0):
            raise ValueError("momentum should be in (0, 1].")
        self.eps, self.momentum = float(eps), float(momentum)


def _merge_last(P):
    """(n0, n1, n2, n3) column-major -> (n0, n1, n2*n3), index k2 + n2*k3."""
    n0, n1, n2, n3 = P.shape
    return P.permute(3, 2, 1, 0).contiguous().reshape(n3 * n2, n1, n0).permute(2, 1, 0)


def _ch_left(m, site, expan, trunc):
    """left_expansion! for CHExpansion (chexpand.jl:20-29, 42-51): bond
    (site-1, site); A1' = [alpha L.A1.W | A1], A2' = [0; A2], truncated SVD
    of A1', then A1 = u, A2 = s v A2'."""
    mpo, mps, hs = m.mpo, m.mps, m.hstorage
    A1, A2 = mps[site - 1], mps[site]
    Dl, d, Dm = A1.shape
    _, d2, Dr = A2.shape
    if expan.eps == 0.0:
        a, nz = as_matrix(A1, 2), 0
    else:
        P = _merge_last(ac_prime_open(A1, mpo[site - 1], hs[site - 1]))   # (Dl, d, wr Dm)
        nz = P.shape[2]
        a = torch.cat([as_matrix(P, 2) * expan.eps, as_matrix(A1, 2)], dim=1)
    u, s, v = _tsvd_mat(a, trunc)
    k = u.shape[1]
    mps[site - 1] = from_matrix(u, (Dl, d, k))
    sv = v[:, nz:] * s.to(v.dtype).reshape(-1, 1)
    mps[site] = from_matrix(gemm(sv.contiguous(), as_matrix(A2, 1)), (k, d2, Dr))
    m.updateleft(site - 1)


def _ch_right(m, site, expan, trunc):
    """right_expansion! for CHExpansion (chexpand.jl:31-40, 52-61): bond
    (site, site+1); A2' = [alpha W.A2.R (left bond (w, l) fused, w fastest);
    A2], A1' = [0 | A1], truncated SVD of A2', then A2 = v, A1 = A1' u s.
    The open-left contraction is ac_prime_open of the bond-transposed
    tensors: x~[r2,s,l] = A2[l,s,r2], W~[w',s',w,s] = W[w,s',w',s], L~ = R."""
    mpo, mps, hs = m.mpo, m.mps, m.hstorage
    A1, A2 = mps[site], mps[site + 1]
    Dl, d, Dm = A1.shape
    _, d2, Dr = A2.shape
    if expan.eps == 0.0:
        b, nz = as_matrix(A2, 1), 0
    else:
        Pt = ac_prime_open(A2.permute(2, 1, 0), mpo[site + 1].permute(2, 1, 0, 3), hs[site + 2])
        Q = _merge_last(Pt).permute(2, 1, 0)                                # (wl Dm, d2, Dr)
        nz = Q.shape[0]
        b = torch.cat([as_matrix(Q, 1) * expan.eps, as_matrix(A2, 1)], dim=0)
    u, s, v = _tsvd_mat(b, trunc)
    k = v.shape[0]
    mps[site + 1] = from_matrix(v, (k, d2, Dr))
    us = u[nz:, :] * s.to(u.dtype).reshape(1, -1)
    mps[site] = from_matrix(gemm(as_matrix(A1, 2), us.contiguous()), (Dl, d, k))
    m.updateright(site + 1)


def right_expansion(m, site, trunc, prec="f64", expan=None):
    """right_expansion!: expand bond (site, site+1) while sweeping right with
    the optimal (default) or the CH scheme."""
    if isinstance(expan, CHExpansion):
        return _ch_right(m, site, expan, trunc)
    return _optimal_right(m, site, trunc, prec)


def left_expansion(m, site, trunc, prec="f64", expan=None):
    """left_expansion!: expand bond (site-1, site) while sweeping left."""
    if isinstance(expan, CHExpansion):
        return _ch_left(m, site, expan, trunc)
    return _optimal_left(m, site, trunc, prec)


def _optimal_right(m, site, trunc, prec="f64"):
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


def _optimal_left(m, site, trunc, prec="f64"):
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
