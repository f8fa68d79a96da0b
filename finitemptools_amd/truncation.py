"""Truncation scheme and truncated SVD of the sweep step.

``MPSTruncation`` restates auxiliary/truncation.jl:8-47: drop singular values
at or below eps * ||s||_2, then keep at most D; the error is the 2-norm of the
dropped values.  ``stable_tsvd`` is auxiliary/stable_svd.jl:12-18 on a
two-site tensor split as (a, s1) | (s2, b).
"""
import numpy as np
import torch

from . import eigen
from ._lib import MHError
from .linalg import as_matrix, from_matrix

DEFAULT_SVD_CUTOFF = 1.0e-8   # truncation.jl:49
DEFAULT_NONSYMMETRIC_D = 100  # truncation.jl:51
# Dense splits whose density matrix has order <= SMALL_EIGH_ORDER use the
# one-launch device solver (mh_heev_small); larger ones the library eigh.
# Measured on MI355X (eigenvalues back on the host included): n = 16 complex
# 0.29 ms against 0.51 ms, even at n = 32, and slower above (n = 64: 3.4
# against 1.8 ms).
SMALL_EIGH = True
SMALL_EIGH_ORDER = 24
# Real density matrices above that order go to the batched device
# eigensolver (eigen.SectorEigh: LDS tridiagonalisation, divide and
# conquer, back-transformation of the kept columns only) instead of the
# library eigh, a chain of launches at these orders.
SECTOR_EIGH = True
# SVDs of two-site matrices too large for the one-workgroup device SVD but up
# to this size (larger extent) run in host LAPACK, which the reference calls
# too: at these orders the device library's SVD is launch-latency bound.
HOST_SVD_MAX = 256
# SVDs whose matrices fit one workgroup's LDS run on the device first
# (mh_svd_small: one launch, one-sided Jacobi); False sends them to the
# branches below.
DEVICE_SVD = True
# Complex density-matrix splits whose matrix fits mh_svd_small take the
# SVD split directly (at the small bonds where these run, a relative cutoff
# of 1e-8 falls inside rho's noise on nearly every bond, so the rho split
# mostly ended in the SVD split anyway: 96% of the bonds of the L = 100
# TFIM quench).
COMPLEX_SVD_SPLIT = True

# A density-matrix eigenvalue of an order-n matrix carries an absolute error
# of about n * eps64 * lambda_max (eps64 = 2^-52); singular values taken as
# sqrt(lambda) are therefore only known to sqrt of that.  RHO_NOISE is the
# safety factor on the bound.
EPS64 = float(np.finfo(np.float64).eps)
RHO_NOISE = 10.0


def rho_cut_resolved(s2, noise, trunc):
    """Whether MPSTruncation's decision on s = sqrt(s2) holds for every
    spectrum within ``noise`` of s2.

    s2: squared singular values from density-matrix eigenvalues, descending;
    noise: absolute error bound of each (squared units).  The relative cutoff
    compares every value with t = eps * ||s|| (truncation.jl:32): the cut is
    resolved when no compared value lies within twice its noise of t^2.
    When D binds instead (more than D values above t), only the D-th value
    meets the cutoff.  Ties between values around a D cut are not tested: an
    SVD breaks those arbitrarily too."""
    s2 = np.asarray(s2, dtype=np.float64)
    noise = np.broadcast_to(np.asarray(noise, dtype=np.float64), s2.shape)
    if s2.size == 0:
        return True
    t2 = trunc.eps ** 2 * float(s2.sum())
    k_eps = int(np.count_nonzero(s2 > t2))
    if k_eps <= trunc.D:
        return not bool(np.any(np.abs(s2 - t2) <= 2.0 * noise))
    i = trunc.D - 1
    return not (abs(float(s2[i]) - t2) <= 2.0 * float(noise[i]))


def split_svd(x, trunc, absorb, normalize=True):
    """The reference's split: stable_tsvd, s normalised (dmrg.jl:105) unless
    normalize=False, absorbed into A2 ("right": u | s v) or A1 ("left")."""
    u, s, v, _ = stable_tsvd(x, trunc)
    if normalize:
        s = s / torch.linalg.vector_norm(s)
    if absorb == "right":
        return u, v * s.to(v.dtype).reshape(-1, 1, 1), s
    return u * s.to(u.dtype).reshape(1, 1, -1), v, s


class MPSTruncation:
    def __init__(self, D=DEFAULT_NONSYMMETRIC_D, eps=DEFAULT_SVD_CUTOFF, verbosity=0):
        self.D, self.eps, self.verbosity = int(D), float(eps), verbosity

    def truncate(self, s):
        """s: descending host array -> (kept count, 2-norm error)."""
        n = float(np.linalg.norm(s))
        k1 = int(np.count_nonzero(s > self.eps * n))
        err1 = float(np.linalg.norm(s[k1:]))
        if k1 <= self.D:
            return k1, err1
        err2 = float(np.linalg.norm(s[self.D:k1]))
        if self.verbosity > 0:
            print(f"sum: {len(s)} -> {k1} -> {self.D}, maximum {self.D}, truncation error: "
                  f"abs={np.hypot(err1, err2)}, rel={np.hypot(err1, err2) / n}.")
        return self.D, float(np.hypot(err1, err2))


def default_truncation():
    return MPSTruncation(DEFAULT_NONSYMMETRIC_D, DEFAULT_SVD_CUTOFF)


def split_rho(x, trunc, absorb, normalize=True):
    """Truncated split of a two-site tensor x[a,s1,s2,b] through the reduced
    density matrix (White's formulation of the DMRG step), FP64 on the
    device: the left sweep diagonalises rho = M M^dag (M = x as (a,s1)|(s2,b))
    for u and forms s v = u^dag M exactly; the right sweep diagonalises
    M^dag M for v and forms u s = M v^dag.  The cut of truncation.jl:27-47 is
    taken on s = sqrt(lambda) when ``rho_cut_resolved`` shows it cannot move
    within the eigenvalues' noise; otherwise (a relative cutoff that falls
    below about sqrt(n eps64) s_max, where sqrt(lambda) is noise) the split
    is the SVD split (``split_svd``).  The returned singular values are the
    norms of the rows of s v (columns of u s), which are accurate to eps64
    s_max where sqrt(lambda) is not; s is normalised as dmrg.jl:105 does
    (normalize=False keeps the raw values, as the TDVP sweeps need).
    Returns (A1, A2, s) with s absorbed into A2 (absorb="right") or A1
    ("left")."""
    Dl, d1, d2, Dr = x.shape
    M = as_matrix(x, 2)
    left = absorb == "right"
    r = min(M.shape)
    n = M.shape[0] if left else M.shape[1]
    se = None
    if (COMPLEX_SVD_SPLIT and DEVICE_SVD and x.is_cuda and M.dtype == torch.complex128
            and eigen.svd_small_fits(M.shape, M.dtype, x.device)):
        # complex matrices that fit the one-launch SVD take the SVD split
        # (the reference's own stable_tsvd) at once: it costs about what the
        # one-launch eigensolver does, and a library eigh more
        return split_svd(x, trunc, absorb, normalize)
    if (SMALL_EIGH and x.is_cuda and n <= SMALL_EIGH_ORDER
            and n <= eigen.small_eigh_max(M.dtype, x.device)):
        # one-launch solver; it takes G column-major: G^T = conj(M) M^T (left)
        # or M^T conj(M) (right); only the top r columns are finished
        Gt = M.conj() @ M.t() if left else M.t() @ M.conj()
        lam, V, _ = eigen.small_eigh(Gt, col0=n - r)
    elif SECTOR_EIGH and x.is_cuda and M.dtype == torch.float64 and n <= eigen.MAX_ORDER:
        # the batched solver on one matrix; vectors only for the kept columns
        se = eigen.SectorEigh([n], x.device)
        if left:
            torch.matmul(M, M.t(), out=se.matrix(0))
        else:
            torch.matmul(M.t(), M, out=se.matrix(0))
        lam = se.eigenvalues()[0]
    else:
        G = M @ M.conj().t() if left else M.conj().t() @ M
        lam, V = torch.linalg.eigh(G)
    lam_h = lam.detach().cpu().numpy()[::-1][:r]
    if se is not None:
        se.check()
    if not np.isfinite(lam_h).all():
        raise MHError(6, "split_rho: the eigensolver did not converge")
    lam_max = max(float(lam_h[0]), 0.0)
    lam_h = np.maximum(lam_h, 0.0)
    if not rho_cut_resolved(lam_h, RHO_NOISE * n * EPS64 * lam_max, trunc):
        return split_svd(x, trunc, absorb, normalize)
    k, _ = trunc.truncate(np.sqrt(lam_h))
    k = max(1, min(k, int(np.count_nonzero(lam_h > 0)) or 1))
    if se is not None:
        V = se.vectors([k])[0]            # n x k, ascending
        Q = V.flip(-1)
    else:
        Q = V[:, n - k:].flip(-1)
    if left:
        B = Q.conj().t() @ M                      # s v, rows of norm s
        s = torch.linalg.vector_norm(B, dim=1)
    else:
        B = M @ Q                                 # u s, columns of norm s
        s = torch.linalg.vector_norm(B, dim=0)
    if normalize:
        norm = torch.linalg.vector_norm(s)
        B = B / norm
        s = s / norm
    if left:
        return from_matrix(Q, (Dl, d1, k)), from_matrix(B, (k, d2, Dr)), s
    return from_matrix(B, (Dl, d1, k)), from_matrix(Q.conj().t(), (k, d2, Dr)), s


def stable_tsvd(x, trunc):
    """x[a, s1, s2, b] -> u[a, s1, k], s[k], v[k, s2, b], err: the device
    one-workgroup SVD where the matrix fits it, else LAPACK gesdd (gesvd if
    it fails) up to HOST_SVD_MAX, else the device library."""
    Dl, d1, d2, Dr = x.shape
    M = as_matrix(x, 2)
    if x.is_cuda and DEVICE_SVD and eigen.svd_small_fits(M.shape, M.dtype, x.device):
        # one launch on the device (one-sided Jacobi, mh_svd_small)
        U, S, Vh = eigen.svd_small(M)
    elif max(M.shape) <= HOST_SVD_MAX:
        # LAPACK gesdd, gesvd if it fails (tsvd(alg=SDD()) and its fallback)
        import scipy.linalg as sla
        Mh = M.detach().cpu().numpy()
        try:
            Uh, Sh, Vhh = sla.svd(Mh, full_matrices=False, lapack_driver="gesdd", check_finite=False)
        except (np.linalg.LinAlgError, ValueError):
            Uh, Sh, Vhh = sla.svd(Mh, full_matrices=False, lapack_driver="gesvd", check_finite=False)
        U, S, Vh = (torch.as_tensor(a, device=x.device) for a in (Uh, Sh, Vhh))
    else:
        try:
            U, S, Vh = torch.linalg.svd(M, full_matrices=False)
        except RuntimeError:
            U, S, Vh = torch.linalg.svd(M, full_matrices=False, driver="gesvd")
    s_host = S.detach().cpu().numpy()
    if not np.isfinite(s_host).all():
        raise MHError(6, "stable_tsvd: the SVD did not converge")
    k, err = trunc.truncate(s_host)
    k = max(k, 1)
    u = from_matrix(U[:, :k], (Dl, d1, k))
    v = from_matrix(Vh[:k], (k, d2, Dr))
    return u, S[:k], v, err
