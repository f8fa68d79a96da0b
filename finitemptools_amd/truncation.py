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
    for u and s^2 and forms s v = u^dag M exactly; the right sweep
    diagonalises M^dag M for v and forms u s = M v^dag.  The kept values and
    the truncated tensor u s v are those of the SVD (truncation.jl:27-47 on
    s = sqrt(lambda)), with s normalised as dmrg.jl:105 does (normalize=False
    keeps the raw values, as the TDVP sweeps need).  Returns (A1, A2, s)
    with s absorbed into A2 (absorb="right") or A1 ("left")."""
    Dl, d1, d2, Dr = x.shape
    M = as_matrix(x, 2)
    left = absorb == "right"
    r = min(M.shape)
    n = M.shape[0] if left else M.shape[1]
    if (SMALL_EIGH and x.is_cuda and n <= SMALL_EIGH_ORDER
            and n <= eigen.small_eigh_max(M.dtype, x.device)):
        # one-launch solver; it takes G column-major: G^T = conj(M) M^T (left)
        # or M^T conj(M) (right); only the top r columns are finished
        Gt = M.conj() @ M.t() if left else M.t() @ M.conj()
        lam, V, _ = eigen.small_eigh(Gt, col0=n - r)
    else:
        G = M @ M.conj().t() if left else M.conj().t() @ M
        lam, V = torch.linalg.eigh(G)
    lam_h = lam.detach().cpu().numpy()[::-1][:r]
    if not np.isfinite(lam_h).all():
        raise MHError(6, "split_rho: the eigensolver did not converge")
    lam_h = np.maximum(lam_h, 0.0)
    s_h = np.sqrt(lam_h)
    k, _ = trunc.truncate(s_h)
    k = max(1, min(k, int(np.count_nonzero(s_h > 0)) or 1))
    norm = float(np.linalg.norm(s_h[:k])) if normalize else 1.0
    Q = V[:, n - k:].flip(-1)
    s = torch.as_tensor(s_h[:k] / norm, device=x.device)  # normalize=False: the raw values
    if left:
        A1 = from_matrix(Q, (Dl, d1, k))
        A2 = from_matrix((Q.conj().t() @ M) / norm, (k, d2, Dr))
    else:
        A2 = from_matrix(Q.conj().t(), (k, d2, Dr))
        A1 = from_matrix((M @ Q) / norm, (Dl, d1, k))
    return A1, A2, s


def stable_tsvd(x, trunc):
    """x[a, s1, s2, b] -> u[a, s1, k], s[k] (host + device), v[k, s2, b], err."""
    Dl, d1, d2, Dr = x.shape
    M = as_matrix(x, 2)
    try:
        U, S, Vh = torch.linalg.svd(M, full_matrices=False)
    except RuntimeError:
        U, S, Vh = torch.linalg.svd(M, full_matrices=False, driver="gesvd")
    s_host = S.detach().cpu().numpy()
    k, err = trunc.truncate(s_host)
    k = max(k, 1)
    u = from_matrix(U[:, :k], (Dl, d1, k))
    v = from_matrix(Vh[:k], (k, d2, Dr))
    return u, S[:k], v, err
