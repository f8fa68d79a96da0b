"""TDVP on the device: algorithms/tdvp.jl (TDVP1 :8-69, TDVP2 :71-142,
TDVP1S :145-213, sweep! :219-222).

One ``sweep`` is a left sweep followed by a right sweep and advances the state
by ``stepsize`` (a real stepsize evolves in imaginary time, -1j*dt in real
time).  Every local evolution runs the device Lanczos exponentiator on the
engine's ``ac2_prime`` / ``ac_prime`` / ``c_prime`` operators
(``_exp_evolve``, tdvp.jl:3-6: tolerance exptol per unit time).
"""
import torch

from .derivatives import HeffOperator
from .krylov import exponentiate
from .linalg import as_matrix, cm_view, from_matrix, gemm, two_site
from .truncation import default_truncation, split_rho, stable_tsvd
from . import transfer

TDVP_EXP_TOL = 1.0e-8  # FiniteMPTools.jl:59 (Defaults.tdvp_exp_tol)
DEFAULT_D = 50         # Defaults.D


def _exp(op, dt, x, alg):
    """_exp_evolve (tdvp.jl:3-6): Lanczos when alg.ishermitian, else Arnoldi."""
    y, _ = exponentiate(op, dt, x, tol=alg.exptol / abs(dt), kdim=alg.kdim,
                        ishermitian=alg.ishermitian)
    return y


def leftorth(A):
    """QR of A[(a,s), b]: returns the isometry Q[a, s, k] and R[k, b]."""
    Dl, d, Dr = A.shape
    q, r = torch.linalg.qr(as_matrix(A, 2), mode="reduced")
    k = q.shape[1]
    return from_matrix(q, (Dl, d, k)), from_matrix(r, (k, Dr))


def rightorth(A):
    """LQ of A[a, (s,b)]: returns L[a, k] and the isometry Q[k, s, b]."""
    Dl, d, Dr = A.shape
    q, r = torch.linalg.qr(as_matrix(A, 1).conj().t(), mode="reduced")
    k = q.shape[1]
    return from_matrix(r.conj().t(), (Dl, k)), from_matrix(q.conj().t(), (k, d, Dr))


def _mul_left(v, A):
    """v[k, a] A[a, s, b]"""
    return cm_view(gemm(v, as_matrix(A, 1)), (v.shape[0],) + tuple(A.shape[1:]))


def _mul_right(A, v):
    """A[a, s, b] v[b, k]"""
    return cm_view(gemm(as_matrix(A, 2), v), tuple(A.shape[:2]) + (v.shape[1],))


def _scale_cols(u, s):
    return u * s.to(u.dtype).reshape(1, 1, -1)


def _scale_rows(v, s):
    return v * s.to(v.dtype).reshape(-1, 1, 1)


class TDVP2:
    """tdvp.jl:71-79 (exptol 1e-8, default truncation, ishermitian = false:
    Arnoldi exponentiate unless the caller declares H Hermitian)."""

    def __init__(self, stepsize, exptol=TDVP_EXP_TOL, trunc=None, prec="f64", kdim=12, verbosity=0,
                 split="rho", ishermitian=False):
        self.ishermitian = bool(ishermitian)
        self.stepsize, self.exptol = complex(stepsize), exptol
        self.trunc = trunc or default_truncation()
        self.prec, self.kdim, self.verbosity = prec, kdim, verbosity
        self.split = split  # "rho": reduced-density-matrix split; "svd": stable_tsvd


def _split2(two, alg, absorb):
    """Truncated split of the evolved two-site tensor with the (unnormalised)
    singular values absorbed into A2 ("right") or A1 ("left")."""
    if getattr(alg, "split", "svd") == "rho":
        return split_rho(two, alg.trunc, absorb, normalize=False)
    u, s, v, _ = stable_tsvd(two, alg.trunc)
    if absorb == "right":
        return u, _scale_rows(v, s), s
    return _scale_cols(u, s), v, s


class TDVP1:
    """tdvp.jl:8-17 (bond dimension D reached by zero-padding, increase_bond!;
    ishermitian = false: Arnoldi exponentiate)."""

    def __init__(self, stepsize, D=DEFAULT_D, exptol=TDVP_EXP_TOL, prec="f64", kdim=12, verbosity=0,
                 ishermitian=False):
        self.ishermitian = bool(ishermitian)
        self.stepsize, self.exptol, self.D = complex(stepsize), exptol, D
        self.prec, self.kdim, self.verbosity = prec, kdim, verbosity


class TDVP1S:
    """tdvp.jl:145-155: one-site TDVP with subspace expansion before every
    forward step (``expan``: OptimalExpansion, the default, or CHExpansion)."""

    def __init__(self, stepsize, exptol=TDVP_EXP_TOL, trunc=None, expan=None, prec="f64",
                 kdim=12, verbosity=0, ishermitian=False):
        from .expansion import OptimalExpansion
        self.ishermitian = bool(ishermitian)
        self.stepsize, self.exptol = complex(stepsize), exptol
        self.trunc = trunc or default_truncation()
        self.expan = expan or OptimalExpansion()
        self.prec, self.kdim, self.verbosity = prec, kdim, verbosity


def _dt(alg, x):
    """stepsize as the tensors' scalar type (a real stepsize on real data)."""
    return alg.stepsize.real if not x.is_complex() and alg.stepsize.imag == 0 else alg.stepsize


def _leftsweep2(m, alg):
    """tdvp.jl:83-117"""
    mpo, mps, hs = m.mpo, m.mps, m.hstorage
    L = len(mps)
    for site in range(L - 2):
        two = two_site(mps[site], mps[site + 1])
        dt = _dt(alg, two)
        op2 = HeffOperator(2, hs[site], hs[site + 2], mpo[site], mpo[site + 1], prec=alg.prec)
        two = _exp(op2, dt / 2, two, alg)
        mps[site], sitemps, s = _split2(two, alg, "right")
        hs[site + 1] = transfer.updateleft(hs[site], mps[site], mpo[site], mps[site], prec=alg.prec)
        op1 = HeffOperator(1, hs[site + 1], hs[site + 2], mpo[site + 1], prec=alg.prec)
        mps[site + 1] = _exp(op1, -dt / 2, sitemps, alg)
    site = L - 2
    two = two_site(mps[site], mps[site + 1])
    dt = _dt(alg, two)
    op2 = HeffOperator(2, hs[site], hs[site + 2], mpo[site], mpo[site + 1], prec=alg.prec)
    two = _exp(op2, dt, two, alg)
    mps[site], mps[site + 1], s = _split2(two, alg, "left")
    hs[site + 1] = transfer.updateright(hs[site + 2], mps[site + 1], mpo[site + 1], mps[site + 1], prec=alg.prec)
    mps.s[site + 1] = s


def _rightsweep2(m, alg):
    """tdvp.jl:120-142"""
    mpo, mps, hs = m.mpo, m.mps, m.hstorage
    L = len(mps)
    for site in range(L - 3, -1, -1):
        dt = _dt(alg, mps[site + 1])
        op1 = HeffOperator(1, hs[site + 1], hs[site + 2], mpo[site + 1], prec=alg.prec)
        mps[site + 1] = _exp(op1, -dt / 2, mps[site + 1], alg)
        two = two_site(mps[site], mps[site + 1])
        op2 = HeffOperator(2, hs[site], hs[site + 2], mpo[site], mpo[site + 1], prec=alg.prec)
        two = _exp(op2, dt / 2, two, alg)
        mps[site], mps[site + 1], s = _split2(two, alg, "left")
        hs[site + 1] = transfer.updateright(hs[site + 2], mps[site + 1], mpo[site + 1], mps[site + 1], prec=alg.prec)
        mps.s[site + 1] = s


def _leftsweep1(m, alg):
    """tdvp.jl:19-45"""
    increase_bond(m, alg.D)
    mpo, mps, hs = m.mpo, m.mps, m.hstorage
    L = len(mps)
    for site in range(L - 1):
        dt = _dt(alg, mps[site])
        op1 = HeffOperator(1, hs[site], hs[site + 1], mpo[site], prec=alg.prec)
        tmp = _exp(op1, dt / 2, mps[site], alg)
        mps[site], v = leftorth(tmp)
        hnew = transfer.updateleft(hs[site], mps[site], mpo[site], mps[site], prec=alg.prec)
        op0 = HeffOperator(0, hnew, hs[site + 1], prec=alg.prec)
        v = _exp(op0, -dt / 2, v, alg)
        mps[site + 1] = _mul_left(v, mps[site + 1])
        hs[site + 1] = hnew
    site = L - 1
    dt = _dt(alg, mps[site])
    op1 = HeffOperator(1, hs[site], hs[site + 1], mpo[site], prec=alg.prec)
    mps[site] = _exp(op1, dt, mps[site], alg)


def _rightsweep1(m, alg):
    """tdvp.jl:47-69"""
    mpo, mps, hs = m.mpo, m.mps, m.hstorage
    L = len(mps)
    for site in range(L - 2, -1, -1):
        dt = _dt(alg, mps[site])
        v, Q = rightorth(mps[site + 1])
        mps[site + 1] = Q
        hnew = transfer.updateright(hs[site + 2], Q, mpo[site + 1], Q, prec=alg.prec)
        op0 = HeffOperator(0, hs[site + 1], hnew, prec=alg.prec)
        v = _exp(op0, -dt / 2, v, alg)
        hs[site + 1] = hnew
        mps[site] = _mul_right(mps[site], v)
        op1 = HeffOperator(1, hs[site], hs[site + 1], mpo[site], prec=alg.prec)
        mps[site] = _exp(op1, dt / 2, mps[site], alg)


def _leftsweep1s(m, alg):
    """tdvp.jl:158-180: expand bond (site, site+1), forward one-site step,
    QR, backward bond step; the last site takes the full step."""
    from .expansion import right_expansion
    mpo, mps, hs = m.mpo, m.mps, m.hstorage
    L = len(mps)
    for site in range(L - 1):
        right_expansion(m, site, alg.trunc, prec=alg.prec, expan=alg.expan)
        dt = _dt(alg, mps[site])
        op1 = HeffOperator(1, hs[site], hs[site + 1], mpo[site], prec=alg.prec)
        tmp = _exp(op1, dt / 2, mps[site], alg)
        mps[site], v = leftorth(tmp)
        hnew = transfer.updateleft(hs[site], mps[site], mpo[site], mps[site], prec=alg.prec)
        op0 = HeffOperator(0, hnew, hs[site + 1], prec=alg.prec)
        v = _exp(op0, -dt / 2, v, alg)
        mps[site + 1] = _mul_left(v, mps[site + 1])
        hs[site + 1] = hnew
    site = L - 1
    dt = _dt(alg, mps[site])
    op1 = HeffOperator(1, hs[site], hs[site + 1], mpo[site], prec=alg.prec)
    mps[site] = _exp(op1, dt, mps[site], alg)


def _rightsweep1s(m, alg):
    """tdvp.jl:182-213: LQ, backward bond step, expand bond (site-1, site),
    forward one-site step."""
    from .expansion import left_expansion
    mpo, mps, hs = m.mpo, m.mps, m.hstorage
    L = len(mps)
    for site in range(L - 2, -1, -1):
        dt = _dt(alg, mps[site])
        v, Q = rightorth(mps[site + 1])
        mps[site + 1] = Q
        hnew = transfer.updateright(hs[site + 2], Q, mpo[site + 1], Q, prec=alg.prec)
        op0 = HeffOperator(0, hs[site + 1], hnew, prec=alg.prec)
        v = _exp(op0, -dt / 2, v, alg)
        hs[site + 1] = hnew
        mps[site] = _mul_right(mps[site], v)
        if site > 0:
            left_expansion(m, site, alg.trunc, prec=alg.prec, expan=alg.expan)
        op1 = HeffOperator(1, hs[site], hs[site + 1], mpo[site], prec=alg.prec)
        mps[site] = _exp(op1, dt / 2, mps[site], alg)


def sweep(m, alg):
    """sweep! (tdvp.jl:219-222): left sweep, then right sweep."""
    if isinstance(alg, TDVP2):
        _leftsweep2(m, alg)
        _rightsweep2(m, alg)
    elif isinstance(alg, TDVP1S):
        _leftsweep1s(m, alg)
        _rightsweep1s(m, alg)
    elif isinstance(alg, TDVP1):
        _leftsweep1(m, alg)
        _rightsweep1(m, alg)
    else:
        raise TypeError("unknown TDVP algorithm")


def increase_bond(m, D):
    """increase_bond! (envs/finiteenv.jl:87-93, mps/finitemps.jl:302-317): zero-pad
    the site tensors to bond dimension min(D, d^i, d^(L-i)), re-canonicalise
    (without normalising) and rebuild the environments."""
    from .mps import virtual_dims
    mps = m.mps
    if mps.bond_dimension() >= D:
        return
    L = len(mps)
    d = mps[0].shape[1]
    dims = virtual_dims(L, d, D)
    for i in range(L):
        A = mps[i]
        sl = max(min(dims[i], D), A.shape[0])
        sr = max(min(dims[i + 1], D), A.shape[2])
        B = torch.zeros((sr, d, sl), dtype=A.dtype, device=A.device).permute(2, 1, 0)
        B[: A.shape[0], :, : A.shape[2]] = A
        mps[i] = B
    mps.right_canonicalize(normalize=False)
    m.init_hstorage(center=0)
