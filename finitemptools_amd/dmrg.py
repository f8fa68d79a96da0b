"""Two-site DMRG on the device: algorithms/dmrg.jl:79-144, the sweep
dispatcher of dmrgexcited.jl:219-223, compute! (dmrg.jl:221-234) and
ground_state! (:240).

Per bond: two-site tensor (MFMA GEMM), lowest eigenpair of ac2_prime by the
device-resident Lanczos (kmax = eig_maxiter, btol = eig_tol), FP64 truncated
split (reduced density matrix by default, or the SVD), normalised singular
values, overlap error, environment update.
``prec`` selects the arithmetic of the H_eff and environment contractions;
the mixed forms add FP64 re-orthogonalisation to the Lanczos recurrence.
"""
import time

import numpy as np
import torch

from .derivatives import HeffOperator
from .envs import environments
from .krylov import simple_lanczos_solver
from .linalg import as_matrix, cm_view, gemm, two_site
from .truncation import MPSTruncation, split_rho, stable_tsvd

MAXITER = 100   # FiniteMPTools.jl:54 (Defaults.maxiter)
TOL = 1e-12     # Defaults.tol
EIG_TOL = 1e-12  # Defaults.dmrg_eig_tol


def _default_reorth(prec):
    """Full re-orthogonalisation of the Lanczos basis by default only for
    split-BF16 matvecs (relative error ~1e-5): with FP32 matvecs (~1e-7) the
    <= eig_maxiter-step recurrence gives the same energies without it
    (D = 2048 Hubbard sweep: dE/L 9.9e-7 against 9.5e-7 from FP64, 1.2e-14
    after one FP64 polish sweep) at two thirds of the Lanczos time, while
    split-BF16 without it doubles its energy error."""
    return prec == "bf16x3"


class DMRG2:
    def __init__(self, maxiter=MAXITER, tol=TOL, eig_maxiter=10, eig_tol=EIG_TOL,
                 verbosity=0, trunc=None, prec="f64", reorth=None, split="rho"):
        self.maxiter, self.tol = maxiter, tol
        self.eig_maxiter, self.eig_tol = eig_maxiter, eig_tol
        self.verbosity = verbosity
        self.trunc = trunc or MPSTruncation()
        self.prec = prec
        self.reorth = _default_reorth(prec) if reorth is None else reorth
        # "rho": reduced-density-matrix split (batched FP64 eigh on the
        # device); "svd": stable_tsvd (the reference's tsvd)
        self.split = split
        self.stats = {"matvec_flops": 0.0, "lanczos_s": 0.0, "svd_s": 0.0, "env_s": 0.0, "bonds": 0}


def _scale_rows(v, s):
    """s[k] * v[k, s2, b]"""
    return v * s.to(v.dtype).reshape(-1, 1, 1)


def _scale_cols(u, s):
    """u[a, s1, k] * s[k]"""
    return u * s.to(u.dtype).reshape(1, 1, -1)


def _bond_step(m, alg, site, absorb):
    """Two-site tensor, lowest eigenpair, truncated split with the normalised
    singular values absorbed into A2 (absorb="right": u | s v) or A1
    ("left": u s | v).  Returns (two, ev, A1, A2, s)."""
    mpo, mps, hs = m.mpo, m.mps, m.hstorage
    two = two_site(mps[site], mps[site + 1])
    op = HeffOperator(2, hs[site], hs[site + 2], mpo[site], mpo[site + 1], prec=alg.prec)
    t0 = time.perf_counter()
    # excited states: the lower states through the overlap environments
    # (dmrgexcited.jl:85-90); the solver projects the start vector once
    proj = m.bond_projectors(site) if hasattr(m, "bond_projectors") else None
    ev, vecs, info = simple_lanczos_solver(op, two, "SR", alg.eig_maxiter, alg.eig_tol,
                                           reorth=alg.reorth, projectors=proj)
    t1 = time.perf_counter()
    alg.stats["lanczos_s"] += t1 - t0
    alg.stats["matvec_flops"] += op.flops() * alg.eig_maxiter
    if getattr(alg, "split", "svd") == "rho":
        A1, A2, s = split_rho(vecs, alg.trunc, absorb)
    else:
        u, s, v, err = stable_tsvd(vecs, alg.trunc)
        s = s / torch.linalg.vector_norm(s)
        A1, A2 = (u, _scale_rows(v, s)) if absorb == "right" else (_scale_cols(u, s), v)
    alg.stats["svd_s"] += time.perf_counter() - t1
    alg.stats["bonds"] += 1
    return two, ev, A1, A2, s


def _overlap_error(two, A1, A2):
    """err_1 = <u s v, twositemps> (dmrg.jl:109-110); returns |1 - |err_1||."""
    usv = gemm(as_matrix(A1, 2), as_matrix(A2, 1))
    e = torch.vdot(usv.t().reshape(-1), as_matrix(two, 2).t().reshape(-1)).item()
    return abs(1.0 - abs(e))


def leftsweep(m, alg):
    """dmrg.jl:89-115"""
    mps = m.mps
    energies, delta = [], 0.0
    for site in range(len(mps) - 2):
        two, ev, A1, A2, s = _bond_step(m, alg, site, "right")
        energies.append(ev)
        mps[site] = A1
        mps[site + 1] = A2
        delta = max(delta, _overlap_error(two, A1, A2))
        t0 = time.perf_counter()
        m.updateleft(site)
        alg.stats["env_s"] += time.perf_counter() - t0
        if alg.verbosity > 2:
            print(f"bond {site}: E = {ev}")
    return energies, delta


def rightsweep(m, alg):
    """dmrg.jl:117-144"""
    mps = m.mps
    energies, delta = [], 0.0
    for site in range(len(mps) - 2, -1, -1):
        two, ev, A1, A2, s = _bond_step(m, alg, site, "left")
        energies.append(ev)
        mps[site] = A1
        mps[site + 1] = A2
        mps.s[site + 1] = s
        delta = max(delta, _overlap_error(two, A1, A2))
        t0 = time.perf_counter()
        m.updateright(site + 1)
        alg.stats["env_s"] += time.perf_counter() - t0
        if alg.verbosity > 2:
            print(f"bond {site}: E = {ev}")
    return energies, delta


def sweep(m, alg):
    """leftsweep! then rightsweep! (dmrgexcited.jl:219-223)"""
    e1, d1 = leftsweep(m, alg)
    e2, d2 = rightsweep(m, alg)
    return e1 + e2, max(d1, d2)


def compute(m, alg, nsweeps=None):
    """compute! (dmrg.jl:221-234); ``nsweeps`` caps the number of sweeps."""
    all_e, it, delta = [], 0, 2 * alg.tol
    maxit = alg.maxiter if nsweeps is None else nsweeps
    while it < maxit and delta > alg.tol:
        e, delta = sweep(m, alg)
        all_e += e
        it += 1
        if alg.verbosity > 2:
            print(f"finish the {it}-th sweep with error {delta}")
    return all_e, delta


def ground_state(state, h, alg=None, nsweeps=None):
    """ground_state!(state, h, alg) (dmrg.jl:240); returns (energies, delta)."""
    alg = alg or DMRG2()
    return compute(environments(h, state, prec=alg.prec), alg, nsweeps)


def excited_state(state, h, lower_states, alg=None, nsweeps=None):
    """The lowest state orthogonal (through the projected solve) to
    ``lower_states``: compute! on environments(h, state, projectors)
    (dmrgexcited.jl, runtests.jl:494-535); returns (energies, delta)."""
    alg = alg or DMRG2()
    return compute(environments(h, state, projectors=list(lower_states), prec=alg.prec), alg,
                   nsweeps)


# ------------------------------------------------------------ one-site DMRG

class DMRG1:
    """dmrg.jl:8-15: one-site DMRG; the bond dimension is raised to D by
    zero-padding (increase_bond!) at the start of every left sweep."""

    def __init__(self, D=50, maxiter=MAXITER, tol=TOL, eig_maxiter=10, eig_tol=EIG_TOL,
                 verbosity=0, prec="f64", reorth=None):
        self.D, self.maxiter, self.tol = D, maxiter, tol
        self.eig_maxiter, self.eig_tol, self.verbosity = eig_maxiter, eig_tol, verbosity
        self.prec = prec
        self.reorth = _default_reorth(prec) if reorth is None else reorth


class DMRG1S:
    """dmrg.jl:148-156: one-site DMRG with subspace expansion (``expan``:
    expansion.OptimalExpansion, the default, or expansion.CHExpansion)."""

    def __init__(self, maxiter=MAXITER, tol=TOL, eig_maxiter=10, eig_tol=EIG_TOL,
                 verbosity=0, trunc=None, prec="f64", reorth=None, expan=None):
        from .expansion import OptimalExpansion
        self.maxiter, self.tol = maxiter, tol
        self.eig_maxiter, self.eig_tol, self.verbosity = eig_maxiter, eig_tol, verbosity
        self.expan = expan or OptimalExpansion()
        self.trunc = trunc or MPSTruncation()
        self.prec = prec
        self.reorth = _default_reorth(prec) if reorth is None else reorth


def calc_galerkin(m, site, prec="f64"):
    """dmrg.jl:17-24: || N_L^dag H_eff A || with N_L the left null space of
    the current site tensor A (right null space when the left one is empty)."""
    from .derivatives import ac_prime
    from .expansion import leftnull, rightnull
    A = m.mps[site]
    y = ac_prime(A, m.mpo[site], m.hstorage[site], m.hstorage[site + 1], prec=prec)
    NL = leftnull(A)
    if NL.shape[1] > 0:
        return torch.linalg.norm(gemm(NL.t(), as_matrix(y, 2), conjA=True)).item()
    NR = rightnull(as_matrix(A, 1))
    return torch.linalg.norm(gemm(as_matrix(y, 1), NR.conj().t().contiguous())).item()


def _site_eig(m, alg, site):
    op = HeffOperator(1, m.hstorage[site], m.hstorage[site + 1], m.mpo[site], prec=alg.prec)
    proj = m.site_projectors(site) if hasattr(m, "site_projectors") else None  # dmrgexcited.jl:22-25
    ev, vecs, _ = simple_lanczos_solver(op, m.mps[site], "SR", alg.eig_maxiter, alg.eig_tol,
                                        reorth=alg.reorth, projectors=proj)
    return ev, vecs


def _leftsweep1(m, alg, expand):
    """dmrg.jl:27-49 (DMRG1) and :158-185 (DMRG1S)"""
    from .expansion import right_expansion
    from .tdvp import increase_bond, leftorth, _mul_left
    if not expand:
        increase_bond(m, alg.D)
    mps = m.mps
    energies, delta = [], 0.0
    for site in range(len(mps) - 1):
        if expand:
            right_expansion(m, site, alg.trunc, prec=alg.prec, expan=alg.expan)
        ev, vecs = _site_eig(m, alg, site)
        energies.append(ev)
        delta = max(delta, calc_galerkin(m, site, alg.prec))
        Q, R = leftorth(vecs)
        mps[site] = Q
        mps[site + 1] = _mul_left(R, mps[site + 1])
        m.updateleft(site)
    return energies, delta


def _rightsweep1(m, alg, expand):
    """dmrg.jl:52-76 (DMRG1) and :187-215 (DMRG1S)"""
    from .expansion import left_expansion
    from .tdvp import rightorth, _mul_right
    mps = m.mps
    energies, delta = [], 0.0
    for site in range(len(mps) - 1, 0, -1):
        if expand:
            left_expansion(m, site, alg.trunc, prec=alg.prec, expan=alg.expan)
        ev, vecs = _site_eig(m, alg, site)
        energies.append(ev)
        delta = max(delta, calc_galerkin(m, site, alg.prec))
        Lf, Q = rightorth(vecs)
        mps[site] = Q
        mps[site - 1] = _mul_right(mps[site - 1], Lf)
        m.updateright(site)
    return energies, delta


_sweep2 = sweep


def sweep(m, alg):
    """sweep! (dmrgexcited.jl:219-223): left sweep then right sweep."""
    if isinstance(alg, DMRG2):
        return _sweep2(m, alg)
    expand = isinstance(alg, DMRG1S)
    e1, d1 = _leftsweep1(m, alg, expand)
    e2, d2 = _rightsweep1(m, alg, expand)
    return e1 + e2, max(d1, d2)
