"""Environment cache on the device: envs/finiteenv.jl:3-93.

``hstorage[i]`` is the left environment of site i and ``hstorage[i+1]`` its
right environment (0-based sites).  ``environments(mpo, mps)`` builds every
slot from the right (centre = first site), as ``init_hstorage_right``.
"""
import numpy as np
import torch

from . import transfer
from .engine import to_device


def _boundary(dtype, device):
    """Strict boundary environment: the 1x1x1 isometry (mpo/arithmetics.jl:253-258)."""
    return to_device(torch.ones((1, 1, 1), dtype=dtype), device)


class ExpectationCache:
    def __init__(self, mpo, mps, prec="f64"):
        if len(mpo) != len(mps):
            raise ValueError("MPO and MPS lengths differ")
        if mpo[0].shape[0] != 1 or mpo[-1].shape[2] != 1:
            raise ValueError("operator must be strict")
        self.mpo, self.mps, self.prec = mpo, mps, prec
        self.hstorage = [None] * (len(mps) + 1)
        self.init_hstorage(center=0)

    # reference property names (finiteenv.jl:59-69)
    @property
    def state(self):
        return self.mps

    @property
    def h(self):
        return self.mpo

    @property
    def env(self):
        return self.hstorage

    def init_hstorage(self, center=0):
        """init_hstorage! (finiteenv.jl:3-21)"""
        L = len(self.mps)
        dt, dev = self.mps[0].dtype, self.mps[0].device
        self.hstorage[L] = _boundary(dt, dev)
        self.hstorage[0] = _boundary(dt, dev)
        for i in range(L - 1, center, -1):
            self.updateright(i)
        for i in range(center):
            self.updateleft(i)

    def updateleft(self, site):
        """updateleft! (finiteenv.jl:72-74)"""
        A = self.mps[site]
        self.hstorage[site + 1] = transfer.updateleft(self.hstorage[site], A, self.mpo[site], A, prec=self.prec)

    def updateright(self, site):
        """updateright! (finiteenv.jl:76-78)"""
        A = self.mps[site]
        self.hstorage[site] = transfer.updateright(self.hstorage[site + 1], A, self.mpo[site], A, prec=self.prec)


def overlap_updateleft(hold, psiA, psiB):
    """hnew[b, b'] = conj(psiA[a,s,b]) hold[a,a'] psiB[a',s,b'] (mps/transfer.jl:23-31),
    two engine GEMMs; psiA and psiB may have different bond dimensions."""
    from .linalg import as_matrix, cm_view, gemm
    Dla, d, Dra = psiA.shape
    U = cm_view(gemm(hold, as_matrix(psiB, 1)), (Dla, d, psiB.shape[2]))       # [a, s, b']
    return gemm(as_matrix(psiA, 2).t(), as_matrix(U, 2), conjA=True)           # [b, b']


def overlap_updateright(hold, psiA, psiB):
    """hnew[a, a'] = conj(psiA[a,s,b]) hold[b,b'] psiB[a',s,b'] (mps/transfer.jl:8-16)."""
    from .linalg import as_matrix, cm_view, gemm
    Dlb, d, Drb = psiB.shape
    U = cm_view(gemm(as_matrix(psiB, 2), hold.t()), (Dlb, d, psiA.shape[2]))  # [a', s, b]
    return gemm(as_matrix(psiA, 1), as_matrix(U, 1).t(), conjA=True)           # [a, a']


def c_proj(x, cleft, cright):
    """tmp[a,s,b] = cleft[a,x] x[x,s,y] cright[b,y] (derivatives.jl:65-67): a
    lower state's site tensor seen through the overlap environments."""
    from .linalg import as_matrix, cm_view, gemm
    Dl, Dr = cleft.shape[0], cright.shape[0]
    t = cm_view(gemm(cleft, as_matrix(x, 1)), (Dl, x.shape[1], x.shape[2]))
    return cm_view(gemm(as_matrix(t, 2), cright.t()), (Dl, x.shape[1], Dr))


def c2_proj(x1, x2, cleft, cright):
    """tmp[a,s1,s2,b] = cleft[a,1] x1[1,s1,2] x2[2,s2,3] cright[b,3]
    (_c2_prime_2, dmrgexcited.jl:70-72)."""
    from .linalg import as_matrix, cm_view, gemm, two_site
    x = two_site(x1, x2)
    _, d1, d2, _ = x.shape
    Dl, Dr = cleft.shape[0], cright.shape[0]
    t = cm_view(gemm(cleft, as_matrix(x, 1)), (Dl, d1, d2, x.shape[3]))
    return cm_view(gemm(as_matrix(t, 3), cright.t()), (Dl, d1, d2, Dr))


class ProjectedExpectationCache(ExpectationCache):
    """Environments of an excited-state search (envs/finiteenv.jl:97-168):
    the MPO environments plus, for every lower state ("projector") l, the
    overlap environments cstorages[l][i] between the state and l
    (cstorages[l][i] is the left environment of site i, [i+1] the right)."""

    def __init__(self, mpo, mps, projectors, prec="f64"):
        self.projectors = list(projectors)
        for p in self.projectors:
            if len(p) != len(mps):
                raise _lib_dm("projector length differs from the state")
        self.cstorages = [[None] * (len(mps) + 1) for _ in self.projectors]
        super().__init__(mpo, mps, prec)

    @property
    def cenvs(self):
        return self.cstorages

    def init_hstorage(self, center=0):
        """init_hstorage_right + init_cstorage_right! (finiteenv.jl:97-117): the
        overlap boundaries (l_LL, r_RR), then the base class's passes, whose
        updateright / updateleft calls (below) fill both kinds of slots."""
        L = len(self.mps)
        dt, dev = self.mps[0].dtype, self.mps[0].device
        for cs in self.cstorages:
            cs[0] = torch.ones((1, 1), dtype=dt, device=dev)   # l_LL
            cs[L] = torch.ones((1, 1), dtype=dt, device=dev)   # r_RR
        super().init_hstorage(center)

    def updateleft(self, site):
        """updateleft! (finiteenv.jl:151-156)"""
        super().updateleft(site)
        for l, cs in enumerate(self.cstorages):
            cs[site + 1] = overlap_updateleft(cs[site], self.mps[site], self.projectors[l][site])

    def updateright(self, site):
        """updateright! (finiteenv.jl:159-164)"""
        super().updateright(site)
        for l, cs in enumerate(self.cstorages):
            cs[site] = overlap_updateright(cs[site + 1], self.mps[site], self.projectors[l][site])

    def site_projectors(self, site):
        """[c_proj(projectors[l][site], cstorages[l][site], cstorages[l][site+1])]"""
        return [c_proj(self.projectors[l][site], cs[site], cs[site + 1])
                for l, cs in enumerate(self.cstorages)]

    def bond_projectors(self, site):
        """[_c2_prime_2(projectors[l][site], projectors[l][site+1], cstorages[l][site],
        cstorages[l][site+2])]"""
        return [c2_proj(self.projectors[l][site], self.projectors[l][site + 1], cs[site], cs[site + 2])
                for l, cs in enumerate(self.cstorages)]


def _lib_dm(msg):
    from . import _lib
    return _lib.DimensionMismatch(2, msg)


def environments(mpo, mps, projectors=None, prec="f64"):
    """environments(mpo, mps) (finiteenv.jl:57) and, with lower states,
    environments(mpo, mps, projectors) (finiteenv.jl:119-120)."""
    if isinstance(projectors, str):   # environments(mpo, mps, prec)
        projectors, prec = None, projectors
    if projectors:
        return ProjectedExpectationCache(mpo, mps, projectors, prec)
    return ExpectationCache(mpo, mps, prec)


def expectation(mpo, mps):
    """<psi|h|psi> by right-to-left environment updates (mpo/arithmetics.jl:267-275)."""
    hold = _boundary(mps[0].dtype, mps[0].device)
    for i in range(len(mps) - 1, -1, -1):
        hold = transfer.updateright(hold, mps[i], mpo[i], mps[i])
    return hold.reshape(-1)[0].item()


def local_expectations(mps, op):
    """<psi|op_i|psi> / <psi|psi> at every site i: one left and one right pass
    of identity-MPO environment updates (mpo/transfer.jl:37-55 with
    hj::Nothing), then per site one update with the 1 x d x 1 x d operator
    tensor closed against the right environment -- O(L) contractions instead
    of one full expectation (mpo/arithmetics.jl:267-275) per site."""
    L, d = len(mps), mps[0].shape[1]
    dt, dev = mps[0].dtype, mps[0].device
    W = np.zeros((1, d, 1, d), dtype=np.complex128 if mps[0].is_complex() else np.float64)
    W[0, :, 0, :] = np.asarray(op)
    Wd = to_device(W, dev, dt)
    left = [_boundary(dt, dev)]
    for i in range(L):
        left.append(transfer.updateleft(left[-1], mps[i], None, mps[i]))
    right = [None] * (L + 1)
    right[L] = _boundary(dt, dev)
    for i in range(L - 1, 0, -1):
        right[i] = transfer.updateright(right[i + 1], mps[i], None, mps[i])
    norm = left[L].reshape(-1)[0]
    vals = torch.stack([torch.sum(transfer.updateleft(left[i], mps[i], Wd, mps[i]) * right[i + 1])
                        for i in range(L)])
    return (vals / norm).real.cpu().numpy()


def mpo_to_device(mpo_host, dtype=torch.float64, device="cuda"):
    return [to_device(W, device, dtype) for W in mpo_host]


def local_expectation(mps, op, site):
    """<psi|op_site|psi> / <psi|psi> for a local d x d operator."""
    from .models import site_operator_mpo
    L, d = len(mps), mps[0].shape[1]
    dt = np.complex128 if mps[0].is_complex() else np.float64
    num = expectation(mpo_to_device(site_operator_mpo(L, site, op, dt), mps[0].dtype, mps[0].device), mps)
    den = expectation(mpo_to_device(site_operator_mpo(L, site, np.eye(d), dt), mps[0].dtype, mps[0].device), mps)
    return num / den


def env_gather(parts):
    """The reference's Vector environment (one (D1, D2) device matrix per MPO
    channel) as the engine's E[a, w, a'] (mh_env_gather)."""
    import ctypes
    from . import _lib
    from .engine import as_colmajor, colmajor_empty, dtype_code, engine, ptr
    cm = [as_colmajor(p) for p in parts]   # column-major storage
    D1, D2 = cm[0].shape
    E = colmajor_empty((D1, len(cm), D2), cm[0].dtype, cm[0].device)
    e = engine(cm[0].device)
    h = e.bind_stream()
    arr = (ctypes.c_void_p * len(cm))(*[ptr(p).value for p in cm])
    _lib.raise_for(_lib.load().mh_env_gather(h, dtype_code(cm[0]), D1, D2, len(cm), arr,
                                             ptr(E)), h, "mh_env_gather")
    return E


def env_scatter(E):
    """E[a, w, a'] as the Vector of per-channel (D1, D2) matrices
    (mh_env_scatter)."""
    import ctypes
    from . import _lib
    from .engine import as_colmajor, colmajor_empty, dtype_code, engine, ptr
    E = as_colmajor(E)
    D1, nw, D2 = E.shape
    parts = [colmajor_empty((D1, D2), E.dtype, E.device) for _ in range(nw)]
    e = engine(E.device)
    h = e.bind_stream()
    arr = (ctypes.c_void_p * nw)(*[ptr(p).value for p in parts])
    _lib.raise_for(_lib.load().mh_env_scatter(h, dtype_code(E), D1, D2, nw, ptr(E), arr), h,
                   "mh_env_scatter")
    return parts
