"""numpy restatement of the two-site DMRG sweep (algorithms/dmrg.jl:79-144),
its environment cache (envs/finiteenv.jl:3-78), the truncation scheme
(auxiliary/truncation.jl:8-47, auxiliary/stable_svd.jl:12-18) and the seeded
random right-canonical initial MPS (mps/initializers.jl:14-26, mps/orth.jl:40-49).
"""
import numpy as np

from . import contractions as C
from .lanczos import simple_lanczos_solver


# ------------------------------------------------------------------ MPS
def virtual_dims(L, d, D):
    """mps/initializers.jl:14-26 for a dense space: min(D, d^i, d^(L-i))."""
    dims = [1]
    for i in range(1, L):
        dims.append(min(D, dims[-1] * d))
    dims.append(1)
    for i in range(L - 1, 0, -1):
        dims[i] = min(dims[i], dims[i + 1] * d)
    return dims


def randommps(L, d, D, seed=0, cplx=False):
    """iid N(0,1) site tensors A[a, s, b] with the dims above (seeded PCG64)."""
    rng = np.random.default_rng(seed)
    dims = virtual_dims(L, d, D)
    mps = []
    for i in range(L):
        shape = (dims[i], d, dims[i + 1])
        a = rng.standard_normal(shape)
        if cplx:
            a = a + 1j * rng.standard_normal(shape)
        mps.append(a)
    return mps


def _mat(A, left):
    Dl, d, Dr = A.shape
    return A.reshape(Dl * d, Dr, order="F") if left else A.reshape(Dl, d * Dr, order="F")


def right_canonicalize(mps, normalize=True):
    """mps/orth.jl:40-49 (leftorth sweep then rightorth sweep), QR/LQ based."""
    mps = [a.copy() for a in mps]
    L = len(mps)
    for i in range(L - 1):
        Dl, d, Dr = mps[i].shape
        q, r = np.linalg.qr(_mat(mps[i], True))
        k = q.shape[1]
        mps[i] = q.reshape(Dl, d, k, order="F")
        mps[i + 1] = np.tensordot(r, mps[i + 1], axes=([1], [0]))
    if normalize:
        mps[-1] = mps[-1] / np.linalg.norm(mps[-1])
    for i in range(L - 1, 0, -1):
        Dl, d, Dr = mps[i].shape
        q, r = np.linalg.qr(_mat(mps[i], False).conj().T)  # M = r^dag q^dag
        k = q.shape[1]
        mps[i] = q.conj().T.reshape(k, d, Dr, order="F")
        mps[i - 1] = np.tensordot(mps[i - 1], r.conj().T, axes=([2], [0]))
    return mps


def to_dense(mps):
    """Full state psi[s1, ..., sL] flattened C-order (site 1 slowest)."""
    t = mps[0][0]  # (d, Dr)
    for A in mps[1:]:
        t = np.tensordot(t, A, axes=([-1], [0]))
    return t.reshape(-1)


def mpo_to_dense(mpo):
    """H[(s1'..sL'), (s1..sL)] by contracting the MPO bonds (C-order, site 1 slowest)."""
    t = mpo[0][0]  # [s', wr, s]
    t = t.transpose(1, 0, 2)  # [w, s', s]
    dout = din = 1
    t = t.reshape(t.shape[0], t.shape[1], t.shape[2])
    for W in mpo[1:]:
        # t[w, S', S] W[w, s', w', s] -> [w', S' s', S s]
        u = np.tensordot(t, W, axes=([0], [0]))  # [S', S, s', w', s]
        u = u.transpose(3, 0, 2, 1, 4)
        dout = u.shape[1] * u.shape[2]
        din = u.shape[3] * u.shape[4]
        t = u.reshape(u.shape[0], dout, din)
    return t[0]


# ----------------------------------------------------------- truncation
class MPSTruncation:
    """auxiliary/truncation.jl:8-47: cut singular values below eps*||s||_2, then
    keep at most D; error = 2-norm of the discarded values."""

    def __init__(self, D=100, eps=1e-8):
        self.D, self.eps = D, eps

    def truncate(self, s):
        n = np.linalg.norm(s)
        keep = s > self.eps * n
        k1 = int(np.count_nonzero(keep))
        err1 = np.linalg.norm(s[k1:])
        if k1 <= self.D:
            return k1, err1
        err2 = np.linalg.norm(s[self.D:k1])
        return self.D, float(np.hypot(err1, err2))


def tsvd_two_site(x, trunc):
    """stable_tsvd!(twositemps, trunc) (stable_svd.jl:12-18) on x[a,s1,s2,b]."""
    Dl, d1, d2, Dr = x.shape
    M = x.reshape(Dl * d1, d2 * Dr, order="F")
    u, s, vh = np.linalg.svd(M, full_matrices=False)
    k, err = trunc.truncate(s)
    k = max(k, 1)
    u = u[:, :k].reshape(Dl, d1, k, order="F")
    v = vh[:k].reshape(k, d2, Dr, order="F")
    return u, s[:k], v, err


# ---------------------------------------------------------- environments
class ExpectationCache:
    """envs/finiteenv.jl:3-21,51-78: hstorage[i] left env of site i, hstorage[i+1]
    its right env (0-based); built from the right (center = first site)."""

    def __init__(self, mpo, mps):
        self.mpo, self.mps = mpo, mps
        L = len(mps)
        dt = np.result_type(mps[0], mpo[0])
        self.hstorage = [None] * (L + 1)
        self.hstorage[L] = C.r_RR(dt)
        self.hstorage[0] = C.l_LL(dt)
        for i in range(L - 1, 0, -1):
            self.updateright(i)

    def updateleft(self, site):
        self.hstorage[site + 1] = C.updateleft(self.hstorage[site], self.mps[site],
                                               self.mpo[site], self.mps[site])

    def updateright(self, site):
        self.hstorage[site] = C.updateright(self.hstorage[site + 1], self.mps[site],
                                            self.mpo[site], self.mps[site])


class DMRG2:
    """dmrg.jl:79-86 defaults (verbosity off)"""

    def __init__(self, maxiter=100, tol=1e-12, eig_maxiter=10, eig_tol=1e-12, trunc=None):
        self.maxiter, self.tol = maxiter, tol
        self.eig_maxiter, self.eig_tol = eig_maxiter, eig_tol
        self.trunc = trunc or MPSTruncation()


def _two_site(mps, site):
    return np.tensordot(mps[site], mps[site + 1], axes=([2], [0]))


def _overlap(two, u, s, v):
    usv = np.tensordot(u * s[None, None, :], v, axes=([2], [0]))
    return np.vdot(usv, two)


def leftsweep(env, alg, counter=None):
    """dmrg.jl:89-115"""
    mpo, mps = env.mpo, env.mps
    energies, delta = [], 0.0
    for site in range(len(mps) - 2):
        two = _two_site(mps, site)
        A = lambda x: C.ac2_prime(x, mpo[site], mpo[site + 1], env.hstorage[site], env.hstorage[site + 2])
        ev, vecs, _, k = simple_lanczos_solver(A, two, "SR", alg.eig_maxiter, alg.eig_tol)
        if counter is not None:
            counter.append(k)
        energies.append(ev)
        u, s, v, _ = tsvd_two_site(vecs, alg.trunc)
        s = s / np.linalg.norm(s)
        mps[site] = u
        mps[site + 1] = s[:, None, None] * v
        delta = max(delta, abs(1 - abs(_overlap(two, u, s, v))))
        env.updateleft(site)
    return energies, delta


def rightsweep(env, alg, counter=None):
    """dmrg.jl:117-144"""
    mpo, mps = env.mpo, env.mps
    energies, delta = [], 0.0
    for site in range(len(mps) - 2, -1, -1):
        two = _two_site(mps, site)
        A = lambda x: C.ac2_prime(x, mpo[site], mpo[site + 1], env.hstorage[site], env.hstorage[site + 2])
        ev, vecs, _, k = simple_lanczos_solver(A, two, "SR", alg.eig_maxiter, alg.eig_tol)
        if counter is not None:
            counter.append(k)
        energies.append(ev)
        u, s, v, _ = tsvd_two_site(vecs, alg.trunc)
        s = s / np.linalg.norm(s)
        mps[site] = u * s[None, None, :]
        mps[site + 1] = v
        delta = max(delta, abs(1 - abs(_overlap(two, u, s, v))))
        env.updateright(site + 1)
    return energies, delta


def sweep(env, alg, counter=None):
    """dmrgexcited.jl:219-223: left sweep then right sweep"""
    e1, d1 = leftsweep(env, alg, counter)
    e2, d2 = rightsweep(env, alg, counter)
    return e1 + e2, max(d1, d2)


def compute(env, alg, nsweeps=None):
    """dmrg.jl:221-234 (nsweeps caps the sweep count when given)"""
    all_e, it, delta = [], 0, 2 * alg.tol
    maxit = alg.maxiter if nsweeps is None else nsweeps
    while it < maxit and delta > alg.tol:
        e, delta = sweep(env, alg)
        all_e += e
        it += 1
    return all_e, delta
