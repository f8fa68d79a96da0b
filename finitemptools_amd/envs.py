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


def environments(mpo, mps, prec="f64"):
    """environments(mpo, mps) (finiteenv.jl:57)"""
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
