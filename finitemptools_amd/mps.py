"""Finite MPS on the device (mps/finitemps.jl:19-22, index convention :130-141).

Site tensors are column-major device tensors A[a, s, b].  ``randommps``
draws iid N(0, 1) entries from a seeded numpy PCG64 stream with the virtual
dimensions min(D, d^i, d^(L-i)) of mps/initializers.jl:14-26; the same seed
always gives the same tensors (Julia's randn stream itself is not reproducible
outside Julia).
"""
import numpy as np
import torch

from .engine import to_device
from .linalg import as_matrix, from_matrix, gemm, cm_view


def virtual_dims(L, d, D):
    dims = [1]
    for _ in range(1, L):
        dims.append(min(D, dims[-1] * d))
    dims.append(1)
    for i in range(L - 1, 0, -1):
        dims[i] = min(dims[i], dims[i + 1] * d)
    return dims


def random_site_tensors(L, d, D, seed=0, cplx=False):
    """Host arrays of the random MPS (numpy, seeded)."""
    rng = np.random.default_rng(seed)
    dims = virtual_dims(L, d, D)
    out = []
    for i in range(L):
        shape = (dims[i], d, dims[i + 1])
        a = rng.standard_normal(shape)
        if cplx:
            a = a + 1j * rng.standard_normal(shape)
        out.append(a)
    return out


class FiniteMPS:
    """List of device site tensors plus the bond singular values ``s``."""

    def __init__(self, tensors):
        self.data = list(tensors)
        self.s = [None] * (len(self.data) + 1)

    def __len__(self):
        return len(self.data)

    def __getitem__(self, i):
        return self.data[i]

    def __setitem__(self, i, v):
        self.data[i] = v

    @property
    def dtype(self):
        return self.data[0].dtype

    def bond_dimension(self):
        return max(t.shape[2] for t in self.data)

    def to_host(self):
        return [t.detach().cpu().numpy() for t in self.data]

    def right_canonicalize(self, normalize=True):
        """mps/orth.jl:40-49: a QR sweep left-to-right, then an LQ sweep
        right-to-left (QR of the adjoint), leaving sites 2..L right-isometric."""
        L = len(self)
        for i in range(L - 1):
            Dl, d, Dr = self.data[i].shape
            q, r = torch.linalg.qr(as_matrix(self.data[i], 2), mode="reduced")
            k = q.shape[1]
            self.data[i] = from_matrix(q, (Dl, d, k))
            nxt = self.data[i + 1]
            self.data[i + 1] = cm_view(gemm(r, as_matrix(nxt, 1)), (k,) + tuple(nxt.shape[1:]))
        if normalize:
            self.data[-1] = self.data[-1] / torch.linalg.vector_norm(self.data[-1])
        for i in range(L - 1, 0, -1):
            Dl, d, Dr = self.data[i].shape
            q, r = torch.linalg.qr(as_matrix(self.data[i], 1).conj().t(), mode="reduced")
            k = q.shape[1]
            self.data[i] = from_matrix(q.conj().t(), (k, d, Dr))
            prv = self.data[i - 1]
            m = gemm(as_matrix(prv, 2), r.conj().t())
            self.data[i - 1] = cm_view(m, tuple(prv.shape[:2]) + (k,))
        return self


def randommps(L, d, D, seed=0, cplx=False, device="cuda"):
    """Seeded random MPS, uploaded (not canonicalised)."""
    dt = torch.complex128 if cplx else torch.float64
    return FiniteMPS([to_device(a, device, dt) for a in random_site_tensors(L, d, D, seed, cplx)])


def prodmps(vectors, dtype=torch.float64, device="cuda"):
    """Product state from one local vector per site (mps/initializers.jl:89-96)."""
    ts = []
    for v in vectors:
        a = np.asarray(v).reshape(1, -1, 1)
        ts.append(to_device(a, device, dtype))
    return FiniteMPS(ts)
