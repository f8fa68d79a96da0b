"""Energy gradients by adjoint contractions: dE/dtheta for MPO couplings and
dE/dA for site tensors.  New functionality: the reference has no
differentiation (Project.toml:6-11 lists no ChainRulesCore / Zygote).

For an MPS psi (any gauge, not necessarily normalised) and an MPO H,

    E = <psi|H|psi> / <psi|psi>,
    dE/dW_i = G_i / <psi|psi>,   G_i = mpo_grad(HL_i, A_i, A_i, HR_{i+1})
    dE/dA_i = 2 (H_eff_i A_i - E N_eff_i A_i) / <psi|psi>      (real tensors)

with HL/HR the MPO environments (updateleft / updateright, mpo/transfer.jl:
26-34), N_eff the same contraction with the identity MPO (norm environments,
mpo/transfer.jl:37-55), H_eff_i = ac_prime (derivatives.jl:2-4) and G_i the
adjoint of ac_prime with respect to the MPO tensor (mh_mpo_grad).  For an MPO
H(theta), dE/dtheta = sum_i <dW_i/dtheta, dE/dW_i>; at an eigenstate this is
the Hellmann-Feynman value <psi|dH/dtheta|psi>.

Environment storage: ``mode="store"`` keeps every right environment (L of
them); ``mode="recompute"`` keeps one every ceil(sqrt(L)) bonds and rebuilds
the others segment by segment during the left-to-right pass (about 2 sqrt(L)
environments resident, one extra pass of right updates).

``MPSEnergy`` wraps the functional as a torch.autograd.Function whose
backward calls these adjoint kernels (the role of a ChainRulesCore rrule).
"""
import math
import time

import numpy as np
import torch

from . import _lib
from .derivatives import ac_prime
from .engine import PREC, as_colmajor, colmajor_empty, dtype_code, engine, ptr, to_device
from .transfer import _identity_mpo, updateleft, updateright


def mpo_grad(hleft, psiA, psiB, hright, prec="f64"):
    """G[w,s',w',s] = hleft[a,w,a'] conj(psiA[a,s',b]) psiB[a',s,b'] hright[b,w',b']"""
    hleft, psiA, psiB, hright = (as_colmajor(t) for t in (hleft, psiA, psiB, hright))
    Dl, d, Dr = psiB.shape
    wl, wr = hleft.shape[1], hright.shape[1]
    if tuple(psiA.shape) != (Dl, d, Dr) or tuple(hleft.shape) != (Dl, wl, Dl) or \
            tuple(hright.shape) != (Dr, wr, Dr):
        raise _lib.DimensionMismatch(2, "mpo_grad: shapes do not match")
    G = colmajor_empty((wl, d, wr, d), psiB.dtype, psiB.device)
    e = engine(psiB.device)
    h = e.bind_stream()
    e.check(e.lib.mh_mpo_grad(h, dtype_code(psiB), PREC[prec], Dl, d, Dr, wl, wr, ptr(hleft),
                              ptr(psiA), ptr(psiB), ptr(hright), ptr(G)), "mpo_grad")
    return G


def _boundary(dtype, device):
    return to_device(torch.ones((1, 1, 1), dtype=dtype), device)


class _RightEnvs:
    """Right environments HR[i] (MPO) and NR[i] (norm) at bonds i = 1..L,
    either all stored or rebuilt from checkpoints."""

    def __init__(self, mpo, mps, mode, prec):
        self.mpo, self.mps, self.mode, self.prec = mpo, mps, mode, prec
        L = len(mps)
        self.L = L
        self.c = max(1, math.ceil(math.sqrt(L))) if mode == "recompute" else 1
        self.updates = 0
        self.peak = 0
        b = _boundary(mps[0].dtype, mps[0].device)
        self.ckpt = {L: (b, b)}
        H, N = b, b
        for i in range(L - 1, 0, -1):
            H, N = self._step(i, H, N)
            if mode == "store" or i % self.c == 0:
                self.ckpt[i] = (H, N)
        self.seg = {}
        self._track()

    def _step(self, i, H, N):
        """environments at bond i from those at bond i+1 (site i)"""
        A = self.mps[i]
        self.updates += 2
        return (updateright(H, A, self.mpo[i], A, prec=self.prec),
                updateright(N, A, None, A, prec=self.prec))

    def _track(self):
        self.peak = max(self.peak, len(self.ckpt) + len(self.seg))

    def get(self, i):
        """(HR[i], NR[i]) for increasing i"""
        if i in self.ckpt:
            return self.ckpt[i]
        if i not in self.seg:
            # rebuild the segment (top checkpoint above i) down to i
            top = min(k for k in self.ckpt if k > i)
            self.seg = {}
            H, N = self.ckpt[top]
            for j in range(top - 1, i - 1, -1):
                H, N = self._step(j, H, N)
                self.seg[j] = (H, N)
            self._track()
        return self.seg[i]

    def env_bytes(self):
        """upper bound: resident count times the largest environment pair"""
        per = max((H.numel() + N.numel()) * H.element_size()
                  for H, N in list(self.ckpt.values()) + list(self.seg.values()))
        return per * self.peak


class GradResult:
    def __init__(self, energy, norm, dW, dA, stats):
        self.energy, self.norm, self.dW, self.dA, self.stats = energy, norm, dW, dA, stats


def energy_gradients(mpo, mps, want_dA=True, mode="store", prec="f64"):
    """E, dE/dW_i (all sites) and dE/dA_i (real tensors) of
    E = <psi|H|psi>/<psi|psi> by one right pass and one left pass."""
    if mode not in ("store", "recompute"):
        raise ValueError("mode is 'store' or 'recompute'")
    L = len(mps)
    if len(mpo) != L:
        raise _lib.DimensionMismatch(2, "MPO and MPS lengths differ")
    if want_dA and mps[0].is_complex():
        raise TypeError("dE/dA is implemented for real site tensors")
    t0 = time.perf_counter()
    R = _RightEnvs(mpo, mps, mode, prec)
    HL = NL = _boundary(mps[0].dtype, mps[0].device)
    G, hA, nA = [], [], []
    Hval = Nval = None
    left_updates = 0
    for i in range(L):
        A, W = mps[i], mpo[i]
        HR, NR = R.get(i + 1)
        G.append(mpo_grad(HL, A, A, HR, prec=prec))
        if want_dA or i == 0:
            d = A.shape[1]
            h = ac_prime(A, W, HL, HR, prec=prec)
            n = ac_prime(A, _identity_mpo(d, 1, A.dtype, A.device), NL, NR, prec=prec)
            if i == 0:
                Hval = torch.vdot(A.reshape(-1), h.reshape(-1)).item()
                Nval = torch.vdot(A.reshape(-1), n.reshape(-1)).item()
            if want_dA:
                hA.append(h)
                nA.append(n)
        if i + 1 < L:
            HL = updateleft(HL, A, W, A, prec=prec)
            NL = updateleft(NL, A, None, A, prec=prec)
            left_updates += 2
    E = (Hval / Nval).real if isinstance(Hval, complex) else Hval / Nval
    Nv = Nval.real if isinstance(Nval, complex) else Nval
    dW = [g / Nv for g in G]
    dA = [2.0 * (h - E * n) / Nv for h, n in zip(hA, nA)] if want_dA else None
    torch.cuda.synchronize(mps[0].device) if mps[0].is_cuda else None
    stats = {"seconds": time.perf_counter() - t0, "right_updates": R.updates,
             "left_updates": left_updates, "resident_right_envs": R.peak,
             "resident_env_bytes": R.env_bytes(), "mode": mode}
    return GradResult(E, Nv, dW, dA, stats)


def dE_dtheta(dW_dtheta, res):
    """sum_i <dW_i/dtheta, dE/dW_i> for an MPO H(theta)."""
    tot = 0.0
    for dWi, g in zip(dW_dtheta, res.dW):
        tot += torch.sum(as_colmajor(dWi) * g).item()
    return tot.real if isinstance(tot, complex) else tot


def hellmann_feynman(dmpo, mps):
    """<psi|dH/dtheta|psi> / <psi|psi> with dH/dtheta given as its own MPO."""
    from .envs import expectation
    num = expectation(dmpo, mps)
    b = _boundary(mps[0].dtype, mps[0].device)
    for i in range(len(mps) - 1, -1, -1):
        b = updateright(b, mps[i], None, mps[i])
    return num / b.reshape(-1)[0].item()


class MPSEnergy(torch.autograd.Function):
    """E = <psi|H|psi>/<psi|psi> as a differentiable function of the MPO
    tensors and the (real) site tensors; backward runs the adjoint kernels.

        E = MPSEnergy.apply(L, mode, *W_list, *A_list)
    """

    @staticmethod
    def forward(ctx, L, mode, *tensors):
        W, A = list(tensors[:L]), list(tensors[L:])
        want_dA = any(t.requires_grad for t in A)
        res = energy_gradients(W, A, want_dA=want_dA, mode=mode)
        ctx.save_for_backward(*res.dW, *(res.dA if want_dA else []))
        ctx.L, ctx.want_dA = L, want_dA
        return torch.tensor(res.energy, dtype=torch.float64, device=A[0].device)

    @staticmethod
    def backward(ctx, gE):
        saved = ctx.saved_tensors
        L = ctx.L
        dW = saved[:L]
        dA = saved[L:] if ctx.want_dA else [None] * L
        g = gE.item()
        return (None, None) + tuple(x * g for x in dW) + \
            tuple(None if x is None else x * g for x in dA)


def mps_energy(mpo, mps, mode="store"):
    return MPSEnergy.apply(len(mps), mode, *mpo, *mps)
