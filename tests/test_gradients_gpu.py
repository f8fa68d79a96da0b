"""Adjoint kernels and energy gradients on the device (no counterpart in the
reference, which has no differentiation): the MPO-tensor adjoint against the
oracle, dE/dJ2 against the Hellmann-Feynman value and central finite
differences, dE/dA against directional finite differences, stored versus
recomputed environments, the autograd wrapper, and dE/dJ2 at converged
J1-J2 ground states (Majumdar-Ghosh closed form at J2 = 1/2)."""
import numpy as np
import pytest
import torch

from conftest import rand
from finitemptools_amd import gradients as GR
from finitemptools_amd import models
from finitemptools_amd.engine import to_device
from finitemptools_amd.envs import mpo_to_device
from oracle import contractions as C
from oracle import dmrg as OD

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("cplx,Dl,Dr,wl,wr,d", [
    (False, 24, 40, 8, 8, 2), (True, 16, 12, 3, 5, 2), (False, 8, 1100, 3, 3, 2),
    (False, 64, 64, 6, 6, 4), (True, 70, 33, 5, 5, 2)])
def test_mpo_grad_matches_oracle(gpu, cplx, Dl, Dr, wl, wr, d):
    rng = np.random.default_rng(Dl + Dr)
    L, R = rand(rng, (Dl, wl, Dl), cplx), rand(rng, (Dr, wr, Dr), cplx)
    A, B = rand(rng, (Dl, d, Dr), cplx), rand(rng, (Dl, d, Dr), cplx)
    G = GR.mpo_grad(*(to_device(t) for t in (L, A, B, R))).cpu().numpy()
    assert rel(G, C.mpo_grad(L, A, B, R)) < 1e-12


def test_mpo_grad_mixed_precision(gpu):
    rng = np.random.default_rng(2)
    L, R = rand(rng, (96, 8, 96)), rand(rng, (96, 8, 96))
    A, B = rand(rng, (96, 2, 96)), rand(rng, (96, 2, 96))
    ref = C.mpo_grad(L, A, B, R)
    G = GR.mpo_grad(*(to_device(t) for t in (L, A, B, R)), prec="f32").cpu().numpy()
    assert 1e-12 < rel(G, ref) < 1e-5


def random_state(L, D, seed):
    return [to_device(A) for A in OD.randommps(L, 2, D, seed=seed)]


def j1j2_tensors(L, J2):
    return mpo_to_device(models.j1j2_mpo(L, 1.0, J2, keep_j2=True))


def dW_dJ2(L):
    W1, W0 = j1j2_tensors(L, 1.0), j1j2_tensors(L, 0.0)
    return [a - b for a, b in zip(W1, W0)]


def test_energy_and_gradients_on_a_random_state(gpu):
    L, D, J2 = 8, 8, 0.3
    mps = random_state(L, D, 3)               # not normalised, not canonical
    mpo = j1j2_tensors(L, J2)
    res = GR.energy_gradients(mpo, mps)
    psi = OD.to_dense([A.cpu().numpy() for A in mps])
    H = OD.mpo_to_dense(models.j1j2_mpo(L, 1.0, J2))
    E_ref = np.vdot(psi, H @ psi) / np.vdot(psi, psi)
    assert abs(res.energy - E_ref) < 1e-11 * abs(E_ref)
    # dE/dJ2 at fixed psi: the expectation of dH/dJ2 (exact identity)
    dH = OD.mpo_to_dense(models.j1j2_dJ2_mpo(L))
    hf_ref = np.vdot(psi, dH @ psi) / np.vdot(psi, psi)
    g = GR.dE_dtheta(dW_dJ2(L), res)
    assert abs(g - hf_ref) < 1e-11
    assert abs(GR.hellmann_feynman(mpo_to_device(models.j1j2_dJ2_mpo(L)), mps) - hf_ref) < 1e-11
    h = 1e-5
    fd = (GR.energy_gradients(j1j2_tensors(L, J2 + h), mps, want_dA=False).energy -
          GR.energy_gradients(j1j2_tensors(L, J2 - h), mps, want_dA=False).energy) / (2 * h)
    assert abs(fd - g) < 1e-8
    # dE/dA along a random direction
    rng = np.random.default_rng(5)
    dirs = [to_device(rng.standard_normal(tuple(A.shape))) for A in mps]
    lin = sum(torch.sum(gA * dA).item() for gA, dA in zip(res.dA, dirs))
    hA = 1e-5
    Ep = GR.energy_gradients(mpo, [A + hA * u for A, u in zip(mps, dirs)], want_dA=False).energy
    Em = GR.energy_gradients(mpo, [A - hA * u for A, u in zip(mps, dirs)], want_dA=False).energy
    assert abs((Ep - Em) / (2 * hA) - lin) < 1e-6 * max(1.0, abs(lin))


def test_recomputed_environments_give_identical_gradients(gpu):
    L, D = 12, 10
    mps = random_state(L, D, 4)
    mpo = j1j2_tensors(L, 0.25)
    a = GR.energy_gradients(mpo, mps, mode="store")
    b = GR.energy_gradients(mpo, mps, mode="recompute")
    assert a.energy == b.energy
    for x, y in zip(a.dW + a.dA, b.dW + b.dA):
        assert torch.equal(x, y)
    assert b.stats["resident_right_envs"] < a.stats["resident_right_envs"]
    assert b.stats["right_updates"] > a.stats["right_updates"]


def test_autograd_wrapper(gpu):
    """torch.autograd through MPSEnergy: the MPO depends on J2 through torch
    ops, the site tensors are leaves; backward runs the adjoint kernels."""
    L, D = 8, 6
    mps = [A.clone().requires_grad_(True) for A in random_state(L, D, 6)]
    W0, W1 = j1j2_tensors(L, 0.0), j1j2_tensors(L, 1.0)
    J2 = torch.tensor(0.4, dtype=torch.float64, device="cuda", requires_grad=True)
    mpo = [a + J2 * (b - a) for a, b in zip(W0, W1)]
    E = GR.mps_energy(mpo, mps)
    E.backward()
    res = GR.energy_gradients([m.detach() for m in mpo], [A.detach() for A in mps])
    assert abs(J2.grad.item() - GR.dE_dtheta(dW_dJ2(L), res)) < 1e-12
    for A, g in zip(mps, res.dA):
        assert torch.allclose(A.grad, g, rtol=0, atol=1e-13)


@pytest.mark.parametrize("J2", [0.0, 0.25, 0.5])
def test_dE_dJ2_at_the_ground_state(gpu, J2):
    """Converged DMRG2 ground state of the J1-J2 chain (L = 12, D = 64 holds
    it exactly): the adjoint dE/dJ2 equals the Hellmann-Feynman value and the
    central finite difference of ground energies (h = 1e-5); at the
    Majumdar-Ghosh point the dimer state gives sum <S_i.S_{i+2}> = 0."""
    from finitemptools_amd import dmrg
    from finitemptools_amd.mps import randommps
    from finitemptools_amd.truncation import MPSTruncation
    L, D = 12, 64

    def ground(j2):
        psi = randommps(L, 2, 16, seed=1).right_canonicalize()
        alg = dmrg.DMRG2(trunc=MPSTruncation(D=D, eps=1e-12))
        e, _ = dmrg.ground_state(psi, j1j2_tensors(L, j2), alg, nsweeps=8)
        return e[-1], psi

    E0, psi = ground(J2)
    sites = [psi[i] for i in range(L)]
    res = GR.energy_gradients(j1j2_tensors(L, J2), sites, want_dA=False)
    g = GR.dE_dtheta(dW_dJ2(L), res)
    hf = GR.hellmann_feynman(mpo_to_device(models.j1j2_dJ2_mpo(L)), sites)
    assert abs(res.energy - E0) < 1e-9
    assert abs(g - hf) < 1e-10
    h = 1e-5
    fd = (ground(J2 + h)[0] - ground(J2 - h)[0]) / (2 * h)
    assert abs(fd - g) < 1e-5
    if J2 == 0.5:
        assert abs(g) < 1e-7
