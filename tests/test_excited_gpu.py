"""Excited states (algorithms/dmrgexcited.jl, envs/finiteenv.jl:97-168,
mps/transfer.jl) on the device: the projected Lanczos solve, the overlap
environments between states of different bond dimensions, and DMRG2 / DMRG1S
on environments(mpo, mps, [ground state]) against the second eigenvalue of
exact diagonalisation (1e-8, the reference's excitation test,
test/runtests.jl:494-535,780-784)."""
import numpy as np
import pytest

from conftest import hermitian_env, rand
from finitemptools_amd import models
from oracle import contractions as C
from oracle import dmrg as OD
from oracle import ed

pytestmark = pytest.mark.gpu


def test_projected_lanczos_second_eigenvalue(gpu):
    """With the ground vector as projector the full-dimension solve returns
    the second eigenvalue of the dense H_eff; the start vector is projected."""
    import torch
    from finitemptools_amd.derivatives import HeffOperator
    from finitemptools_amd.engine import to_device
    from finitemptools_amd.krylov import simple_lanczos_solver
    rng = np.random.default_rng(3)
    W = models.tfim_mpo(4, 1.0, 0.7)[1]   # Hermitian blocks per channel: H_eff symmetric
    w = W.shape[0]
    D = 2
    L = hermitian_env(rng, D, w)
    R = hermitian_env(rng, D, w)
    H = C.dense_heff2(W, W, L, R)
    assert np.allclose(H, H.T, atol=1e-12)
    evals, evecs = np.linalg.eigh(H)
    assert evals[1] - evals[0] > 1e-3
    n = H.shape[0]
    op = HeffOperator(2, to_device(L), to_device(R), to_device(W), to_device(W))
    g = evecs[:, 0].reshape(D, 2, 2, D, order="F")
    x0 = rand(rng, (D, 2, 2, D))
    e, v, _ = simple_lanczos_solver(op, to_device(x0), "SR", n, 1e-12, reorth=True,
                                    projectors=[to_device(g)])
    assert abs(e - evals[1]) < 1e-10
    vh = v.cpu().numpy().reshape(-1, order="F")
    assert abs(np.dot(vh, evecs[:, 0])) < 1e-8


def test_overlap_environments_give_the_overlap(gpu):
    """Left-to-right and right-to-left overlap environment passes between two
    MPS of different bond dimensions both give <psi|phi>."""
    import torch
    from finitemptools_amd.envs import overlap_updateleft, overlap_updateright
    from finitemptools_amd.mps import randommps
    L = 7
    psi = randommps(L, 2, 5, seed=1, cplx=True)
    phi = randommps(L, 2, 3, seed=2, cplx=True)
    ref = np.vdot(OD.to_dense(psi.to_host()), OD.to_dense(phi.to_host()))
    c = torch.ones((1, 1), dtype=torch.complex128, device="cuda")
    for i in range(L):
        c = overlap_updateleft(c, psi[i], phi[i])
    assert abs(c.item() - ref) < 1e-12 * max(1.0, abs(ref))
    c = torch.ones((1, 1), dtype=torch.complex128, device="cuda")
    for i in range(L - 1, -1, -1):
        c = overlap_updateright(c, psi[i], phi[i])
    assert abs(c.item() - ref) < 1e-12 * max(1.0, abs(ref))


@pytest.mark.parametrize("algname", ["DMRG2", "DMRG1S"])
def test_first_excited_state_matches_ed(gpu, algname):
    from finitemptools_amd import dmrg
    from finitemptools_amd.envs import mpo_to_device
    from finitemptools_amd.mps import randommps
    from finitemptools_amd.truncation import MPSTruncation
    L = 8
    params = (1.0, 0.9, 0.1)
    evals = np.linalg.eigvalsh(ed.xxz_hamiltonian(L, *params).toarray())
    mpo = mpo_to_device(models.heisenberg_xxz_mpo(L, *params))
    gs = randommps(L, 2, 4, seed=1).right_canonicalize()
    e0, _ = dmrg.ground_state(gs, mpo, dmrg.DMRG2(trunc=MPSTruncation(D=16, eps=1e-12)), nsweeps=6)
    assert abs(e0[-1] - evals[0]) < 1e-8
    psi = randommps(L, 2, 4, seed=7).right_canonicalize()
    if algname == "DMRG2":
        alg = dmrg.DMRG2(trunc=MPSTruncation(D=16, eps=1e-12))
    else:
        alg = dmrg.DMRG1S(trunc=MPSTruncation(D=16, eps=1e-12))
    e1, _ = dmrg.excited_state(psi, mpo, [gs], alg, nsweeps=10)
    assert abs(e1[-1] - evals[1]) < 1e-8
    ov = np.vdot(OD.to_dense(gs.to_host()), OD.to_dense(psi.to_host()))
    assert abs(ov) < 1e-6
