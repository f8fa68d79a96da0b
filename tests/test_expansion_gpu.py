"""CH subspace expansion (expansion/chexpand.jl) and one-site TDVP with
expansion (TDVP1S, tdvp.jl:145-213) on the device: the open-bond H_eff term
against the oracle (1e-12), the expansion's invariants (state unchanged when
nothing is truncated, isometric site tensor), DMRG1S with CH expansion against
exact diagonalisation (1e-8) and TDVP1S against exact evolution (1e-4, the
reference's TDVP tolerance, test/runtests.jl:550-594)."""
import json
import os

import numpy as np
import pytest

from conftest import rand
from finitemptools_amd import models
from oracle import contractions as C
from oracle import ed
from oracle import dmrg as OD

pytestmark = pytest.mark.gpu

ED = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ed_energies.json")))


def relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("Dl,Dr,Wk", [(34, 20, "xxz"), (7, 9, "xxz"), (12, 12, "hubbard")])
def test_ac_prime_open_matches_oracle(gpu, cplx, Dl, Dr, Wk):
    """Open right bond of ac_prime (the fused kernel at even Dl for real
    data, GEMM + mode product otherwise) and the bond-transposed form the
    right expansion uses."""
    import torch
    from finitemptools_amd.derivatives import ac_prime_open
    from finitemptools_amd.engine import to_device
    from finitemptools_amd.expansion import _merge_last
    rng = np.random.default_rng(Dl + 3 * Dr)
    W = models.heisenberg_xxz_mpo(4)[1] if Wk == "xxz" else models.hubbard_mpo(4)[1]
    if cplx:
        W = W.astype(np.complex128)
    wl, d, wr, _ = W.shape
    A = rand(rng, (Dl, d, Dr), cplx)
    L = rand(rng, (Dl, wl, Dl), cplx)
    R = rand(rng, (Dr, wr, Dr), cplx)
    dt = torch.complex128 if cplx else torch.float64
    P = _merge_last(ac_prime_open(to_device(A, "cuda", dt), to_device(W, "cuda", dt),
                                  to_device(L, "cuda", dt))).cpu().numpy()
    assert relerr(P, C.ch_expansion_term_left(L, A, W, 1.0)) < 1e-12
    Pt = ac_prime_open(to_device(A, "cuda", dt).permute(2, 1, 0),
                       to_device(W, "cuda", dt).permute(2, 1, 0, 3), to_device(R, "cuda", dt))
    Q = _merge_last(Pt).permute(2, 1, 0).cpu().numpy()
    assert relerr(Q, C.ch_expansion_term_right(A, W, R, 1.0)) < 1e-12


@pytest.mark.parametrize("side", ["left", "right"])
def test_ch_expansion_keeps_the_state(gpu, side):
    """With nothing truncated the expansion changes the bond, not the state;
    the site tensor it leaves behind is an isometry and the new bond is the
    rank of [alpha L.A.W | A] (resp. [alpha W.A.R; A])."""
    import torch
    from finitemptools_amd.envs import environments, mpo_to_device
    from finitemptools_amd.expansion import CHExpansion, left_expansion, right_expansion
    from finitemptools_amd.mps import randommps
    from finitemptools_amd.truncation import MPSTruncation
    L = 8
    psi = randommps(L, 2, 4, seed=5).right_canonicalize()
    mpo = mpo_to_device(models.heisenberg_xxz_mpo(L))
    m = environments(mpo, psi)
    site = 3
    before = OD.to_dense(psi.to_host())
    host = psi.to_host()
    Wh = models.heisenberg_xxz_mpo(L)
    if side == "left":
        # the left expansion runs in the right sweep: environments left of
        # site-1 are current; rebuild them so.
        m.init_hstorage(center=site)
        Lh = m.hstorage[site - 1].cpu().numpy()
        term = C.ch_expansion_term_left(Lh, host[site - 1], Wh[site - 1], 0.05)
        mat = np.concatenate([term, host[site - 1]], axis=2)
        rank = np.linalg.matrix_rank(mat.reshape(-1, mat.shape[2], order="F"))
        left_expansion(m, site, MPSTruncation(D=1000, eps=0.0), expan=CHExpansion(eps=0.05))
        iso = psi[site - 1].cpu().numpy()
        M = iso.reshape(-1, iso.shape[2], order="F")
        assert np.allclose(M.T @ M, np.eye(M.shape[1]), atol=1e-12)
        assert iso.shape[2] <= rank
    else:
        m.init_hstorage(center=site)
        Rh = m.hstorage[site + 2].cpu().numpy()
        term = C.ch_expansion_term_right(host[site + 1], Wh[site + 1], Rh, 0.05)
        mat = np.concatenate([term, host[site + 1]], axis=0)
        rank = np.linalg.matrix_rank(mat.reshape(mat.shape[0], -1, order="F"))
        right_expansion(m, site, MPSTruncation(D=1000, eps=0.0), expan=CHExpansion(eps=0.05))
        iso = psi[site + 1].cpu().numpy()
        M = iso.reshape(iso.shape[0], -1, order="F")
        assert np.allclose(M @ M.T, np.eye(M.shape[0]), atol=1e-12)
        assert iso.shape[0] <= rank
    after = OD.to_dense(psi.to_host())
    assert np.allclose(after, before, atol=1e-12 * np.abs(before).max())


def test_dmrg1s_with_ch_expansion_reaches_ed(gpu):
    from finitemptools_amd import dmrg
    from finitemptools_amd.envs import mpo_to_device
    from finitemptools_amd.expansion import CHExpansion
    from finitemptools_amd.mps import randommps
    from finitemptools_amd.truncation import MPSTruncation
    L = 12
    psi = randommps(L, 2, 4, seed=2).right_canonicalize()
    alg = dmrg.DMRG1S(trunc=MPSTruncation(D=64, eps=1e-10), expan=CHExpansion(eps=0.01))
    energies, delta = dmrg.ground_state(psi, mpo_to_device(models.heisenberg_xxz_mpo(L)), alg,
                                        nsweeps=10)
    assert abs(energies[-1] - ED["xxz_L12_J1_D1"]) < 1e-8


def exact_sz(L, h, t):
    import scipy.sparse.linalg as spla
    H = ed.tfim_hamiltonian(L, 1.0, h)
    psi0 = np.full(2 ** L, 2 ** (-L / 2), dtype=np.complex128)
    psi = spla.expm_multiply(-1j * t * H, psi0)
    return np.array([np.vdot(psi, ed._site(ed.SZ, i, L, 2) @ psi).real for i in range(L)])


@pytest.mark.parametrize("expan", ["optimal", "ch"])
def test_tdvp1s_quench_matches_exact(gpu, expan):
    """TDVP1S from the product state |+>^L: the expansion grows the bond
    before every forward step (a pure one-site TDVP could never leave the
    D = 1 manifold)."""
    import torch
    from finitemptools_amd import tdvp
    from finitemptools_amd.envs import environments, local_expectations, mpo_to_device
    from finitemptools_amd.expansion import CHExpansion
    from finitemptools_amd.mps import prodmps
    from finitemptools_amd.truncation import MPSTruncation
    L, h, dt, nsteps = 8, 1.0, 0.05, 8
    psi = prodmps([np.ones(2) / np.sqrt(2)] * L, dtype=torch.complex128)
    mpo = mpo_to_device(models.tfim_mpo(L, 1.0, h, dtype=np.complex128), torch.complex128)
    ex = CHExpansion(eps=0.01) if expan == "ch" else None
    alg = tdvp.TDVP1S(stepsize=-1j * dt, trunc=MPSTruncation(D=16, eps=1e-12), expan=ex)
    env = environments(mpo, psi)
    for _ in range(nsteps):
        tdvp.sweep(env, alg)
    sz = local_expectations(psi, ed.SZ)
    assert max(t.shape[2] for t in psi.data) > 1
    assert np.max(np.abs(sz - exact_sz(L, h, dt * nsteps))) < 1e-4
