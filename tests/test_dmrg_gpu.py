"""Two-site DMRG on the device against the oracle's sweep and exact
diagonalisation ("ground state dmrg algorithms", test/runtests.jl:773-778,
tolerance 1e-8 on the energy)."""
import json
import os

import numpy as np
import pytest

from finitemptools_amd import models
from oracle import dmrg as OD

pytestmark = pytest.mark.gpu

ED = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ed_energies.json")))


def run_gpu_dmrg(mpo_host, L, d, D, nsweeps, seed=0, prec="f64", Dmax=None, eps=1e-10):
    from finitemptools_amd import dmrg
    from finitemptools_amd.envs import mpo_to_device
    from finitemptools_amd.mps import randommps
    from finitemptools_amd.truncation import MPSTruncation
    psi = randommps(L, d, D, seed=seed).right_canonicalize()
    alg = dmrg.DMRG2(trunc=MPSTruncation(D=Dmax or D, eps=eps), prec=prec)
    energies, delta = dmrg.ground_state(psi, mpo_to_device(mpo_host), alg, nsweeps=nsweeps)
    return energies, delta, psi, alg


def test_sweep_trajectory_matches_oracle(gpu):
    """Same seeded start state, same sweep order: every per-bond Lanczos energy
    of two sweeps agrees with the numpy restatement."""
    L, D = 10, 12
    mpo = models.heisenberg_xxz_mpo(L, J=1.0, Jzz=0.8, hz=0.05)
    e_gpu, _, _, _ = run_gpu_dmrg(mpo, L, 2, D, nsweeps=2)
    psi = OD.right_canonicalize(OD.randommps(L, 2, D, seed=0))
    env = OD.ExpectationCache(mpo, psi)
    alg = OD.DMRG2(trunc=OD.MPSTruncation(D=D, eps=1e-10))
    e_ref, _ = OD.compute(env, alg, nsweeps=2)
    assert len(e_gpu) == len(e_ref)
    assert np.max(np.abs(np.array(e_gpu) - np.array(e_ref))) < 1e-9


@pytest.mark.parametrize("name,builder,L,d,D", [
    ("xxz_L12_J1_D1", lambda L: models.heisenberg_xxz_mpo(L), 12, 2, 32),
    ("tfim_L12_J1_h0.7", lambda L: models.tfim_mpo(L, 1.0, 0.7), 12, 2, 24),
    ("j1j2_L12_J2_0.25", lambda L: models.j1j2_mpo(L, 1.0, 0.25), 12, 2, 32),
    ("hubbard_L6_t1_U4", lambda L: models.hubbard_mpo(L, 1.0, 4.0), 6, 4, 48),
])
def test_dmrg2_reaches_ed(gpu, name, builder, L, d, D):
    energies, delta, _, _ = run_gpu_dmrg(builder(L), L, d, D, nsweeps=5)
    assert abs(energies[-1] - ED[name]) < 1e-8


def test_majumdar_ghosh_dimer_energy(gpu):
    """J2 = J1/2 on an even open chain: exact dimer ground state, E = -3/4 J1 L/2."""
    L = 12
    energies, _, _, _ = run_gpu_dmrg(models.j1j2_mpo(L, 1.0, 0.5), L, 2, 16, nsweeps=5)
    assert abs(energies[-1] - (-0.75 * L / 2)) < 1e-8


@pytest.mark.parametrize("algname", ["DMRG1", "DMRG1S"])
def test_one_site_dmrg_reaches_ed(gpu, algname):
    """DMRG1 (zero-padded bond dimension) and DMRG1S (optimal subspace
    expansion) against exact diagonalisation, as the reference tests DMRG1S."""
    from finitemptools_amd import dmrg
    from finitemptools_amd.envs import mpo_to_device
    from finitemptools_amd.mps import randommps
    from finitemptools_amd.truncation import MPSTruncation
    L = 12
    psi = randommps(L, 2, 4, seed=2).right_canonicalize()
    if algname == "DMRG1":
        alg = dmrg.DMRG1(D=32)
    else:
        alg = dmrg.DMRG1S(trunc=MPSTruncation(D=32, eps=1e-10))
    energies, delta = dmrg.ground_state(psi, mpo_to_device(models.heisenberg_xxz_mpo(L)), alg, nsweeps=8)
    assert abs(energies[-1] - ED["xxz_L12_J1_D1"]) < 1e-8
    assert delta < 1e-4  # Galerkin error of the converged state


def test_complex_canonical_form_preserves_state(gpu):
    """Right-canonicalising a complex MPS on the device leaves the state
    unchanged and sites 2..L right-isometric (exercises conjugated factors)."""
    from finitemptools_amd.mps import randommps
    L = 6
    psi = randommps(L, 2, 4, seed=9, cplx=True)
    before = OD.to_dense(psi.to_host())
    psi.right_canonicalize(normalize=False)
    after = OD.to_dense(psi.to_host())
    assert np.allclose(after, before, rtol=1e-12, atol=1e-12 * np.abs(before).max())
    for A in psi.to_host()[1:]:
        M = A.reshape(A.shape[0], -1, order="F")
        assert np.allclose(M @ M.conj().T, np.eye(A.shape[0]), atol=1e-12)


def test_xxz_L20_D64_four_sweeps(gpu):
    """The reference CPU run configuration: XXZ (J = Delta = 1), L = 20, D = 64,
    four sweeps from a seeded random right-canonical MPS."""
    energies, delta, psi, alg = run_gpu_dmrg(models.heisenberg_xxz_mpo(20), 20, 2, 64, nsweeps=4,
                                             eps=1e-8)
    assert abs(energies[-1] - ED["xxz_L20_J1_D1"]) < 1e-8
    assert psi.bond_dimension() <= 64
