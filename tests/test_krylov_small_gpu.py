"""The single-workgroup Krylov steps used for short vectors (at most 16384
doubles) against the multi-block kernels on the same operators: the same
recurrence with a different summation order, so eigenvalues, Ritz vectors
and exponentials agree to rounding.  Both paths are also pinned to the
oracle by the other Krylov tests (short vectors take the single-workgroup
path there, long ones the multi-block path)."""
import numpy as np
import pytest
import torch

from conftest import hermitian_env
from finitemptools_amd import models
from finitemptools_amd.engine import engine, to_device

pytestmark = pytest.mark.gpu


def paths(fn):
    eng = engine(0)
    out = []
    try:
        for on in (True, False):
            eng.set_krylov_small(on)
            out.append(fn())
            torch.cuda.synchronize()
    finally:
        eng.set_krylov_small(True)
    return out


def rel(a, b):
    return (torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b)).item()


def _op(rng, D, w, cplx, W):
    from finitemptools_amd.derivatives import HeffOperator
    L = to_device(hermitian_env(rng, D, w, cplx=cplx, scale=1.0 / D))
    R = to_device(hermitian_env(rng, D, w, cplx=cplx, scale=1.0 / D))
    return HeffOperator(2, L, R, to_device(W), to_device(W))


@pytest.mark.parametrize("reorth", [False, True])
@pytest.mark.parametrize("D", [6, 40, 60])
def test_lanczos_eig_small_steps(gpu, reorth, D):
    from finitemptools_amd.krylov import simple_lanczos_solver
    rng = np.random.default_rng(D)
    W = models.heisenberg_xxz_mpo(4)[1]
    op = _op(rng, D, 5, False, W)
    x = to_device(rng.standard_normal((D, 2, 2, D)))
    (e0, v0, k0), (e1, v1, k1) = paths(lambda: simple_lanczos_solver(op, x, "SR", 12, 1e-12,
                                                                     reorth=reorth))
    assert abs(e0 - e1) <= 1e-12 * abs(e1)
    assert rel(v0, v1) < 1e-9


@pytest.mark.parametrize("herm", [True, False])
@pytest.mark.parametrize("D", [8, 30])
def test_exponentiate_small_steps(gpu, herm, D):
    from finitemptools_amd.krylov import exponentiate
    rng = np.random.default_rng(D + 1)
    W = models.tfim_mpo(4, 1.0, 1.0, dtype=np.complex128)[1]
    op = _op(rng, D, 3, True, W)
    x = to_device(rng.standard_normal((D, 2, 2, D)) + 1j * rng.standard_normal((D, 2, 2, D)))
    (y0, i0), (y1, i1) = paths(lambda: exponentiate(op, -0.3j, x, tol=1e-12, kdim=12,
                                                    ishermitian=herm))
    assert i0 == i1 == 0
    assert rel(y0, y1) < 1e-11


def test_projected_lanczos_small_steps(gpu):
    from finitemptools_amd.krylov import simple_lanczos_solver
    rng = np.random.default_rng(7)
    D = 20
    W = models.tfim_mpo(4, 1.0, 0.7)[1]
    op = _op(rng, D, 3, False, W)
    p = rng.standard_normal((D, 2, 2, D))
    p = to_device(p / np.linalg.norm(p))
    x = to_device(rng.standard_normal((D, 2, 2, D)))
    (e0, v0, _), (e1, v1, _) = paths(lambda: simple_lanczos_solver(op, x, "SR", 12, 1e-12,
                                                                   projectors=[p]))
    assert abs(e0 - e1) <= 1e-12 * abs(e1)
    assert rel(v0, v1) < 1e-9
    assert abs(torch.vdot(p.flatten(), v0.flatten()).item()) < 1e-10


def test_block_lanczos_small_steps(gpu):
    from finitemptools_amd import blockmps as bm
    from finitemptools_amd import blocksparse as bs
    phys = bm.hubbard_charges()
    mpo = bm.BlockMPO(models.hubbard_mpo(8, 1.0, 4.0), phys)
    psi = bm.randommps(8, phys, (4, 4), D=24, seed=5).right_canonicalize()
    env = bm.BlockExpectationCache(mpo, psi)
    for i in range(3):
        env.updateleft(i)
    x = bm.two_site(psi[3], psi[4])
    L, R = env.hstorage[3], env.hstorage[5]
    plan = bs.make_plan("heff2", x.layout, L.layout, R.layout, mpo[3], mpo[4], "f64")
    (e0, v0, _, k0), (e1, v1, _, k1) = paths(lambda: plan.lanczos(L.data, R.data, x.data, 10,
                                                                  1e-12))
    assert k0 == k1
    assert abs(e0 - e1) <= 1e-12 * abs(e1)
    assert rel(v0, v1) < 1e-9
