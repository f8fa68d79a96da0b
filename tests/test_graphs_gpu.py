"""Krylov recurrences captured as hipGraphs give the same numbers as the
eager launches: dense and block-sparse Lanczos eigensolves and the
sub-stepped exponentiate (whose step graph is launched once per sub-step),
each repeated with eager work of the program in between, with the
single-workgroup steps for short vectors and with the multi-block ones."""
import numpy as np
import pytest
import torch

from conftest import hermitian_env
from finitemptools_amd import models
from finitemptools_amd.engine import engine, to_device

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[True, False], ids=["small-steps", "multiblock-steps"])
def small(request):
    eng = engine(0)
    eng.set_krylov_small(request.param)
    yield request.param
    eng.set_krylov_small(True)


def both(fn):
    """fn() with graphs off, then on (twice), with unrelated eager work between;
    capture is forced (mode 2) so the short-vector steps are captured too."""
    eng = engine(0)
    out = []
    try:
        for on in (0, 2, 2):
            eng.set_graphs(on)
            a = torch.randn(4096, dtype=torch.float64, device="cuda")
            (a @ a).item()
            out.append(fn())
            torch.cuda.synchronize()
    finally:
        eng.set_graphs(True)
    return out


def test_dense_lanczos_graph_matches_eager(gpu, small):
    from finitemptools_amd.derivatives import HeffOperator
    from finitemptools_amd.krylov import simple_lanczos_solver
    rng = np.random.default_rng(3)
    D, w = 40, 5
    W = models.heisenberg_xxz_mpo(4)[1]
    L, R = to_device(hermitian_env(rng, D, w)), to_device(hermitian_env(rng, D, w))
    op = HeffOperator(2, L, R, to_device(W), to_device(W))
    for seed in range(3):
        x = to_device(np.random.default_rng(seed).standard_normal((D, 2, 2, D)))
        res = both(lambda: simple_lanczos_solver(op, x, "SR", 10, 1e-12))
        (e0, v0, _), (e1, v1, _), (e2, v2, _) = res
        assert e0 == e1 == e2
        assert torch.equal(v0, v1) and torch.equal(v0, v2)


def test_block_lanczos_graph_matches_eager(gpu, small):
    from finitemptools_amd import blockmps as bm
    from finitemptools_amd import blocksparse as bs
    phys = bm.hubbard_charges()
    mpo = bm.BlockMPO(models.hubbard_mpo(8, 1.0, 4.0), phys)
    psi = bm.randommps(8, phys, (4, 4), D=24, seed=2).right_canonicalize()
    env = bm.BlockExpectationCache(mpo, psi)
    for i in range(3):
        env.updateleft(i)
    x = bm.two_site(psi[3], psi[4])
    L, R = env.hstorage[3], env.hstorage[5]
    plan = bs.Plan(bs.plan_heff2(x.layout, L.layout, R.layout, mpo[3], mpo[4]))
    res = both(lambda: plan.lanczos(L.data, R.data, x.data, 10, 1e-12))
    (e0, v0, _, k0), (e1, v1, _, k1), (e2, v2, _, k2) = res
    assert e0 == e1 == e2 and k0 == k1 == k2
    assert torch.equal(v0, v1) and torch.equal(v0, v2)


def test_exponentiate_graph_matches_eager(gpu, small):
    """A tolerance that needs several sub-steps: the step graph is captured
    once and launched for each of them."""
    from finitemptools_amd.derivatives import HeffOperator
    from finitemptools_amd.krylov import exponentiate
    rng = np.random.default_rng(4)
    D, w = 24, 3
    W = models.tfim_mpo(4, 1.0, 1.0, dtype=np.complex128)[1]
    L = to_device(hermitian_env(rng, D, w, cplx=True, scale=1.0 / D))
    R = to_device(hermitian_env(rng, D, w, cplx=True, scale=1.0 / D))
    op = HeffOperator(2, L, R, to_device(W), to_device(W))
    x = to_device(rng.standard_normal((D, 2, 2, D)) + 1j * rng.standard_normal((D, 2, 2, D)))
    res = both(lambda: exponentiate(op, -1.0j, x, tol=1e-10, kdim=8))
    (y0, i0), (y1, i1), (y2, i2) = res
    assert i0 == i1 == i2 == 0
    assert torch.equal(y0, y1) and torch.equal(y0, y2)


def test_arnoldi_exponentiate_graph_matches_eager(gpu, small):
    """The Arnoldi steps (Hessenberg columns in device memory) captured as a
    graph give bitwise the eager result."""
    from finitemptools_amd.derivatives import HeffOperator
    from finitemptools_amd.krylov import exponentiate
    rng = np.random.default_rng(6)
    D, w = 20, 3
    W = models.tfim_mpo(4, 1.0, 1.0, dtype=np.complex128)[1]
    L = to_device(hermitian_env(rng, D, w, cplx=True, scale=1.0 / D))
    R = to_device(hermitian_env(rng, D, w, cplx=True, scale=1.0 / D))
    op = HeffOperator(2, L, R, to_device(W), to_device(W))
    x = to_device(rng.standard_normal((D, 2, 2, D)) + 1j * rng.standard_normal((D, 2, 2, D)))
    res = both(lambda: exponentiate(op, -1.0j, x, tol=1e-10, kdim=8, ishermitian=False))
    (y0, i0), (y1, i1), (y2, i2) = res
    assert i0 == i1 == i2 == 0
    assert torch.equal(y0, y1) and torch.equal(y0, y2)
