"""One-launch small Hermitian eigensolver (mh_heev_small) against LAPACK
(numpy.linalg.eigh) and the library path of the dense two-site split."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _herm(n, cplx, rng, spectrum=None):
    if cplx:
        X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    else:
        X = rng.standard_normal((n, n))
    if spectrum is None:
        return (X + X.conj().T) / 2
    Q, _ = np.linalg.qr(X)
    return (Q * np.asarray(spectrum)) @ Q.conj().T


def _solve(G, col0=0):
    from finitemptools_amd import eigen
    Gt = torch.as_tensor(np.ascontiguousarray(G.T), device=DEV)   # G column-major
    lam, V, info = eigen.small_eigh(Gt, col0)
    torch.cuda.synchronize()
    return lam.cpu().numpy(), V.cpu().numpy(), int(info.cpu().item())


def _check(G, lam, V, col0=0):
    n = G.shape[0]
    scale = max(np.abs(np.linalg.eigvalsh(G)).max(), 1e-300)
    ref = np.linalg.eigvalsh(G)
    assert np.all(np.diff(lam) >= 0)
    np.testing.assert_allclose(lam, ref, rtol=0, atol=4 * n * 2.2e-16 * scale)
    Vk = V[:, col0:]
    res = np.linalg.norm(G @ Vk - Vk * lam[col0:], axis=0).max()
    assert res <= 8 * n * 2.2e-16 * scale
    ortho = np.abs(Vk.conj().T @ Vk - np.eye(n - col0)).max()
    assert ortho <= 8 * n * 2.2e-16


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("n", [1, 2, 3, 17, 48, 64])
def test_random_hermitian(n, cplx):
    print(f"heev_small random n={n} cplx={cplx}", flush=True)
    rng = np.random.default_rng(100 + n)
    G = _herm(n, cplx, rng)
    lam, V, info = _solve(G)
    assert info == 0
    _check(G, lam, V)


@pytest.mark.parametrize("cplx", [False, True])
def test_largest_order(cplx):
    from finitemptools_amd import eigen
    dt = torch.complex128 if cplx else torch.float64
    nmax = eigen.small_eigh_max(dt, DEV)
    print(f"heev_small largest order {nmax} cplx={cplx}", flush=True)
    assert nmax >= (48 if cplx else 84)
    rng = np.random.default_rng(7)
    G = _herm(nmax, cplx, rng)
    lam, V, info = _solve(G)
    assert info == 0
    _check(G, lam, V)
    with pytest.raises(ValueError):
        eigen.small_eigh(torch.zeros((nmax + 1, nmax + 1), dtype=dt, device=DEV))


@pytest.mark.parametrize("cplx", [False, True])
def test_degenerate_and_rank_deficient(cplx):
    print(f"heev_small degenerate cplx={cplx}", flush=True)
    rng = np.random.default_rng(3)
    n = 40
    spec = [0.0] * 18 + [1e-9] * 2 + [0.3] * 5 + [1.0] * 12 + [2.0, 2.0, 5.0]
    G = _herm(n, cplx, rng, spec)
    lam, V, info = _solve(G)
    assert info == 0
    _check(G, lam, V)
    # a Gram matrix of rank 7: the kept (top) columns only
    M = rng.standard_normal((n, 7)) + (1j * rng.standard_normal((n, 7)) if cplx else 0)
    G = M @ M.conj().T
    lam, V, info = _solve(G, col0=n - 7)
    assert info == 0
    _check(G, lam, V, col0=n - 7)


@pytest.mark.parametrize("cplx", [False, True])
def test_diagonal_zero_and_selected_columns(cplx):
    print(f"heev_small diagonal/zero/col0 cplx={cplx}", flush=True)
    G = np.diag(np.array([3.0, -1.0, 2.0, 0.5, 7.0])).astype(complex if cplx else float)
    lam, V, info = _solve(G)
    assert info == 0
    _check(G, lam, V)
    Z = np.zeros((6, 6), dtype=complex if cplx else float)
    lam, V, info = _solve(Z)
    assert info == 0 and np.all(lam == 0.0)
    np.testing.assert_array_equal(np.abs(V), np.eye(6))
    rng = np.random.default_rng(11)
    G = _herm(33, cplx, rng)
    lam0, V0, _ = _solve(G)
    lam1, V1, _ = _solve(G, col0=20)
    np.testing.assert_array_equal(lam0, lam1)
    np.testing.assert_array_equal(V0[:, 20:], V1[:, 20:])


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("absorb", ["left", "right"])
def test_split_rho_matches_library_path(cplx, absorb):
    """The dense two-site split through the one-launch solver equals the
    library eigh split: same kept values, same truncated product."""
    from finitemptools_amd import truncation
    print(f"split_rho small vs library cplx={cplx} absorb={absorb}", flush=True)
    g = torch.Generator().manual_seed(5)
    dt = torch.complex128 if cplx else torch.float64
    x = torch.randn((12, 2, 2, 9), generator=g, dtype=dt).to(DEV)
    tr = truncation.MPSTruncation(D=20, eps=1e-10)
    out = {}
    for flag in (True, False):
        truncation.SMALL_EIGH = flag
        truncation.COMPLEX_SVD_SPLIT = False     # the density-matrix path itself
        try:
            A1, A2, s = truncation.split_rho(x, tr, absorb)
        finally:
            truncation.SMALL_EIGH = True
            truncation.COMPLEX_SVD_SPLIT = True
        P = torch.einsum("asc,ctb->astb", A1.cpu(), A2.cpu())
        out[flag] = (P, s.cpu())
    np.testing.assert_allclose(out[True][1].numpy(), out[False][1].numpy(), rtol=0, atol=1e-12)
    np.testing.assert_allclose(out[True][0].numpy(), out[False][0].numpy(), rtol=0, atol=1e-12)


@pytest.mark.parametrize("absorb", ["left", "right"])
@pytest.mark.parametrize("D", [8, 24])
def test_split_rho_batched_solver_matches_library_path(absorb, D):
    """Real density matrices above the one-launch solver's order go to the
    batched eigensolver (SectorEigh, kept columns back-transformed only):
    the split equals the library eigh split (kept values, truncated product)
    where D cuts a graded spectrum.  (Kept vectors whose eigenvalues sit near
    the noise floor are fixed only to eps64 lambda_max / gap by either
    solver, so the cut is placed where the gaps are wide: there both agree
    to rounding.)"""
    from finitemptools_amd import truncation
    print(f"split_rho batched vs library absorb={absorb} D={D}", flush=True)
    rng = np.random.default_rng(21)
    Dl, d, Dr = 20, 2, 24                      # orders 40 (left) and 48 (right)
    r = min(Dl * d, d * Dr)
    Q1 = np.linalg.qr(rng.standard_normal((Dl * d, r)))[0]
    Q2 = np.linalg.qr(rng.standard_normal((d * Dr, r)))[0]
    M = (Q1 * np.logspace(0, -6, r)) @ Q2.T
    x = torch.as_tensor(M.reshape(Dl, d, d, Dr), device=DEV)
    tr = truncation.MPSTruncation(D=D, eps=1e-10)
    out = {}
    for flag in (True, False):
        truncation.SECTOR_EIGH = flag
        try:
            A1, A2, s = truncation.split_rho(x, tr, absorb)
        finally:
            truncation.SECTOR_EIGH = True
        P = torch.einsum("asc,ctb->astb", A1.cpu(), A2.cpu())
        out[flag] = (P, s.cpu())
    assert len(out[True][1]) == len(out[False][1])
    np.testing.assert_allclose(out[True][1].numpy(), out[False][1].numpy(), rtol=0, atol=1e-12)
    np.testing.assert_allclose(out[True][0].numpy(), out[False][0].numpy(), rtol=0, atol=1e-12)


def test_graded_rank_deficient_gram():
    """Gram matrices of rank r < n with singular values spread over up to 12
    decades (the reduced density matrices of a converged state): every solve
    converges, at the normwise accuracy of LAPACK."""
    print("heev_small graded rank-deficient Gram matrices", flush=True)
    from finitemptools_amd import eigen
    rng = np.random.default_rng(1)
    nmax = min(128, eigen.small_eigh_max(torch.float64, DEV))
    for trial in range(12):
        n = nmax
        r = int(rng.integers(n // 6, n))
        U, _ = np.linalg.qr(rng.standard_normal((n, r)))
        s = 10.0 ** (-rng.uniform(0, 12) * np.linspace(0, 1, r))
        M = (U * s) @ rng.standard_normal((r, n)) / np.sqrt(n)
        G = M @ M.T
        lam, V, info = _solve(G)
        assert info == 0, trial
        nrm = np.linalg.norm(G, 2)
        assert np.abs(lam - np.linalg.eigvalsh(G)).max() <= 1e-13 * nrm
        assert np.abs(G @ V - V * lam).max() <= 1e-13 * nrm
        assert np.abs(V.T @ V - np.eye(n)).max() <= 1e-13
