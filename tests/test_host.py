"""Host-side logic of the engine package (no GPU): seeded initial states,
truncation scheme and MPO shapes agree with the oracle."""
import numpy as np
import pytest

from finitemptools_amd import models
from finitemptools_amd.mps import random_site_tensors, virtual_dims
from finitemptools_amd.truncation import MPSTruncation
from oracle import dmrg as OD


def test_random_mps_matches_oracle_stream():
    a = random_site_tensors(9, 2, 8, seed=3)
    b = OD.randommps(9, 2, 8, seed=3)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert virtual_dims(9, 2, 8) == OD.virtual_dims(9, 2, 8)


def test_truncation_matches_oracle():
    rng = np.random.default_rng(0)
    for _ in range(20):
        s = np.sort(np.abs(rng.standard_normal(30)) ** 6)[::-1]
        for D, eps in [(5, 1e-8), (40, 1e-3), (10, 1e-12)]:
            assert MPSTruncation(D, eps).truncate(s) == OD.MPSTruncation(D, eps).truncate(s)


def test_mpo_strict_boundaries():
    for mpo in (models.heisenberg_xxz_mpo(5), models.hubbard_mpo(4), models.j1j2_mpo(6, 1.0, 0.3)):
        assert mpo[0].shape[0] == 1 and mpo[-1].shape[2] == 1
        for a, b in zip(mpo[:-1], mpo[1:]):
            assert a.shape[2] == b.shape[0]


def test_conjugate_views_are_materialised_before_the_engine():
    """torch's conj() is a lazy view; .contiguous() keeps the lazy bit on an
    already-contiguous view, so raw pointers would see unconjugated values."""
    import torch
    from finitemptools_amd.engine import as_colmajor
    from finitemptools_amd.linalg import from_matrix
    r = torch.tensor([[1 + 2j, 3 + 4j], [5 + 6j, 7 + 8j]], dtype=torch.complex128)
    for t in (as_colmajor(r.conj()), from_matrix(r.conj().t(), (2, 2))):
        assert not t.is_conj()
        raw = torch.view_as_real(t)  # would raise on a conj view
        assert torch.equal(torch.view_as_complex(raw), t)
    t = from_matrix(r.conj().t(), (2, 2))
    assert torch.equal(t, r.conj().t())


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("absorb", ["left", "right"])
def test_density_matrix_split_equals_truncated_svd(cplx, absorb):
    """split_rho keeps the singular values of stable_tsvd (normalised) and the
    same truncated two-site tensor u s v; the non-absorbing factor is an
    isometry."""
    import torch
    from finitemptools_amd.linalg import as_matrix
    from finitemptools_amd.truncation import MPSTruncation, split_rho, stable_tsvd
    rng = np.random.default_rng(7)
    Dl, d, Dr = 6, 2, 5
    x = rng.standard_normal((Dl, d, d, Dr))
    if cplx:
        x = x + 1j * rng.standard_normal((Dl, d, d, Dr))
    # a decaying spectrum, as a DMRG two-site tensor has
    M = x.reshape(Dl * d, d * Dr, order="F")
    U, S, Vh = np.linalg.svd(M, full_matrices=False)
    M = (U * (0.5 ** np.arange(len(S)))) @ Vh
    xt = torch.as_tensor(M.reshape(Dl, d, d, Dr, order="F"))
    tr = MPSTruncation(D=7, eps=1e-10)
    A1, A2, s = split_rho(xt, tr, absorb)
    u, s_ref, v, _ = stable_tsvd(xt, tr)
    s_ref = s_ref / torch.linalg.vector_norm(s_ref)
    assert np.allclose(s.numpy(), s_ref.numpy(), atol=1e-12)
    rec = (as_matrix(A1, 2) @ as_matrix(A2, 1)).numpy()
    ref = (as_matrix(u, 2) @ torch.diag(s_ref.to(u.dtype)) @ as_matrix(v, 1)).numpy()
    assert np.allclose(rec, ref, atol=1e-12)
    iso = as_matrix(A1, 2) if absorb == "right" else as_matrix(A2, 1).conj().t()
    assert np.allclose((iso.conj().t() @ iso).numpy(), np.eye(iso.shape[1]), atol=1e-12)


def test_sector_eigh_rejects_orders_beyond_the_kernel():
    """Orders outside [1, 2048] are refused on the host, before any device work."""
    import pytest
    from finitemptools_amd import _lib, eigen
    with pytest.raises(_lib.DimensionMismatch):
        eigen.SectorEigh([4, 4096], "cpu")
    with pytest.raises(_lib.DimensionMismatch):
        eigen.SectorEigh([0], "cpu")


def _terms_from_cell(cell):
    """Schur entries of a dense site tensor: identity blocks as scalar
    entries, other nonzero blocks as operator entries."""
    wl, d, wr, _ = cell.shape
    terms = []
    for i in range(wl):
        for j in range(wr):
            B = cell[i, :, j, :]
            if not np.any(B):
                continue
            c = B[0, 0]
            if np.allclose(B, c * np.eye(d), atol=0, rtol=0):
                terms.append((i, j, c, None))
            else:
                terms.append((i, j, 1.0, B))
    return terms


def test_schur_fold_rebuilds_the_model_site_tensors():
    """mh_schur_fold (host only) assembles W[wl, s_out, wr, s_in] from
    (i, j, coef, operator | identity) entries: the bulk site tensors of the
    XXZ (with a field), Hubbard and complex TFIM Schur MPOs come back entry for
    entry; a hand-made case pins the scalar/identity and coefficient rules."""
    from finitemptools_amd import models
    for mpo in (models.heisenberg_xxz_mpo(4, 1.0, 0.7, 0.3), models.hubbard_mpo(4, 1.0, 4.0),
                models.tfim_mpo(4, 1.0, 0.6, dtype=np.complex128)):
        cell = mpo[1]
        wl, d, wr, _ = cell.shape
        W = models.schur_fold(_terms_from_cell(cell), wl, wr, d, dtype=cell.dtype)
        assert W.dtype == cell.dtype and np.array_equal(W, cell)
    O = np.array([[0.0, 2.0], [3.0, 0.0]])
    W = models.schur_fold([(0, 0, 1.0, None), (0, 1, 0.5, O), (1, 1, -2.0, None),
                           (0, 1, 1.0, None)], 2, 2, 2)
    ref = np.zeros((2, 2, 2, 2))
    ref[0, :, 0, :] = np.eye(2)
    ref[0, :, 1, :] = 0.5 * O + np.eye(2)
    ref[1, :, 1, :] = -2.0 * np.eye(2)
    assert np.array_equal(W, ref)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus N > 1 exits with status 2 and says
    why, before touching a device."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=root)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def _graded_matrix(rng, R, C, s, cplx=False):
    """R x C matrix with singular values s (random orthonormal factors)."""
    def orth(n, k):
        A = rng.standard_normal((n, k))
        if cplx:
            A = A + 1j * rng.standard_normal((n, k))
        return np.linalg.qr(A)[0]
    k = len(s)
    return (orth(R, k) * s) @ orth(C, k).conj().T


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("absorb", ["left", "right"])
@pytest.mark.parametrize("eps", [1e-8, 1e-10])
def test_density_matrix_split_at_a_cutoff_inside_its_noise(cplx, absorb, eps):
    """A 24 x 24 two-site matrix whose spectrum runs from 1 down to 1e-12
    straddles the relative cutoff: sqrt(lambda) from rho resolves s only down
    to about sqrt(n eps64) ~ 7e-8, so a kept count taken from the density
    matrix depends on rounding for this spectrum.  split_rho must return
    the SVD split's kept count, singular values and truncation error
    (truncation.jl:27-47 on stable_tsvd)."""
    import torch
    from finitemptools_amd.linalg import as_matrix
    from finitemptools_amd.truncation import MPSTruncation, split_rho, stable_tsvd
    rng = np.random.default_rng(31)
    Dl, d, Dr = 12, 2, 12
    s_true = np.logspace(0, -12, 24)
    M = _graded_matrix(rng, Dl * d, d * Dr, s_true, cplx)
    xt = torch.as_tensor(M.reshape(Dl, d, d, Dr, order="F"))
    tr = MPSTruncation(D=100, eps=eps)
    # what the check guards against: around the cut sqrt(eig(M M^dag)) is
    # noise, so a count taken from it is a matter of rounding
    lam = np.linalg.eigvalsh(M @ M.conj().T)[::-1]
    s_rho = np.sqrt(np.maximum(lam, 0))
    k_svd, err_svd = tr.truncate(np.linalg.svd(M, compute_uv=False))
    near = np.abs(np.log10(s_true) - np.log10(eps)) < 1.0
    assert np.max(np.abs(s_rho[near] - s_true[near]) / s_true[near]) > 0.1
    A1, A2, s = split_rho(xt, tr, absorb)
    u, s_ref, v, err = stable_tsvd(xt, tr)
    assert len(s) == len(s_ref) == k_svd
    # an SVD fixes singular values to a few eps64 * ||M|| absolute (||M|| ~ 1)
    assert np.isclose(err, err_svd, rtol=1e-6, atol=50 * np.finfo(np.float64).eps)
    s_ref = s_ref / torch.linalg.vector_norm(s_ref)
    assert np.allclose(s.numpy(), s_ref.numpy(), rtol=1e-10, atol=50 * np.finfo(np.float64).eps)
    rec = (as_matrix(A1, 2) @ as_matrix(A2, 1)).numpy()
    ref = (as_matrix(u, 2) @ torch.diag(s_ref.to(u.dtype)) @ as_matrix(v, 1)).numpy()
    assert np.allclose(rec, ref, atol=1e-12)


@pytest.mark.parametrize("absorb", ["left", "right"])
def test_density_matrix_split_where_D_binds_keeps_accurate_values(absorb):
    """D binds well above the noise (the middle-bond case of a sweep): the
    density-matrix path is taken, and its singular values, read off the
    projected factor, match the SVD's to the SVD's own accuracy (a few eps64
    of the largest, absolute) even four decades below the largest, where
    sqrt(lambda) alone is off by ~4e-11 absolute."""
    import torch
    from finitemptools_amd import truncation as T
    rng = np.random.default_rng(32)
    Dl, d, Dr = 20, 2, 20
    s_true = np.logspace(0, -12, 40)
    M = _graded_matrix(rng, Dl * d, d * Dr, s_true)
    xt = torch.as_tensor(M.reshape(Dl, d, d, Dr, order="F"))
    tr = T.MPSTruncation(D=14, eps=1e-8)
    lam = np.maximum(np.linalg.eigvalsh(M @ M.T)[::-1], 0)
    assert T.rho_cut_resolved(lam, T.RHO_NOISE * 40 * T.EPS64 * lam[0], tr)
    A1, A2, s = T.split_rho(xt, tr, absorb)
    s_svd = np.linalg.svd(M, compute_uv=False)[:14]
    s_svd = s_svd / np.linalg.norm(s_svd)
    assert np.allclose(s.numpy(), s_svd, rtol=1e-13, atol=50 * T.EPS64 * s_svd[0])
    # the values sqrt(lambda) would give miss that bar
    assert np.max(np.abs(np.sqrt(lam[:14]) / np.linalg.norm(np.sqrt(lam[:14])) - s_svd)) > 1e-13


def test_rho_cut_resolved_rules():
    """The cutoff band: a relative cut inside the noise is unresolved, one far
    above it resolved; with D binding only the D-th value meets the cutoff."""
    from finitemptools_amd.truncation import MPSTruncation, rho_cut_resolved
    s2 = np.array([1.0, 1e-2, 1e-4, 1e-6])
    assert rho_cut_resolved(s2, 1e-15, MPSTruncation(D=10, eps=1e-8))
    # t^2 = 1e-6 * 1.0101 lies 1.01e-8 above the last value: resolved at a
    # noise of 1e-15, not at 1e-8
    assert rho_cut_resolved(s2, 1e-15, MPSTruncation(D=10, eps=1e-3))
    assert not rho_cut_resolved(s2, 1e-8, MPSTruncation(D=10, eps=1e-3))
    assert rho_cut_resolved(np.append(s2, 1e-20), 1e-15, MPSTruncation(D=2, eps=1e-8))
    assert not rho_cut_resolved(np.append(s2, 1e-20), 1e-15, MPSTruncation(D=10, eps=1e-8))
