"""One-workgroup device SVD (mh_svd_small, one-sided Jacobi) against LAPACK
(numpy's gesdd) and the SVD's defining properties.

Pins: singular values against LAPACK to O(n eps s_max) — both methods meet
that bound, LAPACK as a backward-stable bidiagonalisation, Jacobi with its
small values to high relative accuracy on top; the factors by reconstruction
(U diag(S) V^H = M to 1e-13 relative) and orthonormality (U's columns with
S > 0, V), since singular vectors are unique only up to phases and rotations
inside degenerate subspaces.  Cases: tall, wide and square, real and
complex, orders 1 to the LDS limit, a graded spectrum falling to 1e-14, exact
rank deficiency and repeated singular values; then the truncated SVD split
(stable_tsvd) on the device against the same split in host LAPACK."""
import ctypes

import numpy as np
import pytest

from finitemptools_amd import _lib, eigen
from finitemptools_amd.engine import ptr

pytestmark = pytest.mark.gpu


def _mat(rng, m, n, cplx, spectrum=None):
    def rnd(a, b):
        x = rng.standard_normal((a, b))
        return x + 1j * rng.standard_normal((a, b)) if cplx else x
    if spectrum is None:
        return rnd(m, n)
    r = min(m, n)
    Q1 = np.linalg.qr(rnd(m, r))[0]
    Q2 = np.linalg.qr(rnd(n, r))[0]
    return (Q1 * spectrum[:r]) @ Q2.conj().T


def _check(M, tol=1e-13):
    import torch
    Md = torch.from_numpy(M).cuda()
    U, S, Vh = eigen.svd_small(Md)
    U, S, Vh = U.cpu().numpy(), S.cpu().numpy(), Vh.cpu().numpy()
    ref = np.linalg.svd(M, compute_uv=False)
    r = min(M.shape)
    smax = max(ref.max(initial=0.0), 1e-300)
    assert U.shape == (M.shape[0], r) and S.shape == (r,) and Vh.shape == (r, M.shape[1])
    assert np.all(np.diff(S) <= 0), "singular values not descending"
    assert np.max(np.abs(S - ref)) <= 10 * r * np.finfo(float).eps * smax, np.max(np.abs(S - ref)) / smax
    assert np.linalg.norm((U * S) @ Vh - M) <= tol * max(np.linalg.norm(M), 1e-300)
    k = int(np.count_nonzero(S > 1e-12 * smax))
    assert np.max(np.abs(U[:, :k].conj().T @ U[:, :k] - np.eye(k)), initial=0.0) <= tol * max(r, 10)
    assert np.max(np.abs(Vh @ Vh.conj().T - np.eye(r))) <= tol * max(r, 10)
    return S, ref


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("shape", [(1, 1), (3, 1), (1, 3), (7, 5), (5, 7), (33, 33), (40, 17),
                                   (17, 40), (64, 64), (100, 60), (60, 100), (300, 20),
                                   (20, 300)])
def test_random_matrices(gpu, cplx, shape):
    rng = np.random.default_rng(hash((shape, cplx)) % 2**32)
    _check(_mat(rng, *shape, cplx))


def test_largest_orders(gpu):
    import torch
    rng = np.random.default_rng(7)
    for dtype, cplx in ((torch.float64, False), (torch.complex128, True)):
        n = 1
        while eigen.svd_small_fits((n + 1, n + 1), dtype, "cuda"):
            n += 1
        assert n >= (96 if not cplx else 64)
        print(f"largest square order {'complex' if cplx else 'real'}: {n}", flush=True)
        _check(_mat(rng, n, n, cplx))


@pytest.mark.parametrize("cplx", [False, True])
def test_graded_spectrum(gpu, cplx):
    # a two-site matrix's spectrum: values falling over 14 decades
    rng = np.random.default_rng(11)
    m, n = 48, 40
    sv = np.logspace(0, -14, min(m, n))
    S, ref = _check(_mat(rng, m, n, cplx, spectrum=sv))
    # the matrix as stored carries rounding of eps64 in its entries, so its
    # own singular values are those of the spectrum to that absolute level;
    # the large values carry Jacobi's normwise error (a few ulp per sweep),
    # the small ones must stay at the entries' own level
    eps = np.finfo(float).eps
    assert np.max(np.abs(S - sv)) <= 4 * len(sv) * eps
    small = sv < 1e-6
    assert np.max(np.abs(S[small] - sv[small])) <= 1e-16


@pytest.mark.parametrize("cplx", [False, True])
def test_rank_deficient_and_repeated(gpu, cplx):
    rng = np.random.default_rng(13)
    sv = np.array([3.0, 3.0, 3.0, 1.0, 1.0, 0.5, 0.0, 0.0, 0.0, 0.0])
    M = _mat(rng, 24, 10, cplx, spectrum=sv)
    S, _ = _check(M, tol=1e-12)
    np.testing.assert_allclose(S, sv, atol=1e-13 * 3.0)
    _check(np.zeros((6, 4), dtype=complex if cplx else float))


def test_shape_errors(gpu):
    import torch
    h = gpu.bind_stream()
    lib = _lib.load()
    A = torch.zeros(16, dtype=torch.float64, device="cuda")
    S = torch.zeros(4, dtype=torch.float64, device="cuda")
    # wide matrices go through the conjugate transpose, not to the kernel
    assert lib.mh_svd_small(h, _lib.MH_F64, 2, 4, ptr(A), 2, ptr(A), ptr(S), ptr(A), None) == 2
    fits = ctypes.c_int(1)
    assert lib.mh_svd_small_fits(h, _lib.MH_C64, 4096, 4096, ctypes.byref(fits)) == 0
    assert fits.value == 0


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("eps", [1e-8, 1e-10])
def test_device_svd_split_matches_host_lapack(gpu, monkeypatch, cplx, eps):
    """The truncated SVD split (stable_tsvd) on the device equals the host
    LAPACK split: kept count, singular values and the truncated tensor."""
    import torch
    from finitemptools_amd import truncation as tr
    rng = np.random.default_rng(17)
    Dl, d, Dr = 20, 2, 24
    sv = np.logspace(0, -13, min(Dl * d, d * Dr))
    M = _mat(rng, Dl * d, d * Dr, cplx, spectrum=sv)
    dt = torch.complex128 if cplx else torch.float64
    x = torch.from_numpy(M.reshape(Dl, d, d, Dr)).to(dt).cuda()
    trunc = tr.MPSTruncation(D=64, eps=eps)
    out = {}
    for dev in (True, False):
        monkeypatch.setattr(tr, "DEVICE_SVD", dev)
        u, s, v, err = tr.stable_tsvd(x, trunc)
        out[dev] = (u.cpu().numpy(), s.cpu().numpy(), v.cpu().numpy(), err)
    (u1, s1, v1, e1), (u0, s0, v0, e0) = out[True], out[False]
    assert len(s1) == len(s0)
    np.testing.assert_allclose(s1, s0, atol=1e-13, rtol=0)
    k = len(s1)
    t1 = np.einsum("ask,k,ktb->astb", u1, s1, v1)
    t0 = np.einsum("ask,k,ktb->astb", u0, s0, v0)
    assert np.max(np.abs(t1 - t0)) <= 1e-12
    assert abs(e1 - e0) <= 1e-13


@pytest.mark.parametrize("cplx", [False, True])
def test_sweep_count(gpu, cplx):
    """Jacobi converges quadratically: a random 48 x 48 matrix takes about 10
    sweeps (cyclic Jacobi on the CPU takes the same), a graded one about 8
    after the pivoted QR (25 without); a wrong rotation angle still
    converges, only slowly, and shows here first."""
    import torch
    rng = np.random.default_rng(19)
    sw = torch.zeros(1, dtype=torch.int32, device="cuda")
    for spectrum, limit in ((None, 14), (np.logspace(0, -14, 48), 12)):
        M = torch.from_numpy(_mat(rng, 48, 48, cplx, spectrum=spectrum)).cuda()
        eigen.svd_small(M, sweeps=sw)
        n_sw = int(sw.item())
        print(f"{'graded' if spectrum is not None else 'random'}: {n_sw} sweeps", flush=True)
        assert 1 <= n_sw <= limit


@pytest.mark.parametrize("cplx", [False, True])
def test_plain_and_preconditioned_agree(gpu, monkeypatch, cplx):
    """The pivoted-QR preconditioned iteration (the default) and the plain
    one on A (MH_SVD_PLAIN=1) give the same SVD; the preconditioned one
    takes a third of the sweeps on a graded spectrum."""
    import torch
    rng = np.random.default_rng(23)
    M = _mat(rng, 40, 32, cplx, spectrum=np.logspace(0, -14, 32))
    sw = torch.zeros(1, dtype=torch.int32, device="cuda")
    out = {}
    for plain in ("0", "1"):
        monkeypatch.setenv("MH_SVD_PLAIN", plain)
        S, _ = _check(M)
        eigen.svd_small(torch.from_numpy(M).cuda(), sweeps=sw)
        out[plain] = (S, int(sw.item()))
    np.testing.assert_allclose(out["0"][0], out["1"][0], atol=4 * 32 * np.finfo(float).eps, rtol=0)
    assert out["0"][1] <= 12 and out["0"][1] * 2 <= out["1"][1], (out["0"][1], out["1"][1])
