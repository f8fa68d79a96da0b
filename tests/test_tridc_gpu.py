"""Batched divide-and-conquer tridiagonal eigensolver (mh_stedc_dc_batched)
against LAPACK (scipy eigh_tridiagonal) and the spectral theorem.

Pins: eigenvalues against LAPACK's dstemr/dstebz results to O(n eps ||T||);
eigenvectors by the residual ||T v - lambda v|| and orthonormality (D&C's
eigenvectors are unique only up to sign and rotations inside degenerate
eigenspaces, so they are not compared entry by entry).  Cases: the classic
stress matrices of tridiagonal solvers (Toeplitz -1 2 -1, Wilkinson W21+
glued copies, zero diagonal, zero and underflowing couplings, a near-identity
with tiny couplings that deflates everything, clusters), graded tridiagonals
from density matrices, every order 1..70 in one batch (leaf and first merge
boundaries), and the largest order."""
import ctypes

import numpy as np
import pytest
import scipy.linalg as sla

from finitemptools_amd import _lib
from finitemptools_amd.engine import ptr

pytestmark = pytest.mark.gpu


def _run(gpu, cases):
    import torch
    ns = np.array([len(d) for d, _ in cases], dtype=np.int32)
    v_off = np.concatenate([[0], np.cumsum(ns)[:-1]]).astype(np.int64)
    z_off = np.concatenate([[0], np.cumsum(ns.astype(np.int64) ** 2)[:-1]]).astype(np.int64)
    D = np.zeros(int(ns.sum()))
    E = np.zeros(int(ns.sum()))
    for (d, e), o in zip(cases, v_off):
        D[o:o + len(d)] = d
        E[o:o + len(e)] = e
    Dd = torch.from_numpy(D).cuda()
    Ed = torch.from_numpy(E).cuda()
    Zd = torch.full((int((ns.astype(np.int64) ** 2).sum()),), np.nan, dtype=torch.float64,
                    device="cuda")
    h = gpu.bind_stream()
    lib = _lib.load()
    P32, P64 = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int64)
    print(f"mh_stedc_dc_batched: {len(cases)} matrices, orders {sorted(set(ns.tolist()))[:8]}...",
          flush=True)
    _lib.raise_for(lib.mh_stedc_dc_batched(h, len(cases), ns.ctypes.data_as(P32),
                                           v_off.ctypes.data_as(P64), z_off.ctypes.data_as(P64),
                                           ptr(Dd), ptr(Ed), ptr(Zd)), h)
    _lib.raise_for(lib.mh_eig_check(h, 0), h)
    lam, Z = Dd.cpu().numpy(), Zd.cpu().numpy()
    assert np.array_equal(Ed.cpu().numpy(), E), "E must be left untouched"
    out = []
    for (d, e), n, vo, zo in zip(cases, ns, v_off, z_off):
        out.append((lam[vo:vo + n], Z[zo:zo + n * n].reshape(n, n).T))
    return out


def _check(cases, results, c=20.0):
    eps = np.finfo(float).eps
    for (d, e), (lam, V) in zip(cases, results):
        n = len(d)
        T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        nt = max(np.abs(d).max(initial=0.0) + 2 * np.abs(e).max(initial=0.0), 1e-300)
        ref = sla.eigh_tridiagonal(d, e, eigvals_only=True) if n > 1 else d.copy()
        assert np.all(np.diff(lam) >= 0), "eigenvalues not ascending"
        tol = c * max(n, 1) * eps * nt
        assert np.max(np.abs(lam - ref)) <= tol, (n, np.max(np.abs(lam - ref)) / nt)
        assert np.max(np.abs(V.T @ V - np.eye(n))) <= c * max(n, 1) * eps, n
        assert np.max(np.abs(T @ V - V * lam)) <= tol, n


def _classic(rng):
    cases = []
    for n in (1, 2, 3, 31, 32, 33, 64, 65, 100, 257, 511):
        cases.append((rng.standard_normal(n), rng.standard_normal(n - 1)))
    n = 300
    cases.append((np.full(n, 2.0), np.full(n - 1, -1.0)))
    cases.append((np.zeros(n), np.ones(n - 1)))
    w = np.abs(np.arange(-10, 11)).astype(float)
    cases.append((np.tile(w, 12), np.ones(21 * 12 - 1)))
    cases.append((np.tile(w, 12), np.where(np.arange(21 * 12 - 1) % 21 == 20, 1e-14, 1.0)))
    e = rng.standard_normal(n - 1)
    e[[37, 150, 151]] = [0.0, 1e-300, 0.0]
    cases.append((rng.standard_normal(n), e))
    cases.append((np.ones(n), 1e-13 * rng.standard_normal(n - 1)))
    # clusters: three eigenvalue groups of 40 within 1e-12
    cl = np.repeat([1.0, 2.0, 3.0], 40) + 1e-12 * rng.standard_normal(120)
    Q = np.linalg.qr(rng.standard_normal((120, 120)))[0]
    T = (Q * cl) @ Q.T
    Tt = sla.hessenberg(T)
    cases.append((np.diag(Tt).copy(), np.diag(Tt, -1).copy()))
    return cases


def _graded(rng, n, decay):
    M = rng.standard_normal((n, n)) * np.exp(-decay * np.arange(n) / n)
    T = sla.hessenberg(M @ M.T)
    return np.diag(T).copy(), np.diag(T, -1).copy()


def test_classic_tridiagonals(gpu):
    rng = np.random.default_rng(0)
    cases = _classic(rng)
    _check(cases, _run(gpu, cases))


def test_graded_density_matrices(gpu):
    rng = np.random.default_rng(1)
    cases = [_graded(rng, n, dec) for n, dec in ((50, 10.0), (200, 40.0), (767, 40.0), (600, 5.0))]
    _check(cases, _run(gpu, cases))


def test_every_small_order_in_one_batch(gpu):
    rng = np.random.default_rng(2)
    cases = [(rng.standard_normal(n), rng.standard_normal(n - 1)) for n in range(1, 71)]
    _check(cases, _run(gpu, cases))


def test_largest_order(gpu):
    rng = np.random.default_rng(3)
    cases = [(rng.standard_normal(4096), rng.standard_normal(4095)), _graded(rng, 1500, 30.0)]
    _check(cases, _run(gpu, cases), c=40.0)


def test_order_limit_is_an_error(gpu):
    import torch
    h = gpu.bind_stream()
    lib = _lib.load()
    ns = np.array([4097], dtype=np.int32)
    off = np.zeros(1, dtype=np.int64)
    buf = torch.zeros(8, dtype=torch.float64, device="cuda")
    P32, P64 = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int64)
    st = lib.mh_stedc_dc_batched(h, 1, ns.ctypes.data_as(P32), off.ctypes.data_as(P64),
                                 off.ctypes.data_as(P64), ptr(buf), ptr(buf), ptr(buf))
    assert st == 2   # MH_ERR_SHAPE
