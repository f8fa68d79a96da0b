"""Batched FP64 symmetric eigensolver on the device for the sector density
matrices of the graded two-site split.

The reference truncates every charge sector of the two-site tensor by a
LAPACK SVD (`auxiliary/stable_svd.jl:4-18`, cut by `MPSTruncation`,
`auxiliary/truncation.jl:27-47`); the device split takes u and s^2 from the
eigenpairs of rho_c = M_c M_c^T instead.  `SectorEigh` runs those eigenpairs
for all sectors of one bond as

1. `mh_sytrd_batched`: Householder tridiagonalisation of every matrix in one
   cooperative launch (LAPACK dsytrd, uplo = 'L', output format);
2. `mh_stedc_dc_batched`: divide and conquer on all tridiagonals at once
   (every merge of one tree height over all matrices in the same launches;
   TRIDIAG = "library" takes rocSOLVER's dstedc per matrix instead,
   `mh_stedc_batched`);
3. `mh_ormtr_batched`: back-transformation of only the eigenvector columns
   the truncation keeps.

`small_eigh` is the one-launch solver for the small matrices of the dense
split (`mh_heev_small`: one workgroup per matrix, the matrix in LDS).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .engine import engine, ptr

MAX_ORDER = 2048   # largest matrix order of mh_sytrd_batched
# Tridiagonal eigensolver of SectorEigh: "dc" (mh_stedc_dc_batched, batched
# divide and conquer) or "library" (rocSOLVER dstedc per matrix).
TRIDIAG = "dc"


def _i32(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _i64(a):
    a = np.ascontiguousarray(a, dtype=np.int64)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


class SectorEigh:
    """Eigen-decomposition of a batch of symmetric FP64 matrices of orders
    `sizes` on `device`.  Fill `matrix(m)` (an n x n view; symmetric, so row-
    and column-major agree), call `eigenvalues()` (ascending per matrix, one
    device tensor per matrix), then `vectors(keep)` for the top `keep[m]`
    eigenvectors of each matrix (n x keep[m] views, ascending order)."""

    def __init__(self, sizes, device):
        self.sizes = [int(n) for n in sizes]
        if any(n < 1 or n > MAX_ORDER for n in self.sizes):
            raise _lib.DimensionMismatch(
                2, f"SectorEigh: orders must lie in [1, {MAX_ORDER}]: {self.sizes}")
        device = torch.device(device)
        if device.type == "cuda" and device.index is None:
            device = torch.device("cuda", torch.cuda.current_device())
        self.device = device
        n = np.array(self.sizes, dtype=np.int64)
        sq = n * n
        self.a_off = np.concatenate([[0], np.cumsum(sq)[:-1]]).astype(np.int64)
        self.v_off = np.concatenate([[0], np.cumsum(n)[:-1]]).astype(np.int64)
        tot_a, tot_v = int(sq.sum()), int(n.sum())
        self.A = torch.empty(max(tot_a, 1), dtype=torch.float64, device=device)
        self.Z = torch.empty(max(tot_a, 1), dtype=torch.float64, device=device)
        self.D = torch.empty(max(tot_v, 1), dtype=torch.float64, device=device)
        self.E = torch.empty(max(tot_v, 1), dtype=torch.float64, device=device)
        self.tau = torch.empty(max(tot_v, 1), dtype=torch.float64, device=device)
        self._n = _i32(n)
        self._a = _i64(self.a_off)
        self._v = _i64(self.v_off)

    def matrix(self, m):
        n, o = self.sizes[m], int(self.a_off[m])
        return self.A[o:o + n * n].view(n, n)

    def eigenvalues(self):
        """Tridiagonalise and diagonalise every matrix (stream-ordered on the
        engine's stream); returns [lambda_m ascending] as device views."""
        e = engine(self.device)
        h = e.bind_stream()
        lib = _lib.load()
        nm = len(self.sizes)
        _lib.raise_for(lib.mh_sytrd_batched(h, nm, self._n[1], self._a[1], self._v[1],
                                            ptr(self.A), ptr(self.D), ptr(self.E),
                                            ptr(self.tau)), h, "mh_sytrd_batched")
        fn = "mh_stedc_dc_batched" if TRIDIAG == "dc" else "mh_stedc_batched"
        _lib.raise_for(getattr(lib, fn)(h, nm, self._n[1], self._v[1], self._a[1],
                                        ptr(self.D), ptr(self.E), ptr(self.Z)), h, fn)
        return [self.D[int(o):int(o) + n] for o, n in zip(self.v_off, self.sizes)]

    def check(self):
        """Synchronise and raise if a barrier timed out or dstedc failed."""
        e = engine(self.device)
        h = e.bind_stream()
        _lib.raise_for(_lib.load().mh_eig_check(h, len(self.sizes)), h, "mh_eig_check")

    def vectors(self, keep):
        """Back-transform the top keep[m] eigenvectors of every matrix; returns
        [n_m x keep_m] device views (columns in ascending eigenvalue order)."""
        keep = [int(k) for k in keep]
        col0 = np.array([n - k for n, k in zip(self.sizes, keep)], dtype=np.int32)
        e = engine(self.device)
        h = e.bind_stream()
        c0, k_ = _i32(col0), _i32(keep)
        _lib.raise_for(_lib.load().mh_ormtr_batched(
            h, len(self.sizes), self._n[1], self._a[1], self._v[1], ptr(self.A), ptr(self.tau),
            self._a[1], c0[1], k_[1], ptr(self.Z)), h, "mh_ormtr_batched")
        out = []
        for m, (n, k) in enumerate(zip(self.sizes, keep)):
            o = int(self.a_off[m])
            Zm = self.Z[o:o + n * n].view(n, n).t()   # column-major n x n
            out.append(Zm[:, n - k:])
        return out


def sector_eigh(mats, device=None):
    """Full eigen-decomposition of symmetric FP64 device matrices (for tests):
    [(lambda ascending, V)] with V column j the eigenvector of lambda[j]."""
    device = device or mats[0].device
    se = SectorEigh([m.shape[0] for m in mats], device)
    for j, M in enumerate(mats):
        se.matrix(j).copy_(M)
    lams = se.eigenvalues()
    V = se.vectors(se.sizes)
    se.check()
    return [(l.clone(), v.clone()) for l, v in zip(lams, V)]


_SMALL_MAX = {}


def small_eigh_max(dtype, device):
    """Largest order `small_eigh` takes for this dtype on this device."""
    device = torch.device(device)
    if device.type != "cuda":
        return 0
    e = engine(device)
    key = (e.device, dtype)
    if key not in _SMALL_MAX:
        h = e.bind_stream()
        out = ctypes.c_int(0)
        code = _lib.MH_C64 if dtype == torch.complex128 else _lib.MH_F64
        _lib.raise_for(_lib.load().mh_heev_small_max(h, code, ctypes.byref(out)), h,
                       "mh_heev_small_max")
        _SMALL_MAX[key] = out.value
    return _SMALL_MAX[key]


_SVD_FITS = {}


def svd_small_fits(shape, dtype, device):
    """Whether `svd_small` takes a matrix of this shape on this device."""
    device = torch.device(device)
    if device.type != "cuda" or dtype not in (torch.float64, torch.complex128):
        return False
    m, n = max(shape), min(shape)
    e = engine(device)
    key = (e.device, dtype, m, n)
    if key not in _SVD_FITS:
        h = e.bind_stream()
        out = ctypes.c_int(0)
        code = _lib.MH_C64 if dtype == torch.complex128 else _lib.MH_F64
        _lib.raise_for(_lib.load().mh_svd_small_fits(h, code, m, n, ctypes.byref(out)), h,
                       "mh_svd_small_fits")
        _SVD_FITS[key] = bool(out.value)
    return _SVD_FITS[key]


def svd_small(M, sweeps=None):
    """Thin SVD of a device matrix M (m x n, float64 or complex128) in one
    launch (mh_svd_small, one-sided Jacobi): (U m x r, S r descending, Vh
    r x n), r = min(m, n), as torch.linalg.svd(M, full_matrices=False).  A
    wide matrix is factorised through its conjugate transpose.  S is NaN if
    the Jacobi sweeps did not converge.  sweeps: an optional one-element
    device int32 tensor that receives the sweeps run (-1: not converged)."""
    m, n = M.shape
    e = engine(M.device)
    h = e.bind_stream()
    code = _lib.MH_C64 if M.dtype == torch.complex128 else _lib.MH_F64
    tall = m >= n
    # column-major storage of the tall operand: M itself (row-major M^T), or
    # M^H (row-major conj(M))
    mm, nn = (m, n) if tall else (n, m)
    M = M.resolve_conj()
    At = (M.t() if tall else torch.conj_physical(M)).contiguous()
    Ut = torch.empty((nn, mm), dtype=M.dtype, device=M.device)   # U, column-major
    S = torch.empty(nn, dtype=torch.float64, device=M.device)
    Vt = torch.empty((nn, nn), dtype=M.dtype, device=M.device)   # V, column-major
    _lib.raise_for(_lib.load().mh_svd_small(h, code, mm, nn, ptr(At), mm, ptr(Ut), ptr(S),
                                            ptr(Vt), ptr(sweeps) if sweeps is not None else None),
                   h, "mh_svd_small")
    # (conjugates are taken physically: the engine reads raw storage, which
    # a lazily conjugated view would not change)
    U1, V1 = Ut.t(), Vt.t()
    if tall:
        return U1, S, torch.conj_physical(Vt)
    # M^H = U1 S V1^H  ->  M = V1 S U1^H
    return V1, S, torch.conj_physical(Ut)


def small_eigh(Gt, col0=0):
    """Eigenpairs of a Hermitian (complex128) or symmetric (float64) n x n
    device matrix G given as Gt = G^T (row-major), i.e. G in column-major
    storage.  Returns (lam ascending, V, info): V[:, j] is the eigenvector of
    lam[j] for j >= col0 (the columns below col0 are unfinished); info is a
    device int32 flag, 1 when the QL iteration did not converge (lam is then
    NaN).  One launch, stream-ordered on the engine's stream."""
    n = Gt.shape[0]
    if Gt.dim() != 2 or Gt.shape[1] != n:
        raise _lib.DimensionMismatch(2, f"small_eigh: square matrix expected, got {tuple(Gt.shape)}")
    if Gt.dtype not in (torch.float64, torch.complex128):
        raise _lib.MHError(3, f"small_eigh: float64 or complex128, got {Gt.dtype}")
    Gt = Gt.contiguous()
    e = engine(Gt.device)
    h = e.bind_stream()
    lam = torch.empty(n, dtype=torch.float64, device=Gt.device)
    Zt = torch.empty((n, n), dtype=Gt.dtype, device=Gt.device)
    info = torch.empty(1, dtype=torch.int32, device=Gt.device)
    code = _lib.MH_C64 if Gt.dtype == torch.complex128 else _lib.MH_F64
    _lib.raise_for(_lib.load().mh_heev_small(h, code, n, 1, ptr(Gt), ptr(lam), ptr(Zt),
                                             int(col0), ptr(info)), h, "mh_heev_small")
    return lam, Zt.t(), info
