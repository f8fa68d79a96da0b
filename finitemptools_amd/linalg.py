"""Dense tensor glue for the sweeps: products through the engine's MFMA GEMM,
and the FP64 truncated SVD / QR of the sweep step.

The SVD and QR go through torch.linalg on the device (rocSOLVER); they stay in
FP64 in every precision mode (auxiliary/stable_svd.jl, mps/orth.jl).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .engine import as_colmajor, colmajor_empty, dtype_code, engine, ptr, resolve


def gemm(A, B, alpha=1.0, conjA=False, out=None, beta=0.0):
    """C = alpha * op(A) @ B (+ beta C) for 2-D device tensors of any strides
    with one unit stride each, through mh_gemm."""
    M, K = A.shape
    K2, N = B.shape
    if K != K2:
        raise _lib.DimensionMismatch(2, f"gemm: {tuple(A.shape)} x {tuple(B.shape)}")
    A, B = resolve(A), resolve(B)
    if A.stride(0) != 1 and A.stride(1) != 1:
        A = as_colmajor(A)
    if B.stride(0) != 1 and B.stride(1) != 1:
        B = as_colmajor(B)
    C = out if out is not None else colmajor_empty((M, N), A.dtype, A.device)
    e = engine(A.device)
    h = e.bind_stream()
    a = complex(alpha)
    b = complex(beta)
    e.check(e.lib.mh_gemm(h, dtype_code(A), 0, M, N, K, ptr(A), A.stride(0), A.stride(1),
                          1 if conjA else 0, ptr(B), B.stride(0), B.stride(1), ptr(C),
                          C.stride(0), C.stride(1), a.real, a.imag, b.real, b.imag), "gemm")
    return C


def gemm_batched(tasks, alpha=1.0, beta=0.0):
    """C = alpha A B + beta C for every (A, B, C) in `tasks` in one launch
    (mh_gemm_batched): float64 device matrices of any strides (transposed
    views included), all on one device; C must not overlap A or B."""
    tasks = [t for t in tasks if t[2].numel() > 0]
    if not tasks:
        return
    dev = tasks[0][2].device
    desc = np.empty((len(tasks), 12), dtype=np.int64)
    for i, (A, B, C) in enumerate(tasks):
        M, K = A.shape
        K2, N = B.shape
        if K2 != K or tuple(C.shape) != (M, N):
            raise _lib.DimensionMismatch(2, f"gemm_batched: {tuple(A.shape)} x {tuple(B.shape)} -> {tuple(C.shape)}")
        for T in (A, B, C):
            if T.dtype != torch.float64 or T.device != dev:
                raise _lib.MHError(3, "gemm_batched: float64 operands on one device")
        desc[i] = (A.data_ptr(), B.data_ptr(), C.data_ptr(), M, N, K,
                   A.stride(0), A.stride(1), B.stride(0), B.stride(1), C.stride(0), C.stride(1))
    e = engine(dev)
    h = e.bind_stream()
    _lib.raise_for(_lib.load().mh_gemm_batched(
        h, len(tasks), desc.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), float(alpha),
        float(beta)), h, "mh_gemm_batched")


def as_matrix(t, nrow_legs):
    """Column-major tensor -> the (prod first legs) x (prod remaining legs)
    matrix view (TensorMap codomain/domain split)."""
    t = as_colmajor(t)
    m = int(np.prod(t.shape[:nrow_legs]))
    n = int(np.prod(t.shape[nrow_legs:]))
    return t.permute(*reversed(range(t.dim()))).reshape(n, m).t()


def from_matrix(M, shape):
    """(m x n) matrix (any layout) -> column-major tensor of the given legs."""
    Mc = resolve(M.t().contiguous())  # C-order (n, m) == column-major (m, n)
    return Mc.reshape(tuple(reversed(shape))).permute(*reversed(range(len(shape))))


def cm_view(C, shape):
    """A column-major matrix (strides (1, M)) seen as a column-major tensor
    of the given legs (no copy)."""
    if not (C.stride(0) == 1 and (C.shape[1] == 1 or C.stride(1) == C.shape[0])):
        C = as_colmajor(C)
    return C.t().reshape(tuple(reversed(shape))).permute(*reversed(range(len(shape))))


def two_site(A1, A2):
    """x[a, s1, s2, b] = A1[a, s1, k] A2[k, s2, b] (dmrg.jl:98)."""
    Dl, d1, _ = A1.shape
    _, d2, Dr = A2.shape
    return cm_view(gemm(as_matrix(A1, 2), as_matrix(A2, 1)), (Dl, d1, d2, Dr))


def svd(M):
    """Thin SVD of a device matrix in FP64 (torch.linalg -> rocSOLVER)."""
    return torch.linalg.svd(M, full_matrices=False)


def qr(M):
    return torch.linalg.qr(M, mode="reduced")
