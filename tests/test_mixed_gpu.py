"""Mixed-precision forms of the contractions against the FP64 oracle.

Tolerances (relative 2-norm error of the result):
  FP32 operands, FP32 MFMA accumulation            : 1e-5
  split-BF16 x3 (hi*hi + hi*lo + lo*hi), FP32 acc. : 1e-4
The expected errors are ~2^-24*sqrt(K) and ~2^-16 respectively; the bounds
leave an order of magnitude of headroom and still fail for a wrong layout,
a dropped split term (error ~2^-8) or a missing conjugation (error ~1).
"""
import ctypes

import numpy as np
import pytest

from conftest import rand
from oracle import contractions as C

pytestmark = pytest.mark.gpu

TOLS = {"f32": 1e-5, "bf16x3": 1e-4}
PRECS = {"f32": 1, "bf16x3": 2}


def relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def dev(a):
    from finitemptools_amd.engine import to_device
    return to_device(a)


def host(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("prec", ["f32", "bf16x3"])
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("layout", ["NN", "NT", "TN", "TT"])
def test_mixed_gemm_layouts(gpu, prec, cplx, layout):
    import torch
    rng = np.random.default_rng(3)
    M, N, K = 77, 130, 45
    A, B, Cm = rand(rng, (M, K), cplx), rand(rng, (K, N), cplx), rand(rng, (M, N), cplx)
    Ad = dev(A) if layout[0] == "N" else dev(np.ascontiguousarray(A.T)).t()
    Bd = dev(B) if layout[1] == "N" else dev(np.ascontiguousarray(B.T)).t()
    alpha, beta = (0.6 + 0.3j, -0.2 + 0.1j) if cplx else (0.6, -0.2)
    h = gpu.bind_stream()
    for conj in (0, 1):
        Cd = dev(Cm)
        st = gpu.lib.mh_gemm(h, 1 if cplx else 0, PRECS[prec], M, N, K, ctypes.c_void_p(Ad.data_ptr()),
                             Ad.stride(0), Ad.stride(1), conj, ctypes.c_void_p(Bd.data_ptr()),
                             Bd.stride(0), Bd.stride(1), ctypes.c_void_p(Cd.data_ptr()), Cd.stride(0),
                             Cd.stride(1), complex(alpha).real, complex(alpha).imag,
                             complex(beta).real, complex(beta).imag)
        gpu.check(st)
        torch.cuda.synchronize()
        ref = alpha * (A.conj() if conj else A) @ B + beta * Cm
        assert relerr(host(Cd), ref) < TOLS[prec]


@pytest.mark.parametrize("prec", ["f32", "bf16x3"])
@pytest.mark.parametrize("cplx", [False, True])
def test_mixed_heff_and_env(gpu, prec, cplx):
    from finitemptools_amd import models
    from finitemptools_amd.derivatives import ac2_prime, ac_prime, c_prime
    from finitemptools_amd.transfer import updateleft, updateright
    rng = np.random.default_rng(11)
    W = models.heisenberg_xxz_mpo(4, Jzz=0.9, hz=0.1)[1]
    W = W.astype(np.complex128) if cplx else W
    Dl, Dr, w = 40, 36, W.shape[0]
    L = rand(rng, (Dl, w, Dl), cplx) / np.sqrt(Dl * w)
    R = rand(rng, (Dr, w, Dr), cplx) / np.sqrt(Dr * w)
    x2 = rand(rng, (Dl, 2, 2, Dr), cplx)
    x1 = rand(rng, (Dl, 2, Dr), cplx)
    x0 = rand(rng, (Dl, Dr), cplx)
    tol = TOLS[prec]
    assert relerr(host(ac2_prime(dev(x2), dev(W), dev(W), dev(L), dev(R), prec=prec)),
                  C.ac2_prime(x2, W, W, L, R)) < tol
    assert relerr(host(ac_prime(dev(x1), dev(W), dev(L), dev(R), prec=prec)), C.ac_prime(x1, W, L, R)) < tol
    assert relerr(host(c_prime(dev(x0), dev(L), dev(R), prec=prec)), C.c_prime(x0, L, R)) < tol
    A, Bk = rand(rng, (Dl, 2, Dr), cplx), rand(rng, (Dl, 2, Dr), cplx)
    assert relerr(host(updateleft(dev(L), dev(A), dev(W), dev(Bk), prec=prec)), C.updateleft(L, A, W, Bk)) < tol
    assert relerr(host(updateright(dev(R), dev(A), dev(W), dev(Bk), prec=prec)), C.updateright(R, A, W, Bk)) < tol


@pytest.mark.parametrize("prec", ["f32", "bf16x3"])
def test_mixed_benchmark_size(gpu, prec):
    """2-site H_eff at D = 512 (the benchmark shape) in the mixed forms."""
    from finitemptools_amd import models
    from finitemptools_amd.derivatives import HeffOperator
    rng = np.random.default_rng(5)
    W = models.heisenberg_xxz_mpo(4)[1]
    D, w = 512, 5
    L = rng.standard_normal((D, w, D)) / np.sqrt(D * w)
    R = rng.standard_normal((D, w, D)) / np.sqrt(D * w)
    x = rng.standard_normal((D, 2, 2, D))
    op = HeffOperator(2, dev(L), dev(R), dev(W), dev(W), prec=prec)
    assert relerr(host(op(dev(x))), C.ac2_prime(x, W, W, L, R)) < TOLS[prec]


@pytest.mark.parametrize("prec", ["f32", "bf16x3"])
@pytest.mark.parametrize("cplx", [False, True])
def test_every_mixed_gemm_configuration(gpu, prec, cplx):
    """Each configuration of the mixed families (tile, split-K, tile order,
    four or eight waves), forced in turn, within the precision's tolerance;
    configurations a shape does not admit fall back to tuning."""
    import torch
    rng = np.random.default_rng(9)
    n = ctypes.c_int()
    gpu.check(gpu.lib.mh_gemm_candidates(PRECS[prec], ctypes.byref(n)))
    h = gpu.bind_stream()
    try:
        for c in range(n.value):
            gpu.check(gpu.lib.mh_gemm_force_candidate(h, c))
            for M, N, K in [(130, 70, 1100), (64, 192, 40)]:
                A, B = rand(rng, (M, K), cplx), rand(rng, (K, N), cplx)
                Ad, Bd = dev(A), dev(np.ascontiguousarray(B.T)).t()
                Cd = dev(np.zeros((M, N), dtype=A.dtype))
                st = gpu.lib.mh_gemm(h, 1 if cplx else 0, PRECS[prec], M, N, K,
                                     ctypes.c_void_p(Ad.data_ptr()), Ad.stride(0), Ad.stride(1), 0,
                                     ctypes.c_void_p(Bd.data_ptr()), Bd.stride(0), Bd.stride(1),
                                     ctypes.c_void_p(Cd.data_ptr()), Cd.stride(0), Cd.stride(1),
                                     1.0, 0.0, 0.0, 0.0)
                gpu.check(st)
                torch.cuda.synchronize()
                assert relerr(host(Cd), A @ B) < TOLS[prec], (c, M, N, K)
    finally:
        gpu.check(gpu.lib.mh_gemm_force_candidate(h, -1))


FUSED_MX_CASES = [  # (kind, (w, d) or model, Dl, Dr): shapes the packed fused first step takes
    (2, "xxz", 2, 2), (2, "xxz", 34, 50), (2, "xxz", 96, 130), (2, (4, 2), 40, 66),
    (2, (8, 2), 36, 20), (2, (3, 2), 64, 64), (1, "hubbard", 50, 64), (1, (8, 4), 32, 33),
    (1, (6, 2), 130, 70), (1, (3, 4), 2, 7),
]


@pytest.mark.parametrize("prec", ["f32", "bf16x3"])
@pytest.mark.parametrize("kind,Wk,Dl,Dr", FUSED_MX_CASES)
def test_fused_mixed_first_step(gpu, prec, kind, Wk, Dl, Dr):
    """The packed fused GEMM + mode-product kernel (FP64 real operators,
    wl = wr in {3,4,5,6,8}, physical extent 2 or 4, even Dl) and the
    two-launch packed path, each against the FP64 oracle within the
    precision's tolerance; ragged Dl, Dr off the 32 x 64/ds block tile."""
    from finitemptools_amd import models
    from finitemptools_amd.derivatives import ac2_prime, ac_prime
    from finitemptools_amd.engine import engine
    rng = np.random.default_rng(Dl * 37 + Dr)
    if Wk == "xxz":
        W = models.heisenberg_xxz_mpo(4, Jzz=0.9, hz=0.1)[1]
    elif Wk == "hubbard":
        W = models.hubbard_mpo(4, 1.0, 4.0)[1]
    else:
        w, d = Wk
        W = rand(rng, (w, d, w, d))
    w, d = W.shape[0], W.shape[1]
    L = rand(rng, (Dl, w, Dl)) / np.sqrt(Dl * w)
    R = rand(rng, (Dr, w, Dr)) / np.sqrt(Dr * w)
    x = rand(rng, (Dl, d, d, Dr) if kind == 2 else (Dl, d, Dr))
    ref = C.ac2_prime(x, W, W, L, R) if kind == 2 else C.ac_prime(x, W, L, R)
    eng = engine()
    errs, modes = [], []
    eng.profile_enable(True)
    for fused in (True, False):
        eng.set_fusion(fused)
        try:
            eng.profile_reset()
            if kind == 2:
                y = ac2_prime(dev(x), dev(W), dev(W), dev(L), dev(R), prec=prec)
            else:
                y = ac_prime(dev(x), dev(W), dev(L), dev(R), prec=prec)
            errs.append(relerr(host(y), ref))
            modes.append(eng.profile_read(1)[1])   # separate mode-product launches
        finally:
            eng.set_fusion(True)
    eng.profile_enable(False)
    assert modes == [0, 1], modes
    assert errs[0] < TOLS[prec], errs
    assert errs[1] < TOLS[prec], errs
