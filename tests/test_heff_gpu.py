"""Parity of the HIP engine (through the C ABI) with the CPU oracle.

FP64 tolerance: 1e-12 relative in the 2-norm (the bar for the plain path);
complex data uses the same bar.  Shapes cover edge bonds (D = 1, 2), ragged
extents that are not multiples of the 16x16 MFMA tile, the XXZ (w = 5, d = 2)
and Hubbard-like (w = 6, d = 4) MPO shapes, and the full benchmark sizes.
"""
import ctypes

import numpy as np
import pytest

from conftest import hermitian_env, model_env, rand
from oracle import contractions as C
from oracle.lanczos import simple_lanczos_solver as oracle_lanczos

pytestmark = pytest.mark.gpu

TOL = 1e-12


def relerr(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


def dev(a):
    from finitemptools_amd.engine import to_device
    return to_device(a)


def host(t):
    return t.detach().cpu().numpy()


def mpo_tensor(kind, cplx=False, rng=None):
    from finitemptools_amd import models
    if kind == "xxz":
        W = models.heisenberg_xxz_mpo(4, Jzz=0.9, hz=0.1)[1]
    elif kind == "hubbard":
        W = models.hubbard_mpo(4, 1.0, 4.0)[1]
    elif kind == "tfim":
        W = models.tfim_mpo(4, 1.0, 0.7)[1]
    else:
        W = rand(rng, kind, cplx)
    return W.astype(np.complex128) if cplx else W


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("layout", ["NN", "NT", "TN", "TT"])
def test_gemm_strided_layouts(gpu, cplx, layout):
    import torch
    from finitemptools_amd import _lib
    rng = np.random.default_rng(20)
    M, N, K = 37, 53, 19
    A = rand(rng, (M, K), cplx)
    B = rand(rng, (K, N), cplx)
    Cm = rand(rng, (M, N), cplx)
    # device storage: column-major when "N" for A (m fastest) / B (k fastest)
    Ad = dev(A) if layout[0] == "N" else dev(np.ascontiguousarray(A.T)).t()
    Bd = dev(B) if layout[1] == "N" else dev(np.ascontiguousarray(B.T)).t()
    Cd = dev(Cm)
    sam, sak = Ad.stride()
    sbk, sbn = Bd.stride()
    alpha, beta = (0.7 - 0.2j, 0.3 + 0.1j) if cplx else (0.7, 0.3)
    h = gpu.bind_stream()
    for conj in (0, 1):
        Cd = dev(Cm)
        st = gpu.lib.mh_gemm(h, 1 if cplx else 0, 0, M, N, K, ctypes.c_void_p(Ad.data_ptr()), sam, sak, conj,
                             ctypes.c_void_p(Bd.data_ptr()), sbk, sbn, ctypes.c_void_p(Cd.data_ptr()),
                             Cd.stride(0), Cd.stride(1), complex(alpha).real, complex(alpha).imag,
                             complex(beta).real, complex(beta).imag)
        gpu.check(st)
        torch.cuda.synchronize()
        Aop = A.conj() if conj else A
        ref = alpha * Aop @ B + beta * Cm
        assert relerr(host(Cd), ref) < TOL


def test_gemm_beta_on_a_tuned_shape(gpu):
    """C = alpha A B + beta C at a size the tuner times (1024^3 > 2e9 flops)
    on a fresh handle, so the first call meets an uncached shape: the tuning
    runs must not fold their results into C."""
    import torch
    from finitemptools_amd.engine import Engine
    rng = np.random.default_rng(21)
    n = 1024
    A, B, C0 = (rng.standard_normal((n, n)) for _ in range(3))
    e = Engine(0)
    try:
        h = e.bind_stream()
        Ad, Bd = dev(A), dev(B)
        for call in range(2):   # first call tunes (beta = 0 shape unknown), second is cached
            Cd = dev(C0)
            print(f"gemm beta call {call}", flush=True)
            st = e.lib.mh_gemm(h, 0, 0, n, n, n, ctypes.c_void_p(Ad.data_ptr()), 1, n, 0,
                               ctypes.c_void_p(Bd.data_ptr()), 1, n, ctypes.c_void_p(Cd.data_ptr()),
                               1, n, 0.7, 0.0, 0.3, 0.0)
            e.check(st)
            torch.cuda.synchronize()
            assert relerr(host(Cd), 0.7 * A @ B + 0.3 * C0) < TOL
        # a beta = 0 call tunes the shape; a later beta call reuses that choice
        Cd = dev(C0)
        e.check(e.lib.mh_gemm(h, 0, 0, n, n, n, ctypes.c_void_p(Ad.data_ptr()), 1, n, 0,
                              ctypes.c_void_p(Bd.data_ptr()), 1, n, ctypes.c_void_p(Cd.data_ptr()),
                              1, n, 1.0, 0.0, 0.0, 0.0))
        torch.cuda.synchronize()
        assert relerr(host(Cd), A @ B) < TOL
    finally:
        e.close()


SHAPES1 = [  # Dl, d, Dr, W
    (1, 2, 2, "xxz"), (2, 2, 4, "xxz"), (16, 2, 16, "xxz"), (37, 2, 29, "xxz"),
    (64, 2, 64, "xxz"), (24, 4, 20, "hubbard"), (33, 2, 47, "tfim"),
]


def _envs(rng, Dl, wl, Dr, wr, cplx):
    L = rand(rng, (Dl, wl, Dl), cplx) / np.sqrt(Dl * wl)
    R = rand(rng, (Dr, wr, Dr), cplx) / np.sqrt(Dr * wr)
    return L, R


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("Dl,d,Dr,Wk", SHAPES1)
def test_ac_prime_parity(gpu, cplx, Dl, d, Dr, Wk):
    from finitemptools_amd.derivatives import ac_prime
    rng = np.random.default_rng(Dl * 1000 + Dr)
    W = mpo_tensor(Wk, cplx)
    wl, wr = W.shape[0], W.shape[2]
    L, R = _envs(rng, Dl, wl, Dr, wr, cplx)
    x = rand(rng, (Dl, d, Dr), cplx)
    y = ac_prime(dev(x), dev(W), dev(L), dev(R))
    assert relerr(host(y), C.ac_prime(x, W, L, R)) < TOL


SHAPES2 = [  # Dl, Dr, W (two-site uses the same MPO tensor on both sites)
    (1, 4, "xxz"), (4, 8, "xxz"), (16, 16, "xxz"), (45, 31, "xxz"),
    (64, 64, "xxz"), (12, 20, "hubbard"), (40, 40, "hubbard"), (30, 22, "tfim"),
]


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("Dl,Dr,Wk", SHAPES2)
def test_ac2_prime_parity(gpu, cplx, Dl, Dr, Wk):
    from finitemptools_amd.derivatives import ac2_prime
    rng = np.random.default_rng(Dl * 100 + Dr + 7)
    W = mpo_tensor(Wk, cplx)
    d, w = W.shape[1], W.shape[0]
    L, R = _envs(rng, Dl, w, Dr, w, cplx)
    x = rand(rng, (Dl, d, d, Dr), cplx)
    y = ac2_prime(dev(x), dev(W), dev(W), dev(L), dev(R))
    assert relerr(host(y), C.ac2_prime(x, W, W, L, R)) < TOL


def test_ac2_prime_random_mpo_unequal_bonds(gpu):
    """Dense random MPO tensors with wl != wm != wr and d1 != d2."""
    from finitemptools_amd.derivatives import ac2_prime
    rng = np.random.default_rng(31)
    Dl, d1, d2, Dr, wl, wm, wr = 9, 2, 3, 14, 3, 4, 2
    W1, W2 = rand(rng, (wl, d1, wm, d1)), rand(rng, (wm, d2, wr, d2))
    L, R = _envs(rng, Dl, wl, Dr, wr, False)
    x = rand(rng, (Dl, d1, d2, Dr))
    y = ac2_prime(dev(x), dev(W1), dev(W2), dev(L), dev(R))
    assert relerr(host(y), C.ac2_prime(x, W1, W2, L, R)) < TOL


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("Dl,Dr,w", [(1, 1, 1), (3, 5, 4), (64, 48, 5), (100, 77, 3)])
def test_c_prime_parity(gpu, cplx, Dl, Dr, w):
    from finitemptools_amd.derivatives import c_prime
    rng = np.random.default_rng(Dl + 11 * Dr)
    L, R = _envs(rng, Dl, w, Dr, w, cplx)
    x = rand(rng, (Dl, Dr), cplx)
    y = c_prime(dev(x), dev(L), dev(R))
    assert relerr(host(y), C.c_prime(x, L, R)) < TOL


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("Dl,d,Dr,Wk", [(1, 2, 2, "xxz"), (8, 2, 16, "xxz"), (37, 2, 41, "xxz"),
                                        (20, 4, 24, "hubbard"), (64, 2, 64, "tfim")])
def test_env_updates_parity(gpu, cplx, Dl, d, Dr, Wk):
    from finitemptools_amd.transfer import updateleft, updateright
    rng = np.random.default_rng(Dl * 7 + Dr)
    W = mpo_tensor(Wk, cplx)
    wl, wr = W.shape[0], W.shape[2]
    A, B = rand(rng, (Dl, d, Dr), cplx), rand(rng, (Dl, d, Dr), cplx)
    hl, hr = _envs(rng, Dl, wl, Dr, wr, cplx)
    out_l = updateleft(dev(hl), dev(A), dev(W), dev(B))
    assert relerr(host(out_l), C.updateleft(hl, A, W, B)) < TOL
    out_r = updateright(dev(hr), dev(A), dev(W), dev(B))
    assert relerr(host(out_r), C.updateright(hr, A, W, B)) < TOL


def test_env_identity_and_scalar_entries(gpu):
    from finitemptools_amd.transfer import updateleft, updateright
    rng = np.random.default_rng(5)
    A, B = rand(rng, (5, 2, 6), True), rand(rng, (5, 2, 6), True)
    hl, hr = rand(rng, (5, 3, 5), True), rand(rng, (6, 3, 6), True)
    assert relerr(host(updateleft(dev(hl), dev(A), 0.5, dev(B))), C.updateleft(hl, A, 0.5, B)) < TOL
    assert relerr(host(updateright(dev(hr), dev(A), None, dev(B))), C.updateright(hr, A, None, B)) < TOL


def test_shape_mismatch_raises(gpu):
    from finitemptools_amd import _lib
    from finitemptools_amd.derivatives import ac_prime
    rng = np.random.default_rng(0)
    W = mpo_tensor("xxz")
    with pytest.raises(_lib.DimensionMismatch):
        ac_prime(dev(rand(rng, (4, 2, 5))), dev(W), dev(rand(rng, (4, 5, 4))), dev(rand(rng, (6, 5, 6))))


def test_krylov_vector_primitives(gpu):
    from finitemptools_amd.krylov import dot, norm
    rng = np.random.default_rng(6)
    x, y = rand(rng, 10007, True), rand(rng, 10007, True)
    assert abs(dot(dev(x), dev(y)) - np.vdot(x, y)) < 1e-10 * abs(np.vdot(x, y))
    assert abs(norm(dev(x)) - np.linalg.norm(x)) < 1e-12 * np.linalg.norm(x)


@pytest.mark.parametrize("kind,cplx", [(2, False), (1, False), (2, True), (0, True)])
def test_lanczos_eig_matches_oracle(gpu, kind, cplx):
    """Same start vector, same kmax / btol: eigenvalue and Ritz vector of
    simple_lanczos_solver (reference semantics: no re-orthogonalisation)."""
    from finitemptools_amd.derivatives import HeffOperator
    from finitemptools_amd.krylov import simple_lanczos_solver
    rng = np.random.default_rng(40 + kind)
    W = mpo_tensor("xxz", cplx)
    w = W.shape[0]
    Dl, Dr, d = 12, 10, 2
    L = model_env(rng, Dl, w, "xxz", cplx) / Dl
    R = model_env(rng, Dr, w, "xxz", cplx) / Dr
    if kind == 2:
        x = rand(rng, (Dl, d, d, Dr), cplx)
        op = HeffOperator(2, dev(L), dev(R), dev(W), dev(W))
        f = lambda v: C.ac2_prime(v, W, W, L, R)
    elif kind == 1:
        x = rand(rng, (Dl, d, Dr), cplx)
        op = HeffOperator(1, dev(L), dev(R), dev(W))
        f = lambda v: C.ac_prime(v, W, L, R)
    else:
        x = rand(rng, (Dl, Dr), cplx)
        op = HeffOperator(0, dev(L), dev(R))
        f = lambda v: C.c_prime(v, L, R)
    e_ref, v_ref, info_ref, k_ref = oracle_lanczos(f, x, "SR", 10, 1e-12)
    e, v, info = simple_lanczos_solver(op, dev(x), "SR", 10, 1e-12)
    assert abs(e - e_ref) < 1e-10 * max(1.0, abs(e_ref))
    vh = host(v)
    ph = np.vdot(v_ref, vh)
    ph = ph / abs(ph)
    assert relerr(vh, ph * v_ref) < 1e-8
    assert info == info_ref


def test_lanczos_early_exit_small_space(gpu):
    """Krylov space exhausted (dimension 4 < kmax): beta < btol ends the
    recurrence, later steps stay finite, info = 0 (simple_lanczos.jl:104-107)."""
    from finitemptools_amd.derivatives import HeffOperator
    from finitemptools_amd.krylov import simple_lanczos_solver
    rng = np.random.default_rng(8)
    W = mpo_tensor("xxz")
    L = hermitian_env(rng, 1, 1, False)
    R = model_env(rng, 2, 5, "xxz", False)
    W1 = W[0:1]
    x = rand(rng, (1, 2, 2))
    op = HeffOperator(1, dev(L), dev(R), dev(W1))
    H = C.dense_heff1(W1, L, R)
    e, v, info = simple_lanczos_solver(op, dev(x), "SR", 10, 1e-12)
    assert info == 0
    assert abs(e - np.linalg.eigvalsh(H)[0]) < 1e-12 * max(1, abs(e))
    assert np.all(np.isfinite(host(v)))


@pytest.mark.parametrize("dt", [-0.05j, 0.02, -0.4j])
def test_lanczos_expm_matches_dense_expm(gpu, dt):
    import scipy.linalg as sla
    from finitemptools_amd.derivatives import HeffOperator
    from finitemptools_amd.krylov import exponentiate
    rng = np.random.default_rng(9)
    W = mpo_tensor("tfim", True)
    w = W.shape[0]
    Dl, Dr, d = 4, 5, 2
    L = hermitian_env(rng, Dl, w, True)
    R = hermitian_env(rng, Dr, w, True)
    x = rand(rng, (Dl, d, d, Dr), True)
    op = HeffOperator(2, dev(L), dev(R), dev(W), dev(W))
    H = C.dense_heff2(W, W, L, R)
    ref = sla.expm(dt * H) @ C.vec(x)
    y, info = exponentiate(op, dt, dev(x), tol=1e-12, kdim=12)
    assert info == 0
    assert relerr(C.vec(host(y)), ref) < 1e-9


@pytest.mark.parametrize("which", ["ac2_D512", "ac_D1024"])
def test_benchmark_size_parity_and_properties(gpu, which):
    """Full benchmark sizes: one matvec against the oracle, plus linearity and
    Hermiticity (size-independent properties)."""
    from finitemptools_amd.derivatives import HeffOperator
    rng = np.random.default_rng(77)
    W = mpo_tensor("xxz")
    w = W.shape[0]
    D = 512 if which == "ac2_D512" else 1024
    L = model_env(rng, D, w, "xxz") / np.sqrt(D * w)
    R = model_env(rng, D, w, "xxz") / np.sqrt(D * w)
    if which == "ac2_D512":
        shape = (D, 2, 2, D)
        op = HeffOperator(2, dev(L), dev(R), dev(W), dev(W))
        f = lambda v: C.ac2_prime(v, W, W, L, R)
    else:
        shape = (D, 2, D)
        op = HeffOperator(1, dev(L), dev(R), dev(W))
        f = lambda v: C.ac_prime(v, W, L, R)
    x, z = rand(rng, shape), rand(rng, shape)
    hx, hz = op(dev(x)), op(dev(z))
    assert relerr(host(hx), f(x)) < TOL
    hs = op(dev(0.3 * x - 1.7 * z))
    assert relerr(host(hs), 0.3 * host(hx) - 1.7 * host(hz)) < 1e-12
    lhs = np.vdot(z, host(hx))
    rhs = np.vdot(host(hz), x)
    assert abs(lhs - rhs) < 1e-11 * np.linalg.norm(host(hx)) * np.linalg.norm(z)


FUSED_CASES = [  # (kind, Wk or (w, d), Dl, Dr): shapes the fused first step takes
    (2, "xxz", 2, 2), (2, "xxz", 34, 50), (2, "xxz", 96, 130), (2, "tfim", 62, 18),
    (2, (4, 2), 40, 66), (2, (8, 2), 36, 20), (1, "xxz", 130, 70), (1, "hubbard", 50, 64),
    (1, (8, 4), 32, 33), (1, (3, 2), 2, 7),
]


@pytest.mark.parametrize("cplx", [False, True], ids=["f64", "c64"])
@pytest.mark.parametrize("kind,Wk,Dl,Dr", FUSED_CASES)
def test_fused_first_step_matches_unfused_and_oracle(gpu, kind, Wk, Dl, Dr, cplx):
    """The fused GEMM + mode-product kernels (FP64 real: wl = wr in
    {3,4,5,6,8}; ComplexF64: wl = wr in {3,4,5,6}; physical extent 2 or 4,
    even Dl) against the oracle and against the two-launch path; ragged Dl,
    Dr that are not multiples of the block tile (32 x 64 real, 32 x 32
    complex)."""
    from finitemptools_amd.derivatives import ac2_prime, ac_prime
    from finitemptools_amd.engine import engine
    rng = np.random.default_rng(Dl * 31 + Dr)
    if isinstance(Wk, tuple):
        w, d = Wk
        W = rand(rng, (w, d, w, d), cplx)
    else:
        W = mpo_tensor(Wk, cplx)
        w, d = W.shape[0], W.shape[1]
    L, R = _envs(rng, Dl, w, Dr, w, cplx)
    eng = engine()
    ys = []
    for fused in (True, False):
        eng.set_fusion(fused)
        try:
            if kind == 2:
                x = rand(np.random.default_rng(5), (Dl, d, d, Dr), cplx)
                ys.append(host(ac2_prime(dev(x), dev(W), dev(W), dev(L), dev(R))))
            else:
                x = rand(np.random.default_rng(5), (Dl, d, Dr), cplx)
                ys.append(host(ac_prime(dev(x), dev(W), dev(L), dev(R))))
        finally:
            eng.set_fusion(True)
    ref = C.ac2_prime(x, W, W, L, R) if kind == 2 else C.ac_prime(x, W, L, R)
    assert relerr(ys[0], ref) < TOL
    assert relerr(ys[1], ref) < TOL
    assert relerr(ys[0], ys[1]) < 1e-13


@pytest.mark.parametrize("dt,cplx", [(-0.05j, True), (0.03, False), (-0.3j, True)])
def test_arnoldi_expm_non_hermitian(gpu, dt, cplx):
    """exp(dt H_eff) x for a NON-Hermitian H_eff (random MPO tensor and
    environments) by the Arnoldi exponentiator, against scipy's expm of the
    dense H_eff; the Lanczos exponentiator is not valid here and differs."""
    import scipy.linalg as sla
    from finitemptools_amd.derivatives import HeffOperator
    from finitemptools_amd.krylov import exponentiate
    rng = np.random.default_rng(13)
    Dl, Dr, d, w = 4, 3, 2, 3
    W = rand(rng, (w, d, w, d), cplx) * 0.5
    L = rand(rng, (Dl, w, Dl), cplx) * 0.5
    R = rand(rng, (Dr, w, Dr), cplx) * 0.5
    x = rand(rng, (Dl, d, d, Dr), cplx)
    H = C.dense_heff2(W, W, L, R)
    assert np.linalg.norm(H - H.conj().T) > 0.1 * np.linalg.norm(H)
    ref = sla.expm(dt * H) @ C.vec(x)
    op = HeffOperator(2, dev(L), dev(R), dev(W), dev(W))
    y, info = exponentiate(op, dt, dev(x), tol=1e-12, kdim=12, ishermitian=False)
    assert info == 0
    assert relerr(C.vec(host(y)), ref) < 1e-9
    yl, _ = exponentiate(op, dt, dev(x), tol=1e-12, kdim=12, ishermitian=True)
    assert relerr(C.vec(host(yl)), ref) > 1e-6


def test_arnoldi_expm_equals_lanczos_on_hermitian(gpu):
    """On a Hermitian H_eff both exponentiators give exp(dt H) x."""
    import scipy.linalg as sla
    from finitemptools_amd.derivatives import HeffOperator
    from finitemptools_amd.krylov import exponentiate
    rng = np.random.default_rng(21)
    W = mpo_tensor("tfim", True)  # Hermitian blocks per channel: H_eff Hermitian
    w = W.shape[0]
    Dl, Dr, d = 5, 4, 2
    L = hermitian_env(rng, Dl, w, True)
    R = hermitian_env(rng, Dr, w, True)
    x = rand(rng, (Dl, d, d, Dr), True)
    op = HeffOperator(2, dev(L), dev(R), dev(W), dev(W))
    ref = sla.expm(-0.1j * C.dense_heff2(W, W, L, R)) @ C.vec(x)
    ya, _ = exponentiate(op, -0.1j, dev(x), tol=1e-12, kdim=10, ishermitian=False)
    yl, _ = exponentiate(op, -0.1j, dev(x), tol=1e-12, kdim=10, ishermitian=True)
    assert relerr(C.vec(host(ya)), ref) < 1e-9
    assert relerr(C.vec(host(yl)), ref) < 1e-9


@pytest.mark.parametrize("cplx", [False, True])
def test_every_gemm_configuration(gpu, cplx):
    """Each FP64 GEMM configuration the tuner may choose (tile, K depth,
    split-K, tile order, waves, K-grouped waves), forced in turn, on ragged
    shapes in the layouts of the two H_eff steps (C = A B with A m-contiguous
    and B k- or n-contiguous), against numpy; configurations a shape does not
    admit fall back to tuning."""
    import ctypes
    import torch
    from finitemptools_amd.engine import engine, to_device
    from finitemptools_amd.linalg import gemm
    eng = engine()
    n = ctypes.c_int()
    eng.check(eng.lib.mh_gemm_candidates(0, ctypes.byref(n)))
    rng = np.random.default_rng(17)
    dt = torch.complex128 if cplx else torch.float64
    shapes = [(130, 70, 600), (64, 64, 1100), (200, 90, 40)]
    try:
        for c in range(n.value):
            eng.check(eng.lib.mh_gemm_force_candidate(eng.h, c))
            for M, N, K in shapes:
                A = rand(rng, (M, K), cplx)
                B = rand(rng, (K, N), cplx)
                Bt = rand(rng, (N, K), cplx)
                Ad = to_device(A, "cuda", dt)
                C1 = gemm(Ad, to_device(B, "cuda", dt)).cpu().numpy()
                C2 = gemm(Ad, to_device(Bt, "cuda", dt).t()).cpu().numpy()
                assert relerr(C1, A @ B) < 1e-13, (c, M, N, K)
                assert relerr(C2, A @ Bt.T) < 1e-13, (c, M, N, K)
    finally:
        eng.check(eng.lib.mh_gemm_force_candidate(eng.h, -1))


@pytest.mark.parametrize("cplx", [False, True])
def test_engine_two_site_reduces_to_one_site_with_absorbed_env(gpu, cplx):
    """On the device: ac2_prime(A1.A2) contracted with conj(A2) equals
    ac_prime(A1) with site 2 absorbed into the right environment by
    updateright (ragged bond dimensions, real and complex)."""
    import torch
    from finitemptools_amd import models
    from finitemptools_amd.derivatives import ac2_prime, ac_prime
    from finitemptools_amd.engine import to_device
    from finitemptools_amd.transfer import updateright
    rng = np.random.default_rng(31 + cplx)
    W = (models.tfim_mpo(4, 1.0, 0.8, dtype=np.complex128) if cplx
         else models.heisenberg_xxz_mpo(4))[1]
    w, d = W.shape[0], W.shape[1]
    Dl, Dm, Dr = 40, 56, 48
    A1, A2 = rand(rng, (Dl, d, Dm), cplx), rand(rng, (Dm, d, Dr), cplx)
    L, R = rand(rng, (Dl, w, Dl), cplx), rand(rng, (Dr, w, Dr), cplx)
    x = np.einsum("asm,mtb->astb", A1, A2)
    Wd, A2d = to_device(W), to_device(A2)
    y2 = ac2_prime(to_device(x), Wd, Wd, to_device(L), to_device(R)).cpu().numpy()
    lhs = np.einsum("astb,mtb->asm", y2, A2.conj())
    Rp = updateright(to_device(R), A2d, Wd, A2d)
    rhs = ac_prime(to_device(A1), Wd, to_device(L), Rp).cpu().numpy()
    assert np.linalg.norm(lhs - rhs) <= 1e-12 * np.linalg.norm(rhs)
