"""Batched strided FP64 GEMMs (mh_gemm_batched) against numpy: unrelated
shapes in one launch, transposed and strided operand views, alpha / beta,
shapes below and above the 64 x 64 tile and K = 0."""
import numpy as np
import pytest

from finitemptools_amd.linalg import gemm_batched

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (-0.5, 1.5)])
def test_gemm_batched_matches_numpy(gpu, alpha, beta):
    import torch
    rng = np.random.default_rng(5)
    shapes = [(1, 1, 1), (5, 7, 3), (64, 64, 16), (65, 130, 17), (200, 3, 513), (31, 90, 0),
              (128, 128, 128), (17, 1, 40)]
    tasks, ref = [], []
    for j, (M, N, K) in enumerate(shapes):
        A = rng.standard_normal((M, K) if j % 2 == 0 else (K, M))
        B = rng.standard_normal((K, N) if j % 3 else (N, K))
        C = rng.standard_normal((M, N) if j % 2 else (N, M))
        At = torch.from_numpy(A).cuda()
        Bt = torch.from_numpy(B).cuda()
        Ct = torch.from_numpy(C).cuda()
        a = At if j % 2 == 0 else At.t()
        b = Bt if j % 3 else Bt.t()
        c = Ct if j % 2 else Ct.t()
        tasks.append((a, b, c))
        An = A if j % 2 == 0 else A.T
        Bn = B if j % 3 else B.T
        Cn = C if j % 2 else C.T
        ref.append(alpha * An @ Bn + beta * Cn)
    print(f"mh_gemm_batched: {len(tasks)} tasks, alpha {alpha}, beta {beta}", flush=True)
    gemm_batched(tasks, alpha=alpha, beta=beta)
    for (a, b, c), r, (M, N, K) in zip(tasks, ref, shapes):
        got = c.cpu().numpy()
        scale = max(1.0, np.abs(r).max())
        assert np.abs(got - r).max() <= 1e-13 * scale * max(K, 1), (M, N, K)
