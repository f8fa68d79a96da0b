"""mh_ac2_prime_noalloc (static scratch, no allocation) against mh_ac2_prime.

Both entry points are called through the C ABI on the same inputs.  The
allocation-free one takes its GEMM tiles from the heuristic instead of the
tuner, so the bar is the FP64 parity bar of the plain path (1e-12 relative),
not bit equality.  Two calls in a row, the second on a smaller operator, show
that the shared scratch keeps no state that a later call depends on; each
result is also checked against the oracle.
"""
import ctypes

import numpy as np
import pytest

from conftest import model_env, rand
from oracle import contractions as C

pytestmark = pytest.mark.gpu

TOL = 1e-12


def relerr(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _run(gpu, fn, cplx, D1, D2, W, L, R, x):
    import torch
    from finitemptools_amd.engine import colmajor_empty, to_device
    dt = torch.complex128 if cplx else torch.float64
    Ld, Rd, Wd, xd = (to_device(a, dtype=dt) for a in (L, R, W, x))
    y = colmajor_empty(x.shape, dtype=dt)
    y.fill_(float("nan"))
    w, d = W.shape[0], W.shape[1]
    st = fn(gpu.bind_stream(), 1 if cplx else 0, 0, D1, d, d, D2, w, w, w, _vp(xd), _vp(Wd),
            _vp(Wd), _vp(Ld), _vp(Rd), _vp(y))
    gpu.check(st)
    torch.cuda.synchronize()
    return y.cpu().numpy()


def _case(rng, D1, D2, cplx, model="xxz"):
    from finitemptools_amd import models
    if model == "xxz":
        W = models.heisenberg_xxz_mpo(4, Jzz=0.9, hz=0.1)[1]
    else:
        W = models.hubbard_mpo(4, 1.0, 4.0)[1]
    if cplx:
        W = W.astype(np.complex128)
    w, d = W.shape[0], W.shape[1]
    L = model_env(rng, D1, w, model, cplx, 1.0 / D1)
    R = model_env(rng, D2, w, model, cplx, 1.0 / D2)
    x = rand(rng, (D1, d, d, D2), cplx)
    return W, L, R, x


@pytest.mark.parametrize("cplx", [False, True])
def test_noalloc_matches_ac2_prime_twice(gpu, cplx):
    lib = gpu.lib
    rng = np.random.default_rng(7 if cplx else 5)
    # the second call is smaller than the first: nothing of the first call's
    # scratch contents may leak into it
    for D1, D2 in [(192, 160), (37, 29)]:
        print(f"noalloc case cplx={cplx} D1={D1} D2={D2}", flush=True)
        W, L, R, x = _case(rng, D1, D2, cplx)
        y_na = _run(gpu, lib.mh_ac2_prime_noalloc, cplx, D1, D2, W, L, R, x)
        y_ws = _run(gpu, lib.mh_ac2_prime, cplx, D1, D2, W, L, R, x)
        ref = C.ac2_prime(x, W, W, L, R)
        assert relerr(y_na, y_ws) < TOL
        assert relerr(y_na, ref) < TOL


def test_noalloc_hubbard_shape_twice(gpu):
    lib = gpu.lib
    rng = np.random.default_rng(11)
    for D1, D2 in [(96, 128), (40, 24)]:
        print(f"noalloc hubbard D1={D1} D2={D2}", flush=True)
        W, L, R, x = _case(rng, D1, D2, False, "hubbard")
        y_na = _run(gpu, lib.mh_ac2_prime_noalloc, False, D1, D2, W, L, R, x)
        y_ws = _run(gpu, lib.mh_ac2_prime, False, D1, D2, W, L, R, x)
        assert relerr(y_na, y_ws) < TOL
        assert relerr(y_na, C.ac2_prime(x, W, W, L, R)) < TOL


def test_noalloc_rejects_what_it_cannot_hold(gpu):
    """Oversized operators and mixed precisions are errors, raised before any
    launch (the pointers are never dereferenced)."""
    import torch
    lib = gpu.lib
    h = gpu.bind_stream()
    buf = torch.zeros(16, dtype=torch.float64, device="cuda")
    p = _vp(buf)
    # D = 1024, d = 2, w = 5: T2 alone is 1024*4*5*1024*8 B = 168 MB > 64 MiB
    st = lib.mh_ac2_prime_noalloc(h, 0, 0, 1024, 2, 2, 1024, 5, 5, 5, p, p, p, p, p,
                                  _vp(torch.zeros(16, dtype=torch.float64, device="cuda")))
    assert st == 2, st  # MH_ERR_SHAPE
    assert b"static scratch" in lib.mh_last_error(h)
    st = lib.mh_ac2_prime_noalloc(h, 0, 1, 8, 2, 2, 8, 5, 5, 5, p, p, p, p, p,
                                  _vp(torch.zeros(16, dtype=torch.float64, device="cuda")))
    assert st == 7, st  # MH_ERR_UNSUPPORTED
