"""Batched device eigensolver (mh_sytrd_batched / mh_stedc_batched /
mh_ormtr_batched) against LAPACK and the spectral theorem.

Pins: the tridiagonalisation against scipy's LAPACK dsytrd (uplo = 'L': same
Householder convention, so D, E, tau and the stored reflectors agree entry by
entry); eigenvalues against numpy's eigvalsh; eigenvectors by residual
||A v - lambda v|| and orthonormality; the selected-column back-transform
against the full one."""
import numpy as np
import pytest
import scipy.linalg.lapack as lapack

from finitemptools_amd import eigen

pytestmark = pytest.mark.gpu


def density_matrix(rng, n, decay=20.0):
    """rho = M M^T with a graded spectrum, like a sector reduced density matrix."""
    M = rng.standard_normal((n, n)) * np.exp(-decay * np.arange(n) / max(n, 1))
    return M @ M.T


def random_symmetric(rng, n):
    A = rng.standard_normal((n, n))
    return A + A.T


def _mats(rng, sizes):
    out = []
    for j, n in enumerate(sizes):
        out.append(density_matrix(rng, n) if j % 2 == 0 else random_symmetric(rng, n))
    return out


# the largest order of a batch picks the kernel instance: 600 (10 row slots
# per lane, 5 register columns per wavefront), 744 (12 slots, 4 register
# columns), 1100 (18 slots, columns only in LDS)
@pytest.mark.parametrize("sizes", [[1, 2, 3, 5, 17, 64, 100, 257, 600, 33], [744, 3, 700, 65],
                                   [1100, 130, 2]])
def test_sytrd_matches_lapack(gpu, sizes):
    import torch
    rng = np.random.default_rng(1)
    mats = _mats(rng, sizes)
    se = eigen.SectorEigh(sizes, "cuda")
    for j, M in enumerate(mats):
        se.matrix(j).copy_(torch.from_numpy(M))
    from finitemptools_amd import _lib
    from finitemptools_amd.engine import ptr
    h = gpu.bind_stream()
    lib = _lib.load()
    st = lib.mh_sytrd_batched(h, len(sizes), se._n[1], se._a[1], se._v[1], ptr(se.A), ptr(se.D),
                              ptr(se.E), ptr(se.tau))
    _lib.raise_for(st, h)
    _lib.raise_for(lib.mh_eig_check(h, 0), h)
    A = se.A.cpu().numpy()
    D, E, tau = se.D.cpu().numpy(), se.E.cpu().numpy(), se.tau.cpu().numpy()
    for j, (n, M) in enumerate(zip(sizes, mats)):
        a_ref, d_ref, e_ref, tau_ref, info = lapack.dsytrd(M, lower=1)
        assert info == 0
        scale = max(np.abs(M).max(), 1e-300)
        vo, ao = int(se.v_off[j]), int(se.a_off[j])
        # backward-stable reductions agree to O(n eps ||A||)
        tol = 1e-13 * max(n, 10) * scale
        np.testing.assert_allclose(D[vo:vo + n], d_ref, atol=tol, rtol=0)
        np.testing.assert_allclose(E[vo:vo + n - 1], e_ref, atol=tol, rtol=0)
        Am = A[ao:ao + n * n].reshape(n, n).T   # column-major storage
        # a reflector is determined by its column only where that column lies
        # above the rounding level of the trailing block (graded spectra end
        # in columns of pure rounding noise, whose reflectors are arbitrary)
        for i in range(max(n - 2, 0)):
            if abs(e_ref[i]) >= 1e-3 * scale:
                assert abs(tau[vo + i] - tau_ref[i]) <= 1e-8
                np.testing.assert_allclose(Am[i + 2:, i], a_ref[i + 2:, i], atol=1e-8, rtol=0)


def test_exchange_wait_times_out_cleanly(gpu, monkeypatch):
    """The failure path of the tridiagonalisation's exchange between
    workgroups, forced from the host: workgroup 0 withholds its p entries of
    step 5 (MH_SYTRD_DROP) and every wait is bounded by 2000 polls
    (MH_SYTRD_SPIN), so the readers' waits end in the error flag, the launch
    completes and mh_eig_check reports the time-out; with the hooks cleared
    the next reduction equals LAPACK again."""
    import torch
    from finitemptools_amd import _lib
    from finitemptools_amd.engine import ptr
    rng = np.random.default_rng(5)
    sizes = [300, 40]
    mats = _mats(rng, sizes)
    h = gpu.bind_stream()
    lib = _lib.load()

    def run():
        se = eigen.SectorEigh(sizes, "cuda")
        for j, M in enumerate(mats):
            se.matrix(j).copy_(torch.from_numpy(M))
        st = lib.mh_sytrd_batched(h, len(sizes), se._n[1], se._a[1], se._v[1], ptr(se.A),
                                  ptr(se.D), ptr(se.E), ptr(se.tau))
        _lib.raise_for(st, h)
        return se, lib.mh_eig_check(h, 0)

    monkeypatch.setenv("MH_SYTRD_SPIN", "2000")
    monkeypatch.setenv("MH_SYTRD_DROP", "5")
    print("mh_sytrd_batched with a withheld exchange step", flush=True)
    _, st = run()
    assert st == 4  # MH_ERR_HIP
    assert "timed out" in lib.mh_last_error(h).decode()
    monkeypatch.delenv("MH_SYTRD_SPIN")
    monkeypatch.delenv("MH_SYTRD_DROP")
    se, st = run()
    _lib.raise_for(st, h)
    D = se.D.cpu().numpy()
    for j, (n, M) in enumerate(zip(sizes, mats)):
        d_ref = lapack.dsytrd(M, lower=1)[1]
        vo = int(se.v_off[j])
        np.testing.assert_allclose(D[vo:vo + n], d_ref, atol=1e-13 * n * np.abs(M).max(), rtol=0)


@pytest.mark.parametrize("sizes", [[1, 2, 7, 64, 300, 801], list(range(1, 71)), [1100, 63, 1024]])
def test_eigenpairs(gpu, sizes):
    """Eigenvalues vs eigvalsh, residuals and orthonormality; the second case
    spans more matrices than one launch's table holds, the third takes the
    global-memory tridiagonalisation (an order above the LDS form's 1024)."""
    import torch
    rng = np.random.default_rng(len(sizes))
    mats = _mats(rng, sizes)
    res = eigen.sector_eigh([torch.from_numpy(M).cuda() for M in mats])
    for M, (lam, V) in zip(mats, res):
        n = M.shape[0]
        lam, V = lam.cpu().numpy(), V.cpu().numpy()
        scale = max(np.abs(np.linalg.eigvalsh(M)).max(), 1e-300)
        np.testing.assert_allclose(lam, np.linalg.eigvalsh(M), atol=1e-12 * scale, rtol=0)
        assert np.abs(M @ V - V * lam).max() <= 1e-12 * scale * max(1.0, np.sqrt(n))
        assert np.abs(V.T @ V - np.eye(n)).max() <= 1e-12 * n


def test_blocked_back_transform_matches_reflector_by_reflector(gpu, monkeypatch):
    """The blocked (compact WY) back-transformation equals the reflector-by-
    reflector one column for column, for orders around the block size and a
    selection of columns that ends inside a column tile."""
    import torch
    rng = np.random.default_rng(11)
    sizes = [1, 2, 3, 31, 32, 33, 34, 65, 97, 300, 744]
    mats = _mats(rng, sizes)
    out = {}
    for mode in ("blocked", "columns"):
        if mode == "columns":
            monkeypatch.setenv("MH_ORMTR_COLS", "1")
        se = eigen.SectorEigh(sizes, "cuda")
        for j, M in enumerate(mats):
            se.matrix(j).copy_(torch.from_numpy(M))
        se.eigenvalues()
        keep = [max(1, n - 5) for n in sizes]
        out[mode] = [v.cpu().numpy() for v in se.vectors(keep)]
        se.check()
    for n, a, b in zip(sizes, out["blocked"], out["columns"]):
        assert np.abs(a - b).max() <= 1e-13 * max(n, 10), n


def test_selected_columns_and_many_workgroups(gpu):
    """Top-k back-transform equals the same columns of the full one; 40
    matrices of order 1000 need more workgroups than are resident at once,
    so the tridiagonalisation runs as several launches."""
    import torch
    rng = np.random.default_rng(7)
    sizes = [1000] * 40 + [3, 1]
    mats = [density_matrix(rng, n, decay=8.0) for n in sizes]
    full = eigen.sector_eigh([torch.from_numpy(M).cuda() for M in mats])
    se = eigen.SectorEigh(sizes, "cuda")
    for j, M in enumerate(mats):
        se.matrix(j).copy_(torch.from_numpy(M))
    se.eigenvalues()
    keep = [min(n, 250) for n in sizes]
    V = se.vectors(keep)
    se.check()
    for (lam, Vf), Vk, k, M in zip(full, V, keep, mats):
        torch.testing.assert_close(Vk, Vf[:, Vf.shape[1] - k:], atol=0, rtol=0)
        scale = float(lam.abs().max())
        Vk = Vk.cpu().numpy()
        lk = lam.cpu().numpy()[len(lam) - k:]
        assert np.abs(M @ Vk - Vk * lk).max() <= 1e-11 * scale


@pytest.mark.parametrize("absorb", ["left", "right"])
@pytest.mark.parametrize("solver", ["sector", "sector_kept", "library"])
def test_device_split_matches_svd_split(gpu, absorb, solver, monkeypatch):
    """The device density-matrix split (either sector eigensolver) keeps the
    singular values of the host sector SVDs and reproduces the same truncated
    two-site tensor."""
    import torch
    from finitemptools_amd import blockmps as bm
    from finitemptools_amd import blocksparse as bs
    from finitemptools_amd.truncation import MPSTruncation
    from plan_exec import execute
    from test_blocksparse import rand_space, reachable
    monkeypatch.setattr(bm, "EIG_SOLVER", solver)
    rng = np.random.default_rng(31)
    phys = bm.hubbard_charges()
    Sl = rand_space(rng, [(a, b) for a in range(1, 4) for b in range(1, 4)], 4, 30)
    Sr = rand_space(rng, reachable(Sl, phys, 2), 4, 30)
    xl = bs.BlockLayout(Sl, phys, 2, Sr)
    x = rng.standard_normal(xl.size)
    tmap = bm.TwoSiteMap(xl)
    tr = MPSTruncation(D=150, eps=1e-10)
    a1, a2, s_ref = bm.split_two_site(tmap, x, tr, absorb, "cpu")
    b1, b2, s_dev = bm.split_two_site_rho(tmap, torch.as_tensor(x).cuda(), tr, absorb)
    assert sorted(s_ref) == sorted(s_dev)
    for c in s_ref:
        np.testing.assert_allclose(s_dev[c], s_ref[c], atol=1e-12, rtol=0)
    rec = []
    for A1, A2 in ((a1, a2), (b1, b2)):
        tab = bs.plan_two_site(A1.layout, A2.layout, xl)
        rec.append(execute(tab, [A1.data.cpu().numpy(), np.zeros(xl.size), A2.data.cpu().numpy(),
                                 None, None, None])[1])
    np.testing.assert_allclose(rec[1], rec[0], atol=1e-11, rtol=0)


def _graded_bond(seed, top_range, decades, nmin, nmax, dscale=None):
    from finitemptools_amd import blockmps as bm
    from finitemptools_amd import blocksparse as bs
    from test_blocksparse import rand_space, reachable
    rng = np.random.default_rng(seed)
    phys = bm.hubbard_charges()
    Sl = rand_space(rng, [(a, b) for a in range(1, 3) for b in range(1, 3)], nmin, nmax)
    Sr = rand_space(rng, reachable(Sl, phys, 2), nmin, nmax)
    xl = bs.BlockLayout(Sl, phys, 2, Sr)
    tmap = bm.TwoSiteMap(xl)
    cat = np.zeros(xl.size)
    for c in tmap.charges:
        R, C = tmap.shape[c]
        r = min(R, C)
        top = rng.uniform(*top_range)
        s = np.logspace(top, top - decades, r)
        U = np.linalg.qr(rng.standard_normal((R, r)))[0]
        V = np.linalg.qr(rng.standard_normal((C, r)))[0]
        cat[tmap.base[c]:tmap.base[c] + R * C] = ((U * s) @ V.T).ravel()
    x = np.empty(xl.size)
    x[tmap.perm] = cat
    return xl, tmap, x


@pytest.mark.parametrize("absorb", ["left", "right"])
@pytest.mark.parametrize("case", ["noise_cut_1e-8", "noise_cut_1e-10", "D_binds"])
def test_batched_split_matches_sector_svds(gpu, absorb, case):
    """The batched device split (every per-sector product in a few launches,
    both levels of the exact cut) keeps what the reference's sector SVDs keep
    (truncation.jl:27-47 on the merged spectrum), with the same values and the
    same truncated two-site tensor: a cutoff that falls inside the noise of
    sqrt(lambda) (second level, no fallback) and a cut where D binds."""
    import torch
    from finitemptools_amd import blockmps as bm
    from finitemptools_amd import blocksparse as bs
    from finitemptools_amd import truncation as tr_mod
    from finitemptools_amd.truncation import MPSTruncation
    from plan_exec import execute
    if case == "D_binds":
        xl, tmap, x = _graded_bond(14, (-0.3, 0.0), 12, 4, 9)
        tr = MPSTruncation(D=40, eps=1e-8)
    else:
        xl, tmap, x = _graded_bond(13, (-1.0, 0.0), 13, 3, 8)
        tr = MPSTruncation(D=10 ** 6, eps=float(case.split("_")[-1]))
    a1, a2, s_ref = bm.split_two_site(tmap, x, tr, absorb, "cpu")
    out = bm._split_rho_sector(tmap, torch.as_tensor(x).cuda(), tr, absorb)
    assert out is not None, "the batched split handed the bond to the per-sector path"
    b1, b2, s_dev = out
    assert sorted(s_ref) == sorted(s_dev)
    smax = max(float(np.max(v)) for v in s_ref.values())
    for c in s_ref:
        assert len(s_ref[c]) == len(s_dev[c])
        assert np.allclose(s_ref[c], s_dev[c], rtol=1e-10, atol=50 * tr_mod.EPS64 * smax)
    rec = []
    for A1, A2 in ((a1, a2), (b1, b2)):
        tab = bs.plan_two_site(A1.layout, A2.layout, xl)
        rec.append(execute(tab, [A1.data.cpu().numpy(), np.zeros(xl.size), A2.data.cpu().numpy(),
                                 None, None, None])[1])
    assert np.allclose(rec[0], rec[1], atol=1e-11)


def test_block_dmrg_with_sector_eigensolver_matches_ed(gpu, monkeypatch):
    """Two-site DMRG of the Hubbard chain (L = 8, half filling, random start)
    with every bond split by the batched sector eigensolver reaches the
    charge-sector ED energy (the reference's DMRG-vs-ED tolerance, 1e-8) and
    the same energy as the library eigensolver."""
    import json
    import os
    from finitemptools_amd import blockmps as bm
    from finitemptools_amd import models
    from finitemptools_amd.dmrg import DMRG2
    from finitemptools_amd.truncation import MPSTruncation
    ed = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ed_energies.json")))
    phys = bm.hubbard_charges()
    out = {}
    for solver in ("sector", "library"):
        monkeypatch.setattr(bm, "EIG_SOLVER", solver)
        psi = bm.randommps(8, phys, (4, 4), D=16, seed=0).right_canonicalize()
        mpo = bm.BlockMPO(models.hubbard_mpo(8, 1.0, 4.0), phys)
        alg = DMRG2(trunc=MPSTruncation(D=256, eps=1e-10))
        energies, _ = bm.ground_state(psi, mpo, alg, nsweeps=6)
        out[solver] = energies[-1]
    assert abs(out["sector"] - ed["hubbard_L8_t1_U4_N4_4"]) < 1e-8
    assert abs(out["sector"] - out["library"]) < 1e-10


CHILD = r"""
import sys
sys.path.insert(0, {root!r})
import ctypes
import numpy as np
import scipy.linalg as sl
import torch
from finitemptools_amd import _lib
from finitemptools_amd.engine import engine, ptr
rng = np.random.default_rng(3)
sizes = [40, 96, 17]
vo = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
zo = np.concatenate([[0], np.cumsum(np.square(sizes))[:-1]]).astype(np.int64)
d = rng.standard_normal(sum(sizes))
e = rng.standard_normal(sum(sizes))
D, E = torch.as_tensor(d, device="cuda"), torch.as_tensor(e, device="cuda")
Z = torch.empty(int(np.square(sizes).sum()), dtype=torch.float64, device="cuda")
eng = engine(0)
h = eng.bind_stream()
n32 = np.array(sizes, dtype=np.int32)
P32, P64 = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int64)
eng.check(eng.lib.mh_stedc_batched(h, 3, n32.ctypes.data_as(P32), vo.ctypes.data_as(P64),
                                   zo.ctypes.data_as(P64), ptr(D), ptr(E), ptr(Z)))
eng.check(eng.lib.mh_eig_check(h, 3))
lam = D.cpu().numpy()
for n, o in zip(sizes, vo):
    ref = sl.eigh_tridiagonal(d[o:o + n], e[o:o + n - 1], eigvals_only=True)
    assert np.abs(lam[o:o + n] - ref).max() < 1e-12 * np.abs(ref).max()
print("child done", flush=True)
"""


def test_exit_after_sector_eigensolver_is_clean(gpu):
    """A process that ran the batched tridiagonal eigensolver (the side
    streams and rocBLAS handles of mh_stedc_batched) exits with status 0:
    the engines are destroyed from atexit, before the runtime tears down.
    The child runs no cooperative launch (the tridiagonalisation), which the
    device may refuse to a second process while this one holds its
    cooperative-launch resources."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    print("child: batched tridiagonal eigensolver, then interpreter exit", flush=True)
    r = subprocess.run([sys.executable, "-c", CHILD.format(root=root)], capture_output=True,
                       text=True, timeout=240)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    assert "child done" in r.stdout
