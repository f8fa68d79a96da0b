"""Block-sparse planner against the dense oracle (CPU): the dense-embedding
invariant -- a block-sparse contraction equals the dense contraction of the
embedded tensors, and the dense result has no weight outside the allowed
blocks (charge conservation).  The plan tables run through a numpy executor
of the documented stage semantics."""
import numpy as np
import pytest
import torch

from finitemptools_amd import blockmps as bm
from finitemptools_amd import blocksparse as bs
from finitemptools_amd import models
from oracle import contractions as OC
from plan_exec import execute


def rand_space(rng, charges, lo=1, hi=5):
    return bs.Space([(q, int(rng.integers(lo, hi + 1))) for q in charges])


def reachable(S, phys, nphys=1):
    acc = {}
    for q, d in S.sectors:
        stack = [q]
        for _ in range(nphys):
            stack = [bs.qadd(x, p) for x in stack for p in phys]
        for x in stack:
            acc[x] = 1
    return list(acc)


def rand_block(rng, lay):
    return bs.BlockTensor(lay, torch.as_tensor(rng.standard_normal(lay.size)))


def rand_env(rng, lay):
    return bs.BlockEnv(lay, torch.as_tensor(rng.standard_normal(lay.size)))


def hubbard_case(rng, L=6, site=2):
    phys = bm.hubbard_charges()
    mpo = bm.BlockMPO(models.hubbard_mpo(L, 1.0, 4.0), phys)
    Sl = rand_space(rng, [(a, b) for a in range(1, 4) for b in range(1, 4)])
    return phys, mpo, Sl


def xxz_case(rng, L=6):
    phys = bm.sz_charges()
    mpo = bm.BlockMPO(models.heisenberg_xxz_mpo(L, 1.0, 0.7, 0.1), phys)
    return phys, mpo


def check_block_result(lay, flat, dense_ref):
    got = bs.BlockTensor(lay, torch.as_tensor(flat)).to_dense()
    # the reference has no weight outside the blocks
    back = bs.BlockTensor.from_dense(lay, dense_ref, device="cpu").to_dense()
    assert np.allclose(back, dense_ref, atol=1e-13 * np.abs(dense_ref).max())
    assert np.allclose(got, dense_ref, rtol=1e-12, atol=1e-12 * np.abs(dense_ref).max())


def test_channel_charges_conserve():
    phys = bm.hubbard_charges()
    assert sorted(phys) == [(0, 0), (0, 1), (1, 0), (1, 1)]
    mpo = bm.BlockMPO(models.hubbard_mpo(5, 1.0, 4.0), phys)
    qw = mpo.qw
    assert qw[0] == [(0, 0)] and qw[-1] == [(0, 0)]
    assert sorted(set(qw[2])) == [(-1, 0), (0, -1), (0, 0), (0, 1), (1, 0)]
    with pytest.raises(ValueError):
        bad = [W.copy() for W in models.hubbard_mpo(4, 1.0, 4.0)]
        bad[1][0, 0, 0, 3] = 0.5   # |up,dn> -> |empty>: changes the charge
        bm.BlockMPO(bad, phys)


@pytest.mark.parametrize("site", [0, 2, 4])
def test_heff2_plan_matches_dense_oracle_hubbard(site):
    rng = np.random.default_rng(10 + site)
    L = 6
    phys = bm.hubbard_charges()
    mpo = bm.BlockMPO(models.hubbard_mpo(L, 1.0, 4.0), phys)
    if site == 0:
        Sl = bs.Space([((0, 0), 1)])
    else:
        Sl = rand_space(rng, [(a, b) for a in range(1, 4) for b in range(1, 4)])
    if site + 2 == L:
        Sr = bs.Space([((3, 3), 1)])
    else:
        Sr = rand_space(rng, reachable(Sl, phys, 2))
    xl = bs.BlockLayout(Sl, phys, 2, Sr)
    Lay = bs.EnvLayout(Sl, mpo.qw[site])
    Ray = bs.EnvLayout(Sr, mpo.qw[site + 2])
    x, Le, Re = rand_block(rng, xl), rand_env(rng, Lay), rand_env(rng, Ray)
    tab = bs.plan_heff2(xl, Lay, Ray, mpo[site], mpo[site + 1])
    out = execute(tab, [x.data.numpy(), np.full(xl.size, np.nan), Le.data.numpy(),
                        Re.data.numpy(), None, None])
    ref = OC.ac2_prime(x.to_dense(), mpo[site], mpo[site + 1], Le.to_dense(), Re.to_dense())
    check_block_result(xl, out[1], ref)
    assert tab.flops() > 0


@pytest.mark.parametrize("unreached", [False, True], ids=["hubbard", "unreached_rows"])
def test_heff2_plan_skips_structural_zeros(unreached):
    """The plan leaves out the T1 columns (w, s1, s2) no MPO entry reads and
    the R channels no entry reaches for an output block: same result as the
    plan over every charge-allowed block (and as the dense oracle), fewer
    flops, and the full plan's count is what heff2_reference_flops reports.
    With the doubly occupied output state of the first site switched off in
    W1, the rows (qa, 3, s2') receive nothing and come out as zeros; the C
    planner emits the same tables."""
    rng = np.random.default_rng(77)
    L = 6
    phys = bm.hubbard_charges()
    mpo = bm.BlockMPO(models.hubbard_mpo(L, 1.0, 4.0), phys)
    W1, W2 = np.array(mpo[2]), np.array(mpo[3])
    if unreached:
        W1[:, 3, :, :] = 0.0
    Sl = rand_space(rng, [(a, b) for a in range(1, 4) for b in range(1, 4)])
    Sr = rand_space(rng, reachable(Sl, phys, 2))
    xl = bs.BlockLayout(Sl, phys, 2, Sr)
    Lay, Ray = bs.EnvLayout(Sl, mpo.qw[2]), bs.EnvLayout(Sr, mpo.qw[4])
    x, Le, Re = rand_block(rng, xl), rand_env(rng, Lay), rand_env(rng, Ray)
    bufs = [x.data.numpy(), np.full(xl.size, np.nan), Le.data.numpy(), Re.data.numpy(), None, None]
    lean = bs.plan_heff2(xl, Lay, Ray, W1, W2)
    full = bs.plan_heff2(xl, Lay, Ray, W1, W2, skip_zeros=False)
    y_lean = execute(lean, list(bufs))[1]
    y_full = execute(full, list(bufs))[1]
    ref = OC.ac2_prime(x.to_dense(), W1, W2, Le.to_dense(), Re.to_dense())
    check_block_result(xl, y_lean, ref)
    assert np.allclose(y_lean, y_full, rtol=0, atol=1e-12 * np.abs(y_full).max())
    assert lean.flops() < 0.9 * full.flops()
    assert bs.heff2_reference_flops(xl, Lay, Ray, W1, W2) == full.flops()
    assert lean.sizes[4] < full.sizes[4] and lean.sizes[5] < full.sizes[5]
    if unreached:
        for b in xl.blocks:
            if b.p[0] == 3:
                assert not np.any(y_lean[b.off:b.off + b.rows * b.cols])
    assert_same_tables(bs.host_native_tables("heff2", xl, Lay, Ray, W1, W2, "f64"), lean)


def test_heff2_plan_xxz_u1():
    rng = np.random.default_rng(3)
    phys, mpo = xxz_case(rng)
    Sl = rand_space(rng, [(q,) for q in range(-3, 4)])
    Sr = rand_space(rng, reachable(Sl, phys, 2))
    xl = bs.BlockLayout(Sl, phys, 2, Sr)
    Lay, Ray = bs.EnvLayout(Sl, mpo.qw[2]), bs.EnvLayout(Sr, mpo.qw[4])
    x, Le, Re = rand_block(rng, xl), rand_env(rng, Lay), rand_env(rng, Ray)
    tab = bs.plan_heff2(xl, Lay, Ray, mpo[2], mpo[3])
    out = execute(tab, [x.data.numpy(), np.full(xl.size, np.nan), Le.data.numpy(),
                        Re.data.numpy(), None, None])
    check_block_result(xl, out[1], OC.ac2_prime(x.to_dense(), mpo[2], mpo[3], Le.to_dense(),
                                                Re.to_dense()))


@pytest.mark.parametrize("model", ["hubbard", "xxz"])
def test_env_update_plans_match_dense_oracle(model):
    rng = np.random.default_rng(7)
    if model == "hubbard":
        phys = bm.hubbard_charges()
        mpo = bm.BlockMPO(models.hubbard_mpo(6, 1.0, 4.0), phys)
        Sl = rand_space(rng, [(a, b) for a in range(1, 4) for b in range(1, 4)])
    else:
        phys, mpo = xxz_case(rng)
        Sl = rand_space(rng, [(q,) for q in range(-2, 3)])
    Sr = rand_space(rng, reachable(Sl, phys, 1))
    al = bs.BlockLayout(Sl, phys, 1, Sr)
    A, B = rand_block(rng, al), rand_block(rng, al)
    W = mpo[2]
    # left
    Lay, Lnew = bs.EnvLayout(Sl, mpo.qw[2]), bs.EnvLayout(Sr, mpo.qw[3])
    Le = rand_env(rng, Lay)
    tab = bs.plan_updateleft(al, Lay, W, Lnew)
    out = execute(tab, [B.data.numpy(), np.full(Lnew.size, np.nan), Le.data.numpy(),
                        A.data.numpy(), None, None])
    ref = OC.updateleft(Le.to_dense(), A.to_dense(), W, B.to_dense())
    got = bs.BlockEnv(Lnew, torch.as_tensor(out[1])).to_dense()
    back = bs.BlockEnv.from_dense(Lnew, ref, device="cpu").to_dense()
    assert np.allclose(back, ref, atol=1e-13 * np.abs(ref).max())
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    # right
    Ray, Rnew = bs.EnvLayout(Sr, mpo.qw[3]), bs.EnvLayout(Sl, mpo.qw[2])
    Re = rand_env(rng, Ray)
    tab = bs.plan_updateright(al, Ray, W, Rnew)
    out = execute(tab, [B.data.numpy(), np.full(Rnew.size, np.nan), Re.data.numpy(),
                        A.data.numpy(), None, None])
    ref = OC.updateright(Re.to_dense(), A.to_dense(), W, B.to_dense())
    got = bs.BlockEnv(Rnew, torch.as_tensor(out[1])).to_dense()
    back = bs.BlockEnv.from_dense(Rnew, ref, device="cpu").to_dense()
    assert np.allclose(back, ref, atol=1e-13 * np.abs(ref).max())
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_two_site_plan():
    rng = np.random.default_rng(5)
    phys = bm.hubbard_charges()
    Sl = rand_space(rng, [(a, b) for a in range(1, 3) for b in range(1, 3)])
    Sm = rand_space(rng, reachable(Sl, phys, 1))
    Sr = rand_space(rng, reachable(Sm, phys, 1))
    l1, l2 = bs.BlockLayout(Sl, phys, 1, Sm), bs.BlockLayout(Sm, phys, 1, Sr)
    xl = bs.BlockLayout(Sl, phys, 2, Sr)
    A1, A2 = rand_block(rng, l1), rand_block(rng, l2)
    tab = bs.plan_two_site(l1, l2, xl)
    out = execute(tab, [A1.data.numpy(), np.zeros(xl.size), A2.data.numpy(), None, None, None])
    ref = np.einsum("asm,mtb->astb", A1.to_dense(), A2.to_dense())
    check_block_result(xl, out[1], ref)


def test_max_virtual_spaces_counts():
    """Bond sector dimensions of the largest spaces equal the numbers of
    basis states on either side (capped by the smaller side)."""
    from math import comb
    phys = bm.sz_charges()
    L = 8
    sp = bm.max_virtual_spaces(L, phys, (0,))
    mid = sp[4]
    # 4 sites with 2 S^z = m: C(4, (4+m)/2) states; the right half must complete to 0
    for q, d in mid.sectors:
        (m,) = q
        assert d == comb(4, (4 + m) // 2)
    assert sp[0].sectors == [((0,), 1)] and sp[L].sectors == [((0,), 1)]


def test_random_block_mps_lives_in_its_sector_and_canonicalises():
    """randommps in the sector (N_up, N_dn) = (2, 1) of a 4-site Hubbard chain:
    the dense state has weight only on basis states of that sector; right
    canonicalisation keeps the state (up to its norm) and makes sites 1..L-1
    right-isometric."""
    phys = bm.hubbard_charges()
    psi = bm.randommps(4, phys, (2, 1), D=6, seed=3, device="cpu")
    v0 = psi.to_dense()
    q = np.array(phys)
    for idx in np.argwhere(np.abs(v0.reshape((4,) * 4)) > 1e-14):
        assert tuple(q[idx].sum(0)) == (2, 1)
    psi.right_canonicalize()
    v1 = psi.to_dense()
    assert abs(np.linalg.norm(v1) - 1.0) < 1e-12
    assert abs(abs(np.vdot(v0, v1)) / np.linalg.norm(v0) - 1.0) < 1e-12
    for t in psi.sites[1:]:
        M = t.to_dense().reshape(t.lspace.total, -1)
        assert np.allclose(M @ M.T, np.eye(t.lspace.total), atol=1e-12)
    assert psi.bond_dimension() <= 6


@pytest.mark.parametrize("absorb", ["left", "right"])
def test_sector_split_reconstructs_and_truncates_globally(absorb):
    """Split of a two-site tensor by sector SVDs: without truncation u s v is
    x / ||x|| and u (or v) is an isometry; with D smaller than the rank the
    kept singular values are the D largest of the merged spectrum."""
    from finitemptools_amd.truncation import MPSTruncation
    rng = np.random.default_rng(11)
    phys = bm.hubbard_charges()
    Sl = rand_space(rng, [(a, b) for a in range(1, 3) for b in range(1, 3)], 1, 4)
    Sr = rand_space(rng, reachable(Sl, phys, 2), 1, 4)
    xl = bs.BlockLayout(Sl, phys, 2, Sr)
    x = rng.standard_normal(xl.size)
    tmap = bm.TwoSiteMap(xl)
    mats = tmap.matrices(x)
    full = np.sort(np.concatenate([np.linalg.svd(M, compute_uv=False) for M in mats.values()]))[::-1]
    A1, A2, svals = bm.split_two_site(tmap, x, MPSTruncation(D=10 ** 6, eps=0.0), absorb, "cpu")
    tab = bs.plan_two_site(A1.layout, A2.layout, xl)
    rec = execute(tab, [A1.data.numpy(), np.zeros(xl.size), A2.data.numpy(), None, None, None])[1]
    assert np.allclose(rec, x / np.linalg.norm(x), atol=1e-12)
    iso = A1 if absorb == "right" else A2
    M = iso.to_dense()
    M = M.reshape(-1, M.shape[-1]) if absorb == "right" else M.reshape(M.shape[0], -1).T
    assert np.allclose(M.T @ M, np.eye(M.shape[1]), atol=1e-12)
    D = len(full) // 3
    _, _, svals = bm.split_two_site(tmap, x, MPSTruncation(D=D, eps=0.0), absorb, "cpu")
    kept = np.sort(np.concatenate(list(svals.values())))[::-1]
    assert len(kept) == D
    assert np.allclose(kept, full[:D] / np.linalg.norm(full[:D]), atol=1e-12)


@pytest.mark.parametrize("absorb", ["left", "right"])
def test_density_matrix_split_matches_svd_split(absorb):
    """The reduced-density-matrix split keeps the same singular values and
    the same truncated two-site tensor as the sector SVDs."""
    from finitemptools_amd.truncation import MPSTruncation
    rng = np.random.default_rng(12)
    phys = bm.hubbard_charges()
    Sl = rand_space(rng, [(a, b) for a in range(1, 3) for b in range(1, 3)], 2, 6)
    Sr = rand_space(rng, reachable(Sl, phys, 2), 2, 6)
    xl = bs.BlockLayout(Sl, phys, 2, Sr)
    x = rng.standard_normal(xl.size)
    tmap = bm.TwoSiteMap(xl)
    tr = MPSTruncation(D=40, eps=1e-10)
    a1, a2, s_ref = bm.split_two_site(tmap, x, tr, absorb, "cpu")
    b1, b2, s_rho = bm.split_two_site_rho(tmap, torch.as_tensor(x), tr, absorb)
    assert sorted(s_ref) == sorted(s_rho)
    for c in s_ref:
        assert np.allclose(s_ref[c], s_rho[c], atol=1e-12)
    rec = []
    for A1, A2 in ((a1, a2), (b1, b2)):
        tab = bs.plan_two_site(A1.layout, A2.layout, xl)
        rec.append(execute(tab, [A1.data.numpy(), np.zeros(xl.size), A2.data.numpy(),
                                 None, None, None])[1])
    assert np.allclose(rec[0], rec[1], atol=1e-11)


def test_prodmps_is_the_basis_state():
    phys = bm.hubbard_charges()
    states = [0, 3, 0, 3]
    psi = bm.prodmps(phys, states, device="cpu")
    v = psi.to_dense().reshape((4,) * 4)
    assert v[tuple(states)] == 1.0 and np.count_nonzero(v) == 1
    assert psi.total_charge() == tuple(np.array(phys)[states].sum(0))


# ------------------------------------------------ native planner (C, host only)

def assert_same_tables(ta, tb):
    """Identical plans: sizes, stage kinds and buffers, every table entry;
    the MPO coefficients agree to the last bit."""
    assert ta.sizes == tb.sizes
    assert [k for k, _ in ta.stages] == [k for k, _ in tb.stages]
    for (_, a), (_, b) in zip(ta.stages, tb.stages):
        assert sorted(a) == sorted(b)
        for key in a:
            va, vb = np.asarray(a[key]), np.asarray(b[key])
            assert va.shape == vb.shape, key
            assert np.array_equal(va, vb), key


def native_cases():
    """(name, kind, args) over both models, inner and boundary bonds."""
    rng = np.random.default_rng(21)
    out = []
    for model in ("hubbard", "xxz"):
        L = 6
        if model == "hubbard":
            phys = bm.hubbard_charges()
            mpo = bm.BlockMPO(models.hubbard_mpo(L, 1.0, 4.0), phys)
            inner = [(a, b) for a in range(1, 4) for b in range(1, 4)]
            first, last = (0, 0), (3, 3)
        else:
            phys, mpo = xxz_case(rng, L)
            inner = [(q,) for q in range(-3, 4)]
            first, last = (0,), (0,)
        for site in (0, 2, L - 2):
            Sl = bs.Space([(first, 1)]) if site == 0 else rand_space(rng, inner)
            Sr = bs.Space([(last, 1)]) if site + 2 == L else rand_space(rng, reachable(Sl, phys, 2))
            xl = bs.BlockLayout(Sl, phys, 2, Sr)
            Lay, Ray = bs.EnvLayout(Sl, mpo.qw[site]), bs.EnvLayout(Sr, mpo.qw[site + 2])
            out.append((f"{model}-heff2-{site}", "heff2",
                        (xl, Lay, Ray, mpo[site], mpo[site + 1], "f64")))
        Sl = rand_space(rng, inner)
        Sm = rand_space(rng, reachable(Sl, phys, 1))
        Sr = rand_space(rng, reachable(Sm, phys, 1))
        al = bs.BlockLayout(Sl, phys, 1, Sm)
        out.append((f"{model}-updateleft", "updateleft",
                    (al, bs.EnvLayout(Sl, mpo.qw[2]), mpo[2], bs.EnvLayout(Sm, mpo.qw[3]), "f32")))
        out.append((f"{model}-updateright", "updateright",
                    (al, bs.EnvLayout(Sm, mpo.qw[3]), mpo[2], bs.EnvLayout(Sl, mpo.qw[2]), "bf16x3")))
        l2 = bs.BlockLayout(Sm, phys, 1, Sr)
        out.append((f"{model}-two_site", "two_site", (al, l2, bs.BlockLayout(Sl, phys, 2, Sr))))
    return out


NATIVE_CASES = native_cases()


@pytest.mark.parametrize("case", NATIVE_CASES, ids=[c[0] for c in NATIVE_CASES])
def test_native_planner_builds_the_python_tables(case):
    """The C planner (mh_bs_plan_heff2 / updateleft / updateright / two_site),
    run without a device, lays out the same buffers and emits the same stage
    tables as the Python planner, whose tables the dense oracle pins above."""
    _, kind, args = case
    assert_same_tables(bs.host_native_tables(kind, *args), bs._PY_PLANNERS[kind](*args))


def test_native_planner_result_matches_dense_oracle():
    """Native tables run through the stage executor give ac2_prime."""
    rng = np.random.default_rng(4)
    _, kind, args = NATIVE_CASES[1]
    xl, Lay, Ray, W1, W2, _ = args
    x, Le, Re = rand_block(rng, xl), rand_env(rng, Lay), rand_env(rng, Ray)
    tab = bs.host_native_tables(kind, *args)
    out = execute(tab, [x.data.numpy(), np.full(xl.size, np.nan), Le.data.numpy(),
                        Re.data.numpy(), None, None])
    check_block_result(xl, out[1], OC.ac2_prime(x.to_dense(), W1, W2, Le.to_dense(),
                                                Re.to_dense()))


def test_native_planner_rejects_charge_breaking_mpo():
    """An MPO entry that changes the total charge: both planners refuse it."""
    from finitemptools_amd import _lib
    _, kind, args = NATIVE_CASES[1]
    xl, Lay, Ray, W1, W2, prec = args
    bad = W1.copy()
    bad[0, 0, 0, 3] = 0.5    # |up,dn> -> |empty> with the identity channel
    with pytest.raises(_lib.MHError):
        bs.host_native_tables(kind, xl, Lay, Ray, bad, W2, prec)
    with pytest.raises(ValueError):
        bs.plan_heff2(xl, Lay, Ray, bad, W2)


def test_native_planner_rejects_out_of_range_charges():
    """Charge components are packed 16 bits each: sectors, physical charges
    or MPO channel charges outside (-16384, 16384) are refused rather than
    wrapped into another sector."""
    from finitemptools_amd import _lib
    _, kind, args = NATIVE_CASES[7]     # xxz heff2 at an inner bond, U(1) charges
    xl, Lay, Ray, W1, W2, prec = args
    big = 20000
    shifted = bs.Space([((q[0] + big,), d) for q, d in xl.lspace.sectors])
    xbad = bs.BlockLayout(shifted, xl.phys, 2, xl.rspace)
    with pytest.raises(_lib.MHError):
        bs.host_native_tables(kind, xbad, Lay, Ray, W1, W2, prec)
    Lbad = bs.EnvLayout(xl.lspace, [(q[0] + big,) for q in Lay.qw])
    with pytest.raises(_lib.MHError):
        bs.host_native_tables(kind, xl, Lbad, Ray, W1, W2, prec)


@pytest.mark.parametrize("absorb", ["left", "right"])
@pytest.mark.parametrize("eps", [1e-8, 1e-10])
def test_density_matrix_split_at_a_cutoff_inside_its_noise(absorb, eps, monkeypatch):
    """Sector matrices whose merged spectrum falls from 1 to 1e-13, so the
    relative cutoff lies where sqrt(lambda) of the sector density matrices is
    noise: the block split must keep what the sector SVDs keep
    (truncation.jl:27-47 on the merged spectrum), with the same values."""
    from finitemptools_amd import truncation as tr_mod
    from finitemptools_amd.truncation import MPSTruncation
    rng = np.random.default_rng(13)
    phys = bm.hubbard_charges()
    Sl = rand_space(rng, [(a, b) for a in range(1, 3) for b in range(1, 3)], 3, 6)
    Sr = rand_space(rng, reachable(Sl, phys, 2), 3, 6)
    xl = bs.BlockLayout(Sl, phys, 2, Sr)
    tmap = bm.TwoSiteMap(xl)
    cat = np.zeros(xl.size)
    for c in tmap.charges:
        R, C = tmap.shape[c]
        r = min(R, C)
        top = rng.uniform(-1.0, 0.0)
        s = np.logspace(top, top - 13, r)
        U = np.linalg.qr(rng.standard_normal((R, r)))[0]
        V = np.linalg.qr(rng.standard_normal((C, r)))[0]
        cat[tmap.base[c]:tmap.base[c] + R * C] = ((U * s) @ V.T).ravel()
    x = np.empty(xl.size)
    x[tmap.perm] = cat
    tr = MPSTruncation(D=10 ** 6, eps=eps)
    a1, a2, s_ref = bm.split_two_site(tmap, x, tr, absorb, "cpu")
    # the level-1 values cannot decide this cut; the second level must (no
    # SVD fallback), and it must land on the SVD's kept set and values
    xt = torch.as_tensor(x)
    cat = torch.empty_like(xt)
    cat[tmap.dest("cpu")] = xt
    lam = []
    for c in tmap.charges:
        R, C = tmap.shape[c]
        Mc = cat[tmap.base[c]:tmap.base[c] + R * C].view(R, C).numpy()
        G = Mc @ Mc.T if absorb == "right" else Mc.T @ Mc
        lam.append((np.linalg.eigvalsh(G)[::-1][:min(R, C)], G.shape[0]))
    s2 = np.concatenate([np.maximum(l, 0) for l, _ in lam])
    nz = np.concatenate([np.full(len(l), tr_mod.RHO_NOISE * n * tr_mod.EPS64 * max(l[0], 0))
                         for l, n in lam])
    o = np.argsort(-s2)
    assert not tr_mod.rho_cut_resolved(s2[o], nz[o], tr)

    def no_svd(*a, **k):
        raise AssertionError("the two-level split fell back to the SVD")
    monkeypatch.setattr(bm, "split_two_site", no_svd)
    b1, b2, s_rho = bm.split_two_site_rho(tmap, xt, tr, absorb)
    assert sorted(s_ref) == sorted(s_rho)
    # an SVD fixes singular values to a few eps64 * ||M|| absolute
    smax = max(float(np.max(v)) for v in s_ref.values())
    for c in s_ref:
        assert len(s_ref[c]) == len(s_rho[c])
        assert np.allclose(s_ref[c], s_rho[c], rtol=1e-10, atol=50 * tr_mod.EPS64 * smax)
    # the truncated tensors agree too
    rec = []
    for A1, A2 in ((a1, a2), (b1, b2)):
        tab = bs.plan_two_site(A1.layout, A2.layout, xl)
        rec.append(execute(tab, [A1.data.numpy(), np.zeros(xl.size), A2.data.numpy(),
                                 None, None, None])[1])
    assert np.allclose(rec[0], rec[1], atol=1e-11)


@pytest.mark.parametrize("absorb", ["left", "right"])
def test_density_matrix_split_where_D_binds_reads_values_off_the_projection(absorb, monkeypatch):
    """D binds far above the noise: the density-matrix path runs, and the kept
    singular values (read off s v or u s) match the sector SVDs to 1e-12
    relative, also values three decades below the largest."""
    from finitemptools_amd.truncation import MPSTruncation
    rng = np.random.default_rng(14)
    phys = bm.hubbard_charges()
    Sl = rand_space(rng, [(a, b) for a in range(1, 3) for b in range(1, 3)], 4, 7)
    Sr = rand_space(rng, reachable(Sl, phys, 2), 4, 7)
    xl = bs.BlockLayout(Sl, phys, 2, Sr)
    tmap = bm.TwoSiteMap(xl)
    cat = np.zeros(xl.size)
    for c in tmap.charges:
        R, C = tmap.shape[c]
        r = min(R, C)
        s = np.logspace(0, -12, r) * rng.uniform(0.5, 1.0)
        U = np.linalg.qr(rng.standard_normal((R, r)))[0]
        V = np.linalg.qr(rng.standard_normal((C, r)))[0]
        cat[tmap.base[c]:tmap.base[c] + R * C] = ((U * s) @ V.T).ravel()
    x = np.empty(xl.size)
    x[tmap.perm] = cat
    tr = MPSTruncation(D=30, eps=1e-8)
    _, _, s_ref = bm.split_two_site(tmap, x, tr, absorb, "cpu")

    def no_svd(*a, **k):
        raise AssertionError("the density-matrix path fell back to the SVD")
    monkeypatch.setattr(bm, "split_two_site", no_svd)
    _, _, s_rho = bm.split_two_site_rho(tmap, torch.as_tensor(x), tr, absorb)
    assert sorted(s_ref) == sorted(s_rho)
    for c in s_ref:
        assert np.allclose(s_ref[c], s_rho[c], rtol=1e-12, atol=0)


def x_entries(lay):
    """(charge key, offset, rows, cols) of a two-site layout's blocks"""
    return [((b.qa, b.p[0], b.p[1], b.qb), b.off, b.rows, b.cols) for b in lay.blocks]


def env_entries(lay):
    out = []
    for (q, w), off in lay.off.items():
        Dq = lay.space.dim[q]
        Dk = lay.space.dim[bs.qsub(q, lay.qw[w])]
        out.append(((q, w), off, Dq, Dk))
    return out


@pytest.mark.parametrize("site", [0, 2, 4])
def test_per_sector_oracle_matches_dense_oracle(site):
    """oracle.blocksparse.ac2_prime (the per-sector loop, derivatives.jl:21
    on graded tensors) equals the dense oracle on the dense embedding of
    random Hubbard U1 x U1 blocks, at edge and bulk bonds."""
    from oracle import blocksparse as OB
    rng = np.random.default_rng(60 + site)
    L = 6
    phys = bm.hubbard_charges()
    mpo = bm.BlockMPO(models.hubbard_mpo(L, 1.0, 4.0), phys)
    Sl = (bs.Space([((0, 0), 1)]) if site == 0 else
          rand_space(rng, [(a, b) for a in range(1, 4) for b in range(1, 4)]))
    Sr = (bs.Space([((3, 3), 1)]) if site + 2 == L else rand_space(rng, reachable(Sl, phys, 2)))
    xl = bs.BlockLayout(Sl, phys, 2, Sr)
    Lay, Ray = bs.EnvLayout(Sl, mpo.qw[site]), bs.EnvLayout(Sr, mpo.qw[site + 2])
    x, Le, Re = rand_block(rng, xl), rand_env(rng, Lay), rand_env(rng, Ray)
    y = OB.ac2_prime(OB.blocks_from_flat(x.data.numpy(), x_entries(xl)), mpo[site], mpo[site + 1],
                     OB.blocks_from_flat(Le.data.numpy(), env_entries(Lay)),
                     OB.blocks_from_flat(Re.data.numpy(), env_entries(Ray)), phys,
                     mpo.qw[site], mpo.qw[site + 2], dict(Sl.sectors), dict(Sr.sectors))
    flat = OB.flat_from_blocks(y, x_entries(xl), xl.size)
    ref = OC.ac2_prime(x.to_dense(), mpo[site], mpo[site + 1], Le.to_dense(), Re.to_dense())
    check_block_result(xl, flat, ref)


# ------------------------------------------------ coupled-sector form (host)

def test_coupled_info_matches_the_two_site_map():
    """mh_bs_coupled_info on a two-site layout with codomain (a, s1) and
    domain (s2, b) gives the charges and matrix shapes of the split's
    TwoSiteMap (the (a, s1) | (s2, b) matrices of stable_tsvd!, dmrg.jl:104),
    in ascending charge order, contiguous."""
    rng = np.random.default_rng(70)
    phys = bm.hubbard_charges()
    Sl = rand_space(rng, [(a, b) for a in range(1, 4) for b in range(1, 4)], 1, 6)
    Sr = rand_space(rng, reachable(Sl, phys, 2), 1, 6)
    xl = bs.BlockLayout(Sl, phys, 2, Sr)
    tmap = bm.TwoSiteMap(xl)
    info = bs.coupled_info(xl, 2)
    assert [q for q, *_ in info] == sorted(tmap.charges)
    off = 0
    for q, r, c, o in info:
        assert (r, c) == tmap.shape[q]
        assert o == off
        off += r * c
    assert off == xl.size


@pytest.mark.parametrize("ncod", [1, 2, 3])
def test_coupled_info_one_and_two_site_splits(ncod):
    """Every split point: element counts add up to the layout; for a one-site
    tensor with codomain (a, s) the coupled charges are the right sectors and
    each matrix stacks that sector's blocks (finitemps.jl:130-141)."""
    rng = np.random.default_rng(71 + ncod)
    phys = bm.hubbard_charges()
    Sl = rand_space(rng, [(a, b) for a in range(1, 4) for b in range(1, 4)], 1, 5)
    Sr = rand_space(rng, reachable(Sl, phys, 1), 1, 5)
    one = bs.BlockLayout(Sl, phys, 1, Sr)
    if ncod <= 2:
        info = bs.coupled_info(one, ncod)
        assert sum(r * c for _, r, c, _ in info) == one.size
        if ncod == 2:
            assert [q for q, *_ in info] == [q for q, _ in Sr.sectors if q in one.by_right]
            for q, r, c, _ in info:
                assert c == Sr.dim[q] and r == sum(b.rows for b in one.by_right[q])
        else:
            assert [q for q, *_ in info] == [q for q, _ in Sl.sectors if q in one.group]
    Sr2 = rand_space(rng, reachable(Sl, phys, 2), 1, 5)
    two = bs.BlockLayout(Sl, phys, 2, Sr2)
    info = bs.coupled_info(two, ncod)
    assert sum(r * c for _, r, c, _ in info) == two.size


def test_coupled_info_rejects_bad_split():
    from finitemptools_amd import _lib
    phys = bm.hubbard_charges()
    Sl = bs.Space([((1, 1), 2)])
    lay = bs.BlockLayout(Sl, phys, 2, bs.Space([((2, 2), 3)]))
    with pytest.raises(_lib.MHError):
        bs.coupled_info(lay, 4)
