"""Block-sparse plans on the device: grouped GEMM / block mode product against
the numpy execution of the same tables and the dense oracle (1e-12), mixed
precisions within their stated tolerances, the device Lanczos against the
projected dense H_eff, and two-site DMRG on charge-conserving Hubbard and XXZ
chains against exact diagonalisation in the charge sector (1e-8, as the
reference's DMRG tests, test/runtests.jl:423-451)."""
import json
import os

import numpy as np
import pytest
import torch

from finitemptools_amd import blockmps as bm
from finitemptools_amd import blocksparse as bs
from finitemptools_amd import models
from oracle import contractions as OC
from plan_exec import execute

pytestmark = pytest.mark.gpu

ED = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ed_energies.json")))


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def rand_space(rng, charges, lo, hi):
    return bs.Space([(q, int(rng.integers(lo, hi + 1))) for q in charges])


def reachable(S, phys, n):
    acc = {}
    for q, _ in S.sectors:
        st = [q]
        for _ in range(n):
            st = [bs.qadd(x, p) for x in st for p in phys]
        for x in st:
            acc[x] = 1
    return list(acc)


def heff_case(seed, lo, hi):
    rng = np.random.default_rng(seed)
    phys = bm.hubbard_charges()
    mpo = bm.BlockMPO(models.hubbard_mpo(8, 1.0, 4.0), phys)
    Sl = rand_space(rng, [(a, b) for a in range(1, 5) for b in range(1, 5)], lo, hi)
    Sr = rand_space(rng, reachable(Sl, phys, 2), lo, hi)
    xl = bs.BlockLayout(Sl, phys, 2, Sr)
    Lay, Ray = bs.EnvLayout(Sl, mpo.qw[3]), bs.EnvLayout(Sr, mpo.qw[5])
    x = rng.standard_normal(xl.size)
    Le = rng.standard_normal(Lay.size) / np.sqrt(hi * 6)
    Re = rng.standard_normal(Ray.size) / np.sqrt(hi * 6)
    tab = bs.plan_heff2(xl, Lay, Ray, mpo[3], mpo[4])
    return tab, xl, x, Le, Re


def dev(a):
    return torch.as_tensor(a, dtype=torch.float64).cuda()


@pytest.mark.parametrize("lo,hi", [(1, 7), (20, 90)])
def test_heff2_plan_device_matches_tables(gpu, lo, hi):
    """Uneven sector sizes (1..7: every block smaller than one tile; 20..90:
    ragged multi-tile blocks)."""
    tab, xl, x, Le, Re = heff_case(1, lo, hi)
    ref = execute(tab, [x, np.zeros(xl.size), Le, Re, None, None])[1]
    y = torch.full((xl.size,), float("nan"), dtype=torch.float64, device="cuda")
    plan = bs.Plan(tab, "f64")
    plan.apply(dev(x), dev(Le), dev(Re), y)
    assert rel(y.cpu().numpy(), ref) < 1e-12
    assert abs(plan.flops() - tab.flops()) <= 1e-9 * tab.flops()


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("f32", 1e-5), ("bf16x3", 1e-4)])
def test_every_tile_variant(gpu, prec, tol):
    """Each grouped-GEMM tile variant (shape x K depth) the tuner may pick,
    forced on every stage, against the numpy execution of the tables; ragged
    sector sizes so that every shape has partial tiles."""
    tab, xl, x, Le, Re = heff_case(5, 3, 150)
    ref = execute(tab, [x, np.zeros(xl.size), Le, Re, None, None])[1]
    plan = bs.Plan(tab, prec)
    assert plan.variants() == (5 if prec == "f64" else 3)
    for v in range(plan.variants()):
        plan.force_variant(v)
        y = torch.full((xl.size,), float("nan"), dtype=torch.float64, device="cuda")
        plan.apply(dev(x), dev(Le), dev(Re), y)
        assert rel(y.cpu().numpy(), ref) < tol, f"variant {v}"
    plan.force_variant(-1)
    y = torch.zeros(xl.size, dtype=torch.float64, device="cuda")
    plan.apply(dev(x), dev(Le), dev(Re), y)   # tuned
    assert rel(y.cpu().numpy(), ref) < tol


def test_heff2_plan_device_matches_dense_oracle(gpu):
    tab, xl, x, Le, Re = heff_case(2, 2, 9)
    y = torch.zeros(xl.size, dtype=torch.float64, device="cuda")
    bs.Plan(tab).apply(dev(x), dev(Le), dev(Re), y)
    mpo = bm.BlockMPO(models.hubbard_mpo(8, 1.0, 4.0), bm.hubbard_charges())
    Lay, Ray = bs.EnvLayout(xl.lspace, mpo.qw[3]), bs.EnvLayout(xl.rspace, mpo.qw[5])
    xd = bs.BlockTensor(xl, torch.as_tensor(x)).to_dense()
    ref = OC.ac2_prime(xd, mpo[3], mpo[4], bs.BlockEnv(Lay, torch.as_tensor(Le)).to_dense(),
                       bs.BlockEnv(Ray, torch.as_tensor(Re)).to_dense())
    got = bs.BlockTensor(xl, y.cpu()).to_dense()
    assert rel(got, ref) < 1e-12


@pytest.mark.parametrize("prec,tol", [("f32", 1e-5), ("bf16x3", 1e-4)])
def test_heff2_plan_mixed_precision(gpu, prec, tol):
    tab, xl, x, Le, Re = heff_case(3, 20, 90)
    ref = execute(tab, [x, np.zeros(xl.size), Le, Re, None, None])[1]
    y = torch.zeros(xl.size, dtype=torch.float64, device="cuda")
    bs.Plan(tab, prec).apply(dev(x), dev(Le), dev(Re), y)
    err = rel(y.cpu().numpy(), ref)
    assert err < tol
    assert err > 1e-12  # the reduced-precision kernels did run


def test_env_and_two_site_plans_device(gpu):
    rng = np.random.default_rng(4)
    phys = bm.hubbard_charges()
    mpo = bm.BlockMPO(models.hubbard_mpo(8, 1.0, 4.0), phys)
    Sl = rand_space(rng, [(a, b) for a in range(1, 4) for b in range(1, 4)], 3, 70)
    Sr = rand_space(rng, reachable(Sl, phys, 1), 3, 70)
    al = bs.BlockLayout(Sl, phys, 1, Sr)
    A, B = rng.standard_normal(al.size), rng.standard_normal(al.size)
    W = mpo[3]
    Lay, Lnew = bs.EnvLayout(Sl, mpo.qw[3]), bs.EnvLayout(Sr, mpo.qw[4])
    Le = rng.standard_normal(Lay.size)
    tab = bs.plan_updateleft(al, Lay, W, Lnew)
    ref = execute(tab, [B, np.zeros(Lnew.size), Le, A, None, None])[1]
    y = torch.zeros(Lnew.size, dtype=torch.float64, device="cuda")
    bs.Plan(tab).apply(dev(B), dev(Le), dev(A), y)
    assert rel(y.cpu().numpy(), ref) < 1e-12
    Ray, Rnew = bs.EnvLayout(Sr, mpo.qw[4]), bs.EnvLayout(Sl, mpo.qw[3])
    Re = rng.standard_normal(Ray.size)
    tab = bs.plan_updateright(al, Ray, W, Rnew)
    ref = execute(tab, [B, np.zeros(Rnew.size), Re, A, None, None])[1]
    y = torch.zeros(Rnew.size, dtype=torch.float64, device="cuda")
    bs.Plan(tab).apply(dev(B), dev(Re), dev(A), y)
    assert rel(y.cpu().numpy(), ref) < 1e-12
    # two-site product over the bond Sr
    Sr2 = rand_space(rng, reachable(Sr, phys, 1), 3, 70)
    l2 = bs.BlockLayout(Sr, phys, 1, Sr2)
    A2 = rng.standard_normal(l2.size)
    xl = bs.BlockLayout(Sl, phys, 2, Sr2)
    tab = bs.plan_two_site(al, l2, xl)
    ref = execute(tab, [A, np.zeros(xl.size), A2, None, None, None])[1]
    y = torch.zeros(xl.size, dtype=torch.float64, device="cuda")
    bs.Plan(tab).apply(dev(A), dev(A2), None, y)
    assert rel(y.cpu().numpy(), ref) < 1e-12


def test_block_lanczos_matches_projected_dense_heff(gpu):
    """Environments of a random half-filled L=6 state: the device Lanczos on
    the block plan reaches the lowest eigenvalue of the dense H_eff restricted
    to the charge-allowed entries."""
    phys = bm.hubbard_charges()
    mpo = bm.BlockMPO(models.hubbard_mpo(6, 1.0, 4.0), phys)
    psi = bm.randommps(6, phys, (3, 3), D=8, seed=1).right_canonicalize()
    env = bm.BlockExpectationCache(mpo, psi)
    env.updateleft(0)
    env.updateleft(1)
    x = bm.two_site(psi[2], psi[3])
    L, R = env.hstorage[2], env.hstorage[4]
    tab = bs.plan_heff2(x.layout, L.layout, R.layout, mpo[2], mpo[3])
    n = x.layout.size
    ev, vec, info, nit = bs.Plan(tab).lanczos(L.data, R.data, x.data, min(n, 120), 1e-12,
                                              reorth=True)
    H = OC.dense_heff2(mpo[2], mpo[3], L.to_dense(), R.to_dense())
    # positions of the block entries in the oracle's vectorisation
    probe = bs.BlockTensor(x.layout, torch.arange(1, n + 1, dtype=torch.float64)).to_dense()
    pos = np.argsort(OC.vec(probe))[-n:]
    pos = pos[np.argsort(OC.vec(probe)[pos])]
    Hs = H[np.ix_(pos, pos)]
    assert np.allclose(Hs, Hs.T, atol=1e-10 * np.abs(Hs).max())
    e0 = np.linalg.eigvalsh(Hs)[0]
    assert abs(ev - e0) < 1e-9 * max(1.0, abs(e0))


def run_block_dmrg(mpo_host, phys, psi, D, nsweeps, prec="f64", eps=1e-10):
    from finitemptools_amd.dmrg import DMRG2
    from finitemptools_amd.truncation import MPSTruncation
    mpo = bm.BlockMPO(mpo_host, phys)
    alg = DMRG2(trunc=MPSTruncation(D=D, eps=eps), prec=prec)
    energies, delta = bm.ground_state(psi, mpo, alg, nsweeps=nsweeps)
    return energies, delta, mpo, alg


def test_block_dmrg2_hubbard_half_filling_from_product_state(gpu):
    """Charge-density-wave product start (the runtests.jl:324-337 pattern):
    empty and doubly occupied sites alternating, N_up = N_dn = 3."""
    phys = bm.hubbard_charges()
    empty, double = phys.index((0, 0)), phys.index((1, 1))
    psi = bm.prodmps(phys, [empty, double] * 3)
    energies, delta, mpo, _ = run_block_dmrg(models.hubbard_mpo(6, 1.0, 4.0), phys, psi, 64, 6)
    assert abs(energies[-1] - ED["hubbard_L6_t1_U4_N3_3"]) < 1e-8
    assert abs(bm.expectation(mpo, psi) - energies[-1]) < 1e-8


def test_block_dmrg2_hubbard_L8_random_start(gpu):
    phys = bm.hubbard_charges()
    psi = bm.randommps(8, phys, (4, 4), D=16, seed=0).right_canonicalize()
    energies, delta, _, _ = run_block_dmrg(models.hubbard_mpo(8, 1.0, 4.0), phys, psi, 256, 6)
    assert abs(energies[-1] - ED["hubbard_L8_t1_U4_N4_4"]) < 1e-8


def test_block_dmrg2_xxz_u1(gpu):
    phys = bm.sz_charges()
    up, dn = phys.index((1,)), phys.index((-1,))
    psi = bm.prodmps(phys, [up, dn] * 6)
    energies, delta, _, _ = run_block_dmrg(models.heisenberg_xxz_mpo(12), phys, psi, 64, 6)
    assert abs(energies[-1] - ED["xxz_L12_J1_D1"]) < 1e-8


@pytest.mark.parametrize("prec", ["f64", "f32", "bf16x3"])
def test_native_plan_equals_python_plan_on_device(gpu, prec):
    """A plan from the C planner and one uploaded from the Python tables give
    bitwise the same H_eff x (same tables, same tile variant), and match the
    numpy execution of the tables."""
    rng = np.random.default_rng(8)
    phys = bm.hubbard_charges()
    mpo = bm.BlockMPO(models.hubbard_mpo(8, 1.0, 4.0), phys)
    Sl = rand_space(rng, [(a, b) for a in range(1, 5) for b in range(1, 5)], 3, 70)
    Sr = rand_space(rng, reachable(Sl, phys, 2), 3, 70)
    xl = bs.BlockLayout(Sl, phys, 2, Sr)
    Lay, Ray = bs.EnvLayout(Sl, mpo.qw[3]), bs.EnvLayout(Sr, mpo.qw[5])
    args = (xl, Lay, Ray, mpo[3], mpo[4], prec)
    x = rng.standard_normal(xl.size)
    Le = rng.standard_normal(Lay.size) / 20
    Re = rng.standard_normal(Ray.size) / 20
    ys = []
    for planner in ("native", "python"):
        plan = bs.make_plan("heff2", *args, planner=planner)
        plan.force_variant(0)
        y = torch.full((xl.size,), float("nan"), dtype=torch.float64, device="cuda")
        plan.apply(dev(x), dev(Le), dev(Re), y)
        ys.append(y.cpu().numpy())
    assert np.array_equal(ys[0], ys[1])
    ref = execute(bs.plan_heff2(*args[:5]), [x, np.zeros(xl.size), Le, Re, None, None])[1]
    assert rel(ys[0], ref) < {"f64": 1e-12, "f32": 1e-5, "bf16x3": 1e-4}[prec]


def test_native_env_and_two_site_plans_on_device(gpu):
    """updateleft / updateright / two-site plans from the C planner against
    the dense oracle."""
    rng = np.random.default_rng(9)
    phys = bm.hubbard_charges()
    mpo = bm.BlockMPO(models.hubbard_mpo(8, 1.0, 4.0), phys)
    Sl = rand_space(rng, [(a, b) for a in range(1, 5) for b in range(1, 5)], 2, 30)
    Sm = rand_space(rng, reachable(Sl, phys, 1), 2, 30)
    Sr = rand_space(rng, reachable(Sm, phys, 1), 2, 30)
    al, l2 = bs.BlockLayout(Sl, phys, 1, Sm), bs.BlockLayout(Sm, phys, 1, Sr)
    A = bs.BlockTensor(al, dev(rng.standard_normal(al.size)))
    B = bs.BlockTensor(al, dev(rng.standard_normal(al.size)))
    A2 = bs.BlockTensor(l2, dev(rng.standard_normal(l2.size)))
    W = mpo[3]
    Lay, Lnew = bs.EnvLayout(Sl, mpo.qw[3]), bs.EnvLayout(Sm, mpo.qw[4])
    Le = bs.BlockEnv(Lay, dev(rng.standard_normal(Lay.size)))
    y = torch.zeros(Lnew.size, dtype=torch.float64, device="cuda")
    bs.make_plan("updateleft", al, Lay, W, Lnew, "f64", planner="native").apply(
        B.data, Le.data, A.data, y)
    ref = OC.updateleft(Le.to_dense(), A.to_dense(), W, B.to_dense())
    assert rel(bs.BlockEnv(Lnew, y.cpu()).to_dense(), ref) < 1e-12
    Ray, Rnew = bs.EnvLayout(Sm, mpo.qw[4]), bs.EnvLayout(Sl, mpo.qw[3])
    Re = bs.BlockEnv(Ray, dev(rng.standard_normal(Ray.size)))
    y = torch.zeros(Rnew.size, dtype=torch.float64, device="cuda")
    bs.make_plan("updateright", al, Ray, W, Rnew, "f64", planner="native").apply(
        B.data, Re.data, A.data, y)
    ref = OC.updateright(Re.to_dense(), A.to_dense(), W, B.to_dense())
    assert rel(bs.BlockEnv(Rnew, y.cpu()).to_dense(), ref) < 1e-12
    xl = bs.BlockLayout(Sl, phys, 2, Sr)
    y = torch.zeros(xl.size, dtype=torch.float64, device="cuda")
    bs.make_plan("two_site", al, l2, xl, planner="native").apply(A.data, A2.data, None, y)
    ref = np.einsum("asm,mtb->astb", A.to_dense(), A2.to_dense())
    assert rel(bs.BlockTensor(xl, y.cpu()).to_dense(), ref) < 1e-12


def test_plain_per_sector_plans_match_grouped(gpu):
    """The per-sector loop (one strided GEMM launch per output block and
    K-segment) and the grouped launches run the same tables: H_eff, both
    environment updates and the two-site product agree to 1e-12, and against
    the numpy execution of the tables."""
    from finitemptools_amd.engine import engine
    tab, xl, x, Le, Re = heff_case(11, 5, 60)
    ref = execute(tab, [x, np.zeros(xl.size), Le, Re, None, None])[1]
    plan = bs.Plan(tab)
    ys = []
    for plain in (False, True):
        engine().set_bs_plain(plain)
        try:
            y = torch.zeros(xl.size, dtype=torch.float64, device="cuda")
            plan.apply(dev(x), dev(Le), dev(Re), y)
            ys.append(y.cpu().numpy())
        finally:
            engine().set_bs_plain(False)
    assert rel(ys[1], ref) < 1e-12
    assert rel(ys[1], ys[0]) < 1e-12
    rng = np.random.default_rng(8)
    phys = bm.hubbard_charges()
    mpo = bm.BlockMPO(models.hubbard_mpo(8, 1.0, 4.0), phys)
    Sl = rand_space(rng, [(a, b) for a in range(1, 4) for b in range(1, 4)], 3, 50)
    Sr = rand_space(rng, reachable(Sl, phys, 1), 3, 50)
    al = bs.BlockLayout(Sl, phys, 1, Sr)
    A, B = rng.standard_normal(al.size), rng.standard_normal(al.size)
    Lay, Lnew = bs.EnvLayout(Sl, mpo.qw[3]), bs.EnvLayout(Sr, mpo.qw[4])
    Ln = rng.standard_normal(Lay.size)
    tab = bs.plan_updateleft(al, Lay, mpo[3], Lnew)
    ref = execute(tab, [B, np.zeros(Lnew.size), Ln, A, None, None])[1]
    with bm.plain_path():
        y = torch.zeros(Lnew.size, dtype=torch.float64, device="cuda")
        bs.Plan(tab).apply(dev(B), dev(Ln), dev(A), y)
    assert rel(y.cpu().numpy(), ref) < 1e-12
    Sr2 = rand_space(rng, reachable(Sr, phys, 1), 3, 50)
    l2 = bs.BlockLayout(Sr, phys, 1, Sr2)
    A2 = rng.standard_normal(l2.size)
    xl2 = bs.BlockLayout(Sl, phys, 2, Sr2)
    tab = bs.plan_two_site(al, l2, xl2)
    ref = execute(tab, [A, np.zeros(xl2.size), A2, None, None, None])[1]
    with bm.plain_path():
        y = torch.zeros(xl2.size, dtype=torch.float64, device="cuda")
        bs.Plan(tab).apply(dev(A), dev(A2), None, y)
    assert rel(y.cpu().numpy(), ref) < 1e-12


def test_plain_per_sector_path_refuses_mixed_plans(gpu):
    from finitemptools_amd import _lib
    tab, xl, x, Le, Re = heff_case(12, 2, 9)
    plan = bs.Plan(tab, "f32")
    with bm.plain_path():
        y = torch.zeros(xl.size, dtype=torch.float64, device="cuda")
        with pytest.raises(_lib.MHError):
            plan.apply(dev(x), dev(Le), dev(Re), y)


def test_plain_per_sector_dmrg_matches_batched(gpu):
    """Two-site DMRG sweeps with per-sector plans and per-sector eigensolves
    reach the same energy as the batched forms (and the sector ED)."""
    phys = bm.hubbard_charges()
    empty, double = phys.index((0, 0)), phys.index((1, 1))
    psi = bm.prodmps(phys, [empty, double] * 3)
    with bm.plain_path():
        energies, _, _, _ = run_block_dmrg(models.hubbard_mpo(6, 1.0, 4.0), phys, psi, 64, 6)
    assert abs(energies[-1] - ED["hubbard_L6_t1_U4_N3_3"]) < 1e-8


# per-site energy bounds of a plain mixed-precision DMRG2 run (the GEMM
# operands rounded to FP32, or split into BF16 pairs: ~1e-6 / ~1e-5 relative
# per matvec), and the criterion after one FP64 polish sweep
MIXED_DE_PER_SITE = {"f32": 1e-5, "bf16x3": 5e-4}
POLISH_DE_PER_SITE = 1e-9


@pytest.mark.parametrize("prec", ["f32", "bf16x3"])
def test_block_dmrg2_hubbard_L8_mixed_then_f64_polish(gpu, prec):
    """Hubbard L=8 at half filling (U1 x U1, charge-density-wave start): six
    mixed-precision DMRG2 sweeps land within MIXED_DE_PER_SITE of the sector
    ED energy per site, and one FP64 sweep from that state reaches
    |dE|/L <= 1e-9 (the mixed-then-FP64 schedule of the D=2048 benchmark)."""
    from finitemptools_amd.dmrg import DMRG2
    from finitemptools_amd.truncation import MPSTruncation
    L = 8
    phys = bm.hubbard_charges()
    empty, double = phys.index((0, 0)), phys.index((1, 1))
    psi = bm.prodmps(phys, [empty, double] * (L // 2))
    mpo = bm.BlockMPO(models.hubbard_mpo(L, 1.0, 4.0), phys)
    e_ed = ED["hubbard_L8_t1_U4_N4_4"]
    print(f"mixed block DMRG2 {prec}", flush=True)
    alg = DMRG2(trunc=MPSTruncation(D=256, eps=1e-10), prec=prec)
    e_mixed, _ = bm.ground_state(psi, mpo, alg, nsweeps=6)
    de_mixed = abs(e_mixed[-1] - e_ed) / L
    assert de_mixed <= MIXED_DE_PER_SITE[prec], de_mixed
    print(f"  mixed dE/L {de_mixed:.2e}; FP64 polish", flush=True)
    alg = DMRG2(trunc=MPSTruncation(D=256, eps=1e-10), prec="f64")
    e_pol, _ = bm.ground_state(psi, mpo, alg, nsweeps=1)
    de_pol = abs(e_pol[-1] - e_ed) / L
    assert de_pol <= POLISH_DE_PER_SITE, de_pol


def test_coupled_form_matches_the_two_site_matrices(gpu):
    """mh_bs_to_coupled on a Hubbard two-site tensor (codomain (a, s1), domain
    (s2, b)) gives the sector matrices of the split's TwoSiteMap (pinned by
    the SVD-split tests) and the dense tensor's (a, s1) x (s2, b) blocks; the
    way back restores the block buffer bit for bit (one- and two-site, every
    split point)."""
    from tests_helpers_rand import rand_layouts
    rng = np.random.default_rng(72)
    phys = bm.hubbard_charges()
    xl, one = rand_layouts(rng, phys)
    x = bs.BlockTensor(xl, torch.as_tensor(rng.standard_normal(xl.size), device="cuda"))
    print("coupled form: two-site", flush=True)
    mats = bs.to_coupled(x, 2)
    tmap = bm.TwoSiteMap(xl)
    ref = tmap.matrices(x.data.cpu().numpy())
    assert sorted(mats) == sorted(ref)
    for q in ref:
        assert np.array_equal(mats[q].cpu().numpy(), ref[q])
    # against the dense tensor: rows (qa, s1) and columns (s2, qb) of charge q
    dense = x.to_dense()
    Sl, Sr = xl.lspace, xl.rspace
    for q, M in mats.items():
        rows = [(qa, s1) for qa, _ in Sl.sectors for s1 in range(xl.d)
                if bs.qadd(qa, phys[s1]) == q and any(
                    bs.qadd(q, phys[s2]) in Sr.dim for s2 in range(xl.d))]
        cols = [(s2, qb) for qb, _ in Sr.sectors for s2 in range(xl.d)
                if bs.qsub(qb, phys[s2]) == q]
        R = np.vstack([np.hstack([dense[Sl.offset[qa]:Sl.offset[qa] + Sl.dim[qa], s1, s2,
                                        Sr.offset[qb]:Sr.offset[qb] + Sr.dim[qb]]
                                  for s2, qb in cols]) for qa, s1 in rows])
        assert np.array_equal(M.cpu().numpy(), R)
    for lay in (xl, one):
        t = bs.BlockTensor(lay, torch.as_tensor(rng.standard_normal(lay.size), device="cuda"))
        for ncod in range(1, lay.nphys + 2):
            print(f"coupled round trip nphys={lay.nphys} ncod={ncod}", flush=True)
            back = bs.from_coupled(lay, ncod, bs.to_coupled(t, ncod))
            assert torch.equal(back.data, t.data)


BONDS = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                    "hubbard_L64_D2048_bonds.json")))["bonds"]


@pytest.mark.parametrize("site", sorted(BONDS, key=int))
def test_d2048_bond_plan_matches_per_sector_oracle(gpu, site):
    """The FP64 grouped plan at the sector sizes of a D = 2048 Hubbard ramp
    (L = 64, U1 x U1; middle bond 47 sectors up to 205, multi-tile blocks)
    against the independent per-coupled-sector restatement
    (oracle/blocksparse.py, derivatives.jl:21 on graded tensors): 1e-12."""
    from oracle import blocksparse as OB
    b = BONDS[site]

    def sp(lst):
        return bs.Space([(tuple(q), int(n)) for q, n in lst])

    phys = bm.hubbard_charges()
    xl = bs.BlockLayout(sp(b["x_l"]), phys, 2, sp(b["x_r"]))
    Lay = bs.EnvLayout(sp(b["L_space"]), [tuple(q) for q in b["L_qw"]])
    Ray = bs.EnvLayout(sp(b["R_space"]), [tuple(q) for q in b["R_qw"]])
    mpo = models.hubbard_mpo(64, 1.0, 4.0)
    W1, W2 = mpo[int(site)], mpo[int(site) + 1]
    rng = np.random.default_rng(int(site))
    Le = rng.standard_normal(Lay.size) / np.sqrt(2048 * 6)
    Re = rng.standard_normal(Ray.size) / np.sqrt(2048 * 6)
    x = rng.standard_normal(xl.size)
    print(f"D=2048 bond {site}: {len(xl.lspace.sectors)} sectors, vector {xl.size}", flush=True)
    plan = bs.make_plan("heff2", xl, Lay, Ray, W1, W2, "f64")
    y = torch.empty(xl.size, dtype=torch.float64, device="cuda")
    plan.apply(dev(x), dev(Le), dev(Re), y)
    got = y.cpu().numpy()

    def env_entries(lay):
        return [((q, w), off, lay.space.dim[q], lay.space.dim[bs.qsub(q, lay.qw[w])])
                for (q, w), off in lay.off.items()]

    xe = [((blk.qa, blk.p[0], blk.p[1], blk.qb), blk.off, blk.rows, blk.cols)
          for blk in xl.blocks]
    yb = OB.ac2_prime(OB.blocks_from_flat(x, xe), W1, W2,
                      OB.blocks_from_flat(Le, env_entries(Lay)),
                      OB.blocks_from_flat(Re, env_entries(Ray)), phys, Lay.qw, Ray.qw,
                      dict(xl.lspace.sectors), dict(xl.rspace.sectors))
    ref = OB.flat_from_blocks(yb, xe, xl.size)
    assert rel(got, ref) < 1e-12


def test_process_exits_cleanly_with_plans_alive(gpu):
    """Plans still referenced at interpreter exit (a module-global cache
    whose last bond prefetched its plans) are collected after the engines'
    atexit teardown; they must not pass the freed handle on (that aborted
    the process at exit): a child process that ends this way exits 0."""
    import subprocess
    import sys
    import textwrap
    code = textwrap.dedent("""
        from finitemptools_amd import blockmps as bm, models
        from finitemptools_amd.dmrg import DMRG2
        from finitemptools_amd.truncation import MPSTruncation
        # the library eigensolver: the parent process (this test run) holds
        # the device's cooperative-launch state, which the child cannot take
        bm.EIG_SOLVER = "library"
        L = 6
        phys = bm.hubbard_charges()
        psi = bm.prodmps(phys, [phys.index((0, 0)), phys.index((1, 1))] * (L // 2))
        env = bm.BlockExpectationCache(bm.BlockMPO(models.hubbard_mpo(L, 1.0, 4.0), phys), psi)
        e, _ = bm.sweep(env, DMRG2(trunc=MPSTruncation(D=16, eps=1e-10)))
        assert env._pre, "the sweep left prefetched plans behind"
        print("swept", e[-1], flush=True)
    """)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    print("child process: block sweep with a module-global cache, then exit", flush=True)
    r = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    assert "swept" in r.stdout
