"""Pins the CPU oracle against closed forms, brute force and exact diagonalisation.

The reference (Julia + TensorKit) cannot run here and stores no golden vectors
(test/runtests.jl computes every check at run time), so each oracle function is
pinned by an independent construction chosen so that a dropped term, a
transposed operand or a missing conjugation fails.
"""
import numpy as np
import pytest
import scipy.linalg as sla

from conftest import hermitian_env, rand
from oracle import contractions as C
from oracle import dmrg as OD
from oracle import ed
from oracle.lanczos import simple_lanczos_solver
from finitemptools_amd import models


@pytest.mark.parametrize("cplx", [False, True])
def test_ac_prime_matches_kronecker_bruteforce(cplx):
    rng = np.random.default_rng(1)
    Dl, d, Dr, wl, wr = 3, 2, 4, 3, 2
    L, R = rand(rng, (Dl, wl, Dl), cplx), rand(rng, (Dr, wr, Dr), cplx)
    W, x = rand(rng, (wl, d, wr, d), cplx), rand(rng, (Dl, d, Dr), cplx)
    H = C.dense_heff1(W, L, R)
    assert np.allclose(C.vec(C.ac_prime(x, W, L, R)), H @ C.vec(x), atol=1e-12)


@pytest.mark.parametrize("cplx", [False, True])
def test_ac2_prime_matches_kronecker_bruteforce(cplx):
    rng = np.random.default_rng(2)
    Dl, d1, d2, Dr, wl, wm, wr = 2, 2, 3, 3, 2, 3, 2
    L, R = rand(rng, (Dl, wl, Dl), cplx), rand(rng, (Dr, wr, Dr), cplx)
    W1, W2 = rand(rng, (wl, d1, wm, d1), cplx), rand(rng, (wm, d2, wr, d2), cplx)
    x = rand(rng, (Dl, d1, d2, Dr), cplx)
    H = C.dense_heff2(W1, W2, L, R)
    assert np.allclose(C.vec(C.ac2_prime(x, W1, W2, L, R)), H @ C.vec(x), atol=1e-12)


def test_c_prime_matches_kronecker_bruteforce():
    rng = np.random.default_rng(3)
    L, R, x = rand(rng, (3, 4, 3), True), rand(rng, (5, 4, 5), True), rand(rng, (3, 5), True)
    H = C.dense_heff0(L, R)
    assert np.allclose(C.vec(C.c_prime(x, L, R)), H @ C.vec(x), atol=1e-12)


def test_w_d_one_reduces_to_two_gemms():
    """w = d = 1: y = L x R^T, the textbook matrix product."""
    rng = np.random.default_rng(4)
    L, R, x = rand(rng, (5, 1, 5)), rand(rng, (6, 1, 6)), rand(rng, (5, 1, 6))
    W = np.full((1, 1, 1, 1), 2.5)
    y = C.ac_prime(x, W, L, R)
    assert np.allclose(y[:, 0, :], 2.5 * L[:, 0, :] @ x[:, 0, :] @ R[:, 0, :].T)


def test_identity_heff_is_identity():
    D, d = 4, 3
    L = np.eye(D)[:, None, :]
    R = np.eye(D)[:, None, :]
    W = np.eye(d)[None, :, None, :]
    x = np.random.default_rng(5).standard_normal((D, d, D))
    assert np.allclose(C.ac_prime(x, W, L, R), x)


@pytest.mark.parametrize("cplx", [False, True])
def test_env_updates_give_expectation_value(cplx):
    """<psiA|H|psiB> three ways: right-to-left updateright, left-to-right
    updateleft, and the dense vectors against the dense MPO."""
    L = 5
    mpo = models.heisenberg_xxz_mpo(L, J=1.0, Jzz=0.7, hz=0.3)
    a = OD.randommps(L, 2, 4, seed=7, cplx=cplx)
    b = OD.randommps(L, 2, 4, seed=8, cplx=cplx)
    e_right = C.expectation(a, mpo, b)
    hold = C.l_LL(np.complex128 if cplx else np.float64)
    for i in range(L):
        hold = C.updateleft(hold, a[i], mpo[i], b[i])
    e_left = hold.reshape(-1)[0]
    H = OD.mpo_to_dense(mpo)
    e_dense = np.vdot(OD.to_dense(a), H @ OD.to_dense(b))
    assert np.allclose(e_right, e_dense, rtol=1e-12)
    assert np.allclose(e_left, e_dense, rtol=1e-12)


def test_env_identity_and_scalar_variants():
    rng = np.random.default_rng(9)
    A, B = rand(rng, (3, 2, 4), True), rand(rng, (3, 2, 4), True)
    hold = rand(rng, (3, 2, 3), True)
    Wid = np.zeros((2, 2, 2, 2))
    for k in range(2):
        Wid[k, :, k, :] = np.eye(2)
    assert np.allclose(C.updateleft(hold, A, 1.5, B), 1.5 * C.updateleft(hold, A, Wid, B))
    holdr = rand(rng, (4, 2, 4), True)
    assert np.allclose(C.updateright(holdr, A, None, B), C.updateright(holdr, A, Wid, B))


def test_heff_at_center_is_energy():
    """<x|ac_prime(x)> with envs from updateleft/updateright = <psi|H|psi>."""
    L, c = 6, 3
    mpo = models.heisenberg_xxz_mpo(L, Jzz=0.5)
    psi = OD.right_canonicalize(OD.randommps(L, 2, 4, seed=11))
    for i in range(c):  # move the orthogonality centre to site c by QR
        Dl, d, Dr = psi[i].shape
        q, r = np.linalg.qr(psi[i].reshape(Dl * d, Dr, order="F"))
        psi[i] = q.reshape(Dl, d, q.shape[1], order="F")
        psi[i + 1] = np.tensordot(r, psi[i + 1], axes=([1], [0]))
    env = OD.ExpectationCache(mpo, psi)
    for i in range(c):
        env.updateleft(i)
    v = OD.to_dense(psi)
    e_dense = np.vdot(v, OD.mpo_to_dense(mpo) @ v)
    y = C.ac_prime(psi[c], mpo[c], env.hstorage[c], env.hstorage[c + 1])
    assert np.isclose(np.vdot(psi[c], y), e_dense, rtol=1e-10)
    assert np.isclose(np.linalg.norm(psi[c]), np.linalg.norm(v))


def test_lanczos_full_dimension_matches_eigh():
    rng = np.random.default_rng(12)
    n = 8
    M = rng.standard_normal((n, n))
    M = M + M.T
    ev, vec, info, k = simple_lanczos_solver(lambda x: M @ x, rng.standard_normal(n), "SR", n, 1e-14)
    w, U = np.linalg.eigh(M)
    assert abs(ev - w[0]) < 1e-10
    assert abs(abs(np.dot(vec, U[:, 0])) - 1) < 1e-8


def test_lanczos_early_exit_on_invariant_subspace():
    """A start vector inside a 2-dim invariant subspace stops after 2 steps
    (beta < btol, simple_lanczos.jl:104-107) with info = 0."""
    M = np.diag([1.0, 2.0, 3.0, 4.0])
    v = np.array([1.0, 1.0, 0.0, 0.0])
    ev, vec, info, k = simple_lanczos_solver(lambda x: M @ x, v, "SR", 10, 1e-12)
    assert k == 2 and info == 0 and abs(ev - 1.0) < 1e-12


@pytest.mark.parametrize("builder,edh,L", [
    (lambda L: models.heisenberg_xxz_mpo(L, J=1.0, Jzz=0.8, hz=0.2), lambda L: ed.xxz_hamiltonian(L, 1.0, 0.8, 0.2), 6),
    (lambda L: models.tfim_mpo(L, J=1.0, h=0.7), lambda L: ed.tfim_hamiltonian(L, 1.0, 0.7), 6),
    (lambda L: models.j1j2_mpo(L, 1.0, 0.35), lambda L: ed.j1j2_hamiltonian(L, 1.0, 0.35), 6),
    (lambda L: models.hubbard_mpo(L, 1.0, 4.0), lambda L: ed.hubbard_hamiltonian(L, 1.0, 4.0), 3),
])
def test_mpo_builders_match_exact_hamiltonians(builder, edh, L):
    H_mpo = OD.mpo_to_dense(builder(L))
    H_ed = edh(L).toarray()
    assert np.allclose(H_mpo, H_ed, atol=1e-12)


def test_mpo_bond_dimensions():
    assert models.heisenberg_xxz_mpo(4)[1].shape[0] == 5
    assert models.tfim_mpo(4)[1].shape[0] == 3
    assert models.j1j2_mpo(4, 1.0, 0.5)[1].shape[0] == 8
    assert models.hubbard_mpo(4)[1].shape[0] == 6


def test_truncation_semantics():
    """truncation.jl:27-47: relative cutoff first, then D; 2-norm error."""
    s = np.array([1.0, 0.5, 1e-3, 1e-10, 1e-12])
    k, err = OD.MPSTruncation(D=10, eps=1e-8).truncate(s)
    assert k == 3 and np.isclose(err, np.hypot(1e-10, 1e-12))
    k, err = OD.MPSTruncation(D=2, eps=1e-8).truncate(s)
    assert k == 2 and np.isclose(err, np.linalg.norm(s[2:]))


def test_right_canonical_form():
    psi = OD.right_canonicalize(OD.randommps(8, 2, 6, seed=3))
    for A in psi[1:]:
        M = A.reshape(A.shape[0], -1, order="F")
        assert np.allclose(M @ M.conj().T, np.eye(A.shape[0]), atol=1e-12)
    assert np.isclose(np.linalg.norm(OD.to_dense(psi)), 1.0)


def test_oracle_dmrg2_reaches_ed_energy():
    """Mirrors "ground state dmrg algorithms" (runtests.jl:773-778): DMRG2 vs
    exact diagonalisation to 1e-8."""
    L = 8
    mpo = models.heisenberg_xxz_mpo(L, J=1.0, Jzz=1.0)
    psi = OD.right_canonicalize(OD.randommps(L, 2, 16, seed=0))
    env = OD.ExpectationCache(mpo, psi)
    alg = OD.DMRG2(trunc=OD.MPSTruncation(D=16, eps=1e-10))
    energies, delta = OD.compute(env, alg, nsweeps=4)
    e0 = ed.ground_energy(ed.xxz_hamiltonian(L))[0]
    assert abs(energies[-1] - e0) < 1e-8


def test_exponentiate_closed_form_reference():
    """scipy's expm on the dense H_eff is the closed form the engine's
    Lanczos exponentiate is compared with."""
    rng = np.random.default_rng(13)
    L = hermitian_env(rng, 3, 2, True)
    R = hermitian_env(rng, 3, 2, True)
    W = np.zeros((2, 2, 2, 2), complex)
    W[0, :, 0, :] = np.eye(2)
    W[1, :, 1, :] = np.array([[0.3, 0.1], [0.1, -0.2]])
    H = C.dense_heff1(W, L, R)
    assert np.allclose(H, H.conj().T)
    x = rand(rng, 18, True)
    y = sla.expm(-0.05j * H) @ x
    assert np.isclose(np.linalg.norm(y), np.linalg.norm(x))


@pytest.mark.parametrize("cplx", [False, True])
def test_mpo_grad_is_the_derivative_of_the_heff_bilinear_form(cplx):
    """G[w,s',w',s] = d <A| ac_prime(B, W) > / dW[w,s',w',s]: brute force over
    unit MPO tensors (pins the index order), and linearity sum G.W."""
    rng = np.random.default_rng(21)
    Dl, d, Dr, wl, wr = 3, 2, 4, 3, 2
    L, R = rand(rng, (Dl, wl, Dl), cplx), rand(rng, (Dr, wr, Dr), cplx)
    A, B = rand(rng, (Dl, d, Dr), cplx), rand(rng, (Dl, d, Dr), cplx)
    G = C.mpo_grad(L, A, B, R)
    assert G.shape == (wl, d, wr, d)
    for idx in np.ndindex(G.shape):
        E = np.zeros((wl, d, wr, d))
        E[idx] = 1.0
        assert abs(G[idx] - np.vdot(A, C.ac_prime(B, E, L, R))) < 1e-12
    W = rand(rng, (wl, d, wr, d), cplx)
    assert abs(np.sum(G * W) - np.vdot(A, C.ac_prime(B, W, L, R))) < 1e-11


def test_hubbard_sector_energy_minimum_is_the_ground_energy():
    """The sector restriction partitions the Hilbert space: the lowest sector
    energy is the full ground energy, and sector sizes add to 4^L."""
    from math import comb
    L = 4
    es = [ed.hubbard_sector_energy(L, a, b) for a in range(L + 1) for b in range(L + 1)]
    full = ed.ground_energy(ed.hubbard_hamiltonian(L))[0]
    assert abs(min(es) - full) < 1e-12
    assert sum(comb(L, a) * comb(L, b) for a in range(L + 1) for b in range(L + 1)) == 4 ** L
    # one electron: the lowest free-particle band energy -2t cos(pi/(L+1))
    assert abs(ed.hubbard_sector_energy(L, 1, 0) + 2 * np.cos(np.pi / (L + 1))) < 1e-12


@pytest.mark.parametrize("h", [1.0, 0.4])
def test_tfim_quench_closed_form_matches_ed(h):
    """The free-fermion <sz_j(t)> of the TFIM quench from |+>^L (the oracle for
    the L = 100 TDVP measurement) equals exact evolution of the L = 8 state
    vector; at t = 0 it vanishes; a flipped field sign moves it."""
    import scipy.sparse.linalg as spla
    from oracle.freefermion import tfim_quench_sz
    L = 8
    H = ed.tfim_hamiltonian(L, 1.0, h)
    psi0 = np.full(2 ** L, 2 ** (-L / 2), dtype=np.complex128)
    times = [0.0, 0.3, 1.0, 2.5]
    ff = tfim_quench_sz(L, 1.0, h, times)
    for t, row in zip(times, ff):
        psi = spla.expm_multiply(-1j * t * H, psi0)
        ref = np.array([np.vdot(psi, ed._site(ed.SZ, i, L, 2) @ psi).real for i in range(L)])
        assert np.allclose(row, ref, atol=1e-12)
    assert np.abs(ff[0]).max() < 1e-14
    assert np.abs(tfim_quench_sz(L, 1.0, -h, [1.0])[0] - ff[2]).max() > 0.1


@pytest.mark.parametrize("cplx", [False, True])
def test_ch_expansion_terms_pinned_by_ac_prime(cplx):
    """The CH expansion terms are ac_prime with one environment left open:
    closing the fused bond of the left term with any R[b, w', b'] (and the
    right term with any L[a, w, a']) gives alpha * ac_prime(A, W, L, R); a
    closing environment of full rank pins every entry."""
    rng = np.random.default_rng(17)
    Dl, d, Dr, wl, wr, alpha = 3, 2, 4, 3, 5, 0.37
    A = rand(rng, (Dl, d, Dr), cplx)
    W = rand(rng, (wl, d, wr, d), cplx)
    L = rand(rng, (Dl, wl, Dl), cplx)
    R = rand(rng, (Dr, wr, Dr), cplx)
    P = C.ch_expansion_term_left(L, A, W, alpha)
    assert P.shape == (Dl, d, wr * Dr)
    Rfull = rand(rng, (wr * Dr, wr, Dr), cplx)              # rank wr*Dr on the fused bond
    ref = alpha * C.ac_prime(A, W, L, Rfull)
    closed = np.einsum("asf,bf->asb", P, Rfull.reshape(wr * Dr, wr * Dr, order="F"))
    assert np.allclose(closed, ref, atol=1e-12)
    Q = C.ch_expansion_term_right(A, W, R, alpha)
    assert Q.shape == (wl * Dl, d, Dr)
    Lfull = rand(rng, (wl * Dl, wl, Dl), cplx)
    ref = alpha * C.ac_prime(A, W, Lfull, R)
    closed = np.einsum("af,fsb->asb", Lfull.reshape(wl * Dl, wl * Dl, order="F"), Q)
    assert np.allclose(closed, ref, atol=1e-12)


@pytest.mark.parametrize("cplx", [False, True])
def test_two_site_heff_reduces_to_one_site_with_absorbed_env(cplx):
    """ac2_prime(A1.A2, W1, W2, L, R) contracted with conj(A2) over (s2', b)
    is ac_prime(A1, W1, L, R') with R' = updateright(R, A2, W2, A2): the
    two-site, one-site and environment-update restatements agree (any A2)."""
    rng = np.random.default_rng(23 + cplx)
    W = (models.tfim_mpo(4, 1.0, 0.8, dtype=np.complex128) if cplx
         else models.heisenberg_xxz_mpo(4))[1]
    w, d = W.shape[0], W.shape[1]
    Dl, Dm, Dr = 3, 5, 4

    def rnd(*shape):
        a = rng.standard_normal(shape)
        return (a + 1j * rng.standard_normal(shape)) if cplx else a

    A1, A2 = rnd(Dl, d, Dm), rnd(Dm, d, Dr)
    L, R = rnd(Dl, w, Dl), rnd(Dr, w, Dr)
    x = np.einsum("asm,mtb->astb", A1, A2)
    y2 = C.ac2_prime(x, W, W, L, R)
    lhs = np.einsum("astb,mtb->asm", y2, A2.conj())
    rhs = C.ac_prime(A1, W, L, C.updateright(R, A2, W, A2))
    assert np.allclose(lhs, rhs, rtol=1e-12, atol=1e-12 * np.abs(rhs).max())


# ---------------------------------------------------------- Schur (Vector) forms

def schur_entries(cell, rng=None):
    """Schur site tensor of a dense cell: identity blocks as ("scal", c),
    other nonzero blocks as ("op", O[1, d, 1, d]); rng rescales every entry
    (so two sites differ and no coefficient is 1)."""
    wl, d, wr, _ = cell.shape
    out = {}
    for i in range(wl):
        for j in range(wr):
            B = cell[i, :, j, :]
            if not np.any(B):
                continue
            f = rng.uniform(0.5, 1.5) if rng is not None else 1.0
            c = B[0, 0]
            if np.array_equal(B, c * np.eye(d)):
                out[(i, j)] = ("scal", f * c)
            else:
                out[(i, j)] = ("op", (f * B).reshape(1, d, 1, d))
    return out


SCHUR_CELLS = {
    "xxz_field": lambda: OD_models().heisenberg_xxz_mpo(4, 1.0, 0.7, 0.3)[1],
    "hubbard": lambda: OD_models().hubbard_mpo(4, 1.0, 4.0)[1],
    "tfim_c64": lambda: OD_models().tfim_mpo(4, 1.0, 0.6, dtype=np.complex128)[1],
}


def OD_models():
    from finitemptools_amd import models
    return models


@pytest.mark.parametrize("name", sorted(SCHUR_CELLS))
def test_schur_ac2_prime_four_branches_match_kronecker(name):
    """oracle.schur.ac2_prime (derivatives.jl:24-49, the four branches) equals
    the Kronecker-product H_eff of the folded site tensors (independent of
    the index strings); every branch is exercised."""
    from oracle import schur as OS
    rng = np.random.default_rng(41)
    cell = SCHUR_CELLS[name]()
    cplx = np.iscomplexobj(cell)
    wl, d, wr, _ = cell.shape
    h1, h2 = schur_entries(cell, rng), schur_entries(cell, rng)
    branches = set()
    for (i, j), e1 in h1.items():
        for (j2, k), e2 in h2.items():
            if j2 == j:
                branches.add((e1[0], e2[0]))
    assert branches == {("scal", "scal"), ("scal", "op"), ("op", "scal"), ("op", "op")}
    Dl, Dr = 3, 4
    hl = [rand(rng, (Dl, 1, Dl), cplx) for _ in range(wl)]
    hr = [rand(rng, (Dr, 1, Dr), cplx) for _ in range(wr)]
    x = rand(rng, (Dl, d, d, Dr), cplx)
    y = OS.ac2_prime(x, h1, h2, hl, hr)
    W1, W2 = OS.fold(h1, wl, wr, d), OS.fold(h2, wl, wr, d)
    L = np.concatenate(hl, axis=1)
    R = np.concatenate(hr, axis=1)
    H = C.dense_heff2(W1, W2, L, R)
    assert np.allclose(C.vec(y), H @ C.vec(x), rtol=0, atol=1e-12 * np.abs(H).max() * x.size)
    # and the one-site Vector form (derivatives.jl:7-17) likewise
    x1 = rand(rng, (Dl, d, Dr), cplx)
    y1 = OS.ac_prime(x1, h1, hl, hr)
    H1 = C.dense_heff1(W1, L, R)
    assert np.allclose(C.vec(y1), H1 @ C.vec(x1), rtol=0, atol=1e-12 * np.abs(H1).max() * x1.size)


def test_schur_ac2_prime_detects_a_dropped_branch():
    """A restatement missing one branch fails the Kronecker check above:
    each branch carries weight in this case."""
    from oracle import schur as OS
    rng = np.random.default_rng(42)
    cell = SCHUR_CELLS["xxz_field"]()
    wl, d, wr, _ = cell.shape
    h1, h2 = schur_entries(cell, rng), schur_entries(cell, rng)
    hl = [rand(rng, (3, 1, 3)) for _ in range(wl)]
    hr = [rand(rng, (4, 1, 4)) for _ in range(wr)]
    x = rand(rng, (3, d, d, 4))
    full = OS.ac2_prime(x, h1, h2, hl, hr)
    for kind in ("scal", "op"):
        h1_cut = {k: v for k, v in h1.items() if v[0] != kind}
        assert np.linalg.norm(OS.ac2_prime(x, h1_cut, h2, hl, hr) - full) > 1e-3 * np.linalg.norm(full)


@pytest.mark.parametrize("name", sorted(SCHUR_CELLS))
def test_schur_env_updates_match_folded_dense(name):
    """mpohamiltonian/transfer.jl:48-75 term by term equals the dense
    updateleft / updateright (mpo/transfer.jl:26-34, pinned three ways
    above) on the folded site tensor and the gathered environment."""
    from oracle import schur as OS
    rng = np.random.default_rng(43)
    cell = SCHUR_CELLS[name]()
    cplx = np.iscomplexobj(cell)
    wl, d, wr, _ = cell.shape
    h = schur_entries(cell, rng)
    W = OS.fold(h, wl, wr, d)
    Dl, Dr = 4, 5
    A, B = rand(rng, (Dl, d, Dr), cplx), rand(rng, (Dl, d, Dr), cplx)
    hold_l = [rand(rng, (Dl, 1, Dl), cplx) for _ in range(wl)]
    got = OS.updateleft(hold_l, A, h, B, wr)
    ref = C.updateleft(np.concatenate(hold_l, axis=1), A, W, B)
    for j in range(wr):
        assert np.allclose(got[j][:, 0, :], ref[:, j, :], atol=1e-12 * np.abs(ref).max())
    hold_r = [rand(rng, (Dr, 1, Dr), cplx) for _ in range(wr)]
    got = OS.updateright(hold_r, A, h, B, wl)
    ref = C.updateright(np.concatenate(hold_r, axis=1), A, W, B)
    for i in range(wl):
        assert np.allclose(got[i][:, 0, :], ref[:, i, :], atol=1e-12 * np.abs(ref).max())


def test_schur_updateright_chain_gives_the_energy():
    """<psi|H|psi> of an XXZ chain (field included) by Schur updateright from
    the right boundary (ones on the last channel, mpohamiltonian/transfer.jl:3-24)
    equals the exact-diagonalisation Hamiltonian's expectation."""
    from finitemptools_amd import models
    from oracle import schur as OS
    L = 5
    mpo = models.heisenberg_xxz_mpo(L, 1.0, 0.7, 0.3)
    psi = OD.randommps(L, 2, 4, seed=5)
    hold = None
    for i in range(L - 1, -1, -1):
        cell = mpo[i]
        wl, d, wr, _ = cell.shape
        h = schur_entries(cell)
        if hold is None:
            hold = [np.ones((1, 1, 1)) if j == wr - 1 else np.zeros((1, 1, 1)) for j in range(wr)]
        hold = OS.updateright(hold, psi[i], h, psi[i], wl)
        hold = [np.zeros_like(hold[0]) if t is None else t for t in hold]
    E = hold[0].reshape(-1)[0]
    v = OD.to_dense(psi)
    H = ed.xxz_hamiltonian(L, 1.0, 0.7, 0.3)
    assert abs(E - v @ (H @ v)) < 1e-12 * abs(v @ v) * np.abs(H).max() * 4
