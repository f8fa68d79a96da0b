"""The Schur-form boundary: the reference's ac_prime on an
AbstractSiteMPOTensor with Vector environments (algorithms/derivatives.jl:7-17)
is the sum over the site tensor's entries of L_i . x . W_ij . R_j (operator
entries) and c_ij L_i . x . R_j (scalar entries).  Evaluated term by term with
the oracle on single-channel tensors, it must equal the engine's dense
ac_prime on the site tensor mh_schur_fold builds and the environments
mh_env_gather assembles from the per-channel matrices."""
import numpy as np
import pytest
import torch

from finitemptools_amd import models
from finitemptools_amd.engine import to_device
from oracle import contractions as OC

pytestmark = pytest.mark.gpu


def _terms(cell):
    wl, d, wr, _ = cell.shape
    out = []
    for i in range(wl):
        for j in range(wr):
            B = cell[i, :, j, :]
            if not np.any(B):
                continue
            c = B[0, 0]
            out.append((i, j, c, None) if np.array_equal(B, c * np.eye(d)) else (i, j, 1.0, B))
    return out


@pytest.mark.parametrize("cplx", [False, True], ids=["real", "complex"])
def test_vector_form_ac_prime_matches_folded_dense(gpu, cplx):
    from finitemptools_amd.derivatives import ac_prime
    from finitemptools_amd.envs import env_gather, env_scatter
    rng = np.random.default_rng(17 + cplx)
    cell = (models.tfim_mpo(4, 1.0, 0.6, dtype=np.complex128) if cplx
            else models.heisenberg_xxz_mpo(4, 1.0, 0.7, 0.3))[1]
    wl, d, wr, _ = cell.shape
    terms = _terms(cell)
    Dl, Dr = 7, 9
    dt = np.complex128 if cplx else np.float64

    def rnd(*shape):
        a = rng.standard_normal(shape)
        return (a + 1j * rng.standard_normal(shape)) if cplx else a

    Lv = [rnd(Dl, Dl) for _ in range(wl)]       # hleft::Vector, one matrix per channel
    Rv = [rnd(Dr, Dr) for _ in range(wr)]
    x = rnd(Dl, d, Dr)
    # derivatives.jl:7-17, one term at a time (single-channel tensors)
    y_ref = np.zeros((Dl, d, Dr), dtype=dt)
    for i, j, c, O in terms:
        Wt = (c * (np.eye(d) if O is None else O)).reshape(1, d, 1, d).astype(dt)
        y_ref += OC.ac_prime(x, Wt, Lv[i].reshape(Dl, 1, Dl), Rv[j].reshape(Dr, 1, Dr))
    # engine: folded site tensor, gathered environments
    W = models.schur_fold(terms, wl, wr, d, dtype=dt)
    L = env_gather([to_device(a) for a in Lv])
    R = env_gather([to_device(a) for a in Rv])
    y = ac_prime(to_device(x), to_device(W), L, R).cpu().numpy()
    assert np.linalg.norm(y - y_ref) <= 1e-12 * np.linalg.norm(y_ref)
    # the gathered layout is E[a, w, a'] and scatters back bit for bit
    E = L.cpu().numpy()
    for k in range(wl):
        assert np.array_equal(E[:, k, :], Lv[k])
    back = env_scatter(L)
    for k in range(wl):
        assert torch.equal(back[k].cpu(), torch.as_tensor(Lv[k]))


@pytest.mark.parametrize("cplx", [False, True], ids=["real", "complex"])
def test_vector_form_env_update_matches_folded_dense(gpu, cplx):
    """mpohamiltonian/transfer.jl:48-75: hnew[j] += updateleft(hold[i], A,
    W_ij, B) over the site tensor's entries (identity entries contract only
    the MPS legs) equals the engine's updateleft on the folded site tensor and
    gathered environment, scattered back to one matrix per channel."""
    from finitemptools_amd.envs import env_gather, env_scatter
    from finitemptools_amd.transfer import updateleft
    rng = np.random.default_rng(29 + cplx)
    cell = (models.tfim_mpo(4, 1.0, 0.6, dtype=np.complex128) if cplx
            else models.heisenberg_xxz_mpo(4, 1.0, 0.7, 0.3))[1]
    wl, d, wr, _ = cell.shape
    terms = _terms(cell)
    Dl, Dr = 6, 8
    dt = np.complex128 if cplx else np.float64

    def rnd(*shape):
        a = rng.standard_normal(shape)
        return (a + 1j * rng.standard_normal(shape)) if cplx else a

    hold = [rnd(Dl, Dl) for _ in range(wl)]
    A, B = rnd(Dl, d, Dr), rnd(Dl, d, Dr)
    ref = [np.zeros((Dr, Dr), dtype=dt) for _ in range(wr)]
    for i, j, c, O in terms:
        Wt = (c * (np.eye(d) if O is None else O)).reshape(1, d, 1, d).astype(dt)
        ref[j] += OC.updateleft(hold[i].reshape(Dl, 1, Dl), A, Wt, B).reshape(Dr, Dr)
    W = models.schur_fold(terms, wl, wr, d, dtype=dt)
    E = updateleft(env_gather([to_device(h) for h in hold]), to_device(A), to_device(W),
                   to_device(B))
    got = [m.cpu().numpy() for m in env_scatter(E)]
    scale = max(np.linalg.norm(r) for r in ref)
    for g, r in zip(got, ref):
        assert np.linalg.norm(g - r) <= 1e-12 * scale


def _entries(cell, rng):
    """oracle Schur entries of a cell, every coefficient rescaled"""
    wl, d, wr, _ = cell.shape
    out = {}
    for i in range(wl):
        for j in range(wr):
            B = cell[i, :, j, :]
            if not np.any(B):
                continue
            f = rng.uniform(0.5, 1.5)
            c = B[0, 0]
            out[(i, j)] = (("scal", f * c) if np.array_equal(B, c * np.eye(d))
                           else ("op", (f * B).reshape(1, d, 1, d)))
    return out


def _terms_of(entries, d):
    return [(i, j, v, None) if k == "scal" else (i, j, 1.0, v[0, :, 0, :])
            for (i, j), (k, v) in entries.items()]


CELLS = {
    "xxz_field": lambda: models.heisenberg_xxz_mpo(4, 1.0, 0.7, 0.3)[1],
    "hubbard": lambda: models.hubbard_mpo(4, 1.0, 4.0)[1],
    "tfim_c64": lambda: models.tfim_mpo(4, 1.0, 0.6, dtype=np.complex128)[1],
}


@pytest.mark.parametrize("name", sorted(CELLS))
def test_two_site_schur_form_matches_term_by_term(gpu, name):
    """ac2_prime(x, h1::AbstractSiteMPOTensor, h2, hleft::Vector, hright::Vector)
    (derivatives.jl:24-49, all four branches, oracle.schur) equals the
    engine's ac2_prime on the site tensors mh_schur_fold builds and the
    environments mh_env_gather assembles."""
    from finitemptools_amd.derivatives import ac2_prime
    from finitemptools_amd.envs import env_gather
    from oracle import schur as OS
    rng = np.random.default_rng(51)
    cell = CELLS[name]()
    cplx = np.iscomplexobj(cell)
    dt = np.complex128 if cplx else np.float64
    wl, d, wr, _ = cell.shape
    h1, h2 = _entries(cell, rng), _entries(cell, rng)
    Dl, Dr = 11, 13

    def rnd(*shape):
        a = rng.standard_normal(shape)
        return (a + 1j * rng.standard_normal(shape)) if cplx else a

    hl = [rnd(Dl, 1, Dl) for _ in range(wl)]
    hr = [rnd(Dr, 1, Dr) for _ in range(wr)]
    x = rnd(Dl, d, d, Dr)
    y_ref = OS.ac2_prime(x, h1, h2, hl, hr)
    W1 = models.schur_fold(_terms_of(h1, d), wl, wr, d, dtype=dt)
    W2 = models.schur_fold(_terms_of(h2, d), wl, wr, d, dtype=dt)
    L = env_gather([to_device(a[:, 0, :]) for a in hl])
    R = env_gather([to_device(a[:, 0, :]) for a in hr])
    y = ac2_prime(to_device(x), to_device(W1), to_device(W2), L, R).cpu().numpy()
    assert np.linalg.norm(y - y_ref) <= 1e-12 * np.linalg.norm(y_ref)


@pytest.mark.parametrize("name", sorted(CELLS))
def test_schur_updateright_matches_term_by_term(gpu, name):
    """mpohamiltonian/transfer.jl:48-60: hnew[i] += updateright(hold[j], A,
    h[i,j] | Os[i,j] * nothing, B) (oracle.schur) equals the engine's
    updateright on the folded site tensor, scattered to one matrix per
    channel; the left update (:62-75) likewise."""
    from finitemptools_amd.envs import env_gather, env_scatter
    from finitemptools_amd.transfer import updateleft, updateright
    from oracle import schur as OS
    rng = np.random.default_rng(52)
    cell = CELLS[name]()
    cplx = np.iscomplexobj(cell)
    dt = np.complex128 if cplx else np.float64
    wl, d, wr, _ = cell.shape
    h = _entries(cell, rng)
    W = models.schur_fold(_terms_of(h, d), wl, wr, d, dtype=dt)
    Dl, Dr = 9, 12

    def rnd(*shape):
        a = rng.standard_normal(shape)
        return (a + 1j * rng.standard_normal(shape)) if cplx else a

    A, B = rnd(Dl, d, Dr), rnd(Dl, d, Dr)
    hold_r = [rnd(Dr, 1, Dr) for _ in range(wr)]
    ref = OS.updateright(hold_r, A, h, B, wl)
    E = updateright(env_gather([to_device(a[:, 0, :]) for a in hold_r]), to_device(A),
                    to_device(W), to_device(B))
    got = [m.cpu().numpy() for m in env_scatter(E)]
    scale = max(np.linalg.norm(r) for r in ref if r is not None)
    for g, r in zip(got, ref):
        r = np.zeros_like(g) if r is None else r[:, 0, :]
        assert np.linalg.norm(g - r) <= 1e-12 * scale
    hold_l = [rnd(Dl, 1, Dl) for _ in range(wl)]
    ref = OS.updateleft(hold_l, A, h, B, wr)
    E = updateleft(env_gather([to_device(a[:, 0, :]) for a in hold_l]), to_device(A),
                   to_device(W), to_device(B))
    got = [m.cpu().numpy() for m in env_scatter(E)]
    scale = max(np.linalg.norm(r) for r in ref if r is not None)
    for g, r in zip(got, ref):
        r = np.zeros_like(g) if r is None else r[:, 0, :]
        assert np.linalg.norm(g - r) <= 1e-12 * scale
