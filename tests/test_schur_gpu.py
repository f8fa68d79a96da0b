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
