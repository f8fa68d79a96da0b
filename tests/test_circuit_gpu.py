"""Gate application and operator-string expectations on the device
(circuit/apply_gates.jl, mps/orth.jl, operators/expecs.jl) against the
dense state vector: canonical form with bond singular values = Schmidt
values, one-/two-/three-site gates on adjacent and distant sites (swap
gates) without truncation to 1e-10, and <c O_i O_j> of the dense state."""
import numpy as np
import pytest

from conftest import rand
from oracle import dmrg as OD
from oracle import ed

pytestmark = pytest.mark.gpu


def dense_apply(v, L, d, positions, G):
    """Apply G[s1'..sn', s1..sn] to the dense C-order state (site 0 slowest)."""
    n = len(positions)
    T = v.reshape((d,) * L)
    Gt = np.asarray(G).reshape((d,) * (2 * n), order="F") if np.asarray(G).ndim == 2 else np.asarray(G)
    # move the gate's sites to the front, contract, move back
    rest = [i for i in range(L) if i not in positions]
    T = np.transpose(T, list(positions) + rest)
    Tm = T.reshape(d ** n, -1, order="C")
    Gm = Gt.reshape(d ** n, d ** n, order="F")     # rows (s1' + d s2' ..), cols (s1 + d s2 ..)
    # C-order flattening of (s1..sn) is s1 slowest: convert to the gate's order
    perm = [int(np.ravel_multi_index(np.unravel_index(k, (d,) * n, order="C"), (d,) * n, order="F"))
            for k in range(d ** n)]
    P = np.zeros((d ** n, d ** n))
    P[np.arange(d ** n), perm] = 1.0
    Gc = P.T @ Gm @ P
    Tm = Gc @ Tm
    T = Tm.reshape((d,) * L, order="C")
    T = np.transpose(T, np.argsort(list(positions) + rest))
    return T.reshape(-1)


def random_unitary(rng, n):
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    q, r = np.linalg.qr(A)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def test_mode_product_entry_point(gpu):
    import torch
    from finitemptools_amd.circuit import _mode
    from finitemptools_amd.engine import to_device
    rng = np.random.default_rng(1)
    x = rand(rng, (7, 6, 5), True)
    Mt = rand(rng, (6, 4), True)
    y = _mode(to_device(x, "cuda", torch.complex128), to_device(Mt, "cuda", torch.complex128), 4)
    assert np.allclose(y.cpu().numpy(), np.einsum("piq,ij->pjq", x, Mt), atol=1e-13)


def test_canonicalize_gives_schmidt_values(gpu):
    from finitemptools_amd.circuit import canonicalize
    from finitemptools_amd.mps import randommps
    L = 6
    psi = randommps(L, 2, 4, seed=3, cplx=True)
    v0 = OD.to_dense(psi.to_host())
    canonicalize(psi, normalize=True)
    v1 = OD.to_dense(psi.to_host())
    assert np.allclose(v1, v0 / np.linalg.norm(v0), atol=1e-12)
    for i in range(1, L):
        sv = np.linalg.svd(v1.reshape(2 ** i, -1), compute_uv=False)
        k = psi.s[i].shape[0]
        assert np.allclose(psi.s[i].cpu().numpy(), sv[:k], atol=1e-12)
        assert np.all(sv[k:] < 1e-12)
    for A in psi.to_host()[1:]:
        M = A.reshape(A.shape[0], -1, order="F")
        assert np.allclose(M @ M.conj().T, np.eye(A.shape[0]), atol=1e-12)


def test_gates_match_dense_state(gpu):
    from finitemptools_amd.circuit import apply_circuit, canonicalize
    from finitemptools_amd.mps import randommps
    from finitemptools_amd.truncation import MPSTruncation
    rng = np.random.default_rng(5)
    L, d = 6, 2
    psi = randommps(L, d, 4, seed=4, cplx=True)
    canonicalize(psi, normalize=True)
    v = OD.to_dense(psi.to_host())
    gates = [((2,), random_unitary(rng, 2)),
             ((1, 2), random_unitary(rng, 4)),
             ((0, 4), random_unitary(rng, 4)),
             ((2, 3, 4), random_unitary(rng, 8)),
             ((0, 2, 5), random_unitary(rng, 8)),
             ((3, 5), rng.standard_normal((4, 4)))]        # not unitary: still exact
    errs = apply_circuit(psi, gates, MPSTruncation(D=1 << 20, eps=0.0))
    for pos, G in gates:
        v = dense_apply(v, L, d, pos, G)
    got = OD.to_dense(psi.to_host())
    assert max(errs) < 1e-12
    assert np.allclose(got, v, atol=1e-10 * np.abs(v).max())


def test_expectation_of_operator_strings(gpu):
    from finitemptools_amd.circuit import canonicalize, expectation_term
    from finitemptools_amd.mps import randommps
    L = 7
    psi = randommps(L, 2, 4, seed=8, cplx=True)
    canonicalize(psi, normalize=True)
    v = OD.to_dense(psi.to_host())
    for pos, ops, c in [((1, 4), [ed.SZ, ed.SX], 0.5), ((0,), [ed.SP], 1.0),
                        ((2, 3, 6), [ed.SX, ed.SZ, ed.SM], -2.0)]:
        Od = None
        for p, o in zip(pos, ops):
            Od = ed._site(o, p, L, 2) if Od is None else Od @ ed._site(o, p, L, 2)
        ref = c * np.vdot(v, Od @ v)
        got = expectation_term(psi, pos, ops, c)
        assert abs(got - ref) < 1e-12
