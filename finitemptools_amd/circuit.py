"""Gate application on a finite MPS (TEBD-style bond evolution):
circuit/apply_gates.jl and the canonical form with bond singular values that
it needs (mps/orth.jl:5-49), plus expectation values of operator strings
(operators/expecs.jl:28-55).

The state is kept right-canonical with the singular values of every bond in
``psi.s`` (``psi.s[i]``: the bond left of site i, 0-based; ``psi.s[0]`` and
``psi.s[L]`` are [1]).  A gate on sites (j-1, j) contracts the two site
tensors (engine GEMM), applies the gate to the physical legs (engine mode
product), scales by the left bond's singular values, splits by a truncated
FP64 SVD and recomputes the left factor as (gate . two-site tensor) . v^dag,
so that both factors stay right-isometric (bond_evolution,
apply_gates.jl:9-16).  Gates on non-adjacent sites are brought together by
swap gates and moved back (_move!, apply_gates.jl:50-71).
"""
import numpy as np
import torch

from . import _lib
from .engine import as_colmajor, colmajor_empty, dtype_code, engine, ptr, to_device
from .linalg import as_matrix, cm_view, from_matrix, gemm, two_site
from .truncation import MPSTruncation, stable_tsvd


def _mode(x3, Mt, nj):
    """out[p, j, q] = sum_i x3[p, i, q] Mt[i, j] on the engine."""
    x3 = as_colmajor(x3)
    P, ni, Q = x3.shape
    out = colmajor_empty((P, nj, Q), x3.dtype, x3.device)
    e = engine(x3.device)
    h = e.bind_stream()
    e.check(e.lib.mh_mode_product(h, dtype_code(x3), P, ni, nj, Q, ptr(x3), ptr(Mt), ptr(out)),
            "mode_product")
    return out


def _gate_matrix(G, nsites, d, dtype, device):
    """G[s1'..sn', s1..sn] (numpy or torch, physical dimension d) -> the
    device matrix Mt = G_mat^T with G_mat[(s1' + d s2' ..), (s1 + d s2 ..)]."""
    Gh = np.asarray(G.detach().cpu().numpy() if torch.is_tensor(G) else G)
    n = d ** nsites
    if Gh.shape == (n, n):
        Gm = Gh
    else:
        Gm = Gh.reshape(n, n, order="F")
    return to_device(np.asfortranarray(Gm.T), device, dtype)


def apply_gate_tensor(x, G):
    """y[a, s', b] = sum_s G[s', s] x[a, s, b] for an n-site tensor x[a, s1..sn, b]."""
    n = x.dim() - 2
    d = x.shape[1]
    Mt = _gate_matrix(G, n, d, x.dtype, x.device)
    return _unflatten_mid(_mode(_flatten_mid(x), Mt, d ** n), tuple(x.shape))


def _flatten_mid(x):
    """x[a, s1..sn, b] (column-major) -> (a, s, b) with s = s1 + d s2 + ..."""
    x = as_colmajor(x)
    a, b = x.shape[0], x.shape[-1]
    m = int(np.prod(x.shape[1:-1]))
    return x.permute(*reversed(range(x.dim()))).reshape(b, m, a).permute(2, 1, 0)


def _unflatten_mid(y, shape):
    return y.permute(2, 1, 0).reshape(tuple(reversed(shape))).permute(*reversed(range(len(shape))))


def _scale_left(x, s):
    """s[a] x[a, ...]"""
    return x * s.to(x.dtype).reshape((-1,) + (1,) * (x.dim() - 1))


def canonicalize(psi, normalize=False, trunc=None):
    """right_canonicalize! (mps/orth.jl:5-47): a left-to-right SVD sweep, then
    a right-to-left one, storing the singular values of every bond in
    psi.s.  Returns the truncation errors."""
    L = len(psi)
    dev = psi[0].device
    one = torch.ones(1, dtype=torch.float64, device=dev)
    psi.s[0], psi.s[L] = one, one.clone()
    errs = []
    for i in range(L - 1):                       # leftorth!
        u, s, v, err = _tsvd_legs(psi[i], 2, trunc)
        psi[i] = u
        sv = v * s.to(v.dtype).reshape(-1, 1)
        nxt = psi[i + 1]
        psi[i + 1] = cm_view(gemm(sv, as_matrix(nxt, 1)), (sv.shape[0],) + tuple(nxt.shape[1:]))
        psi.s[i + 1] = s
        errs.append(err)
    if normalize:
        psi[L - 1] = psi[L - 1] / torch.linalg.vector_norm(psi[L - 1])
    for i in range(L - 1, 0, -1):                # rightorth!
        u, s, v, err = _tsvd_legs(psi[i], 1, trunc)
        psi[i] = v
        us = u * s.to(u.dtype).reshape(1, -1)
        prv = psi[i - 1]
        psi[i - 1] = cm_view(gemm(as_matrix(prv, 2), us), tuple(prv.shape[:2]) + (us.shape[1],))
        psi.s[i] = s
        errs.append(err)
    return errs


def _tsvd_legs(A, nrow, trunc):
    """Truncated SVD of A split after nrow legs (trunc None: keep every
    singular value, TensorKit's NoTruncation): u (matrix for nrow = 1, else
    tensor), s, v (tensor for nrow = 1, else matrix), err."""
    M = as_matrix(A, nrow)
    try:
        U, S, Vh = torch.linalg.svd(M, full_matrices=False)
    except RuntimeError:
        U, S, Vh = torch.linalg.svd(M, full_matrices=False, driver="gesvd")
    if trunc is None:
        k, err = S.shape[0], 0.0
    else:
        k, err = trunc.truncate(S.detach().cpu().numpy())
    k = max(k, 1)
    U, S, Vh = U[:, :k], S[:k], Vh[:k]
    if nrow == 1:
        return U, S, from_matrix(Vh, (k,) + tuple(A.shape[1:])), err
    return from_matrix(U, tuple(A.shape[:nrow]) + (k,)), S, Vh, err


def bond_evolution(G, s1, A1, A2, trunc):
    """bond_evolution (apply_gates.jl:9-16): returns (u, s, v, err)."""
    two = two_site(A1, A2)
    tg = apply_gate_tensor(two, G)                                  # [a, s1', s2', b]
    u, s, v, err = stable_tsvd(_scale_left(tg, s1), trunc)
    k = s.shape[0]
    Dl, d1, d2, Dr = tg.shape
    um = gemm(as_matrix(tg, 2), as_matrix(v, 1).conj().t())        # tg . v^dag
    return cm_view(um, (Dl, d1, k)), s, v, err


def bond_evolution3(G, s1, A1, A2, A3, trunc):
    """bond_evolution3 (apply_gates.jl:19-31): (u, s, u2, s2, v2, err)."""
    three = cm_view(gemm(as_matrix(two_site(A1, A2), 3), as_matrix(A3, 1)),
                    (A1.shape[0], A1.shape[1], A2.shape[1], A3.shape[1], A3.shape[2]))
    tg = apply_gate_tensor(three, G)                                # [a, s1', s2', s3', b]
    Dl, d1, d2, d3, Dr = tg.shape
    M = as_matrix(_scale_left(tg, s1), 2)
    U, S, Vh = torch.linalg.svd(M, full_matrices=False)
    k, err1 = trunc.truncate(S.detach().cpu().numpy())
    k = max(k, 1)
    S, Vh = S[:k], Vh[:k]
    u = cm_view(gemm(as_matrix(tg, 2), Vh.conj().t()), (Dl, d1, k))
    v = from_matrix(Vh, (k, d2, d3, Dr))
    v1 = _scale_left(v, S)
    u2, s2, v2, err2 = stable_tsvd(v1, trunc)
    k2 = s2.shape[0]
    u2 = cm_view(gemm(as_matrix(v, 2), as_matrix(v2, 1).conj().t()), (k, d2, k2))
    return u, S, u2, s2, v2, max(err1, err2)


def _swap_matrix(d):
    G = np.zeros((d, d, d, d))
    for a in range(d):
        for b in range(d):
            G[b, a, a, b] = 1.0
    return G


def _move(psi, i, j, trunc):
    """_move! (apply_gates.jl:50-71): carry site i's tensor to position j by
    swap gates."""
    err = 0.0
    if i == j:
        return err
    d = psi[i].shape[1]
    S = _swap_matrix(d)
    if i < j:
        for k in range(i, j):
            psi[k], psi.s[k + 1], psi[k + 1], e = bond_evolution(S, psi.s[k], psi[k], psi[k + 1], trunc)
            err = max(err, e)
    else:
        for k in range(i, j, -1):
            psi[k - 1], psi.s[k], psi[k], e = bond_evolution(S, psi.s[k - 1], psi[k - 1], psi[k], trunc)
            err = max(err, e)
    return err


def _ensure_s(psi):
    if any(s is None for s in psi.s):
        canonicalize(psi)


def apply_gate(psi, positions, G, trunc=None):
    """apply!(gate, mps) (apply_gates.jl:74-158): positions are 0-based and
    increasing; one-, two- and three-site gates.  Returns the truncation
    error."""
    trunc = trunc or MPSTruncation()
    _ensure_s(psi)
    pos = tuple(int(p) for p in positions)
    if any(b <= a for a, b in zip(pos, pos[1:])):
        raise ValueError("gate positions must be increasing")
    if len(pos) == 1:
        (k,) = pos
        psi[k] = apply_gate_tensor(psi[k], G)
        return 0.0
    if len(pos) == 2:
        i, j = pos
        err = _move(psi, i, j - 1, trunc)
        psi[j - 1], psi.s[j], psi[j], e = bond_evolution(G, psi.s[j - 1], psi[j - 1], psi[j], trunc)
        err = max(err, e)
        return max(err, _move(psi, j - 1, i, trunc))
    if len(pos) == 3:
        i, j, k = pos
        err = max(_move(psi, i, j - 1, trunc), _move(psi, k, j + 1, trunc))
        psi[j - 1], psi.s[j], psi[j], psi.s[j + 1], psi[j + 1], e = bond_evolution3(
            G, psi.s[j - 1], psi[j - 1], psi[j], psi[j + 1], trunc)
        err = max(err, e)
        return max(err, _move(psi, j - 1, i, trunc), _move(psi, j + 1, k, trunc))
    raise _lib.MHError(7, "gates on at most three sites are supported")


def apply_circuit(psi, gates, trunc=None):
    """apply!(circuit, mps) (apply_gates.jl:160-167): gates = [(positions, G)]."""
    return [apply_gate(psi, pos, G, trunc) for pos, G in gates]


def expectation_term(psi, positions, ops, coeff=1.0):
    """<psi| c * O_1 ... O_n |psi> for single-site operators at 0-based
    positions (expectation(psiA, m::QTerm, psiB), operators/expecs.jl:28-55):
    one right-to-left pass of environment updates with the operator's 1 x d
    x 1 x d tensor on its sites and the identity elsewhere."""
    from . import transfer
    from .envs import _boundary
    L, d = len(psi), psi[0].shape[1]
    cplx = psi[0].is_complex() or any(np.iscomplexobj(o) for o in ops)
    dt = torch.complex128 if cplx else torch.float64
    at = dict(zip((int(p) for p in positions), ops))
    hold = _boundary(dt, psi[0].device)
    for i in range(L - 1, -1, -1):
        A = psi[i].to(dt)
        if i in at:
            W = np.zeros((1, d, 1, d), dtype=np.complex128 if cplx else np.float64)
            W[0, :, 0, :] = np.asarray(at[i])
            hold = transfer.updateright(hold, A, to_device(W, A.device, dt), A)
        else:
            hold = transfer.updateright(hold, A, None, A)
    return coeff * hold.reshape(-1)[0].item()
