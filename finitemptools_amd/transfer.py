"""Environment updates on the device: the drop-in for mpo/transfer.jl
``updateleft`` / ``updateright`` with an MPO tensor (mpo/transfer.jl:26-34)
and the identity / scalar MPO-entry variants (:37-55).

``hold`` is an environment E[a_bra, w, a_ket]; ``psiA`` is the bra (conjugated),
``psiB`` the ket, ``hj`` the MPO tensor W[wl, s', wr, s], a number, or None.
"""
from . import _lib
from .engine import PREC, as_colmajor, colmajor_empty, dtype_code, engine, ptr


def _prep(hold, psiA, hj, psiB):
    ts = [hold, psiA, psiB] + ([hj] if hj is not None and not isinstance(hj, (int, float, complex)) else [])
    dt = hold.dtype
    for t in ts:
        if t.dtype != dt or t.device != hold.device:
            raise TypeError("updateleft/updateright: mixed dtypes or devices")
    if tuple(psiA.shape) != tuple(psiB.shape):
        raise _lib.DimensionMismatch(2, "bra and ket site tensors differ in shape")
    return [as_colmajor(t) for t in ts]


_IDENTITY = {}


def _identity_mpo(d, w, dtype, device, scale=1.0):
    """W[w, s', w, s] = scale * delta(s', s) delta(w_l, w_r): the reference's
    ``hj::Nothing`` / ``hj::Number`` branches (mpo/transfer.jl:37-55) keep the
    MPO leg of ``hold`` and contract only the physical legs.  Built on the
    device (column-major, as the engine reads it) once per (d, w, dtype,
    device, scale) and cached."""
    import torch
    key = (int(d), int(w), dtype, str(device), complex(scale))
    W = _IDENTITY.get(key)
    if W is None:
        Id = torch.eye(d, dtype=dtype, device=device) * scale
        Iw = torch.eye(w, dtype=dtype, device=device)
        dense = torch.einsum("ab,st->asbt", Iw, Id)          # [wl, s', wr, s]
        W = dense.permute(3, 2, 1, 0).contiguous().permute(3, 2, 1, 0)
        _IDENTITY[key] = W
    return W


def updateleft(hold, psiA, hj, psiB, prec="f64"):
    """hnew[b,w',b'] = hold[a,w,a'] psiB[a',s,b'] hj[w,s',w',s] conj(psiA[a,s',b])
    (mpo/transfer.jl:31-34)."""
    Dl, d, Dr = psiB.shape
    if hj is None or isinstance(hj, (int, float, complex)):
        scale = 1.0 if hj is None else hj
        hj = _identity_mpo(d, hold.shape[1], hold.dtype, hold.device, scale)
    hold, psiA, psiB, hj = _prep(hold, psiA, hj, psiB)
    wl, wr = hj.shape[0], hj.shape[2]
    if tuple(hold.shape) != (Dl, wl, Dl) or tuple(hj.shape) != (wl, d, wr, d):
        raise _lib.DimensionMismatch(2, f"updateleft: hold {tuple(hold.shape)}, W {tuple(hj.shape)}, psi {tuple(psiB.shape)}")
    out = colmajor_empty((Dr, wr, Dr), hold.dtype, hold.device)
    e = engine(hold.device)
    h = e.bind_stream()
    e.check(e.lib.mh_updateleft(h, dtype_code(hold), PREC[prec], Dl, d, Dr, wl, wr,
                                ptr(hold), ptr(psiA), ptr(hj), ptr(psiB), ptr(out)),
            "updateleft")
    return out


def updateright(hold, psiA, hj, psiB, prec="f64"):
    """hnew[a,w,a'] = hold[b,w',b'] psiB[a',s,b'] hj[w,s',w',s] conj(psiA[a,s',b])
    (mpo/transfer.jl:26-29)."""
    Dl, d, Dr = psiB.shape
    if hj is None or isinstance(hj, (int, float, complex)):
        scale = 1.0 if hj is None else hj
        hj = _identity_mpo(d, hold.shape[1], hold.dtype, hold.device, scale)
    hold, psiA, psiB, hj = _prep(hold, psiA, hj, psiB)
    wl, wr = hj.shape[0], hj.shape[2]
    if tuple(hold.shape) != (Dr, wr, Dr) or tuple(hj.shape) != (wl, d, wr, d):
        raise _lib.DimensionMismatch(2, f"updateright: hold {tuple(hold.shape)}, W {tuple(hj.shape)}, psi {tuple(psiB.shape)}")
    out = colmajor_empty((Dl, wl, Dl), hold.dtype, hold.device)
    e = engine(hold.device)
    h = e.bind_stream()
    e.check(e.lib.mh_updateright(h, dtype_code(hold), PREC[prec], Dl, d, Dr, wl, wr,
                                 ptr(hold), ptr(psiA), ptr(hj), ptr(psiB), ptr(out)),
            "updateright")
    return out
