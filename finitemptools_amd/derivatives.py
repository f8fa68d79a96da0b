"""Effective Hamiltonians on the device: the drop-in for
algorithms/derivatives.jl (``ac_prime``, ``ac2_prime``, ``c_prime``).

Argument order and meaning follow the reference:
``ac_prime(x, m, hleft, hright)`` (derivatives.jl:2),
``ac2_prime(x, h1, h2, hleft, hright)`` (:20), ``c_prime(x, hleft, hright)`` (:52).
Shape mismatches raise ``DimensionMismatch`` like TensorKit's space checks.
"""
import ctypes

from . import _lib
from .engine import (PREC, as_colmajor, colmajor_empty, dtype_code, engine, ptr)


def _same(dev_tensors):
    t0 = dev_tensors[0]
    for t in dev_tensors[1:]:
        if t.dtype != t0.dtype:
            raise TypeError("all tensors must share one dtype")
        if t.device != t0.device:
            raise ValueError("all tensors must live on one device")
    if t0.device.type != "cuda":
        raise ValueError("engine tensors must be on a HIP device")


def _dm(msg):
    return _lib.DimensionMismatch(2, msg)


class HeffOperator:
    """H_eff of one site (kind 1), two sites (kind 2) or one bond (kind 0),
    with its environments and MPO tensors kept on the device.  It is what the
    reference's solvers receive as a closure ``x -> ac2_prime(x, ...)``
    (dmrg.jl:100, tdvp.jl:96)."""

    def __init__(self, kind, L, R, W1=None, W2=None, prec="f64"):
        self.kind = kind
        self.prec = prec
        ts = [t for t in (L, R, W1, W2) if t is not None]
        _same(ts)
        self.L = as_colmajor(L)
        self.R = as_colmajor(R)
        self.W1 = as_colmajor(W1) if W1 is not None else None
        self.W2 = as_colmajor(W2) if W2 is not None else None
        self.dtype = L.dtype
        Dl, wl, Dl2 = L.shape
        Dr, wr, Dr2 = R.shape
        if Dl != Dl2 or Dr != Dr2:
            raise _dm("environments must be (D, w, D)")
        op = _lib.HeffOp()
        op.kind, op.dtype, op.prec = kind, dtype_code(L), PREC[prec]
        op.Dl, op.Dr, op.wl, op.wr = Dl, Dr, wl, wr
        op.d1, op.d2, op.wm = 1, 1, 0
        if kind == 0:
            if wl != wr:
                raise _dm("c_prime: hleft and hright MPO legs differ")
            self.shape = (Dl, Dr)
        elif kind == 1:
            if W1.shape[0] != wl or W1.shape[2] != wr or W1.shape[1] != W1.shape[3]:
                raise _dm(f"ac_prime: MPO {tuple(W1.shape)} vs envs wl={wl} wr={wr}")
            op.d1 = W1.shape[1]
            self.shape = (Dl, op.d1, Dr)
        elif kind == 2:
            if W1.shape[0] != wl or W2.shape[2] != wr or W1.shape[2] != W2.shape[0]:
                raise _dm("ac2_prime: MPO bonds do not chain")
            op.d1, op.d2, op.wm = W1.shape[1], W2.shape[1], W1.shape[2]
            self.shape = (Dl, op.d1, op.d2, Dr)
        else:
            raise ValueError("kind must be 0, 1 or 2")
        op.L, op.R = self.L.data_ptr(), self.R.data_ptr()
        op.W1 = self.W1.data_ptr() if self.W1 is not None else None
        op.W2 = self.W2.data_ptr() if self.W2 is not None else None
        self.op = op
        self.engine = engine(L.device)

    @property
    def size(self):
        n = 1
        for s in self.shape:
            n *= s
        return n

    def flops(self):
        """Algorithmic real flops of one application (complex x4)."""
        o = self.op
        c = 4.0 if o.dtype == _lib.MH_C64 else 1.0
        if self.kind == 0:
            f = 2.0 * o.Dl * o.Dr * o.wl * (o.Dl + o.Dr)
        elif self.kind == 1:
            d = o.d1
            f = 2.0 * o.Dl * o.Dr * d * (o.Dl * o.wl + o.Dr * o.wr) + 2.0 * o.Dl * o.Dr * d * d * o.wl * o.wr
        else:
            d1, d2 = o.d1, o.d2
            f = (2.0 * o.Dl * o.Dr * d1 * d2 * (o.Dl * o.wl + o.Dr * o.wr)
                 + 2.0 * o.Dl * o.Dr * d1 * d2 * (d1 * o.wl * o.wm + d2 * o.wm * o.wr))
        return c * f

    def check_vector(self, x):
        if tuple(x.shape) != tuple(self.shape):
            raise _dm(f"vector shape {tuple(x.shape)} != operator shape {self.shape}")
        if x.dtype != self.dtype or x.device != self.L.device:
            raise TypeError("vector dtype/device differs from the operator's")
        return as_colmajor(x)

    def apply(self, x, out=None):
        x = self.check_vector(x)
        y = out if out is not None else colmajor_empty(self.shape, self.dtype, x.device)
        e = self.engine
        h = e.bind_stream()
        e.check(e.lib.mh_heff_apply(h, ctypes.byref(self.op), ptr(x), ptr(y)), "heff_apply")
        return y

    __call__ = apply


def ac_prime(x, m, hleft, hright, prec="f64"):
    """y[a,s',b] = L[a,w,a'] x[a',s,b'] W[w,s',w',s] R[b,w',b'] (derivatives.jl:2-4)"""
    _same([x, m, hleft, hright])
    Dl, d, Dr = x.shape
    if tuple(hleft.shape[::2]) != (Dl, Dl) or tuple(hright.shape[::2]) != (Dr, Dr):
        raise _dm("ac_prime: environment bonds do not match x")
    return HeffOperator(1, hleft, hright, m, prec=prec).apply(x)


def ac_prime_open(x, m, hleft):
    """P[a,s',w',b'] = L[a,w,a'] x[a',s,b'] W[w,s',w',s]: ac_prime with its
    right environment left open (FP64 / C64), the expansion term of
    expansion/chexpand.jl:23-25.  Returns a (Dl, d, wr, Dr) tensor."""
    _same([x, m, hleft])
    x, m, hleft = as_colmajor(x), as_colmajor(m), as_colmajor(hleft)
    Dl, d, Dr = x.shape
    wl, _, wr, _ = m.shape
    if tuple(hleft.shape) != (Dl, wl, Dl) or tuple(m.shape) != (wl, d, wr, d):
        raise _dm(f"ac_prime_open: L {tuple(hleft.shape)}, W {tuple(m.shape)}, x {tuple(x.shape)}")
    P = colmajor_empty((Dl, d, wr, Dr), x.dtype, x.device)
    e = engine(x.device)
    h = e.bind_stream()
    e.check(e.lib.mh_ac_prime_open(h, dtype_code(x), Dl, d, Dr, wl, wr, ptr(x), ptr(m),
                                   ptr(hleft), ptr(P)), "ac_prime_open")
    return P


def ac2_prime(x, h1, h2, hleft, hright, prec="f64"):
    """y = (((L x) W1) W2) R on a two-site tensor x[a,s1,s2,b] (derivatives.jl:20-22)"""
    _same([x, h1, h2, hleft, hright])
    Dl, d1, d2, Dr = x.shape
    if tuple(hleft.shape[::2]) != (Dl, Dl) or tuple(hright.shape[::2]) != (Dr, Dr):
        raise _dm("ac2_prime: environment bonds do not match x")
    return HeffOperator(2, hleft, hright, h1, h2, prec=prec).apply(x)


def c_prime(x, hleft, hright, prec="f64"):
    """y[a,b] = L[a,w,a'] x[a',b'] R[b,w,b'] (derivatives.jl:52-54)"""
    _same([x, hleft, hright])
    Dl, Dr = x.shape
    if tuple(hleft.shape[::2]) != (Dl, Dl) or tuple(hright.shape[::2]) != (Dr, Dr):
        raise _dm("c_prime: environment bonds do not match x")
    return HeffOperator(0, hleft, hright, prec=prec).apply(x)
