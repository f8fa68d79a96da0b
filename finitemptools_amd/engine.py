"""Engine handle and column-major tensor helpers.

Tensors handed to the engine are torch tensors on a HIP device whose memory
is column-major in the index order of the reference (Julia/TensorKit): the
first index is fastest.  ``colmajor_empty((Dl, d, Dr))`` allocates such a
tensor; indexing ``A[a, s, b]`` then means what it means in the reference.
"""
import atexit
import ctypes

import torch

from . import _lib

PREC = _lib.PREC


def colmajor_empty(shape, dtype=torch.float64, device="cuda"):
    shape = tuple(int(s) for s in shape)
    n = len(shape)
    return torch.empty(tuple(reversed(shape)), dtype=dtype, device=device).permute(
        *reversed(range(n)))


def colmajor_zeros(shape, dtype=torch.float64, device="cuda"):
    t = colmajor_empty(shape, dtype, device)
    t.zero_()
    return t


def is_colmajor(t):
    s = 1
    for ext, st in zip(t.shape, t.stride()):
        if ext != 1 and st != s:
            return False
        s *= ext
    return True


def resolve(t):
    """Materialise torch's lazy conjugate / negative views: the engine reads
    raw memory, which for such a view still holds the unconjugated values
    (``.contiguous()`` keeps the lazy bit on an already-contiguous view)."""
    if t.is_conj():
        t = t.resolve_conj()
    if t.is_neg():
        t = t.resolve_neg()
    return t


def as_colmajor(t):
    t = resolve(t)
    if is_colmajor(t):
        return t
    n = t.dim()
    rev = tuple(reversed(range(n)))
    return t.permute(*rev).contiguous().permute(*rev)


def to_device(a, device="cuda", dtype=None):
    """numpy array (any layout) -> column-major device tensor"""
    t = torch.as_tensor(a)
    if dtype is not None:
        t = t.to(dtype)
    n = t.dim()
    rev = tuple(reversed(range(n)))
    return t.permute(*rev).contiguous().to(device).permute(*rev)


def to_numpy(t):
    return t.detach().cpu().numpy()


def dtype_code(t):
    if t.dtype == torch.float64:
        return _lib.MH_F64
    if t.dtype == torch.complex128:
        return _lib.MH_C64
    raise TypeError(f"engine tensors are float64 or complex128, got {t.dtype}")


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class Engine:
    """One ``mh_handle`` bound to one HIP device."""

    def __init__(self, device=0):
        if not torch.cuda.is_available():
            raise RuntimeError("the MI355X engine needs a HIP device")
        self.lib = _lib.load()
        self.device = int(device)
        h = ctypes.c_void_p()
        st = self.lib.mh_create(self.device, ctypes.byref(h))
        if st != 0:
            raise _lib.MHError(st, f"mh_create({self.device}) failed")
        self.h = h

    def bind_stream(self):
        s = torch.cuda.current_stream(self.device).cuda_stream
        self.lib.mh_set_stream(self.h, ctypes.c_void_p(s))
        return self.h

    def check(self, status, what=""):
        _lib.raise_for(status, self.h, what)

    def profile_enable(self, on=True):
        self.check(self.lib.mh_profile_enable(self.h, 1 if on else 0))

    def profile_reset(self):
        self.check(self.lib.mh_profile_reset(self.h))

    def set_graphs(self, mode=1):
        """Krylov recurrences as hipGraphs: 1 (default, True) captures the
        multi-block steps and runs the short-vector steps eagerly, 2 captures
        both, 0 (False) captures nothing."""
        self.check(self.lib.mh_set_graphs(self.h, int(mode)))

    def set_krylov_small(self, on=True):
        """Single-workgroup Krylov steps for short vectors (default) or the
        multi-block vector kernels at every size."""
        self.check(self.lib.mh_set_krylov_small(self.h, 1 if on else 0))

    def set_bs_plain(self, on=True):
        """Block-sparse GEMM stages as one launch per sector product (the
        plain per-sector loop) instead of the grouped launch (default)."""
        self.check(self.lib.mh_set_bs_plain(self.h, 1 if on else 0))

    def set_fusion(self, on=True):
        """Fused first H_eff step (default) or GEMM + separate mode product."""
        self.check(self.lib.mh_set_fusion(self.h, 1 if on else 0))

    def profile_read(self, kclass):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        fl = ctypes.c_double()
        self.check(self.lib.mh_profile_read(self.h, kclass, ctypes.byref(ms),
                                            ctypes.byref(n), ctypes.byref(fl)))
        return ms.value, n.value, fl.value

    def close(self):
        if getattr(self, "h", None) is not None:
            self.lib.mh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_ENGINES = {}
_ATEXIT = False


def close_all():
    """Tear the engines down while the HIP runtime and every library are
    still alive: first the block-split worker pools (their threads launch on
    side streams), then every handle (mh_destroy waits for its streams and
    destroys the eigensolver's side streams and library handles).  Runs from
    ``atexit`` (registered with the first engine, so before torch's own exit
    hooks) instead of from ``__del__`` during interpreter teardown, when the
    order against runtime and profiler shutdown is not defined."""
    import sys
    bm = sys.modules.get(__package__ + ".blockmps")
    if bm is not None:
        bm.shutdown_workers()
    for dev in list(_ENGINES):
        e = _ENGINES.pop(dev)
        try:
            torch.cuda.synchronize(dev)
        except Exception:
            pass
        e.close()


def engine(device=None):
    """The process-wide engine of ``device`` (default: torch's current device)."""
    if device is None:
        device = torch.cuda.current_device()
    if isinstance(device, torch.device):
        device = device.index if device.index is not None else torch.cuda.current_device()
    global _ATEXIT
    e = _ENGINES.get(device)
    if e is None:
        e = Engine(device)
        _ENGINES[device] = e
        if not _ATEXIT:
            atexit.register(close_all)
            _ATEXIT = True
    return e
