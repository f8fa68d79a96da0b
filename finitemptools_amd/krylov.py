"""Device-resident Krylov solvers over an ``HeffOperator``.

``simple_lanczos_solver`` mirrors auxiliary/simple_lanczos.jl:83-120 (lowest
eigenpair, three-term recurrence, kmax steps, exit when beta < btol);
``exponentiate`` stands in for KrylovKit.exponentiate as called by
``_exp_evolve`` (algorithms/tdvp.jl:3-6).  Krylov vectors never leave HBM.
"""
import ctypes

from . import _lib
from .engine import as_colmajor, colmajor_empty, dtype_code, ptr


def simple_lanczos_solver(A, v, wh="SR", kmax=10, btol=1.0e-8, reorth=False, verbosity=0,
                          projectors=None):
    """Lowest eigenpair of the operator ``A`` (an HeffOperator) started from ``v``.
    Returns (eigenvalue, eigenvector, info) with info = 0 when beta < btol stopped
    the recurrence and -1 when kmax steps ran (simple_lanczos.jl:108-113).
    ``projectors``: vectors p_l of v's shape; the operator becomes
    x -> P(A x) and the start vector P(v), with P(y): y -= dot(y, p_l) p_l in
    order (the excited-state solve, dmrgexcited.jl:1-6)."""
    if wh not in ("SR", "SA"):
        raise ValueError("only the lowest eigenpair ('SR'/'SA') is supported")
    v = A.check_vector(v)
    out = colmajor_empty(A.shape, A.dtype, v.device)
    e = A.engine
    h = e.bind_stream()
    ev = ctypes.c_double()
    info = ctypes.c_int()
    nit = ctypes.c_int()
    if projectors:
        ps = [A.check_vector(p) for p in projectors]
        arr = (ctypes.c_void_p * len(ps))(*[ctypes.c_void_p(p.data_ptr()) for p in ps])
        e.check(e.lib.mh_lanczos_eig_proj(h, ctypes.byref(A.op), len(ps), arr, ptr(v), int(kmax),
                                          float(btol), 1 if reorth else 0, ctypes.byref(ev),
                                          ptr(out), ctypes.byref(info), ctypes.byref(nit)),
                "lanczos_eig_proj")
    else:
        e.check(e.lib.mh_lanczos_eig(h, ctypes.byref(A.op), ptr(v), int(kmax), float(btol),
                                     1 if reorth else 0, ctypes.byref(ev), ptr(out),
                                     ctypes.byref(info), ctypes.byref(nit)), "lanczos_eig")
    if verbosity >= 2:
        print(f"lanczos: {nit.value} steps, info {info.value}")
    return ev.value, out, info.value


def eig_solver(A, x, maxiter, tol, reorth=False):
    """``_eig_solver`` (algorithms/dmrg.jl:3-6)."""
    val, vec, _ = simple_lanczos_solver(A, x, "SR", maxiter, tol, reorth=reorth)
    return val, vec


def exponentiate(A, dt, x, tol=1e-8, kdim=12, ishermitian=True):
    """exp(dt * A) x with a Krylov space of dimension <= kdim, sub-stepped so
    that each sub-step's Krylov error estimate is below tol * |sub-step|:
    Lanczos for a Hermitian A, Arnoldi otherwise (KrylovKit's choice behind
    exponentiate(...; ishermitian), tdvp.jl:4)."""
    x = A.check_vector(x)
    dt = complex(dt)
    out = colmajor_empty(A.shape, A.dtype, x.device)
    e = A.engine
    h = e.bind_stream()
    info = ctypes.c_int()
    fn = e.lib.mh_lanczos_expm if ishermitian else e.lib.mh_arnoldi_expm
    e.check(fn(h, ctypes.byref(A.op), ptr(x), dt.real, dt.imag, int(kdim), float(tol), ptr(out),
               ctypes.byref(info)), "lanczos_expm" if ishermitian else "arnoldi_expm")
    return out, info.value


def exp_evolve(A, dt, x, tol, ishermitian=True):
    """``_exp_evolve(f, dt, x, isherm, tol)`` (tdvp.jl:3-6) passes tol/|dt| to
    KrylovKit as the accuracy per unit time; the engine takes the same."""
    y, _ = exponentiate(A, dt, x, tol=tol / abs(dt) if dt != 0 else tol, ishermitian=ishermitian)
    return y


def dot(x, y):
    """<x, y>, conjugate-linear in x (Julia ``dot``)."""
    from .engine import engine
    x = as_colmajor(x)
    y = as_colmajor(y)
    e = engine(x.device)
    h = e.bind_stream()
    out = (ctypes.c_double * 2)()
    e.check(e.lib.mh_dot(h, dtype_code(x), x.numel(), ptr(x), ptr(y), out), "dot")
    return complex(out[0], out[1]) if x.is_complex() else out[0]


def norm(x):
    from .engine import engine
    x = as_colmajor(x)
    e = engine(x.device)
    h = e.bind_stream()
    out = ctypes.c_double()
    e.check(e.lib.mh_nrm2(h, dtype_code(x), x.numel(), ptr(x), ctypes.byref(out)), "nrm2")
    return out.value
