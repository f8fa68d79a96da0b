"""numpy restatement of auxiliary/simple_lanczos.jl (lowest eigenpair)."""
import numpy as np

LANCZOS_ORTH_TOL = 1.0e-12  # simple_lanczos.jl:1


def _dot(a, b):
    return np.vdot(a, b)  # conjugate-linear in a, like Julia dot


def lanczos_make_first_step(A, v):
    """simple_lanczos.jl:27-42"""
    vm = np.linalg.norm(v)
    q0 = v / vm
    q1 = A(q0)
    a = _dot(q0, q1)
    a = float(np.real(a))
    q1 = q1 - q0 * a
    b = np.linalg.norm(q1)
    if b != 0.0:
        q1 = q1 / b
    return a, b, q0, q1, vm


def lanczos_make_step(A, qim1, qi, bold):
    """simple_lanczos.jl:49-68"""
    qip1 = A(qi)
    a = float(np.real(_dot(qi, qip1)))
    qip1 = qip1 - (qi * a + qim1 * bold)
    b = np.linalg.norm(qip1)
    if b != 0.0:
        qip1 = qip1 / b
    return a, b, qi, qip1


def simple_lanczos_solver(A, v, wh="SA", kmax=10, btol=1.0e-8):
    """simple_lanczos.jl:83-120; returns (eigval, eigvec, info, k)."""
    info = 0
    a, b, qim1, qi, vm = lanczos_make_first_step(A, v)
    if b < btol:
        return a, qim1, info, 1
    V = [qim1]
    T = np.zeros((kmax, kmax))
    T[0, 0] = a
    k = 1
    converged = False
    while k < kmax:
        T[k - 1, k] = b
        T[k, k - 1] = b
        a, b, qim1, qi = lanczos_make_step(A, qim1, qi, b)
        V.append(qim1)
        T[k, k] = a
        k += 1
        if b < btol:
            converged = True
            break
    if not converged:
        info = -1
    w, U = np.linalg.eigh(T[:k, :k])
    pos = 0  # "SR"/"SA": smallest (find_lanczos, :5-17)
    c = U[:, pos]
    evec = V[-1] * c[-1]
    for i in range(len(V) - 1):
        evec = evec + V[i] * c[i]
    return w[pos], evec, info, k
