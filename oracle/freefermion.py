"""Closed form of the transverse-field Ising quench on an open chain (oracle,
test infrastructure only).

H = -J sum_i sx_i sx_{i+1} - h sum_i sz_i  (the chain of models.tfim_mpo; the
reference's TDVP tests evolve such quenches, test/runtests.jl:537-594) from
|+>^L, the h = 0 ground state with all sx = +1.

Jordan-Wigner Majoranas a_j = (prod_{k<j} sz_k) sx_j, b_j = (prod_{k<j} sz_k) sy_j
with sx sy = i sz give sz_j = -i a_j b_j and sx_j sx_{j+1} = -i b_j a_{j+1}, so
H = i J sum b_j a_{j+1} + i h sum a_j b_j = (i/4) sum_kl A_kl g_k g_l with A
real antisymmetric; the Heisenberg equations are dg/dt = A g, g(t) = e^{At} g.
The covariance Gamma_kl = (i/2) <[g_k, g_l]> evolves as O Gamma O^T with
O = e^{At}, and <sz_j(t)> = -Gamma_{a_j b_j}(t).

|+>^L is the equal superposition of the two parity states of the h = 0 open
chain (the edge zero mode a_1, b_L empty or filled); sz and its evolution are
parity-even quadratic forms, so its expectation is that of the averaged
covariance: Gamma_{b_j a_{j+1}} = -1 (sx_j sx_{j+1} = 1), every other
independent entry 0 (including the zero-mode entry a_1 b_L).
"""
import numpy as np
import scipy.linalg as sla


def tfim_quench_sz(L, J, h, times):
    """<sz_j(t)> for every t in ``times``: array (len(times), L)."""
    n = 2 * L
    a = lambda j: 2 * j
    b = lambda j: 2 * j + 1
    A = np.zeros((n, n))
    for j in range(L - 1):          # i J b_j a_{j+1}
        A[b(j), a(j + 1)] += 2 * J
        A[a(j + 1), b(j)] -= 2 * J
    for j in range(L):              # i h a_j b_j
        A[a(j), b(j)] += 2 * h
        A[b(j), a(j)] -= 2 * h
    G0 = np.zeros((n, n))
    for j in range(L - 1):
        G0[b(j), a(j + 1)] = -1.0
        G0[a(j + 1), b(j)] = 1.0
    out = []
    for t in np.atleast_1d(times):
        O = sla.expm(A * t)
        G = O @ G0 @ O.T
        out.append([-G[a(j), b(j)] for j in range(L)])
    return np.array(out)
