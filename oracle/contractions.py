"""numpy restatement of the H_eff and environment contractions.

Index conventions (mps/finitemps.jl:130-141, mpo/finitempo.jl:35-43,
algorithms/derivatives.jl:2): A[a, s, b]; W[wl, s_out, wr, s_in];
environments E[a_bra, w, a_ket].  Pairwise contraction order follows the
parentheses of the reference's @tensor expressions.
"""
import numpy as np


def ac_prime(x, m, hleft, hright):
    """derivatives.jl:2-4:  ((hleft[-1,1,2] x[2,3,4]) m[1,-2,5,3]) hright[-3,5,4]"""
    t = np.tensordot(hleft, x, axes=([2], [0]))            # [a, w, s, b']
    t = np.tensordot(t, m, axes=([1, 2], [0, 3]))          # [a, b', s', w']
    return np.tensordot(t, hright, axes=([1, 3], [2, 1]))  # [a, s', b]


def ch_expansion_term_left(hleft, A, m, alpha):
    """expansion/chexpand.jl:23-25:
    alpha * hleft[-1,1,2] A[2,3,4] m[1,-2,5,3] ft[5,4,-3] with ft the
    isomorphism (w' (x) b') -> fused index w' + wr*b'.  Returns (Dl, d, wr*Dr)."""
    t = np.tensordot(hleft, A, axes=([2], [0]))            # [a, w, s, b']
    t = np.tensordot(t, m, axes=([1, 2], [0, 3]))          # [a, b', s', w']
    t = alpha * t.transpose(0, 2, 3, 1)                    # [a, s', w', b']
    return t.reshape(t.shape[0], t.shape[1], -1, order="F")


def ch_expansion_term_right(A, m, hright, alpha):
    """expansion/chexpand.jl:34-36:
    alpha * ft[-1,1,2] A[2,3,4] m[1,-2,5,3] hright[-3,5,4] with ft the
    isomorphism fused index w + wl*a <- (w (x) a).  Returns (wl*Dl, d, Dr)."""
    t = np.tensordot(A, hright, axes=([2], [2]))           # [a, s, b, w']
    t = np.tensordot(m, t, axes=([2, 3], [3, 1]))          # [w, s', a, b]
    t = alpha * t.transpose(0, 2, 1, 3)                    # [w, a, s', b]
    return t.reshape(-1, t.shape[2], t.shape[3], order="F")


def ac2_prime(x, h1, h2, hleft, hright):
    """derivatives.jl:20-22:
    (((hleft[-1,1,2] x[2,3,4,5]) h1[1,-2,6,3]) h2[6,-3,7,4]) hright[-4,7,5]"""
    t = np.tensordot(hleft, x, axes=([2], [0]))            # [a, w, s1, s2, b']
    t = np.tensordot(t, h1, axes=([1, 2], [0, 3]))         # [a, s2, b', s1', m]
    t = np.tensordot(t, h2, axes=([4, 1], [0, 3]))         # [a, b', s1', s2', w']
    return np.tensordot(t, hright, axes=([1, 4], [2, 1]))  # [a, s1', s2', b]


def c_prime(x, hleft, hright):
    """derivatives.jl:52-54:  (hleft[-1,1,2] x[2,3]) hright[-2,1,3]"""
    t = np.tensordot(hleft, x, axes=([2], [0]))            # [a, w, b']
    return np.tensordot(t, hright, axes=([1, 2], [1, 2]))  # [a, b]


def updateright(hold, psiA, hj, psiB):
    """mpo/transfer.jl:26-29:
    hnew[-1;-2 -3] = hold[1,2,3] psiB[-3,4,3] hj[-2,5,2,4] conj(psiA[-1,5,1])
    hj None / number: identity MPO entry (:37-55)."""
    if hj is None or np.isscalar(hj):
        c = 1.0 if hj is None else hj
        t = np.tensordot(hold, psiB, axes=([2], [2]))          # [b, w, a', s]
        return c * np.einsum("bwcs,asb->awc", t, psiA.conj())
    t = np.tensordot(hold, psiB, axes=([2], [2]))              # [b, w', a', s]
    t = np.tensordot(t, hj, axes=([1, 3], [2, 3]))             # [b, a', w, s']
    t = np.tensordot(psiA.conj(), t, axes=([1, 2], [3, 0]))    # [a, a', w]
    return t.transpose(0, 2, 1)


def updateleft(hold, psiA, hj, psiB):
    """mpo/transfer.jl:31-34:
    hnew[-1;-2 -3] = hold[1,2,3] psiB[3,4,-3] hj[2,5,-2,4] conj(psiA[1,5,-1])"""
    if hj is None or np.isscalar(hj):
        c = 1.0 if hj is None else hj
        t = np.tensordot(hold, psiB, axes=([2], [0]))          # [a, w, s, b']
        return c * np.einsum("awsc,asb->bwc", t, psiA.conj())
    t = np.tensordot(hold, psiB, axes=([2], [0]))              # [a, w, s, b']
    t = np.tensordot(t, hj, axes=([1, 2], [0, 3]))             # [a, b', s', w']
    t = np.tensordot(psiA.conj(), t, axes=([0, 1], [0, 2]))    # [b, b', w']
    return t.transpose(0, 2, 1)


def mpo_grad(hleft, psiA, psiB, hright):
    """Derivative of <psiA| ac_prime(psiB, W, hleft, hright) > with respect to
    the MPO tensor W[w, s', w', s] (no counterpart in the reference):
    G = hleft[a,w,a'] conj(psiA[a,s',b]) psiB[a',s,b'] hright[b,w',b'].
    Pinned by sum(G * W) == vdot(psiA, ac_prime(psiB, W, hleft, hright))."""
    t = np.tensordot(hleft, psiB, axes=([2], [0]))             # [a, w, s, b']
    u = np.tensordot(psiA.conj(), hright, axes=([2], [0]))     # [a, s', w', b']
    return np.einsum("awsc,atvc->wtvs", t, u)


def l_LL(dtype=np.float64):
    """Strict boundary env (mpo/arithmetics.jl:257-258): 1x1x1 isometry."""
    return np.ones((1, 1, 1), dtype=dtype)


def r_RR(dtype=np.float64):
    """mpo/arithmetics.jl:253-254"""
    return np.ones((1, 1, 1), dtype=dtype)


def expectation(psiA, mpo, psiB):
    """<psiA|h|psiB> by right-to-left environment updates
    (mpo/arithmetics.jl:267-275)."""
    dt = np.result_type(psiA[0], mpo[0], psiB[0])
    hold = r_RR(dt)
    for i in range(len(psiA) - 1, -1, -1):
        hold = updateright(hold, psiA[i], mpo[i], psiB[i])
    return hold.reshape(-1)[0]


def dense_heff1(m, hleft, hright):
    """Brute force: the (Dl d Dr)^2 matrix of ac_prime as a sum of Kronecker
    products, H = sum_{w,w'} R[:,w',:] (x) W[w,:,w',:] (x) L[:,w,:], acting on
    column-major vectors (a fastest).  Independent of the pairwise order."""
    wl, wr = m.shape[0], m.shape[2]
    H = 0
    for w in range(wl):
        for wp in range(wr):
            H = H + np.kron(hright[:, wp, :], np.kron(m[w, :, wp, :], hleft[:, w, :]))
    return H


def dense_heff2(h1, h2, hleft, hright):
    """Kronecker-product form of ac2_prime (column-major two-site vectors)."""
    wl, wm, wr = h1.shape[0], h1.shape[2], h2.shape[2]
    H = 0
    for w in range(wl):
        for m in range(wm):
            for wp in range(wr):
                H = H + np.kron(hright[:, wp, :], np.kron(h2[m, :, wp, :],
                                np.kron(h1[w, :, m, :], hleft[:, w, :])))
    return H


def dense_heff0(hleft, hright):
    w = hleft.shape[1]
    return sum(np.kron(hright[:, k, :], hleft[:, k, :]) for k in range(w))


def vec(x):
    """column-major flattening (Julia vec)"""
    return np.asarray(x).reshape(-1, order="F")
