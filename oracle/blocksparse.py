"""numpy restatement of the two-site H_eff on charge-conserving (U(1)^n)
tensors, sector by sector (test infrastructure and the CPU baseline only).

The reference evaluates ac2_prime (algorithms/derivatives.jl:20-22) with
@tensor on TensorKit GradedSpace tensors, which loops over the charge sectors
and calls BLAS on each sector's block.  This restatement keeps that shape:
for every pair (left charge qa of the output, right charge qb' of the input)
it performs the three steps of derivatives.jl:21 in order,

  1. T1[a, w, s1, s2, b'] = L[a, w, a'] x[a', s1, s2, b']     (one GEMM per block)
  2. T2[a, s1', s2', w', b'] = sum_{w, s1, s2} (W1 W2)[...] T1  (the MPO pair)
  3. y[a, s1', s2', b] += T2[a, s1', s2', w', b'] R[b, w', b']^T  (one GEMM per block)

with the physical and MPO indices dense and the bond indices block-sparse.

Blocks are dicts of 2-D arrays:
  x[(qa, s1, s2, qb)]   (D_qa, D_qb), qa + q(s1) + q(s2) = qb
  L[(q, w)]             (D_q, D_k) with k = q - qwl[w] (environment bra, ket)
  R[(q, w)]             (D_q, D_k) with k = q - qwr[w]
Charges are tuples; MPO tensors W[wl, s_out, wr, s_in] dense.
"""
import numpy as np


def _add(a, b):
    return tuple(x + y for x, y in zip(a, b))


def _sub(a, b):
    return tuple(x - y for x, y in zip(a, b))


def ac2_prime(x, W1, W2, L, R, phys, qwl, qwr, left, right):
    """y = ac2_prime(x, W1, W2, L, R) on block tensors.  left / right:
    {charge: dimension} of the output's bond spaces.  Returns the dict of
    output blocks (every charge-allowed block, zeros where nothing lands).

    Like @tensor on graded tensors, each pairwise contraction permutes its
    operands to matrices and multiplies them one coupled sector at a time:
    step 1 one GEMM per charge of a' (rows (a, w), columns (s1, s2, b')),
    step 3 one GEMM per charge of b (rows (a, s1', s2'), inner (w', b')).
    The MPO pair acts between them on the charge-allowed slots only."""
    d = len(phys)
    wl, wr = W1.shape[0], W2.shape[2]
    # the MPO pair, Mt[(w, s1, s2), (s1', s2', w')] = sum_m W1[w,s1',m,s1] W2[m,s2',w',s2]
    Mt = np.einsum("wamb,mcvd->wbdacv", W1, W2).reshape(wl * d * d, d * d * wr)
    # ---- step 1: T1[qk] = [L(qa, w) stacked over rows] @ [x(qk, ...) side by side]
    lrows, xcols = {}, {}
    for (qa, w), Lb in L.items():
        lrows.setdefault(_sub(qa, qwl[w]), []).append((qa, w, Lb))
    for (qk, s1, s2, qbp), xb in x.items():
        xcols.setdefault(qk, []).append((s1, s2, qbp, xb))
    T1, roff, coff = {}, {}, {}
    for qk, cols in xcols.items():
        rows = lrows.get(qk)
        if not rows:
            continue
        T1[qk] = np.vstack([Lb for _, _, Lb in rows]) @ np.hstack([xb for *_, xb in cols])
        r0, c0 = 0, 0
        ro, co = {}, {}
        for qa, w, Lb in rows:
            ro[(qa, w)] = (r0, Lb.shape[0])
            r0 += Lb.shape[0]
        for s1, s2, qbp, xb in cols:
            co.setdefault(qbp, []).append((s1, s2, c0, xb.shape[1]))
            c0 += xb.shape[1]
        roff[qk], coff[qk] = ro, co
    # ---- the layout of step 3: per charge qb, rows (qa, s1', s2'), inner (w', qb')
    rights = {qbp: xb.shape[1] for (_, _, _, qbp), xb in x.items()}
    r3, k3, T2 = {}, {}, {}
    for qb, Db in right.items():
        ro, r0 = {}, 0
        for qa, Da in left.items():
            for s1p in range(d):
                for s2p in range(d):
                    if _add(_add(qa, phys[s1p]), phys[s2p]) == qb:
                        ro[(qa, s1p, s2p)] = r0
                        r0 += Da
        ko, k0 = {}, 0
        for wp in range(wr):
            qbp = _sub(qb, qwr[wp])
            if (qb, wp) in R and qbp in rights:
                ko[wp] = (k0, qbp)
                k0 += rights[qbp]
        if r0 and k0:
            r3[qb], k3[qb] = ro, ko
            T2[qb] = np.zeros((r0, k0))
    # ---- step 2: T2 blocks (qa, s1', s2' | w', qb') from the T1 blocks of each (qa, qb')
    for qa, Da in left.items():
        for qbp, Dbp in rights.items():
            slots, parts = [], []
            for w in range(wl):
                qk = _sub(qa, qwl[w])
                if qk not in T1 or (qa, w) not in roff[qk]:
                    continue
                r0, _ = roff[qk][(qa, w)]
                for s1, s2, c0, _ in coff[qk].get(qbp, ()):
                    slots.append((w * d + s1) * d + s2)
                    parts.append(T1[qk][r0:r0 + Da, c0:c0 + Dbp])
            if not slots:
                continue
            outs = []
            for s1p in range(d):
                for s2p in range(d):
                    qb = _add(_add(qa, phys[s1p]), phys[s2p])
                    for wp, (k0, q) in k3.get(qb, {}).items():
                        if q == qbp:
                            outs.append(((s1p * d + s2p) * wr + wp, qb, r3[qb][(qa, s1p, s2p)], k0))
            if not outs:
                continue
            blk = np.tensordot(np.stack(parts, axis=-1), Mt[np.ix_(slots, [o[0] for o in outs])],
                               axes=([2], [0]))                       # (Da, Db', nJ)
            for t, (_, qb, r0, k0) in enumerate(outs):
                T2[qb][r0:r0 + Da, k0:k0 + Dbp] = blk[:, :, t]
    # ---- step 3: y rows of qb = T2[qb] @ [R(qb, w')^T stacked over w']
    y = {}
    for qa, Da in left.items():
        for s1 in range(d):
            for s2 in range(d):
                qb = _add(_add(qa, phys[s1]), phys[s2])
                if qb in right:
                    y[(qa, s1, s2, qb)] = np.zeros((Da, right[qb]))
    for qb, M2 in T2.items():
        Y = M2 @ np.vstack([R[(qb, wp)].T for wp in k3[qb]])
        for (qa, s1p, s2p), r0 in r3[qb].items():
            y[(qa, s1p, s2p, qb)] = Y[r0:r0 + left[qa]]
    return y


def blocks_from_flat(flat, entries):
    """{key: (rows x cols) block} from a flat buffer of column-major blocks;
    entries: (key, offset, rows, cols)."""
    return {k: np.asarray(flat[o:o + r * c]).reshape(c, r).T for k, o, r, c in entries}


def flat_from_blocks(blocks, entries, size):
    """The inverse of blocks_from_flat (zeros outside the listed blocks)."""
    out = np.zeros(size, dtype=np.result_type(*blocks.values()) if blocks else np.float64)
    for k, o, r, c in entries:
        out[o:o + r * c] = blocks[k].T.reshape(-1)
    return out
