"""numpy restatement of the Schur-form (MPOHamiltonian) contractions, term by
term (test infrastructure: only the tests call it).

A Schur site tensor is a dict over its nonzero entries (i, j):
  ("scal", c)  an identity entry with coefficient c (the reference's scalkeys,
               isscal(h, i, j), coefficient h.Os[i, j]);
  ("op", O)    an operator entry O[wi, s_out, wj, s_in] (opkeys, h[i, j]), an
               MPOTensor whose virtual legs are the channel spaces (usually
               of dimension 1).
Environments are Vectors: hleft[i][a, wi, a'] and hright[j][b, wj, b'] (one
TensorMap{1,2} per channel).  Index strings follow the reference's @tensor
expressions; TensorOperations contracts them pairwise, which einsum does in
an order of its own (the sum is the same).
"""
import numpy as np


def _scal_or_op(entry):
    kind, val = entry
    if kind not in ("scal", "op"):
        raise ValueError(f"unknown Schur entry kind {kind!r}")
    return kind == "scal", val


def ac_prime(x, ham, hleft, hright):
    """algorithms/derivatives.jl:7-17 (ac_prime on an AbstractSiteMPOTensor):
    opkeys:   tmp[-1,-2,-3] += hleft[i][-1,5,4] x[4,2,1] ham[i,j][5,-2,3,2] hright[j][-3,3,1]
    scalkeys: tmp[-1,-2,-3] += hleft[i][-1,5,4] (scal*x)[4,-2,1] hright[j][-3,5,1]"""
    tmp = np.zeros(x.shape, dtype=np.result_type(x, *hleft, *hright,
                                                 *[v for _, v in ham.values()]))
    for (i, j), entry in ham.items():
        scal, v = _scal_or_op(entry)
        if scal:
            tmp += np.einsum("AEd,dSa,BEa->ASB", hleft[i], v * x, hright[j])
        else:
            tmp += np.einsum("AEd,dsa,ESCs,BCa->ASB", hleft[i], x, v, hright[j])
    return tmp


def ac2_prime(x, h1, h2, hleft, hright):
    """algorithms/derivatives.jl:24-49, the four branches over the entries
    (i, j) of h1 and k of h2 with contains(h2, j, k):
      scal x scal: tmp += (s1 s2 hleft[i])[-1,7,6] x[6,-2,-3,1] hright[k][-4,7,1]
      scal x op:   tmp += (s1 hleft[i])[-1,7,6] x[6,-2,3,1] h2[j,k][7,-3,2,3] hright[k][-4,2,1]
      op x scal:   tmp += (s2 hleft[i])[-1,7,6] x[6,5,-3,1] h1[i,j][7,-2,2,5] hright[k][-4,2,1]
      op x op:     tmp += hleft[i][-1,7,6] x[6,5,3,1] h1[i,j][7,-2,4,5] h2[j,k][4,-3,2,3] hright[k][-4,2,1]"""
    vals = [v for _, v in h1.values()] + [v for _, v in h2.values()]
    tmp = np.zeros(x.shape, dtype=np.result_type(x, *hleft, *hright, *vals))
    for (i, j), e1 in h1.items():
        for (j2, k), e2 in h2.items():
            if j2 != j:
                continue
            sc1, v1 = _scal_or_op(e1)
            sc2, v2 = _scal_or_op(e2)
            if sc1 and sc2:
                tmp += np.einsum("AGf,fSTa,BGa->ASTB", (v1 * v2) * hleft[i], x, hright[k])
            elif sc1:
                tmp += np.einsum("AGf,fSca,GTbc,Bba->ASTB", v1 * hleft[i], x, v2, hright[k])
            elif sc2:
                tmp += np.einsum("AGf,feTa,GSbe,Bba->ASTB", v2 * hleft[i], x, v1, hright[k])
            else:
                tmp += np.einsum("AGf,feca,GSde,dTbc,Bba->ASTB", hleft[i], x, v1, v2, hright[k])
    return tmp


def _updateleft1(hold, psiA, hj, psiB):
    """mpo/transfer.jl:31-34 (hj an MPOTensor) and :42-45 (hj nothing):
    hnew[-1;-2 -3] = hold[1,2,3] psiB[3,4,-3] hj[2,5,-2,4] conj(psiA[1,5,-1]);
    with no operator the MPO leg passes through: hold[1,-2,3] psiB[3,4,-3] conj(psiA[1,4,-1])."""
    if hj is None:
        return np.einsum("awc,csd,asb->bwd", hold, psiB, psiA.conj())
    return np.einsum("awc,csd,wtvs,atb->bvd", hold, psiB, hj, psiA.conj())


def _updateright1(hold, psiA, hj, psiB):
    """mpo/transfer.jl:26-29 (hj an MPOTensor) and :37-40 (hj nothing):
    hnew[-1;-2 -3] = hold[1,2,3] psiB[-3,4,3] hj[-2,5,2,4] conj(psiA[-1,5,1]);
    with no operator: hold[1,-2,3] psiB[-3,4,3] conj(psiA[-1,4,1])."""
    if hj is None:
        return np.einsum("bwc,dsc,asb->awd", hold, psiB, psiA.conj())
    return np.einsum("bvc,dsc,wtvs,atb->awd", hold, psiB, hj, psiA.conj())


def updateleft(hold, psiA, hj, psiB, wr):
    """mpohamiltonian/transfer.jl:62-75: hnew[j] += (isscal ? Os[i,j] *
    updateleft(hold[i], A, nothing, B) : updateleft(hold[i], A, h[i,j], B)),
    one environment per right channel j = 0..wr-1 (zero where no entry ends)."""
    out = [None] * wr
    for (i, j), entry in hj.items():
        scal, v = _scal_or_op(entry)
        t = v * _updateleft1(hold[i], psiA, None, psiB) if scal else \
            _updateleft1(hold[i], psiA, v, psiB)
        out[j] = t if out[j] is None else out[j] + t
    return out


def updateright(hold, psiA, hj, psiB, wl):
    """mpohamiltonian/transfer.jl:48-60: hnew[i] += (isscal ? Os[i,j] *
    updateright(hold[j], A, nothing, B) : updateright(hold[j], A, h[i,j], B)),
    one environment per left channel i = 0..wl-1 (zero where no entry starts)."""
    out = [None] * wl
    for (i, j), entry in hj.items():
        scal, v = _scal_or_op(entry)
        t = v * _updateright1(hold[j], psiA, None, psiB) if scal else \
            _updateright1(hold[j], psiA, v, psiB)
        out[i] = t if out[i] is None else out[i] + t
    return out


def fold(ham, wl, wr, d):
    """Dense W[wl, s_out, wr, s_in] of a Schur site tensor whose channels all
    have dimension 1 (identity entries become c * I)."""
    vals = [v for _, v in ham.values()]
    W = np.zeros((wl, d, wr, d), dtype=np.result_type(np.float64, *vals))
    for (i, j), entry in ham.items():
        scal, v = _scal_or_op(entry)
        W[i, :, j, :] += v * np.eye(d) if scal else v[0, :, 0, :]
    return W
