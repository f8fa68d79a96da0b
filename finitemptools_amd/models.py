"""Host-side MPO construction for the benchmark models (numpy, setup time).

Conventions follow the reference: Pauli-convention spin operators
(utilities/spin_siteops.jl:4-11: sigma^+ = [0 0; 1 0], sigma^z = diag(-1, 1)),
MPO tensors W[wl, s_out, wr, s_in] (mpo/finitempo.jl:35-43), and the
upper-triangular Schur layout of mpohamiltonian/constructor.jl:140-167
(channel 0 = "nothing placed yet", last channel = "term complete").

``schur_mpo`` builds the finite MPO from a list of local operator strings.
Strings that share a prefix share MPO channels (a prefix trie), so the XXZ
chain gets w = 5, the transverse-field Ising chain w = 3, the J1-J2 chain
w = 8 and the Fermi-Hubbard chain w = 6.
"""
import numpy as np


def spin_half_matrices():
    """spin_siteops.jl:4-11"""
    sp = np.array([[0.0, 0.0], [1.0, 0.0]])
    sm = np.array([[0.0, 1.0], [0.0, 0.0]])
    z = np.array([[-1.0, 0.0], [0.0, 1.0]])
    return {"x": sp + sm, "y": -1j * (sp - sm), "z": z, "+": sp, "-": sm,
            "I": np.eye(2)}


def spinal_fermion_ops_dense():
    """Spin-1/2 fermion site operators, d = 4, following the construction of
    utilities/fermion_siteops.jl (dense branch): with the spin-1/2 Pauli
    matrices, c^dag_up = kron(sigma^+, I), c^dag_dn = kron(JW, sigma^+) with
    JW = -sigma^z, so the on-site ordering (up before down) is built in.
    Returns the creation/annihilation operators, the number operators and the
    site parity F = JW (x) JW = (-1)^n."""
    p = spin_half_matrices()
    sp, z = p["+"], p["z"]
    I2 = np.eye(2)
    jw = -z
    cup_dag = np.kron(sp, I2)
    cdn_dag = np.kron(jw, sp)
    cup = cup_dag.T.copy()
    cdn = cdn_dag.T.copy()
    nup = cup_dag @ cup
    ndn = cdn_dag @ cdn
    F = np.kron(jw, jw)
    return {"cup+": cup_dag, "cup": cup, "cdn+": cdn_dag, "cdn": cdn,
            "nup": nup, "ndn": ndn, "F": F, "I": np.eye(4)}


def schur_mpo(L, ops, onsite, strings, dtype=np.float64):
    """Finite MPO of  H = sum_i onsite_i + sum_i sum_t coef_t O^t_0(i) O^t_1(i+1) ...

    ops:     name -> (d x d) matrix, O[s_out, s_in]
    onsite:  list of (coef, name)
    strings: list of (coef, [name_0, name_1, ..., name_k]), k >= 1, acting on
             consecutive sites i .. i+k (write identities/JW strings explicitly)
    Returns a list of L tensors W[wl, s_out, wr, s_in] (first wl = 1, last wr = 1).
    """
    d = next(iter(ops.values())).shape[0]
    chan = {}  # prefix tuple -> channel index (1-based among intermediates)
    for _, names in strings:
        for j in range(1, len(names)):
            pre = tuple(names[:j])
            if pre not in chan:
                chan[pre] = len(chan) + 1
    w = len(chan) + 2
    end = w - 1
    cell = np.zeros((w, d, w, d), dtype=dtype)
    eye = np.eye(d)
    cell[0, :, 0, :] = eye
    cell[end, :, end, :] = eye
    for coef, name in onsite:
        cell[0, :, end, :] += coef * ops[name]
    for coef, names in strings:
        k = len(names) - 1
        for j in range(1, k + 1):
            pre = tuple(names[:j])
            src = 0 if j == 1 else chan[tuple(names[:j - 1])]
            cell[src, :, chan[pre], :] = ops[names[j - 1]]
        cell[chan[tuple(names[:k])], :, end, :] += coef * ops[names[k]]
    mpo = []
    for i in range(L):
        W = cell
        if i == 0:
            W = W[0:1]
        if i == L - 1:
            W = W[:, :, end:end + 1]
        mpo.append(np.ascontiguousarray(W))
    return mpo


def heisenberg_xxz_mpo(L, J=1.0, Jzz=1.0, hz=0.0):
    """models.jl:27-42: H = sum hz sz_i + sum [2J (s+_i s-_{i+1} + s-_i s+_{i+1}) + Jzz sz_i sz_{i+1}]
    in the Pauli convention; w = 5."""
    p = spin_half_matrices()
    ops = {k: p[k].real for k in ("+", "-", "z")}
    onsite = [(hz, "z")] if hz != 0.0 else []
    strings = [(2 * J, ["+", "-"]), (2 * J, ["-", "+"]), (Jzz, ["z", "z"])]
    return schur_mpo(L, ops, onsite, strings)


def tfim_mpo(L, J=1.0, h=1.0, dtype=np.float64):
    """H = -J sum sx_i sx_{i+1} - h sum sz_i (Pauli matrices); w = 3."""
    p = spin_half_matrices()
    ops = {"x": p["x"].real, "z": p["z"].real}
    return schur_mpo(L, ops, [(-h, "z")], [(-J, ["x", "x"])], dtype=dtype)


def j1j2_mpo(L, J1=1.0, J2=0.0, keep_j2=False):
    """H = J1 sum S_i.S_{i+1} + J2 sum S_i.S_{i+2} with S = sigma/2; w = 8
    (the J2 strings share their first channel with the J1 strings).
    keep_j2: keep the J2 channels at J2 = 0, so that the tensors are affine in
    J2 with one structure (what a derivative with respect to J2 needs)."""
    p = spin_half_matrices()
    ops = {"+": p["+"].real, "-": p["-"].real, "z": p["z"].real, "I": np.eye(2)}
    strings = [(J1 / 2, ["+", "-"]), (J1 / 2, ["-", "+"]), (J1 / 4, ["z", "z"])]
    if J2 != 0.0 or keep_j2:
        strings += [(J2 / 2, ["+", "I", "-"]), (J2 / 2, ["-", "I", "+"]),
                    (J2 / 4, ["z", "I", "z"])]
    return schur_mpo(L, ops, [], strings)


def j1j2_dJ2_mpo(L):
    """dH/dJ2 = sum S_i.S_{i+2}: the Hellmann-Feynman operator of the J1-J2 chain."""
    p = spin_half_matrices()
    ops = {"+": p["+"].real, "-": p["-"].real, "z": p["z"].real, "I": np.eye(2)}
    strings = [(0.5, ["+", "I", "-"]), (0.5, ["-", "I", "+"]), (0.25, ["z", "I", "z"])]
    return schur_mpo(L, ops, [], strings)


def hubbard_mpo(L, t=1.0, U=4.0, mu=0.0):
    """H = -t sum_{i,s} (c+_{i s} c_{i+1 s} + h.c.) + U sum n_up n_dn - mu sum n,
    Jordan-Wigner strings folded into the site operators:
    c+_i c_{i+1} = (c+ F)_i (c)_{i+1},  c+_{i+1} c_i = (F c)_i (c+)_{i+1}; w = 6."""
    f = spinal_fermion_ops_dense()
    ops = dict(f)
    ops["cup+F"] = f["cup+"] @ f["F"]
    ops["cdn+F"] = f["cdn+"] @ f["F"]
    ops["Fcup"] = f["F"] @ f["cup"]
    ops["Fcdn"] = f["F"] @ f["cdn"]
    ops["nn"] = f["nup"] @ f["ndn"]
    ops["n"] = f["nup"] + f["ndn"]
    onsite = [(U, "nn")] + ([(-mu, "n")] if mu != 0.0 else [])
    strings = [(-t, ["cup+F", "cup"]), (-t, ["cdn+F", "cdn"]),
               (-t, ["Fcup", "cup+"]), (-t, ["Fcdn", "cdn+"])]
    return schur_mpo(L, ops, onsite, strings)


def site_operator_mpo(L, i, op, dtype=np.float64):
    """Bond-dimension-1 MPO of a single local operator op (d x d) on site i
    (identity elsewhere): used for local expectation values."""
    d = op.shape[0]
    mpo = []
    for k in range(L):
        m = op if k == i else np.eye(d)
        mpo.append(np.asarray(m, dtype=dtype).reshape(1, d, 1, d))
    return mpo


def schur_fold(terms, wl, wr, d, dtype=np.float64):
    """Dense site tensor W[wl, s_out, wr, s_in] from the entries of a
    Schur-form MPO site tensor, through mh_schur_fold (host only): terms are
    (i, j, coef, O) with O a (d x d) matrix O[s_out, s_in], or None for a
    scalar entry (coef times the identity) -- the opkeys and scalkeys of the
    reference's AbstractSiteMPOTensor (algorithms/derivatives.jl:7-17)."""
    import ctypes
    from . import _lib
    cplx = np.dtype(dtype) == np.complex128
    n = len(terms)
    ti = (ctypes.c_int * max(n, 1))(*[int(t[0]) for t in terms])
    tj = (ctypes.c_int * max(n, 1))(*[int(t[1]) for t in terms])
    kind = (ctypes.c_int * max(n, 1))(*[0 if t[3] is None else 1 for t in terms])
    coef = np.zeros(2 * max(n, 1))
    for k, t in enumerate(terms):
        coef[2 * k], coef[2 * k + 1] = complex(t[2]).real, complex(t[2]).imag
    mats = [np.asarray(t[3], dtype=np.complex128 if cplx else np.float64) for t in terms
            if t[3] is not None]
    # each O column-major (O[s_out, s_in] at s_out + d s_in): C order of O^T
    ops = np.ascontiguousarray(np.stack([m.T for m in mats])) if mats else np.zeros(1)
    ops = ops.view(np.float64) if cplx else ops
    W = np.zeros(wl * d * wr * d, dtype=np.complex128 if cplx else np.float64)
    st = _lib.load().mh_schur_fold(_lib.MH_C64 if cplx else _lib.MH_F64, wl, wr, d, n, ti, tj,
                                   kind, coef.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                   ops.ctypes.data_as(ctypes.c_void_p),
                                   W.ctypes.data_as(ctypes.c_void_p))
    _lib.raise_for(st, None, "mh_schur_fold")
    return W.reshape((wl, d, wr, d), order="F")
