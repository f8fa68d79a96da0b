"""Charge-conserving (block-sparse) finite MPS, environments and two-site DMRG
on the device.

Restates, for U(1)-type graded spaces, what FiniteMPTools does through
TensorKit: ``prodmps`` / ``randommps`` (mps/initializers.jl:34-73), right
canonical form (mps/orth.jl:40-49), ``ExpectationCache`` (envs/finiteenv.jl),
the DMRG2 sweeps (algorithms/dmrg.jl:89-144) with ``stable_tsvd`` truncation
applied globally across sectors (auxiliary/truncation.jl:27-47).

Contractions (two-site tensor, H_eff, environment updates) run as grouped
plans (blocksparse.py); the Lanczos recurrence stays on the device.  The
two-site split diagonalises the sector reduced density matrices with batched
FP64 eigh on the device (``split_two_site_rho``, the default) or takes the
sector SVDs with host LAPACK (``split_two_site``, ``alg.split = "svd"``); the
sector QRs of the canonical form are FP64 torch.linalg calls on the device.
"""
import time

import numpy as np
import torch

from . import blocksparse as bs
from . import eigen
from .blocksparse import (BlockEnv, BlockLayout, BlockTensor, EnvLayout, Space, qadd,
                          qsub, qzero)
from . import truncation as tr
from .truncation import MPSTruncation


# ------------------------------------------------------------------ models

def hubbard_charges():
    """(n_up, n_dn) of the four local states of the dense spin-1/2 fermion
    basis (models.spinal_fermion_ops_dense): read off the number operators."""
    from .models import spinal_fermion_ops_dense
    f = spinal_fermion_ops_dense()
    return [(int(round(f["nup"][s, s])), int(round(f["ndn"][s, s]))) for s in range(4)]


def sz_charges():
    """2 S^z of the spin-1/2 basis (index 0 = up, models.spin_half_matrices)."""
    from .models import spin_half_matrices
    z = spin_half_matrices()["z"].real
    return [(int(round(z[s, s])),) for s in range(2)]


class BlockMPO:
    """Host MPO tensors with their channel charges (one list per MPO bond)."""

    def __init__(self, mpo, phys):
        self.W = [np.asarray(W, dtype=np.float64) for W in mpo]
        self.phys = [tuple(q) for q in phys]
        self.qw = bs.mpo_channel_charges(self.W, self.phys)

    def __len__(self):
        return len(self.W)

    def __getitem__(self, i):
        return self.W[i]


# ------------------------------------------------------------------ states

class BlockMPS:
    def __init__(self, sites, phys):
        self.sites = list(sites)
        self.phys = [tuple(q) for q in phys]
        self.s = [None] * (len(sites) + 1)

    def __len__(self):
        return len(self.sites)

    def __getitem__(self, i):
        return self.sites[i]

    def __setitem__(self, i, v):
        self.sites[i] = v

    def bond_space(self, i):
        return self.sites[i].lspace if i < len(self.sites) else self.sites[-1].rspace

    def bond_dimension(self):
        return max(t.rspace.total for t in self.sites)

    def total_charge(self):
        (q, _), = self.sites[-1].rspace.sectors
        return q

    def to_dense(self):
        """Full state vector (small L only), site 0 slowest as in the oracle."""
        v = self.sites[0].to_dense()  # (1, d, D1)
        v = v.reshape(-1, v.shape[-1])
        for t in self.sites[1:]:
            A = t.to_dense()
            v = np.einsum("xa,asb->xsb", v, A).reshape(-1, A.shape[-1])
        return v.reshape(-1)

    def right_canonicalize(self, normalize=True):
        """Sites 1..L-1 right-isometric by sector-wise LQ (mps/orth.jl:40-49)."""
        for i in range(len(self) - 1, 0, -1):
            A = self.sites[i]
            lay = A.layout
            new_dims, Ls, Qs = {}, {}, {}
            for qa, Da in lay.lspace.sectors:
                if qa not in lay.group:
                    continue
                G = A.group_matrix(qa)                 # Da x ncols
                Q, R = torch.linalg.qr(G.t(), mode="reduced")   # ncols x k, k x Da
                new_dims[qa] = Q.shape[1]
                Ls[qa] = R.t()                         # Da x k
                Qs[qa] = Q
            ls = Space(new_dims.items())
            nl = BlockLayout(ls, lay.phys, 1, lay.rspace)
            data = torch.empty(nl.size, dtype=torch.float64, device=A.data.device)
            for qa in ls.dim:
                g0, ncols = nl.group[qa]
                data[g0:g0 + ls.dim[qa] * ncols] = Qs[qa].contiguous().reshape(-1)
            self.sites[i] = BlockTensor(nl, data)
            # absorb L into site i-1 (blocks ending in qa)
            P = self.sites[i - 1]
            pl = BlockLayout(P.lspace, P.layout.phys, 1, ls)
            pdata = torch.zeros(pl.size, dtype=torch.float64, device=P.data.device)
            for b in pl.blocks:
                ob = P.layout.index[(b.qa, b.p)]
                blk = P.block(ob) @ Ls[b.qb]
                pdata[b.off:b.off + b.rows * b.cols] = blk.t().reshape(-1)
            self.sites[i - 1] = BlockTensor(pl, pdata)
        if normalize:
            A = self.sites[0]
            A.data /= torch.linalg.vector_norm(A.data)
        return self


def _prune(spaces_left, spaces_right):
    """Sectors present in both the left-reachable and right-reachable sets,
    with the smaller of the two dimensions."""
    out = []
    for sl, sr in zip(spaces_left, spaces_right):
        out.append(Space([(q, min(d, sr.dim[q])) for q, d in sl.sectors if q in sr]))
    return out


def max_virtual_spaces(L, phys, total, cap=None):
    """Largest bond spaces compatible with the physical charges and a total
    charge: sector q of bond i has the number of ways to reach it from the
    left and to complete it to ``total`` on the right (capped at ``cap``)."""
    phys = [tuple(q) for q in phys]
    n = len(phys[0])
    left = [Space([(qzero(n), 1)])]
    for i in range(L):
        acc = {}
        for q, d in left[-1].sectors:
            for p in phys:
                qq = qadd(q, p)
                acc[qq] = acc.get(qq, 0) + d
        left.append(Space([(q, min(d, cap) if cap else d) for q, d in acc.items()]))
    right = [Space([(tuple(total), 1)])]
    for i in range(L):
        acc = {}
        for q, d in right[0].sectors:
            for p in phys:
                qq = qsub(q, p)
                acc[qq] = acc.get(qq, 0) + d
        right.insert(0, Space([(q, min(d, cap) if cap else d) for q, d in acc.items()]))
    return _prune(left, right)


def prodmps(phys, states, device="cuda"):
    """Product state with local basis states ``states[i]`` (prodmps for
    graded spaces, initializers.jl:34-56): every bond carries one sector of
    dimension 1 and every site tensor one block equal to 1."""
    phys = [tuple(q) for q in phys]
    n = len(phys[0])
    q = qzero(n)
    sites = []
    for s in states:
        qn = qadd(q, phys[s])
        lay = BlockLayout(Space([(q, 1)]), phys, 1, Space([(qn, 1)]))
        data = torch.zeros(lay.size, dtype=torch.float64, device=device)
        data[lay.index[(q, (s,))].off] = 1.0
        sites.append(BlockTensor(lay, data))
        q = qn
    return BlockMPS(sites, phys)


def randommps(L, phys, total, D, seed=0, device="cuda"):
    """Random MPS in the sector ``total`` with bond dimension <= D
    (randommps(T, physpaces; sector, D), initializers.jl:58-73): site tensors
    of iid N(0,1) blocks on the largest allowed spaces, each reduced to D by a
    truncated SVD from the left (numpy PCG64 stream)."""
    gen = np.random.Generator(np.random.PCG64(seed))
    vmax = max_virtual_spaces(L, phys, total)
    trunc = MPSTruncation(D=D, eps=0.0)
    sites = []
    left = vmax[0]
    for i in range(L):
        acc = {}
        for q, d in left.sectors:
            for p in phys:
                qq = qadd(q, tuple(p))
                if qq in vmax[i + 1]:
                    acc[qq] = acc.get(qq, 0) + d
        right = Space([(q, min(d, vmax[i + 1].dim[q])) for q, d in acc.items()])
        lay = BlockLayout(left, phys, 1, right)
        A = BlockTensor.randn(lay, gen, device)
        if i < L - 1:
            U, S, _ = sector_svd_left(A, trunc)
            A = U
        sites.append(A)
        left = A.rspace
    return BlockMPS(sites, phys)


def sector_svd_left(A, trunc):
    """Left-isometric factor of a one-site tensor A[(a,s), b] by sector SVDs
    with global truncation (used to build random states)."""
    lay = A.layout
    mats, rows = {}, {}
    for qb, Db in lay.rspace.sectors:
        blocks = lay.by_right.get(qb, [])
        if not blocks:
            continue
        mats[qb] = torch.cat([A.block(b) for b in blocks], dim=0)
        rows[qb] = blocks
    svd = {}
    for qb, M in mats.items():
        svd[qb] = torch.linalg.svd(M, full_matrices=False)
    keep = _global_keep(svd, trunc)
    ns = Space([(qb, k) for qb, k in keep.items()])
    nl = BlockLayout(lay.lspace, lay.phys, 1, ns)
    data = torch.zeros(nl.size, dtype=torch.float64, device=A.data.device)
    for qb, k in keep.items():
        U = svd[qb][0][:, :k]
        r = 0
        for b in rows[qb]:
            nb = nl.index[(b.qa, b.p)]
            data[nb.off:nb.off + nb.rows * nb.cols] = U[r:r + b.rows].t().reshape(-1)
            r += b.rows
    return BlockTensor(nl, data), None, None


def _global_keep(svd, trunc):
    """Kept count per sector when the truncation acts on all singular values
    together (truncation.jl:27-47 on the sector-merged spectrum)."""
    keys = list(svd.keys())
    vals = [svd[q][1].detach().cpu().numpy() for q in keys]
    allv = np.concatenate(vals) if vals else np.zeros(0)
    owner = np.concatenate([np.full(len(v), j) for j, v in enumerate(vals)]) if vals else np.zeros(0, int)
    order = np.argsort(-allv, kind="stable")
    k, _ = trunc.truncate(allv[order])
    k = max(k, 1)
    counts = np.bincount(owner[order[:k]], minlength=len(keys))
    return {q: int(c) for q, c in zip(keys, counts) if c > 0}


# ------------------------------------------------------------------ environments

def boundary_env(space, nq, device):
    (q, d), = space.sectors
    lay = EnvLayout(space, [qzero(nq)])
    return BlockEnv(lay, torch.ones(lay.size, dtype=torch.float64, device=device))


class BlockExpectationCache:
    """ExpectationCache (envs/finiteenv.jl:3-93) for block-sparse states:
    hstorage[i] is the environment at bond i; plans are built per update."""

    def __init__(self, mpo, mps, prec="f64"):
        if len(mpo) != len(mps):
            raise ValueError("MPO and MPS lengths differ")
        self.mpo, self.mps, self.prec = mpo, mps, prec
        self.hstorage = [None] * (len(mps) + 1)
        self.env_time = 0.0
        self.init_hstorage()

    @property
    def state(self):
        return self.mps

    @property
    def h(self):
        return self.mpo

    @property
    def env(self):
        return self.hstorage

    def init_hstorage(self, center=0):
        L = len(self.mps)
        dev = self.mps[0].data.device
        nq = len(self.mps.phys[0])
        self.hstorage[0] = boundary_env(self.mps[0].lspace, nq, dev)
        self.hstorage[L] = boundary_env(self.mps[L - 1].rspace, nq, dev)
        for i in range(L - 1, center, -1):
            self.updateright(i)
        for i in range(center):
            self.updateleft(i)

    def _prefetched_env(self, A, old):
        """(new layout, plan) made for exactly these layouts by the sweep's
        prefetch, or None."""
        got = _prefetched(self).get("env")
        if got and got[0] is A.layout and got[1] is old.layout and got[4] == self.prec:
            return got[2], got[3]
        return None

    def updateleft(self, site):
        """hstorage[site+1] = updateleft(hstorage[site], A, W, A)"""
        t0 = time.perf_counter()
        A = self.mps[site]
        old = self.hstorage[site]
        got = self._prefetched_env(A, old)
        if got:
            new_lay, plan = got
        else:
            new_lay = EnvLayout(A.rspace, self.mpo.qw[site + 1])
            plan = bs.make_plan("updateleft", A.layout, old.layout, self.mpo[site], new_lay,
                                self.prec)
        out = torch.zeros(new_lay.size, dtype=torch.float64, device=A.data.device)
        plan.apply(A.data, old.data, A.data, out)
        self.hstorage[site + 1] = BlockEnv(new_lay, out)
        self.env_time += time.perf_counter() - t0

    def updateright(self, site):
        """hstorage[site] = updateright(hstorage[site+1], A, W, A)"""
        t0 = time.perf_counter()
        A = self.mps[site]
        old = self.hstorage[site + 1]
        got = self._prefetched_env(A, old)
        if got:
            new_lay, plan = got
        else:
            new_lay = EnvLayout(A.lspace, self.mpo.qw[site])
            plan = bs.make_plan("updateright", A.layout, old.layout, self.mpo[site], new_lay,
                                self.prec)
        out = torch.zeros(new_lay.size, dtype=torch.float64, device=A.data.device)
        plan.apply(A.data, old.data, A.data, out)
        self.hstorage[site] = BlockEnv(new_lay, out)
        self.env_time += time.perf_counter() - t0


def expectation(mpo, mps):
    """<psi|H|psi> by right-to-left environment updates."""
    c = BlockExpectationCache(mpo, mps)
    L = len(mps)
    c.updateright(0)
    return c.hstorage[0].data[0].item()


# ------------------------------------------------------------------ two-site step

class TwoSiteMap:
    """Permutation between a two-site tensor's block buffer and its matrices
    per middle charge c, M_c[(qa, s1, i), (s2, qb, j)] = x[(qa, s1, s2, qb)][i, j]
    (the (a, s1) | (s2, b) split of stable_tsvd!, dmrg.jl:104): ``flat[perm]``
    is the row-major concatenation of all M_c."""

    def __init__(self, xl):
        self.xl = xl
        Sl, Sr, phys = xl.lspace, xl.rspace, xl.phys
        rows, cols = {}, {}
        for qa, Da in Sl.sectors:
            for s1 in range(xl.d):
                rows.setdefault(qadd(qa, phys[s1]), []).append((qa, s1, Da))
        for qb, Db in Sr.sectors:
            for s2 in range(xl.d):
                cols.setdefault(qsub(qb, phys[s2]), []).append((s2, qb, Db))
        self.charges = [c for c in sorted(rows) if c in cols]
        self.rows = {c: rows[c] for c in self.charges}
        self.cols = {c: cols[c] for c in self.charges}
        self.shape, self.base = {}, {}
        pieces = []
        base = 0
        for c in self.charges:
            R = sum(r[2] for r in self.rows[c])
            C = sum(r[2] for r in self.cols[c])
            self.shape[c], self.base[c] = (R, C), base
            ro = 0
            for qa, s1, Da in self.rows[c]:
                co = 0
                for s2, qb, Db in self.cols[c]:
                    b = xl.index[(qa, (s1, s2))]
                    pieces.append((b.off, Da, Db, base, C, ro, co))
                    co += Db
                ro += Da
            base += R * C
        assert base == xl.size
        self.pieces = np.array(pieces, dtype=np.int64).reshape(-1, 7)
        self._perm = None
        self._dev = {}

    @property
    def perm(self):
        """host permutation: flat[perm] is the row-major concatenation of all M_c"""
        if self._perm is None:
            blk, mat = piece_index_host(self.pieces)
            self._perm = np.empty(self.xl.size, dtype=np.int64)
            self._perm[mat] = blk
        return self._perm

    def dest(self, device):
        """device index: cat[dest] = flat scatters the blocks into the matrices"""
        key = str(device)
        if key not in self._dev:
            blk, mat = piece_index(self.pieces, device)
            d = torch.empty(self.xl.size, dtype=torch.int64, device=device)
            d[blk] = mat
            self._dev[key] = d
        return self._dev[key]

    def matrices(self, flat_host):
        cat = flat_host[self.perm]
        return {c: cat[self.base[c]:self.base[c] + self.shape[c][0] * self.shape[c][1]]
                .reshape(self.shape[c]) for c in self.charges}


def piece_index_host(pieces):
    """Element indices of block <-> matrix pieces.  Row p of ``pieces`` =
    (blk_off, nr, nc, mat_base, ncols, r0, c0): element (i, j) of the piece is
    at blk_off + i + j*nr in a column-major block buffer and at
    mat_base + (r0 + i)*ncols + c0 + j in a row-major matrix buffer.
    Returns (blk_idx, mat_idx) over all pieces."""
    if len(pieces) == 0:
        z = np.zeros(0, dtype=np.int64)
        return z, z
    P = pieces
    sizes = P[:, 1] * P[:, 2]
    pid = np.repeat(np.arange(len(P)), sizes)
    start = np.cumsum(sizes) - sizes
    local = np.arange(int(sizes.sum())) - start[pid]
    nr = P[pid, 1]
    i, j = local % nr, local // nr
    return P[pid, 0] + local, P[pid, 3] + (P[pid, 5] + i) * P[pid, 4] + P[pid, 6] + j


def _upload(a, device):
    """Host integer indices -> an int64 tensor on ``device`` without draining
    the stream: a pageable host-to-device copy waits for every launch queued
    before it, so device copies go through pinned host memory, whose block
    torch's host allocator keeps until the copy has run."""
    t = torch.from_numpy(np.ascontiguousarray(a, dtype=np.int64))
    device = torch.device(device)
    if device.type != "cuda":
        return t.to(device)
    return t.pin_memory().to(device, non_blocking=True)


def piece_index(pieces, device):
    """piece_index_host computed on the device (one pass, no host loop)."""
    if len(pieces) == 0:
        z = torch.zeros(0, dtype=torch.int64, device=device)
        return z, z
    P = _upload(pieces, device)
    sizes = P[:, 1] * P[:, 2]
    total = int((pieces[:, 1] * pieces[:, 2]).sum())
    pid = torch.repeat_interleave(torch.arange(len(pieces), device=device), sizes,
                                  output_size=total)
    start = torch.cumsum(sizes, 0) - sizes
    local = torch.arange(total, device=device) - start[pid]
    Pp = P[pid]
    nr = Pp[:, 1]
    i, j = local % nr, local // nr
    return Pp[:, 0] + local, Pp[:, 3] + (Pp[:, 5] + i) * Pp[:, 4] + Pp[:, 6] + j


def two_site_layout(A1, A2):
    return BlockLayout(A1.lspace, A1.layout.phys, 2, A2.rspace)


def two_site(A1, A2, xl=None, plan=None):
    """x = A1 * A2 over the middle bond (dmrg.jl:98) by one grouped launch
    (plan: one made for these layouts beforehand)."""
    xl = xl or two_site_layout(A1, A2)
    plan = plan or bs.make_plan("two_site", A1.layout, A2.layout, xl)
    out = torch.zeros(xl.size, dtype=torch.float64, device=A1.data.device)
    plan.apply(A1.data, A2.data, None, out)
    return BlockTensor(xl, out)


def _svd(M):
    """stable_tsvd (auxiliary/stable_svd.jl:4-18): divide-and-conquer LAPACK
    SVD, falling back to the QR-iteration driver if it fails."""
    import scipy.linalg as sla
    try:
        return sla.svd(M, full_matrices=False, lapack_driver="gesdd", check_finite=False)
    except (np.linalg.LinAlgError, ValueError):
        return sla.svd(M, full_matrices=False, lapack_driver="gesvd", check_finite=False)


def _sector_svds(mats):
    """FP64 SVD of every sector matrix on the host cores (LAPACK threads
    limited to one, sectors spread over a thread pool, largest first)."""
    from concurrent.futures import ThreadPoolExecutor
    from threadpoolctl import threadpool_limits
    keys = sorted(mats, key=lambda c: -mats[c].size)
    workers = max(1, min(16, len(keys), (__import__("os").cpu_count() or 1)))
    with threadpool_limits(1):
        with ThreadPoolExecutor(workers) as ex:
            res = list(ex.map(lambda c: _svd(mats[c]), keys))
    return dict(zip(keys, res))


def split_two_site(tmap, vec_host, trunc, absorb, device):
    """Sector SVDs of the two-site matrices, truncation on the merged
    spectrum (truncation.jl:27-47), normalised singular values (dmrg.jl:105);
    s is absorbed into A2 ("right": u | s v) or A1 ("left": u s | v).
    Returns (A1, A2, {c: s_c})."""
    xl = tmap.xl
    svd = _sector_svds(tmap.matrices(vec_host))
    keep = _global_keep_host(svd, trunc)
    norm = float(np.sqrt(sum(float((svd[c][1][:k] ** 2).sum()) for c, k in keep.items())))
    mid = Space(keep.items())
    l1 = BlockLayout(xl.lspace, xl.phys, 1, mid)
    l2 = BlockLayout(mid, xl.phys, 1, xl.rspace)
    d1 = np.zeros(l1.size)
    d2 = np.zeros(l2.size)
    svals = {}
    for c, k in keep.items():
        U, S, Vh = svd[c]
        s = S[:k] / norm
        svals[c] = s
        u, v = U[:, :k], Vh[:k]
        if absorb == "right":
            v = v * s[:, None]
        else:
            u = u * s[None, :]
        ro = 0
        for qa, s1, Da in tmap.rows[c]:
            b = l1.index[(qa, (s1,))]
            d1[b.off:b.off + Da * k] = u[ro:ro + Da].T.ravel()
            ro += Da
        co = 0
        for s2, qb, Db in tmap.cols[c]:
            b = l2.index[(c, (s2,))]
            d2[b.off:b.off + k * Db] = v[:, co:co + Db].T.ravel()
            co += Db
    A1 = BlockTensor(l1, torch.from_numpy(d1).to(device))
    A2 = BlockTensor(l2, torch.from_numpy(d2).to(device))
    return A1, A2, svals


_EIG_POOL = {}


def _eig_workers(dev, n):
    """n side streams (default priority) and a thread pool for the batched
    eigensolves of one device, created once."""
    import concurrent.futures
    key = (str(dev), n)
    if key not in _EIG_POOL:
        _EIG_POOL[key] = (concurrent.futures.ThreadPoolExecutor(max_workers=n),
                          [torch.cuda.Stream(device=dev) for _ in range(n)])
    return _EIG_POOL[key]


def shutdown_workers():
    """Join the eigensolve worker threads and drop their side streams, and
    join the plan-prefetch thread."""
    for pool, streams in list(_EIG_POOL.values()):
        pool.shutdown(wait=True)
        for s in streams:
            s.synchronize()
    _EIG_POOL.clear()
    for pool in _POOL:
        pool.shutdown(wait=True)
    _POOL.clear()


EIG_THREADS = 4
# A new size class of the batched eigensolves starts below this fraction of
# the current class's largest order.
EIG_CLASS_RATIO = 0.5

# Sector eigensolver of the density-matrix split: "sector" (eigen.SectorEigh:
# one tridiagonalisation launch for all sectors of the bond, one batched
# divide and conquer over all their tridiagonals, one back-transformation
# launch) or "library" (one batched library eigh per size class on side
# streams).  Measured at a D = 2048 Hubbard bond (63 sectors up to 744):
# 13 ms against 34 ms.
EIG_SOLVER = "sector"

# Plain per-sector execution (the baseline the batched forms are measured
# against): one GEMM launch per sector product in every block plan and one
# eigensolve per sector, in sequence, in the two-site split.
PLAIN = False


class plain_path:
    """Context manager: run block plans and splits as the plain per-sector
    loop (FP64, Krylov steps launched eagerly) on the engine of ``device``,
    restoring the batched forms and step graphs on exit."""

    def __init__(self, device=None):
        from .engine import engine
        self.eng = engine(device)

    def __enter__(self):
        global PLAIN
        self.eng.set_bs_plain(True)
        self.eng.set_graphs(False)
        PLAIN = True
        return self

    def __exit__(self, *exc):
        global PLAIN
        self.eng.set_bs_plain(False)
        self.eng.set_graphs(True)
        PLAIN = False
        return False


def _plain_eigh(grams):
    """One eigh per sector Gram matrix, in sequence: [(n, [charge], lam, V)]."""
    out = []
    for c in sorted(grams, key=lambda c: -grams[c].shape[0]):
        lam, V = torch.linalg.eigh(grams[c])
        out.append((grams[c].shape[0], [c], lam.unsqueeze(0), V.unsqueeze(0)))
    return out


def _sector_eigh(grams, dev):
    """Eigen-decompositions of the sector Gram matrices by eigen.SectorEigh
    (all sectors in the same launches), in the format of _batched_eigh: one
    (n, [charge], lam[1, n], V[1, n, n]) per sector, ascending."""
    keys = list(grams)
    se = eigen.SectorEigh([grams[c].shape[0] for c in keys], dev)
    for j, c in enumerate(keys):
        se.matrix(j).copy_(grams[c])
    lams = se.eigenvalues()
    vecs = se.vectors(se.sizes)
    se.check()
    return [(n, [c], lam.unsqueeze(0), V.unsqueeze(0))
            for n, c, lam, V in zip(se.sizes, keys, lams, vecs)]


def _batched_eigh(grams, dev):
    """Eigen-decompositions of the sector Gram matrices: with EIG_SOLVER =
    "sector" (on a GPU, orders up to eigen.MAX_ORDER) by _sector_eigh,
    otherwise one batched library eigh per
    size class (zero padding).  A batched call costs a latency that grows
    with n plus a little per matrix, so the classes are geometric (a new
    class starts below 0.5 x the largest of the current one) and, on a GPU,
    run concurrently on side streams from a small thread pool (the solver is
    launch-latency bound).  Returns [(nb, charges, lam, V)]."""
    if (EIG_SOLVER == "sector" and dev.type == "cuda" and grams
            and max(g.shape[0] for g in grams.values()) <= eigen.MAX_ORDER):
        return _sector_eigh(grams, dev)
    classes = {}
    cur = None
    for c in sorted(grams, key=lambda c: -grams[c].shape[0]):
        n = grams[c].shape[0]
        if cur is None or n < EIG_CLASS_RATIO * cur:
            cur = -(-n // 32) * 32
        classes.setdefault(cur, []).append(c)

    def solve(nb, cs):
        Gb = torch.zeros((len(cs), nb, nb), dtype=torch.float64, device=dev)
        for j, c in enumerate(cs):
            n = grams[c].shape[0]
            Gb[j, :n, :n] = grams[c]
            if n < nb:
                # padding eigenvalues below the (PSD) spectrum: they come first
                # and the true eigenvectors stay on the first n coordinates
                pad = torch.arange(n, nb, device=dev)
                Gb[j, pad, pad] = -(1.0 + grams[c].diagonal().abs().sum())
        lam, V = torch.linalg.eigh(Gb)
        return (nb, cs, lam, V)

    items = list(classes.items())
    if dev.type != "cuda" or len(items) < 2 or EIG_THREADS < 2:
        return [solve(nb, cs) for nb, cs in items]
    pool, streams = _eig_workers(dev, EIG_THREADS)
    main = torch.cuda.current_stream(dev)
    ready = torch.cuda.Event()
    ready.record(main)

    def task(k):
        s = streams[k % len(streams)]
        with torch.cuda.stream(s):
            s.wait_event(ready)
            out = solve(*items[k])
            done = torch.cuda.Event()
            done.record(s)
        return out, done

    # classes on one stream run in order; the largest go first
    results = list(pool.map(task, range(len(items))))
    for _, done in results:
        main.wait_event(done)
    return [r for r, _ in results]


def _strided_pieces(pieces, device):
    """Element indices of rectangular pieces copied between two strided
    buffers: row p = (blk_off, ldb, nr, nc, src_base, sr, sc, r0, c0); element
    (i, j) of the piece (i < nr, j < nc) is blk_off + i + j*ldb on the block
    side and src_base + (r0 + i)*sr + (c0 + j)*sc on the source side (strides
    may be negative).  Returns (blk_idx, src_idx) on `device`."""
    if len(pieces) == 0:
        z = torch.zeros(0, dtype=torch.int64, device=device)
        return z, z
    Ph = np.asarray(pieces, dtype=np.int64).reshape(len(pieces), -1)
    P = _upload(Ph, device)
    sizes = P[:, 2] * P[:, 3]
    total = int((Ph[:, 2] * Ph[:, 3]).sum())
    pid = torch.repeat_interleave(torch.arange(len(pieces), device=device), sizes,
                                  output_size=total)
    start = torch.cumsum(sizes, 0) - sizes
    local = torch.arange(total, device=device) - start[pid]
    Pp = P[pid]
    i, j = local % Pp[:, 2], local // Pp[:, 2]
    return (Pp[:, 0] + i + j * Pp[:, 1],
            Pp[:, 4] + (Pp[:, 7] + i) * Pp[:, 5] + (Pp[:, 8] + j) * Pp[:, 6])


def _gemms(descs, alpha=1.0, beta=0.0):
    """One mh_gemm_batched launch over raw (pointer, extent, stride) rows."""
    if not descs:
        return
    from ._lib import load, raise_for
    from .engine import engine
    import ctypes
    d = np.asarray(descs, dtype=np.int64).reshape(-1, 12)
    e = engine(torch.device("cuda", torch.cuda.current_device()))
    h = e.bind_stream()
    raise_for(load().mh_gemm_batched(h, len(d), d.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                     float(alpha), float(beta)), h, "mh_gemm_batched")


def _split_rho_sector(tmap, vec, trunc, absorb, hook=None):
    """split_two_site_rho with every per-sector product batched: the sector
    density matrices, the projections on eigenvector blocks, both levels of
    the exact cut and the assembly of the two new site tensors run as a few
    launches per bond (mh_gemm_batched, the batched eigensolver, gathers over
    flat buffers) instead of a handful of library calls per sector.  Same
    arithmetic as the per-sector path, in the same order of operations per
    sector.  Returns None when a bond needs the per-sector path: a merged
    spectrum whose kept values are not signal-then-tail within a sector, or a
    cut the second level cannot resolve (the caller then takes that path, or
    the SVD)."""
    xl = tmap.xl
    dev = vec.device
    left = absorb == "right"
    charges = list(tmap.charges)
    cat = torch.empty_like(vec)
    cat[tmap.dest(dev)] = vec
    pc = cat.data_ptr()
    # sector c: X_c = M_c (left sweep, n = R rows of width w = C) or M_c^T
    # (right sweep, n = C, w = R), both read in place from the row-major M_c
    sec = []
    for c in charges:
        R, C = tmap.shape[c]
        b = tmap.base[c]
        if left:
            sec.append((c, R, C, b, C, 1, R, C))
        else:
            sec.append((c, C, R, b, 1, C, R, C))
    ns = [t[1] for t in sec]
    se = eigen.SectorEigh(ns, dev)
    pA, pZ = se.A.data_ptr(), se.Z.data_ptr()
    # level 1: rho_c = X X^T straight into the eigensolver's matrices
    _gemms([(pc + 8 * b, pc + 8 * b, pA + 8 * int(se.a_off[j]), n, n, w, sr, sc_, sc_, sr, 1, n)
            for j, (c, n, w, b, sr, sc_, R, C) in enumerate(sec)])
    lams = se.eigenvalues()
    lam_all = torch.cat(lams).cpu().numpy()
    se.check()
    lam_host, pos = {}, 0
    for j, (c, n, w, b, sr, sc_, R, C) in enumerate(sec):
        r = min(R, C)
        lam_host[c] = np.sqrt(np.maximum(lam_all[pos:pos + n][::-1][:r], 0.0))
        pos += n
    keys = charges
    s2 = np.concatenate([lam_host[c] ** 2 for c in keys])
    noise = np.concatenate([np.full(len(lam_host[c]), tr.RHO_NOISE * ns[j] * tr.EPS64 *
                                    float(lam_host[c][0] ** 2 if len(lam_host[c]) else 0.0))
                            for j, c in enumerate(keys)])
    srt = np.argsort(-s2, kind="stable")
    wmax = max(t[2] for t in sec)
    z_off = [int(o) for o in se.a_off]
    if tr.rho_cut_resolved(s2[srt], noise[srt], trunc):
        keep = _global_keep_host({c: (None, sv) for c, sv in lam_host.items()}, trunc)
        keep = {c: k for c, k in keep.items() if lam_host[c][k - 1] > 0.0}
        se.vectors([keep.get(c, 0) for c in charges])
        # per kept sector: nsig = k columns of descending eigenvectors, no tail
        plan = {c: (k, 0) for c, k in keep.items()}
        Tbuf, tb_off, pT = None, {}, 0
    else:
        out = _refine_sector(tmap, sec, se, lam_host, lam_all, trunc, dev, pc, wmax)
        if out is None:
            return None
        plan, Tbuf, tb_off = out
        pT = Tbuf.data_ptr() if Tbuf is not None else 0
    # projections of X on the kept vectors: rows of s v (left) or of (u s)^T
    # (right), signal rows from the descending eigenvectors (read backwards
    # from the eigensolver's columns), tail rows from the tail vectors
    order = [c for c in charges if c in plan]
    kept = {c: plan[c][0] + plan[c][1] for c in order}
    rowb, acc = {}, 0
    for c in order:
        rowb[c] = acc
        acc += kept[c]
    Bk = torch.zeros((max(acc, 1), wmax), dtype=torch.float64, device=dev)
    pB = Bk.data_ptr()
    descs = []
    info = {t[0]: (j, t) for j, t in enumerate(sec)}
    for c in order:
        j, (_, n, w, b, sr, sc_, R, C) = info[c]
        ks, kt = plan[c]
        zlast = z_off[j] + (n - 1) * n   # column n-1 of Z_j: the largest eigenvalue
        if ks:
            descs.append((pZ + 8 * zlast, pc + 8 * b, pB + 8 * rowb[c] * wmax,
                          ks, w, n, -n, 1, sr, sc_, wmax, 1))
        if kt:
            descs.append((pT + 8 * tb_off[c], pc + 8 * b, pB + 8 * (rowb[c] + ks) * wmax,
                          kt, w, n, n, 1, sr, sc_, wmax, 1))
    _gemms(descs)
    svec = torch.linalg.vector_norm(Bk[:acc], dim=1)
    norm = torch.linalg.vector_norm(svec)
    mid = Space(kept.items())
    l1 = BlockLayout(xl.lspace, xl.phys, 1, mid)
    l2 = BlockLayout(mid, xl.phys, 1, xl.rspace)
    if hook is not None:
        hook(l1, l2)   # host work while the device runs the products above
    s_host = (svec / norm).cpu().numpy()
    svals = {c: s_host[rowb[c]:rowb[c] + kept[c]] for c in order}
    # assembly: the eigenvector side straight from Z (and the tail vectors),
    # the projected side from Bk scaled by 1/norm
    zp, tp, bp = [], [], []
    for c in order:
        j, (_, n, w, b, sr, sc_, R, C) = info[c]
        ks, kt = plan[c]
        k = ks + kt
        zlast = z_off[j] + (n - 1) * n
        ro = 0
        for qa, s1, Da in tmap.rows[c]:
            blk = l1.index[(qa, (s1,))].off
            if left:   # u[ro + i, col] = eigenvector col at row ro + i
                if ks:
                    zp.append((blk, Da, Da, ks, zlast, 1, -n, ro, 0))
                if kt:
                    tp.append((blk + ks * Da, Da, Da, kt, tb_off[c], 1, n, ro, 0))
            else:      # (u s)[ro + i, col] = Bk[row col, ro + i]
                bp.append((blk, Da, Da, k, rowb[c] * wmax, 1, wmax, ro, 0))
            ro += Da
        co = 0
        for s2_, qb, Db in tmap.cols[c]:
            blk = l2.index[(c, (s2_,))].off
            if left:   # (s v)[col, co + jj] = Bk[row col, co + jj]
                bp.append((blk, k, k, Db, rowb[c] * wmax, wmax, 1, 0, co))
            else:      # v[col, co + jj] = eigenvector col at row co + jj
                if ks:
                    zp.append((blk, k, ks, Db, zlast, -n, 1, 0, co))
                if kt:
                    tp.append((blk + ks, k, kt, Db, tb_off[c], n, 1, 0, co))
            co += Db
    d1 = torch.zeros(l1.size, dtype=torch.float64, device=dev)
    d2 = torch.zeros(l2.size, dtype=torch.float64, device=dev)
    dz, dp = (d1, d2) if left else (d2, d1)
    for pieces, src, dst, scale in ((zp, se.Z, dz, None), (tp, Tbuf, dz, None),
                                    (bp, Bk.reshape(-1), dp, norm)):
        if not pieces:
            continue
        bi, si = _strided_pieces(pieces, dev)
        vals = src[si]
        dst[bi] = vals / scale if scale is not None else vals
    return BlockTensor(l1, d1), BlockTensor(l2, d2), svals


def _refine_sector(tmap, sec, se, lam_host, lam_all, trunc, dev, pc, wmax):
    """Second level of the exact cut for _split_rho_sector, batched over the
    sectors (the per-sector form and its reasoning: _refine_tail).  Returns
    ({c: (signal kept, tail kept)}, tail vector buffer, {c: its offset}) or
    None when the bond needs the per-sector path."""
    se.vectors(se.sizes)
    pZ = se.Z.data_ptr()
    z_off = [int(o) for o in se.a_off]
    parts = []   # (j, c, n, w, b, sr, sc, r, ks)
    for j, (c, n, w, b, sr, sc_, R, C) in enumerate(sec):
        lam = lam_host[c] ** 2
        lam_max = float(lam[0]) if len(lam) else 0.0
        bound = tr.RHO_NOISE * n * tr.EPS64 * lam_max
        ks = int(np.count_nonzero(lam > SIGNAL_MARGIN * bound))
        parts.append((j, c, n, w, b, sr, sc_, min(R, C), ks, lam_max))
    # B = Q^T X for every sector (rows descending), signal rows first
    rb, acc = [], 0
    for p in parts:
        rb.append(acc)
        acc += p[2]
    Bp = torch.zeros((max(acc, 1), wmax), dtype=torch.float64, device=dev)
    pB = Bp.data_ptr()
    _gemms([(pZ + 8 * (z_off[j] + (n - 1) * n), pc + 8 * b, pB + 8 * rb[q] * wmax,
             n, w, n, -n, 1, sr, sc_, wmax, 1)
            for q, (j, c, n, w, b, sr, sc_, r, ks, lm) in enumerate(parts)])
    bnorm = torch.linalg.vector_norm(Bp, dim=1)
    # normalised signal rows, two Newton-Schulz steps (a batched product each)
    sig_rows = np.concatenate([np.arange(rb[q], rb[q] + p[8]) for q, p in enumerate(parts)]
                              + [np.zeros(0, dtype=np.int64)]).astype(np.int64)
    vb, acc_v = [], 0
    for p in parts:
        vb.append(acc_v)
        acc_v += p[8]
    tails = [q for q, p in enumerate(parts) if p[8] < p[7]]
    if acc_v:
        sr_idx = _upload(sig_rows, dev)
        Vs = Bp[sr_idx] / bnorm[sr_idx, None]
        gsz = [p[8] * p[8] for p in parts]
        go = np.concatenate([[0], np.cumsum(gsz)[:-1]]).astype(np.int64)
        G = torch.empty(max(int(sum(gsz)), 1), dtype=torch.float64, device=dev)
        pG = G.data_ptr()
        for _ in range(2):
            pV = Vs.data_ptr()
            _gemms([(pV + 8 * vb[q] * wmax, pV + 8 * vb[q] * wmax, pG + 8 * int(go[q]),
                     p[8], p[8], p[3], wmax, 1, 1, wmax, 1, p[8])
                    for q, p in enumerate(parts) if p[8]])
            Vn = 1.5 * Vs
            pN = Vn.data_ptr()
            _gemms([(pG + 8 * int(go[q]), pV + 8 * vb[q] * wmax, pN + 8 * vb[q] * wmax,
                     p[8], p[3], p[8], 1, p[8], wmax, 1, wmax, 1)
                    for q, p in enumerate(parts) if p[8]], alpha=-0.5, beta=1.0)
            Vs = Vn
        # tail rows projected off the signal rows, twice
        psz = [(p[2] - p[8]) * p[8] for p in parts]
        po = np.concatenate([[0], np.cumsum(psz)[:-1]]).astype(np.int64)
        P = torch.empty(max(int(sum(psz)), 1), dtype=torch.float64, device=dev)
        pP, pV = P.data_ptr(), Vs.data_ptr()
        proj = [q for q in tails if parts[q][8]]
        for _ in range(2):
            _gemms([(pB + 8 * (rb[q] + parts[q][8]) * wmax, pV + 8 * vb[q] * wmax,
                     pP + 8 * int(po[q]), parts[q][2] - parts[q][8], parts[q][8], parts[q][3],
                     wmax, 1, 1, wmax, 1, parts[q][2] - parts[q][8]) for q in proj])
            _gemms([(pP + 8 * int(po[q]), pV + 8 * vb[q] * wmax,
                     pB + 8 * (rb[q] + parts[q][8]) * wmax, parts[q][2] - parts[q][8],
                     parts[q][3], parts[q][8], 1, parts[q][2] - parts[q][8], wmax, 1, wmax, 1)
                    for q in proj], alpha=-1.0, beta=1.0)
    # the tail Gram matrices straight into a second batched eigensolver
    se2 = eigen.SectorEigh([parts[q][2] - parts[q][8] for q in tails], dev) if tails else None
    if se2 is not None:
        pA2 = se2.A.data_ptr()
        _gemms([(pB + 8 * (rb[q] + parts[q][8]) * wmax, pB + 8 * (rb[q] + parts[q][8]) * wmax,
                 pA2 + 8 * int(se2.a_off[t]), parts[q][2] - parts[q][8],
                 parts[q][2] - parts[q][8], parts[q][3], wmax, 1, 1, wmax, 1,
                 parts[q][2] - parts[q][8]) for t, q in enumerate(tails)])
        mus = se2.eigenvalues()
        flat = [m.reshape(-1) for m in mus]
    else:
        flat = []
    ss_dev = bnorm[_upload(sig_rows, dev)] if len(sig_rows) else bnorm[:0]
    host = torch.cat(flat + [ss_dev]).cpu().numpy()
    if se2 is not None:
        se2.check()
    mu_host, pos = {}, 0
    for t, q in enumerate(tails):
        m = parts[q][2] - parts[q][8]
        mu_host[q] = host[pos:pos + m]
        pos += m
    ss_host = {}
    for q, p in enumerate(parts):
        ss_host[q] = host[pos:pos + p[8]]
        pos += p[8]
    vals, s2_all, nz_all = {}, [], []
    for q, p in enumerate(parts):
        j, c, n, w, b, sr, sc_, r, ks, lam_max = p
        tail = np.zeros(0)
        if q in mu_host:
            m = n - ks
            tail = np.maximum(mu_host[q][::-1][:min(r - ks, m)], 0.0)
        sv = np.concatenate([ss_host[q] ** 2, tail])
        vals[c] = (sv, ks)
        s2_all.append(sv)
        nz_all.append(np.full(len(sv), (tr.RHO_NOISE * n * tr.EPS64) ** 2 * lam_max))
    s2 = np.concatenate(s2_all)
    nz = np.concatenate(nz_all)
    srt = np.argsort(-s2, kind="stable")
    if not tr.rho_cut_resolved(s2[srt], nz[srt], trunc):
        return None
    keep = _global_keep_host({c: (None, np.sqrt(np.sort(v[0])[::-1])) for c, v in vals.items()},
                             trunc)
    plan = {}
    for q, p in enumerate(parts):
        c, ks = p[1], p[8]
        k = keep.get(c, 0)
        if k == 0:
            continue
        sv = vals[c][0]
        # the kept values of a sector must be its signal values (descending)
        # followed by its largest tail values: anything else -> per-sector path
        sig = sv[:ks]
        if np.any(np.diff(sig) > 0) or (len(sv) > ks and ks and sv[ks:].max() > sig[-1]):
            return None
        if np.sqrt(np.sort(sv)[::-1][k - 1]) <= 0.0:
            continue
        plan[c] = (min(k, ks), max(0, k - ks))
    # tail vectors Q_n W (top kept columns of W, descending) for the kept tails
    need = {q: plan[parts[q][1]][1] for q in tails if parts[q][1] in plan and plan[parts[q][1]][1]}
    Tbuf, tb_off = None, {}
    if need:
        kt_all = [need.get(q, 0) for q in tails]
        se2.vectors(kt_all)
        pZ2 = se2.Z.data_ptr()
        acc_t = 0
        for q in need:
            tb_off[parts[q][1]] = acc_t
            acc_t += parts[q][2] * need[q]
        Tbuf = torch.empty(acc_t, dtype=torch.float64, device=dev)
        pT = Tbuf.data_ptr()
        descs = []
        for t, q in enumerate(tails):
            if q not in need:
                continue
            j, c, n, w, b, sr, sc_, r, ks, lm = parts[q]
            m, kt = n - ks, need[q]
            zq = z_off[j] + (n - 1 - ks) * n         # Q column ks (descending)
            zw = int(se2.a_off[t]) + (m - 1) * m     # W column m-1 (largest)
            descs.append((pZ + 8 * zq, pZ2 + 8 * zw, pT + 8 * tb_off[c], n, kt, m,
                          1, -n, 1, -m, 1, n))
        _gemms(descs)
    return plan, Tbuf, tb_off


def split_two_site_rho(tmap, vec, trunc, absorb, hook=None):
    """Device split by the sector reduced density matrices (White's original
    DMRG step): for the left sweep rho_c = M_c M_c^T gives u and s^2, and
    s v = u^T M_c is an exact projection; for the right sweep rho_c = M_c^T M_c
    gives v and u s = M_c v^T.  Where the truncation's relative cutoff falls
    into the eigenvalues' noise (``truncation.rho_cut_resolved``) the bond is
    split by the reference's sector SVDs instead (``split_two_site``); the
    kept singular values are the row (column) norms of s v (u s).  The eigenpairs of all sectors come from one
    batched library eigh per size class (EIG_SOLVER = "library") or from
    `eigen.SectorEigh` (EIG_SOLVER = "sector": one tridiagonalisation launch
    for every sector, divide and conquer per sector, back-transform of the
    kept columns only); the truncation acts on the merged spectrum
    s = sqrt(lambda) exactly as for the SVD.  Returns (A1, A2, {c: s_c})."""
    if (EIG_SOLVER == "sector" and not PLAIN and vec.device.type == "cuda"
            and all(max(tmap.shape[c]) <= eigen.MAX_ORDER for c in tmap.charges)):
        out = _split_rho_sector(tmap, vec, trunc, absorb, hook)
        if out is not None:
            return out
    xl = tmap.xl
    dev = vec.device
    cat = torch.empty_like(vec)
    cat[tmap.dest(dev)] = vec
    left = absorb == "right"
    charges = list(tmap.charges)
    mats = {}
    for c in charges:
        R, C = tmap.shape[c]
        mats[c] = cat[tmap.base[c]:tmap.base[c] + R * C].view(R, C)
    order = {c: (tmap.shape[c][0] if left else tmap.shape[c][1]) for c in charges}
    # the kept-columns-only form of the sector solver (no refinement level)
    use_sector = (not PLAIN and EIG_SOLVER == "sector_kept" and dev.type == "cuda"
                  and all(n <= eigen.MAX_ORDER for n in order.values()))
    lam_host = {}
    if use_sector:
        se = eigen.SectorEigh([order[c] for c in charges], dev)
        for j, c in enumerate(charges):
            M = mats[c]
            if left:
                torch.matmul(M, M.t(), out=se.matrix(j))
            else:
                torch.matmul(M.t(), M, out=se.matrix(j))
        lams = se.eigenvalues()
        lam_all = torch.cat(lams).cpu().numpy()
        se.check()
        pos = 0
        for c in charges:
            n = order[c]
            R, C = tmap.shape[c]
            l = np.maximum(lam_all[pos:pos + n][::-1][:min(R, C)], 0.0)
            lam_host[c] = np.sqrt(l)
            pos += n
    else:
        grams = {c: (mats[c] @ mats[c].t() if left else mats[c].t() @ mats[c]) for c in charges}
        pending = _plain_eigh(grams) if PLAIN else _batched_eigh(grams, dev)
        eig = {}
        lam_all = torch.cat([p[2].reshape(-1) for p in pending]).cpu().numpy() if pending else None
        pos = 0
        for nb, cs, lam, V in pending:
            lh = lam_all[pos:pos + lam.numel()].reshape(lam.shape)
            pos += lam.numel()
            for j, c in enumerate(cs):
                R, C = tmap.shape[c]
                r = min(R, C)
                # descending; only the top min(R, C) can be nonzero
                l = np.maximum(lh[j, ::-1][:r], 0.0)
                lam_host[c] = np.sqrt(l)
                eig[c] = (V, j, nb)
    # the merged spectrum and the noise of each value: a sector's eigenvalues
    # carry about n_c eps64 lambda_max,c of absolute error (truncation.py)
    keys = list(lam_host)
    s2 = np.concatenate([lam_host[c] ** 2 for c in keys])
    noise = np.concatenate([np.full(len(lam_host[c]), tr.RHO_NOISE * order[c] * tr.EPS64 *
                                    float(lam_host[c][0] ** 2 if len(lam_host[c]) else 0.0))
                            for c in keys])
    srt = np.argsort(-s2, kind="stable")
    vecs = None
    if not tr.rho_cut_resolved(s2[srt], noise[srt], trunc):
        # the cut falls where sqrt(lambda) is noise
        refined = None
        if not use_sector and not PLAIN:
            refined = _refine_tail(mats, left, order, lam_host, eig, dev, trunc)
        if refined is None:
            # the reference's sector SVDs
            return split_two_site(tmap, vec.detach().cpu().numpy(), trunc, absorb, dev)
        lam_host, vecs = refined
    keep = _global_keep_host({c: (None, s) for c, s in lam_host.items()}, trunc)
    keep = {c: k for c, k in keep.items() if lam_host[c][k - 1] > 0.0}
    # top-k eigenvectors of each kept sector, descending
    if vecs is not None:
        top = {c: vecs[c][:, :k] for c, k in keep.items()}
    elif use_sector:
        vk = se.vectors([keep.get(c, 0) for c in charges])
        top = {c: vk[j].flip(-1) for j, c in enumerate(charges) if c in keep}
    else:
        top = {}
        for c, k in keep.items():
            V, j, nb = eig[c]
            top[c] = V[j, :order[c], nb - k:].flip(-1)
    mid = Space(keep.items())
    l1 = BlockLayout(xl.lspace, xl.phys, 1, mid)
    l2 = BlockLayout(mid, xl.phys, 1, xl.rspace)
    # s v = Q^T M (left) or u s = M Q (right); the norms of its rows
    # (columns) are the singular values, accurate where sqrt(lambda) is not
    proj, s_dev = {}, []
    for c in keep:
        B = top[c].t() @ mats[c] if left else mats[c] @ top[c]
        proj[c] = B
        s_dev.append(torch.linalg.vector_norm(B, dim=1 if left else 0))
    s_all = torch.cat(s_dev)
    norm = torch.linalg.vector_norm(s_all)
    if hook is not None:
        hook(l1, l2)   # host work while the device runs the products above
    s_host = (s_all / norm).cpu().numpy()
    svals, us, vs, upieces, vpieces = {}, [], [], [], []
    ubase = vbase = spos = 0
    for c, k in keep.items():
        R, C = tmap.shape[c]
        Q = top[c]
        svals[c] = s_host[spos:spos + k]
        spos += k
        if left:
            u = Q                                # R x k
            v = proj[c] / norm                   # s v, normalised s
        else:
            v = Q.t()                            # k x C
            u = proj[c] / norm                   # u s
        us.append(u.reshape(-1))
        vs.append(v.reshape(-1))
        ro = 0
        for qa, s1, Da in tmap.rows[c]:
            upieces.append((l1.index[(qa, (s1,))].off, Da, k, ubase, k, ro, 0))
            ro += Da
        co = 0
        for s2, qb, Db in tmap.cols[c]:
            vpieces.append((l2.index[(c, (s2,))].off, k, Db, vbase, C, 0, co))
            co += Db
        ubase += R * k
        vbase += k * C
    d1 = torch.zeros(l1.size, dtype=torch.float64, device=dev)
    d2 = torch.zeros(l2.size, dtype=torch.float64, device=dev)
    blk, mat = piece_index(np.array(upieces, dtype=np.int64).reshape(-1, 7), dev)
    d1[blk] = torch.cat(us)[mat]
    blk, mat = piece_index(np.array(vpieces, dtype=np.int64).reshape(-1, 7), dev)
    d2[blk] = torch.cat(vs)[mat]
    return BlockTensor(l1, d1), BlockTensor(l2, d2), svals


# Level-1 eigenpairs at least SIGNAL_MARGIN times above their noise bound
# are kept as they are; the rest of each sector is diagonalised again.
SIGNAL_MARGIN = 1.0e3


def _refine_tail(mats, left, order, lam_host, eig, dev, trunc):
    """Second level of the density-matrix split, for a bond whose relative
    cutoff falls below the resolution of sqrt(lambda).

    Per sector, with X = M_c (left sweep) or M_c^T (right sweep) and the
    level-1 eigenvectors Q = [Q_s | Q_n] of X X^T (Q_s: eigenvalues above
    SIGNAL_MARGIN times the noise bound):
      - s_i = ||q_i^T X|| for the signal rows (accurate to ~eps64 s_max);
      - V_s = the rows of Q_s^T X normalised and orthonormalised (two
        Newton-Schulz steps); the rows of Q_n^T X are projected off span(V_s)
        (twice), which removes the signal admixture that Q_n's rounding
        carries, so the block's Gram matrix G = P P^T has eigenvalues at its
        own scale (its noise is eps64 times the tail's largest value, not the
        sector's);
      - the tail's values and vectors come from G's eigenpairs, W, with the
        tail vectors Q_n W exactly orthogonal to Q_s.
    Returns ({c: descending values}, {c: n_c x r_c vectors}) or None when the
    cut is still not resolved (the caller then takes the SVD split)."""
    tails, parts = {}, {}
    vals, noise2 = {}, {}
    for c, (V, j, nb) in eig.items():
        n = order[c]
        R, C = mats[c].shape
        r = min(R, C)
        X = mats[c] if left else mats[c].t()
        lam = lam_host[c] ** 2
        lam_max = float(lam[0]) if len(lam) else 0.0
        bound = tr.RHO_NOISE * n * tr.EPS64 * lam_max
        k_s = int(np.count_nonzero(lam > SIGNAL_MARGIN * bound))
        Q = V[j, :n, nb - n:].flip(-1)                      # n x n, descending
        Qs, Qn = Q[:, :k_s], Q[:, k_s:]
        Bs = Qs.t() @ X
        ss = torch.linalg.vector_norm(Bs, dim=1)
        parts[c] = (Qs, Qn, ss, r, k_s)
        # the level-2 values' noise: the tail Gram's own rounding plus the
        # projection's, eps64 s_max on every entry
        noise2[c] = (tr.RHO_NOISE * n * tr.EPS64) ** 2 * lam_max
        if k_s >= r:
            continue
        Bn = Qn.t() @ X
        if k_s > 0:
            Vs = Bs / ss[:, None]
            for _ in range(2):
                Vs = 1.5 * Vs - 0.5 * (Vs @ Vs.t()) @ Vs
            for _ in range(2):
                Bn = Bn - (Bn @ Vs.t()) @ Vs
        tails[c] = Bn @ Bn.t()
    pending = _batched_eigh(tails, dev) if tails else []
    # one transfer for the tail eigenvalues and the signal row norms
    keys = list(parts)
    flat = [p[2].reshape(-1) for p in pending] + [parts[c][2] for c in keys]
    host = torch.cat(flat).cpu().numpy()
    ntail = sum(p[2].numel() for p in pending)
    lam_all = host[:ntail]
    ss_host, pos = {}, ntail
    for c in keys:
        k_s = parts[c][4]
        ss_host[c] = host[pos:pos + k_s]
        pos += k_s
    tail_host, tail_vec = {}, {}
    pos = 0
    for nb, cs, lam, W in pending:
        lh = lam_all[pos:pos + lam.numel()].reshape(lam.shape)
        pos += lam.numel()
        for j, c in enumerate(cs):
            Qs, Qn, ss, r, k_s = parts[c]
            m = Qn.shape[1]
            t = min(r - k_s, m)
            mu = np.maximum(lh[j, ::-1][:t], 0.0)
            tail_host[c] = mu
            tail_vec[c] = Qn @ W[j, :m, nb - t:].flip(-1)
    out_vals, out_vecs = {}, {}
    s2_all, nz_all = [], []
    for c, (Qs, Qn, ss, r, k_s) in parts.items():
        sv = np.concatenate([ss_host[c] ** 2, tail_host.get(c, np.zeros(0))])
        vv = torch.cat([Qs, tail_vec[c]], dim=1) if c in tail_vec else Qs
        o = np.argsort(-sv, kind="stable")
        out_vals[c] = np.sqrt(sv[o])
        out_vecs[c] = vv[:, _upload(o, vv.device)]
        s2_all.append(sv)
        nz_all.append(np.full(len(sv), noise2[c]))
    s2 = np.concatenate(s2_all)
    nz = np.concatenate(nz_all)
    srt = np.argsort(-s2, kind="stable")
    if not tr.rho_cut_resolved(s2[srt], nz[srt], trunc):
        return None
    return out_vals, out_vecs


def _global_keep_host(svd, trunc):
    keys = list(svd.keys())
    vals = [svd[q][1] for q in keys]
    allv = np.concatenate(vals)
    owner = np.concatenate([np.full(len(v), j) for j, v in enumerate(vals)])
    order = np.argsort(-allv, kind="stable")
    k, _ = trunc.truncate(allv[order])
    k = max(k, 1)
    counts = np.bincount(owner[order[:k]], minlength=len(keys))
    return {q: int(c) for q, c in zip(keys, counts) if c > 0}


def _prefetch(m, alg, site, absorb, l1, l2, xl0, dev):
    """Plans the sweep needs next, made from the new site layouts l1, l2 of
    the bond (site, site + 1) as soon as its cut is known, on a host thread
    beside the rest of the split: the overlap x' = A1' A2' (x layout xl0),
    the environment update that follows the bond, and the next bond's
    two-site tensor and H_eff.  Returns them by name; each consumer checks
    that the layouts it meets are these very objects, so a plan is only ever
    the one it would have made itself."""
    t0 = time.perf_counter()
    mps, hs, mpo = m.mps, m.hstorage, m.mpo

    def make(*a):
        return bs.make_plan(*a, device=dev)

    pre = {"overlap": (l1, l2, make("two_site", l1, l2, xl0))}
    if absorb == "right":       # left sweep: hs[site+1] from l1, then bond site+1
        old = hs[site]
        new_lay = EnvLayout(l1.rspace, mpo.qw[site + 1])
        pre["env"] = (l1, old.layout, new_lay,
                      make("updateleft", l1, old.layout, mpo[site], new_lay, m.prec), m.prec)
        nxt = site + 1
        if nxt + 1 < len(mps):
            B = mps[nxt + 1]
            xl = BlockLayout(l2.lspace, l2.phys, 2, B.rspace)
            pre["two_site"] = (l2, B.layout, xl, make("two_site", l2, B.layout, xl))
            R = hs[nxt + 2]
            pre["heff2"] = (xl, new_lay, R.layout, alg.prec,
                            make("heff2", xl, new_lay, R.layout, mpo[nxt], mpo[nxt + 1],
                                 alg.prec))
    else:                       # right sweep: hs[site+1] from l2, then bond site-1
        old = hs[site + 2]
        new_lay = EnvLayout(l2.lspace, mpo.qw[site + 1])
        pre["env"] = (l2, old.layout, new_lay,
                      make("updateright", l2, old.layout, mpo[site + 1], new_lay, m.prec),
                      m.prec)
        nxt = site - 1
        if nxt >= 0:
            A = mps[nxt]
            xl = BlockLayout(A.lspace, A.layout.phys, 2, l1.rspace)
            pre["two_site"] = (A.layout, l1, xl, make("two_site", A.layout, l1, xl))
            Lh = hs[nxt]
            pre["heff2"] = (xl, Lh.layout, new_lay, alg.prec,
                            make("heff2", xl, Lh.layout, new_lay, mpo[nxt], mpo[nxt + 1],
                                 alg.prec))
    pre["seconds"] = time.perf_counter() - t0
    return pre


_POOL = []


def _planner_thread():
    """The one host thread the sweeps' plan prefetch runs on."""
    if not _POOL:
        from concurrent.futures import ThreadPoolExecutor
        _POOL.append(ThreadPoolExecutor(max_workers=1, thread_name_prefix="mh-plans"))
    return _POOL[0]


def _prefetched(m):
    """The plans prefetched for m (waiting for the planner thread), or {}."""
    fut = getattr(m, "_pre_future", None)
    if fut is not None:
        m._pre_future = None
        m._pre = fut.result()
    return getattr(m, "_pre", {})


# Make the next plans on a host thread once the split's cut is known
# (_prefetch); False makes every plan where it is used.
PREFETCH = True


def _bond_step(m, alg, site, absorb):
    mps, hs = m.mps, m.hstorage
    A1, A2 = mps[site], mps[site + 1]
    dev = A1.data.device
    pre = _prefetched(m)
    m._pre = {}
    t0 = time.perf_counter()
    got = pre.get("two_site")
    if got and got[0] is A1.layout and got[1] is A2.layout:
        x = two_site(A1, A2, got[2], got[3])
    else:
        x = two_site(A1, A2)
    L, R = hs[site], hs[site + 2]
    got = pre.get("heff2")
    if (got and got[0] is x.layout and got[1] is L.layout and got[2] is R.layout
            and got[3] == alg.prec):
        plan = got[4]
    else:
        plan = bs.make_plan("heff2", x.layout, L.layout, R.layout, m.mpo[site],
                            m.mpo[site + 1], alg.prec)
    t1 = time.perf_counter()
    ev, vec, info, nit = plan.lanczos(L.data, R.data, x.data, alg.eig_maxiter, alg.eig_tol,
                                      reorth=alg.reorth)
    t2 = time.perf_counter()
    tmap = TwoSiteMap(x.layout)
    t2a = time.perf_counter()
    if getattr(alg, "split", "rho") == "svd":
        n1, n2, svals = split_two_site(tmap, vec.cpu().numpy(), alg.trunc, absorb, dev)
    else:
        def hook(l1, l2):
            m._pre_future = _planner_thread().submit(_prefetch, m, alg, site, absorb, l1, l2,
                                                     x.layout, dev)

        n1, n2, svals = split_two_site_rho(tmap, vec, alg.trunc, absorb,
                                           hook if PREFETCH else None)
    t2b = time.perf_counter()
    # err_1 = <u s v, twositemps> (dmrg.jl:109-110), left on the device: the
    # sweep reads all of them at its end
    got = _prefetched(m).get("overlap")
    oplan = got[2] if got and got[0] is n1.layout and got[1] is n2.layout else None
    st = alg.stats
    st["prefetch_s"] = st.get("prefetch_s", 0.0) + m._pre.get("seconds", 0.0)
    e = torch.dot(two_site(n1, n2, x.layout, oplan).data, x.data)
    t3 = time.perf_counter()
    st["map_s"] = st.get("map_s", 0.0) + (t2a - t2)
    st["split_s"] = st.get("split_s", 0.0) + (t2b - t2a)
    st["overlap_s"] = st.get("overlap_s", 0.0) + (t3 - t2b)
    st["plan_s"] = st.get("plan_s", 0.0) + (t1 - t0)
    st["lanczos_s"] += t2 - t1
    st["svd_s"] += t3 - t2
    st["matvec_flops"] += plan.flops() * nit
    st["bonds"] += 1
    return ev, n1, n2, svals, e


def _truncation_error(overlaps):
    """max over bonds of |1 - |<u s v, x>|| from the device overlaps (one
    host read)."""
    if not overlaps:
        return 0.0
    vals = torch.stack(overlaps).cpu().numpy()
    return float(np.max(np.abs(1.0 - np.abs(vals))))


def leftsweep(m, alg):
    """dmrg.jl:89-115 on block-sparse tensors"""
    mps = m.mps
    energies, overlaps = [], []
    for site in range(len(mps) - 2):
        ev, n1, n2, svals, e = _bond_step(m, alg, site, "right")
        energies.append(ev)
        mps[site], mps[site + 1] = n1, n2
        overlaps.append(e)
        t0 = time.perf_counter()
        m.updateleft(site)
        alg.stats["env_s"] += time.perf_counter() - t0
    return energies, _truncation_error(overlaps)


def rightsweep(m, alg):
    """dmrg.jl:117-144 on block-sparse tensors"""
    mps = m.mps
    energies, overlaps = [], []
    for site in range(len(mps) - 2, -1, -1):
        ev, n1, n2, svals, e = _bond_step(m, alg, site, "left")
        energies.append(ev)
        mps[site], mps[site + 1] = n1, n2
        mps.s[site + 1] = svals
        overlaps.append(e)
        t0 = time.perf_counter()
        m.updateright(site + 1)
        alg.stats["env_s"] += time.perf_counter() - t0
    return energies, _truncation_error(overlaps)


def sweep(m, alg):
    e1, d1 = leftsweep(m, alg)
    e2, d2 = rightsweep(m, alg)
    return e1 + e2, max(d1, d2)


def compute(m, alg, nsweeps=None):
    all_e, it, delta = [], 0, 2 * alg.tol
    maxit = alg.maxiter if nsweeps is None else nsweeps
    while it < maxit and delta > alg.tol:
        e, delta = sweep(m, alg)
        all_e += e
        it += 1
    return all_e, delta


def ground_state(state, mpo, alg, nsweeps=None):
    """ground_state!(state, h, alg) for a block-sparse state: returns
    (energies, delta) and leaves the optimised state in ``state``."""
    return compute(BlockExpectationCache(mpo, state, alg.prec), alg, nsweeps)
