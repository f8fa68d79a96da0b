"""Abelian-symmetric (block-sparse) MPS tensors, environments and the planner
that lowers their contractions to grouped-GEMM plans.

Stands in for TensorKit ``GradedSpace`` tensors as FiniteMPTools uses them on
charge-conserving models: the Hubbard chain with U1xU1 sectors
(utilities/fermion_siteops.jl:37-62, test state runtests.jl:324-337), graded
``prodmps`` / ``randommps`` (mps/initializers.jl:34-73).  Every ``@tensor`` on
such tensors runs TensorKit's per-sector loop; here one contraction is lowered
once to a *plan* whose stages each run as ONE device launch over all sectors
(grouped GEMM, block mode product; see include/mpsheff.h).

Charges are tuples of ints (one entry per U(1) factor), added component-wise.

Storage (all blocks column-major, one flat FP64 buffer per tensor):

* MPS-type tensor ``A[a, s1..sn, b]`` with ``q(a) + q(s1) + .. + q(sn) = q(b)``:
  blocks ``(qa, p=(s1..sn))`` ordered by left sector, then ``p``; the blocks of
  one left sector are adjacent, so that sector is one ``D_qa x ncols`` matrix
  (its *group*) with leading dimension ``D_qa``.
* Environment ``L[a_bra, w, a_ket]`` (and ``R[b_bra, w', b_ket]``) with
  ``q(a_bra) = q(w) + q(a_ket)``: blocks ``(q, w)`` of size ``D_q x D_{q-q(w)}``
  ordered by bra sector, then channel; the blocks of one bra sector form one
  ``D_q x K_q`` matrix.
* MPO channel charges follow from the MPO entries:
  ``q(w_right) = q(w_left) + q(s_out) - q(s_in)`` for every nonzero
  ``W[w_left, s_out, w_right, s_in]``.
"""
import itertools

import numpy as np
import torch

from . import _lib


def qadd(a, b):
    if len(a) == 2:  # U(1) x U(1), the common case: no generator
        return (a[0] + b[0], a[1] + b[1])
    return tuple(x + y for x, y in zip(a, b))


def qsub(a, b):
    if len(a) == 2:
        return (a[0] - b[0], a[1] - b[1])
    return tuple(x - y for x, y in zip(a, b))


def qzero(n):
    return (0,) * n


class Space:
    """Graded space: sectors (charge -> dimension), kept sorted by charge."""

    def __init__(self, sectors):
        acc = {}
        for q, d in sectors:
            if int(d) > 0:
                acc[tuple(int(v) for v in q)] = acc.get(tuple(q), 0) + int(d)
        self.sectors = sorted(acc.items())
        self.dim = dict(self.sectors)
        self.offset = {}
        o = 0
        for q, d in self.sectors:
            self.offset[q] = o
            o += d
        self.total = o

    def __contains__(self, q):
        return q in self.dim

    def __eq__(self, other):
        return isinstance(other, Space) and self.sectors == other.sectors

    def __repr__(self):
        return "Space(" + ", ".join(f"{q}=>{d}" for q, d in self.sectors) + ")"


class Block:
    __slots__ = ("qa", "p", "qb", "off", "rows", "cols", "colstart")

    def __init__(self, qa, p, qb, off, rows, cols, colstart):
        self.qa, self.p, self.qb = qa, p, qb
        self.off, self.rows, self.cols, self.colstart = off, rows, cols, colstart


# Block structures by (left sectors, physical charges, nphys, right
# sectors): a sweep rebuilds the same layouts several times per bond, and a
# layout is never modified after construction, so they are shared.
_LAYOUTS = {}
_LAYOUTS_MAX = 4096


class BlockLayout:
    """Block structure of an MPS-type tensor with ``nphys`` physical legs,
    each carrying the charges ``phys[s]``."""

    def __init__(self, lspace, phys, nphys, rspace):
        self.lspace, self.rspace = lspace, rspace
        self.phys = [tuple(q) for q in phys]
        self.nphys = nphys
        self.d = len(phys)
        key = (tuple(lspace.sectors), tuple(self.phys), nphys, tuple(rspace.sectors))
        tables = _LAYOUTS.get(key)
        if tables is None:
            tables = self._build(lspace, rspace)
            if len(_LAYOUTS) >= _LAYOUTS_MAX:
                _LAYOUTS.clear()
            _LAYOUTS[key] = tables
        # blocks; (qa, p) -> Block; qa -> (offset, ncols); qb -> [Block] and
        # qa -> [Block] in storage order; total size
        self.blocks, self.index, self.group, self.by_right, self.by_left, self.size = tables

    def _build(self, lspace, rspace):
        blocks, index, group, by_right, by_left = [], {}, {}, {}, {}
        combos = []
        for p in itertools.product(range(self.d), repeat=self.nphys):
            qs = None   # total charge of the physical indices p
            for s in p:
                qs = self.phys[s] if qs is None else qadd(qs, self.phys[s])
            combos.append((p, qs))
        off = 0
        rdim = rspace.dim
        two = all(len(q) == 2 for _, q in combos if q is not None)
        for qa, Da in lspace.sectors:
            g0, ncols = off, 0
            if two:
                a0, a1 = qa
                hits = [(p, qb, rdim.get(qb)) for p, qb in
                        ((p, qa if qs is None else (a0 + qs[0], a1 + qs[1])) for p, qs in combos)]
            else:
                hits = [(p, qb, rdim.get(qb)) for p, qb in
                        ((p, qa if qs is None else qadd(qa, qs)) for p, qs in combos)]
            row = []
            for p, qb, Db in hits:
                if Db is None:
                    continue
                b = Block(qa, p, qb, off, Da, Db, ncols)
                row.append(b)
                index[(qa, p)] = b
                rl = by_right.get(qb)
                if rl is None:
                    by_right[qb] = [b]
                else:
                    rl.append(b)
                off += Da * Db
                ncols += Db
            if row:
                blocks.extend(row)
                by_left[qa] = row
            if ncols:
                group[qa] = (g0, ncols)
        return blocks, index, group, by_right, by_left, off

    def dense_shape(self):
        return (self.lspace.total,) + (self.d,) * self.nphys + (self.rspace.total,)

    def unreachable(self):
        """Sectors of either bond that no block touches."""
        left = [q for q in self.lspace.dim if q not in self.group]
        right = [q for q in self.rspace.dim if q not in self.by_right]
        return left, right


class BlockTensor:
    """MPS-type block-sparse tensor: a layout and its flat FP64 buffer."""

    def __init__(self, layout, data):
        if data.dim() != 1 or data.numel() != layout.size or data.dtype != torch.float64:
            raise _lib.DimensionMismatch(2, f"buffer of {data.numel()} for layout of {layout.size}")
        self.layout, self.data = layout, data

    @classmethod
    def zeros(cls, layout, device="cuda"):
        return cls(layout, torch.zeros(layout.size, dtype=torch.float64, device=device))

    @classmethod
    def randn(cls, layout, gen, device="cuda"):
        v = gen.standard_normal(layout.size)
        return cls(layout, torch.as_tensor(v, dtype=torch.float64).to(device))

    @property
    def lspace(self):
        return self.layout.lspace

    @property
    def rspace(self):
        return self.layout.rspace

    def block(self, b):
        """View of one block as a (rows x cols) column-major matrix."""
        return self.data[b.off:b.off + b.rows * b.cols].view(b.cols, b.rows).t()

    def group_matrix(self, qa):
        g0, ncols = self.layout.group[qa]
        Da = self.layout.lspace.dim[qa]
        return self.data[g0:g0 + Da * ncols].view(ncols, Da).t()

    def to_dense(self):
        """Dense embedding (numpy), sectors at their offsets in each bond."""
        lay = self.layout
        out = np.zeros(lay.dense_shape())
        host = self.data.detach().cpu().numpy()
        lo, ro = lay.lspace.offset, lay.rspace.offset
        for b in lay.blocks:
            blk = host[b.off:b.off + b.rows * b.cols].reshape(b.cols, b.rows).T
            idx = (slice(lo[b.qa], lo[b.qa] + b.rows),) + tuple(b.p) + \
                  (slice(ro[b.qb], ro[b.qb] + b.cols),)
            out[idx] = blk
        return out

    @classmethod
    def from_dense(cls, layout, dense, device="cuda"):
        """Blocks of a dense embedding (entries outside the blocks are ignored)."""
        flat = np.zeros(layout.size)
        lo, ro = layout.lspace.offset, layout.rspace.offset
        for b in layout.blocks:
            idx = (slice(lo[b.qa], lo[b.qa] + b.rows),) + tuple(b.p) + \
                  (slice(ro[b.qb], ro[b.qb] + b.cols),)
            flat[b.off:b.off + b.rows * b.cols] = dense[idx].T.reshape(-1)
        return cls(layout, torch.as_tensor(flat).to(device))

    def norm(self):
        return torch.linalg.vector_norm(self.data).item()


class EnvLayout:
    """Block structure of an environment over ``space`` with MPO channels of
    charges ``qw``: blocks (q, w) of size D_q x D_{q - qw[w]}."""

    def __init__(self, space, qw):
        self.space = space
        self.qw = [tuple(q) for q in qw]
        self.w = len(qw)
        self.off = {}       # (q, w) -> offset
        self.colstart = {}  # (q, w) -> column start inside the sector's matrix
        self.group = {}     # q -> (offset, K_q)
        off = 0
        for q, Dq in space.sectors:
            g0, K = off, 0
            for w in range(self.w):
                qk = qsub(q, self.qw[w])
                if qk not in space:
                    continue
                self.off[(q, w)] = off
                self.colstart[(q, w)] = K
                Dk = space.dim[qk]
                off += Dq * Dk
                K += Dk
            if K:
                self.group[q] = (g0, K)
        self.size = off

    def dense_shape(self):
        return (self.space.total, self.w, self.space.total)


class BlockEnv:
    def __init__(self, layout, data):
        if data.numel() != layout.size:
            raise _lib.DimensionMismatch(2, "environment buffer size")
        self.layout, self.data = layout, data

    def to_dense(self):
        lay = self.layout
        out = np.zeros(lay.dense_shape())
        host = self.data.detach().cpu().numpy()
        o = lay.space.offset
        for (q, w), off in lay.off.items():
            Dq = lay.space.dim[q]
            qk = qsub(q, lay.qw[w])
            Dk = lay.space.dim[qk]
            out[o[q]:o[q] + Dq, w, o[qk]:o[qk] + Dk] = host[off:off + Dq * Dk].reshape(Dk, Dq).T
        return out

    @classmethod
    def from_dense(cls, layout, dense, device="cuda"):
        flat = np.zeros(layout.size)
        o = layout.space.offset
        for (q, w), off in layout.off.items():
            Dq = layout.space.dim[q]
            qk = qsub(q, layout.qw[w])
            Dk = layout.space.dim[qk]
            flat[off:off + Dq * Dk] = dense[o[q]:o[q] + Dq, w, o[qk]:o[qk] + Dk].T.reshape(-1)
        return cls(layout, torch.as_tensor(flat).to(device))


# ------------------------------------------------------------------ MPO charges

def mpo_channel_charges(mpo, phys):
    """Channel charges of every MPO bond (L+1 lists) from the nonzero entries;
    raises if an entry does not conserve the charges."""
    phys = [tuple(q) for q in phys]
    n = len(phys[0])
    if mpo[0].shape[0] != 1 or mpo[-1].shape[2] != 1:
        raise _lib.DimensionMismatch(2, "MPO boundaries must have dimension 1")
    bonds = [[qzero(n)]]
    for i, W in enumerate(mpo):
        wl, d, wr, d2 = W.shape
        if d != len(phys) or d2 != len(phys):
            raise _lib.DimensionMismatch(2, f"site {i}: physical dimension {d} vs {len(phys)} charges")
        qr = [None] * wr
        for a, so, b, si in zip(*np.nonzero(W)):
            q = qsub(qadd(bonds[i][a], phys[so]), phys[si])
            if qr[b] is None:
                qr[b] = q
            elif qr[b] != q:
                raise ValueError(f"MPO site {i} channel {b}: entries carry charges {qr[b]} and {q}")
        if i == len(mpo) - 1 and qr[0] not in (None, qzero(n)):
            raise ValueError("the MPO does not conserve the total charge")
        bonds.append([q if q is not None else qzero(n) for q in qr])
    return bonds


def mode_entries_two(W1, W2):
    """Nonzero entries of the two-site MPO fold
    Mt[w, s1', s2', w', s1, s2] = sum_m W1[w, s1', m, s1] W2[m, s2', w', s2],
    keyed by the output (s1', s2', w') -> [(w, s1, s2, coef)]."""
    M = np.einsum("aibj,bkcl->aikcjl", W1, W2)
    out = {}
    for w, s1o, s2o, wo, s1, s2 in zip(*np.nonzero(M)):
        out.setdefault((int(s1o), int(s2o), int(wo)), []).append(
            (int(w), int(s1), int(s2), float(M[w, s1o, s2o, wo, s1, s2])))
    return out


# ------------------------------------------------------------------ plans

class PlanTables:
    """Host description of a plan: buffer sizes and the stage tables, in the
    form mh_bs_plan_add_gemm / mh_bs_plan_add_mode take them."""

    def __init__(self, sizes):
        self.sizes = list(int(s) for s in sizes)
        self.stages = []

    def add_gemm(self, a_src, b_src, c_src, amc, bkc, tasks):
        """tasks: [(M, N, c_off, ldc, [(a_off, b_off, K, lda, ldb), ...])]"""
        ti, tc, so, si = [], [], [], []
        for M, N, c_off, ldc, sl in tasks:
            ti.append((M, N, len(so), len(sl), ldc))
            tc.append(c_off)
            for a_off, b_off, K, lda, ldb in sl:
                so.append((a_off, b_off))
                si.append((K, lda, ldb))
        self.stages.append(("gemm", dict(
            a_src=a_src, b_src=b_src, c_src=c_src, amc=int(amc), bkc=int(bkc),
            task_i=np.array(ti, dtype=np.int32).reshape(-1, 5),
            task_coff=np.array(tc, dtype=np.int64),
            seg_off=np.array(so, dtype=np.int64).reshape(-1, 2),
            seg_i=np.array(si, dtype=np.int32).reshape(-1, 3))))

    def add_mode(self, in_src, out_src, dests):
        """dests: [(dst_off, rows, cols, ld, [(src_off, src_ld, coef), ...])]"""
        do, di, so, sl, cf = [], [], [], [], []
        for off, rows, cols, ld, cl in dests:
            do.append(off)
            di.append((rows, cols, ld, len(so), len(cl)))
            for s_off, s_ld, c in cl:
                so.append(s_off)
                sl.append(s_ld)
                cf.append(c)
        self.stages.append(("mode", dict(
            in_src=in_src, out_src=out_src,
            dst_off=np.array(do, dtype=np.int64),
            dst_i=np.array(di, dtype=np.int32).reshape(-1, 5),
            src_off=np.array(so, dtype=np.int64),
            src_ld=np.array(sl, dtype=np.int32),
            coef=np.array(cf, dtype=np.float64))))

    def flops(self):
        f = 0.0
        for kind, t in self.stages:
            if kind == "gemm":
                for M, N, s0, ns, _ in t["task_i"]:
                    f += 2.0 * M * N * float(t["seg_i"][s0:s0 + ns, 0].sum())
            else:
                for rows, cols, _, _, nc in t["dst_i"]:
                    f += 2.0 * rows * cols * nc
        return f


def _p(a):
    return a.ctypes.data_as(_lib.c_void_p) if a.size else None


class Plan:
    """A plan on one engine handle (freed with the object); its tables go to
    the device on first use."""

    def __init__(self, tables, prec="f64", device=None, _native=None):
        from .engine import engine
        self.tables = tables
        self.eng = engine(device)
        self.lib = self.eng.lib
        # building a plan launches nothing (apply and lanczos bind the
        # caller's stream), so any thread may build one
        self.h = self.eng.h
        self.prec = prec
        if _native is not None:
            # built by a native planner: _native(h, &ptr) -> status
            ptr = _lib.c_void_p()
            self.eng.check(_native(self.h, _lib.ctypes.byref(ptr)), "native planner")
            self.ptr = ptr
            return
        sizes = (_lib.c_int64 * 6)(*tables.sizes)
        ptr = _lib.c_void_p()
        self.eng.check(self.lib.mh_bs_plan_create(self.h, _lib.PREC[prec], sizes,
                                                  _lib.ctypes.byref(ptr)), "bs plan")
        self.ptr = ptr
        for kind, t in tables.stages:
            if kind == "gemm":
                st = self.lib.mh_bs_plan_add_gemm(
                    self.h, self.ptr, t["a_src"], t["b_src"], t["c_src"], t["amc"], t["bkc"],
                    len(t["task_coff"]), _p(t["task_i"]), _p(t["task_coff"]),
                    len(t["seg_off"]), _p(t["seg_off"]), _p(t["seg_i"]))
            else:
                st = self.lib.mh_bs_plan_add_mode(
                    self.h, self.ptr, t["in_src"], t["out_src"], len(t["dst_off"]),
                    _p(t["dst_off"]), _p(t["dst_i"]), len(t["src_off"]), _p(t["src_off"]),
                    _p(t["src_ld"]), _p(t["coef"]))
            self.eng.check(st, "bs plan stage")

    def flops(self):
        f = _lib.c_double()
        self.eng.check(self.lib.mh_bs_plan_flops(self.ptr, _lib.ctypes.byref(f)), "flops")
        return f.value

    def sizes(self):
        """elements of the plan's six buffers (x, y, p1, p2, t1, t2)"""
        out = (_lib.c_int64 * 6)()
        self.eng.check(self.lib.mh_bs_plan_sizes(self.ptr, out), "sizes")
        return list(out)

    def variants(self):
        n = _lib.c_int()
        self.eng.check(self.lib.mh_bs_plan_variants(self.ptr, _lib.ctypes.byref(n)), "variants")
        return n.value

    def force_variant(self, v):
        """fix the grouped-GEMM tile variant of every stage (-1: tune)"""
        self.eng.check(self.lib.mh_bs_plan_force_variant(self.ptr, int(v)), "force_variant")

    def apply(self, x, p1, p2, y):
        h = self.eng.bind_stream()
        self.eng.check(self.lib.mh_bs_apply(h, self.ptr, _dp(x), _dp(p1), _dp(p2), _dp(y)),
                       "bs apply")
        return y

    def lanczos(self, p1, p2, x0, kmax, btol, reorth=False):
        h = self.eng.bind_stream()
        out = torch.empty_like(x0)
        ev = _lib.c_double()
        info = _lib.c_int()
        nit = _lib.c_int()
        self.eng.check(self.lib.mh_bs_lanczos_eig(
            h, self.ptr, _dp(p1), _dp(p2), _dp(x0), kmax, btol, 1 if reorth else 0,
            _lib.ctypes.byref(ev), _dp(out), _lib.ctypes.byref(info), _lib.ctypes.byref(nit)),
            "bs lanczos")
        return ev.value, out, info.value, nit.value

    def __del__(self):
        try:
            if getattr(self, "ptr", None) is not None and self.ptr.value:
                # the engine's live handle: once the engine is closed (at
                # exit, before interpreter teardown collects what is left)
                # its handle is freed and must not be passed on; the process
                # returns the tables' memory then
                h = getattr(self.eng, "h", None)
                if h is not None:
                    self.lib.mh_bs_plan_destroy(h, self.ptr)
                self.ptr = None
        except Exception:
            pass


# ------------------------------------------------------------ native planner

class _Sectors:
    """An ``mh_bs_sectors`` for a Space (sectors in the Space's order); keeps
    its arrays alive."""

    def __init__(self, space):
        q = [c for c, _ in space.sectors]
        nq = len(q[0]) if q else 1
        self.charges = np.ascontiguousarray(np.array(q, dtype=np.int32).reshape(-1, nq))
        self.dims = np.ascontiguousarray(np.array([d for _, d in space.sectors], dtype=np.int32))
        self.s = _lib.BsSectors(len(q), nq,
                                self.charges.ctypes.data_as(_lib.P(_lib.c_int32)),
                                self.dims.ctypes.data_as(_lib.P(_lib.c_int32)))

    def ref(self):
        return _lib.ctypes.byref(self.s)


def _charges(qs, nq):
    return np.ascontiguousarray(np.array(qs, dtype=np.int32).reshape(-1, nq))


def _wcm(W):
    return np.asfortranarray(np.asarray(W, dtype=np.float64))


def _hp(a):
    return a.ctypes.data_as(_lib.c_void_p)


def native_builder(kind, *args):
    """(call(h, out_ptr) -> status, keep-alive list) of a native planner:
    kind "heff2": (xl, Lay, Ray, W1, W2, prec); "updateleft": (al, Lay, W,
    Lay_new, prec); "updateright": (al, Ray, W, Ray_new, prec); "two_site":
    (l1, l2, xl).  The layouts are those of the Python planner above."""
    lib = _lib.load()
    if kind == "heff2":
        xl, Lay, Ray, W1, W2, prec = args
        if xl.nphys != 2:
            raise _lib.DimensionMismatch(2, "plan_heff2 needs a two-site layout")
        nq = len(xl.phys[0])
        L, R = _Sectors(xl.lspace), _Sectors(xl.rspace)
        ph, ql, qr = _charges(xl.phys, nq), _charges(Lay.qw, nq), _charges(Ray.qw, nq)
        w1, w2 = _wcm(W1), _wcm(W2)
        if w1.shape[0] != Lay.w or w2.shape[2] != Ray.w or w1.shape[1] != xl.d:
            raise _lib.DimensionMismatch(2, "MPO dimensions do not match the environments")
        keep = [L, R, ph, ql, qr, w1, w2]
        return (lambda h, out: lib.mh_bs_plan_heff2(
            h, _lib.PREC[prec], L.ref(), R.ref(), xl.d, _hp(ph), Lay.w, _hp(ql), w1.shape[2],
            Ray.w, _hp(qr), _hp(w1), _hp(w2), out)), keep
    if kind in ("updateleft", "updateright"):
        al, E, W, E_new, prec = args
        if al.nphys != 1:
            raise _lib.DimensionMismatch(2, f"{kind} needs a one-site layout")
        nq = len(al.phys[0])
        L, R = _Sectors(al.lspace), _Sectors(al.rspace)
        ql, qr = (E.qw, E_new.qw) if kind == "updateleft" else (E_new.qw, E.qw)
        ph, cl, cr, w = _charges(al.phys, nq), _charges(ql, nq), _charges(qr, nq), _wcm(W)
        if w.shape[0] != len(ql) or w.shape[2] != len(qr) or w.shape[1] != al.d:
            raise _lib.DimensionMismatch(2, f"{kind}: MPO does not match the channels")
        fn = getattr(lib, "mh_bs_plan_" + kind)
        keep = [L, R, ph, cl, cr, w]
        return (lambda h, out: fn(h, _lib.PREC[prec], L.ref(), R.ref(), al.d, _hp(ph),
                                  len(ql), _hp(cl), len(qr), _hp(cr), _hp(w), out)), keep
    if kind == "two_site":
        l1, l2, xl = args
        nq = len(xl.phys[0])
        L, M, R = _Sectors(l1.lspace), _Sectors(l1.rspace), _Sectors(l2.rspace)
        ph = _charges(xl.phys, nq)
        keep = [L, M, R, ph]
        return (lambda h, out: lib.mh_bs_plan_two_site(h, L.ref(), M.ref(), R.ref(), xl.d,
                                                       _hp(ph), out)), keep
    raise ValueError(kind)


def native_plan(kind, *args, device=None):
    """A Plan built by the native planner (same tables as the Python one)."""
    call, keep = native_builder(kind, *args)
    prec = args[-1] if kind != "two_site" else "f64"
    return Plan(None, prec, device, _native=call)


_PY_PLANNERS = {"heff2": lambda xl, La, Ra, W1, W2, prec: plan_heff2(xl, La, Ra, W1, W2),
                "updateleft": lambda al, E, W, En, prec: plan_updateleft(al, E, W, En),
                "updateright": lambda al, E, W, En, prec: plan_updateright(al, E, W, En),
                "two_site": lambda l1, l2, xl: plan_two_site(l1, l2, xl)}

# "native": plans from the C planner (mh_bs_plan_heff2 ...); "python": the
# table builders above, uploaded stage by stage.  Both give the same tables.
PLANNER = "native"


def make_plan(kind, *args, planner=None, device=None):
    """Plan of one contraction (kind and arguments as native_builder) on the
    engine of ``device`` (default: the current device)."""
    if (planner or PLANNER) == "native":
        return native_plan(kind, *args, device=device)
    prec = args[-1] if kind != "two_site" else "f64"
    return Plan(_PY_PLANNERS[kind](*args), prec, device)


def read_tables(ptr, lib=None):
    """PlanTables of a plan pointer (host tables, no device needed)."""
    lib = lib or _lib.load()
    sizes = (_lib.c_int64 * 6)()
    _lib.raise_for(lib.mh_bs_plan_sizes(ptr, sizes), None, "sizes")
    tab = PlanTables(list(sizes))
    n = _lib.c_int()
    _lib.raise_for(lib.mh_bs_plan_stage_count(ptr, _lib.ctypes.byref(n)), None, "stages")
    for i in range(n.value):
        info = (_lib.c_int32 * 8)()
        _lib.raise_for(lib.mh_bs_plan_stage_info(ptr, i, info), None, "stage info")
        kind, a, b, c, amc, bkc, n1, n2 = list(info)
        if kind == 0:
            ti = np.zeros((n1, 5), np.int32)
            tc = np.zeros(n1, np.int64)
            so = np.zeros((n2, 2), np.int64)
            si = np.zeros((n2, 3), np.int32)
            _lib.raise_for(lib.mh_bs_plan_stage_gemm(ptr, i, _hp(ti), _hp(tc), _hp(so), _hp(si)),
                           None, "stage tables")
            tab.stages.append(("gemm", dict(a_src=a, b_src=b, c_src=c, amc=amc, bkc=bkc,
                                            task_i=ti, task_coff=tc, seg_off=so, seg_i=si)))
        else:
            do = np.zeros(n1, np.int64)
            di = np.zeros((n1, 5), np.int32)
            so = np.zeros(n2, np.int64)
            sl = np.zeros(n2, np.int32)
            cf = np.zeros(n2, np.float64)
            _lib.raise_for(lib.mh_bs_plan_stage_mode(ptr, i, _hp(do), _hp(di), _hp(so), _hp(sl),
                                                     _hp(cf)), None, "stage tables")
            tab.stages.append(("mode", dict(in_src=a, out_src=b, dst_off=do, dst_i=di,
                                            src_off=so, src_ld=sl, coef=cf)))
    return tab


def host_native_tables(kind, *args):
    """Tables the native planner builds, without a device (h = NULL)."""
    lib = _lib.load()
    call, keep = native_builder(kind, *args)
    ptr = _lib.c_void_p()
    _lib.raise_for(call(None, _lib.ctypes.byref(ptr)), None, "native planner")
    try:
        return read_tables(ptr, lib)
    finally:
        lib.mh_bs_plan_destroy(None, ptr)


def _dp(t):
    if t is None:
        return None
    if not t.is_contiguous() or t.dtype != torch.float64:
        raise TypeError("block buffers are contiguous float64 tensors")
    return _lib.c_void_p(t.data_ptr())


X, Y, P1, P2, T1, T2 = range(6)


def _gemm1_env_vec(Lay, xl, t1_off):
    """T1[(qa, w)] = L[(qa, w)] * x_group[qa - q(w)] (one task per (qa, w)).
    Returns the tasks and fills t1_off[(qa, w)]; also returns T1's size."""
    tasks, t1 = [], 0
    S = Lay.space
    for qa, Da in S.sectors:
        for w in range(Lay.w):
            if (qa, w) not in Lay.off:
                continue
            qk = qsub(qa, Lay.qw[w])
            if qk not in xl.group:
                continue
            g0, ncols = xl.group[qk]
            Dk = S.dim[qk]
            t1_off[(qa, w)] = t1
            tasks.append((Da, ncols, t1, Da, [(Lay.off[(qa, w)], g0, Dk, Da, Dk)]))
            t1 += Da * ncols
    return tasks, t1


def _gemm1_used(Lay, xl, used):
    """Stage 1 of the two-site H_eff restricted to the input slots the MPO
    pair reads: T1[(qa, w)] = L[(qa, w)] * (the columns of x_group[qa - q(w)]
    whose block (s1, s2) has (w, s1, s2) in ``used``), stored compactly (the
    used blocks side by side in group order).  One task per run of adjacent
    used blocks.  Returns (tasks, t1_off[(qa, w)], t1_col[(qa, w, p)], size)."""
    tasks, t1_off, t1_col, t1 = [], {}, {}, 0
    S = Lay.space
    for qa, Da in S.sectors:
        for w in range(Lay.w):
            if (qa, w) not in Lay.off:
                continue
            qk = qsub(qa, Lay.qw[w])
            if qk not in xl.group:
                continue
            g0, _ = xl.group[qk]
            Dk = S.dim[qk]
            runs, c = [], 0
            for b in xl.by_left[qk]:
                if (w,) + tuple(b.p) not in used:
                    continue
                t1_col[(qa, w, b.p)] = c
                if runs and runs[-1][0] + runs[-1][1] == b.colstart:
                    runs[-1][1] += b.cols
                else:
                    runs.append([b.colstart, b.cols, c])
                c += b.cols
            if c == 0:
                continue
            t1_off[(qa, w)] = t1
            for x0, n, c0 in runs:
                tasks.append((Da, n, t1 + Da * c0, Da, [(Lay.off[(qa, w)], g0 + Dk * x0, Dk, Da, Dk)]))
            t1 += Da * c
    return tasks, t1_off, t1_col, t1


def plan_heff2(xl, Lay, Ray, W1, W2, skip_zeros=True):
    """Tables of y = ac2_prime(x, W1, W2, L, R) (derivatives.jl:20-22) for an
    x of layout ``xl`` (two physical legs).  Buffers: X = x, Y = y,
    P1 = L (layout Lay over xl.lspace), P2 = R (layout Ray over xl.rspace).

    Stage 1: T1[(qa,w)] = L x over the blocks (s1, s2) of x's group qa-q(w)
             that the MPO pair reads from channel w.
    Stage 2: for every right sector qb, its output blocks (qa, s1', s2') are
             grouped by the set of channels w' that receive MPO entries; per
             group a matrix T2 with rows the group's blocks and columns the
             (w', b') of those channels (in the order of R's sector-qb
             matrix):  T2 = sum Mt-coefficients * T1 blocks.
    Stage 3: y block (qa, s1', s2', qb) = T2[rows of the block] R_qb^T over
             the group's channels.
    The slots and channels left out hold only structural zeros of the MPO
    pair (zero coefficients inside charge-allowed blocks), which the
    reference's graded contraction multiplies; ``skip_zeros=False`` keeps
    them (every (w, s1, s2) column of T1, every channel of R) and gives the
    flop count of that contraction."""
    if xl.nphys != 2:
        raise _lib.DimensionMismatch(2, "plan_heff2 needs a two-site layout")
    if Lay.space != xl.lspace or Ray.space != xl.rspace:
        raise _lib.DimensionMismatch(2, "environment spaces do not match the vector")
    wl, d, wm, _ = W1.shape
    _, _, wr, _ = W2.shape
    if wl != Lay.w or wr != Ray.w or d != xl.d:
        raise _lib.DimensionMismatch(2, "MPO dimensions do not match the environments")
    Sr = xl.rspace
    entries = mode_entries_two(W1, W2)
    if skip_zeros:
        used = {(w, s1, s2) for lst in entries.values() for w, s1, s2, _ in lst}
    else:
        used = {(w, s1, s2) for w in range(wl) for s1 in range(d) for s2 in range(d)}
    g1, t1_off, t1_col, t1 = _gemm1_used(Lay, xl, used)
    dests, g3, t2 = [], [], 0
    for qb, Db in Sr.sectors:
        if qb not in xl.by_right:
            continue
        if qb not in Ray.group:
            raise ValueError(f"right environment has no block for sector {qb}")
        chans = [w2 for w2 in range(wr) if (qb, w2) in Ray.off]
        groups, order = {}, []
        for b in xl.by_right[qb]:
            qa, Da = b.qa, b.rows
            per = {}
            for w2 in chans:
                qbk = qsub(qb, Ray.qw[w2])
                cl = []
                for w, s1, s2, c in entries.get((b.p[0], b.p[1], w2), ()):
                    if (qa, w) not in t1_off:
                        continue
                    qk = qsub(qa, Lay.qw[w])
                    src = xl.index.get((qk, (s1, s2)))
                    if src is None:
                        continue
                    if src.qb != qbk:
                        raise ValueError("MPO entry does not conserve the charges")
                    cl.append((t1_off[(qa, w)] + Da * t1_col[(qa, w, (s1, s2))], Da, c))
                if cl or not skip_zeros:
                    per[w2] = cl
            if not per:
                per = {chans[0]: []}   # a block nothing reaches: one zero channel
            pat = tuple(sorted(per))
            if pat not in groups:
                groups[pat] = []
                order.append(pat)
            groups[pat].append((b, per))
        g0R = Ray.group[qb][0]
        for pat in order:
            members = groups[pat]
            Rg = sum(b.rows for b, _ in members)
            col, c = {}, 0
            for w2 in pat:
                col[w2] = c
                c += Sr.dim[qsub(qb, Ray.qw[w2])]
            base, t2 = t2, t2 + Rg * c
            # K-segments of stage 3: runs of channels adjacent in R's columns
            runs = []
            for w2 in pat:
                K = Sr.dim[qsub(qb, Ray.qw[w2])]
                rc = Ray.colstart[(qb, w2)]
                if runs and runs[-1][1] + runs[-1][2] == rc:
                    runs[-1][2] += K
                else:
                    runs.append([col[w2], rc, K])
            r = 0
            for b, per in members:
                for w2 in pat:
                    dests.append((base + r + Rg * col[w2], b.rows, Sr.dim[qsub(qb, Ray.qw[w2])],
                                  Rg, per[w2]))
                g3.append((b.rows, b.cols, b.off, b.rows,
                           [(base + r + Rg * c0, g0R + Db * rc, K, Rg, b.cols) for c0, rc, K in runs]))
                r += b.rows
    tab = PlanTables([xl.size, xl.size, Lay.size, Ray.size, t1, t2])
    tab.add_gemm(P1, X, T1, True, True, g1)
    tab.add_mode(T1, T2, dests)
    tab.add_gemm(T2, P2, Y, True, False, g3)
    return tab


def heff2_reference_flops(xl, Lay, Ray, W1, W2):
    """Flops of the two-site H_eff as the reference's graded contraction does
    them (every charge-allowed block, MPO zeros included): the count of
    plan_heff2(..., skip_zeros=False)."""
    return plan_heff2(xl, Lay, Ray, W1, W2, skip_zeros=False).flops()


def plan_updateleft(al, Lay, W, Lay_new):
    """Tables of L' = updateleft(L, A, W, B) (mpo/transfer.jl:31-34) for bra
    and ket of layout ``al`` (one physical leg).  Buffers: X = ket B,
    P1 = L, P2 = bra A, Y = L' (layout Lay_new over al.rspace).

    Stage 1: T1[(qa,w)] = L[(qa,w)] * B_group[qa-q(w)].
    Stage 2: V_(qb,w') (rows: bra blocks (qa,s') ending in qb; cols: b' in
             sector qb-q(w')) = sum W-coefficients * T1 blocks.
    Stage 3: L'[(qb,w')] = sum_{(qa,s')} A_block^T V_(qb,w')[rows of block]."""
    if al.nphys != 1:
        raise _lib.DimensionMismatch(2, "updateleft needs a one-site layout")
    wl, d, wr, _ = W.shape
    if Lay.space != al.lspace or Lay_new.space != al.rspace or Lay.w != wl or Lay_new.w != wr:
        raise _lib.DimensionMismatch(2, "updateleft: spaces or channels do not match")
    Sr = al.rspace
    t1_off = {}
    g1, t1 = _gemm1_env_vec(Lay, al, t1_off)
    ent = {}
    for w, so, w2, si in zip(*np.nonzero(W)):
        ent.setdefault((int(so), int(w2)), []).append((int(w), int(si), float(W[w, so, w2, si])))
    v_off, rowoff, RA, t2 = {}, {}, {}, 0
    for qb, Db in Sr.sectors:
        r = 0
        for b in al.by_right.get(qb, ()):
            rowoff[(b.qa, b.p)] = r
            r += b.rows
        RA[qb] = r
    dests, g3 = [], []
    for (qb, w2), loff in Lay_new.off.items():
        qbk = qsub(qb, Lay_new.qw[w2])
        cols = Sr.dim[qbk]
        if RA[qb] == 0:
            raise ValueError(f"sector {qb} of the right bond is unreachable")
        v_off[(qb, w2)] = t2
        segs = []
        for b in al.by_right[qb]:
            qa, Da = b.qa, b.rows
            (so,) = b.p
            cl = []
            for w, si, c in ent.get((so, w2), ()):
                if (qa, w) not in t1_off:
                    continue
                qk = qsub(qa, Lay.qw[w])
                src = al.index.get((qk, (si,)))
                if src is None:
                    continue
                if src.qb != qbk:
                    raise ValueError("MPO entry does not conserve the charges")
                cl.append((t1_off[(qa, w)] + Da * src.colstart, Da, c))
            dests.append((t2 + rowoff[(qa, b.p)], Da, cols, RA[qb], cl))
            segs.append((b.off, t2 + rowoff[(qa, b.p)], Da, Da, RA[qb]))
        g3.append((Sr.dim[qb], cols, loff, Sr.dim[qb], segs))
        t2 += RA[qb] * cols
    tab = PlanTables([al.size, Lay_new.size, Lay.size, al.size, t1, t2])
    tab.add_gemm(P1, X, T1, True, True, g1)
    tab.add_mode(T1, T2, dests)
    tab.add_gemm(P2, T2, Y, False, True, g3)
    return tab


def plan_updateright(al, Ray, W, Ray_new):
    """Tables of R' = updateright(R, A, W, B) (mpo/transfer.jl:26-29).
    Buffers: X = ket B, P1 = R (layout Ray over al.rspace), P2 = bra A,
    Y = R' (layout Ray_new over al.lspace).

    Stage 1: U_(qa,s') (D_qa x K_qb, qb = qa+q(s')) = A_block R_qb.
    Stage 2: V_(qa,w) (D_qa x ncols of B's group qa-q(w)) = sum W-coefficients
             * U column blocks.
    Stage 3: R'[(qa,w)] = V_(qa,w) B_group^T."""
    if al.nphys != 1:
        raise _lib.DimensionMismatch(2, "updateright needs a one-site layout")
    wl, d, wr, _ = W.shape
    if Ray.space != al.rspace or Ray_new.space != al.lspace or Ray.w != wr or Ray_new.w != wl:
        raise _lib.DimensionMismatch(2, "updateright: spaces or channels do not match")
    Sl = al.lspace
    u_off, g1, t1 = {}, [], 0
    for b in al.blocks:
        if b.qb not in Ray.group:
            continue
        g0, K = Ray.group[b.qb]
        u_off[(b.qa, b.p)] = t1
        g1.append((b.rows, K, t1, b.rows, [(b.off, g0, b.cols, b.rows, b.cols)]))
        t1 += b.rows * K
    ent = {}
    for w, so, w2, si in zip(*np.nonzero(W)):
        ent.setdefault((int(w), int(si)), []).append((int(so), int(w2), float(W[w, so, w2, si])))
    dests, g3, t2 = [], [], 0
    for (qa, w), roff in Ray_new.off.items():
        Da = Sl.dim[qa]
        qk = qsub(qa, Ray_new.qw[w])
        if qk not in al.group:
            raise ValueError(f"sector {qk} of the left bond is unreachable")
        bg0, ncols = al.group[qk]
        for b in al.by_left[qk]:
            (si,) = b.p
            cl = []
            for so, w2, c in ent.get((w, si), ()):
                qb = qadd(qa, al.phys[so])
                if (qa, (so,)) not in u_off or (qb, w2) not in Ray.off:
                    continue
                if qsub(qb, Ray.qw[w2]) != b.qb:
                    raise ValueError("MPO entry does not conserve the charges")
                cl.append((u_off[(qa, (so,))] + Da * Ray.colstart[(qb, w2)], Da, c))
            dests.append((t2 + Da * b.colstart, Da, b.cols, Da, cl))
        Dk = Sl.dim[qk]
        g3.append((Da, Dk, roff, Da, [(t2, bg0, ncols, Da, Dk)]))
        t2 += Da * ncols
    tab = PlanTables([al.size, Ray_new.size, Ray.size, al.size, t1, t2])
    tab.add_gemm(P2, P1, T1, True, True, g1)
    tab.add_mode(T1, T2, dests)
    tab.add_gemm(T2, X, Y, True, False, g3)
    return tab


def plan_two_site(l1, l2, xl):
    """Tables of x[a,s1,s2,b] = A1[a,s1,m] A2[m,s2,b] (dmrg.jl:98):
    one task per x block.  Buffers: X = A1, P1 = A2, Y = x."""
    tasks = []
    for b in xl.blocks:
        s1, s2 = b.p
        b1 = l1.index.get((b.qa, (s1,)))
        if b1 is None:
            continue
        b2 = l2.index.get((b1.qb, (s2,)))
        if b2 is None:
            continue
        tasks.append((b.rows, b.cols, b.off, b.rows, [(b1.off, b2.off, b1.cols, b1.rows, b2.rows)]))
    tab = PlanTables([l1.size, xl.size, l2.size, 0, 0, 0])
    tab.add_gemm(X, P1, Y, True, True, tasks)
    return tab


# ------------------------------------------------ TensorKit coupled-sector form

def coupled_info(layout, ncod):
    """The coupled-sector structure of a BlockLayout seen as a TensorMap with
    codomain (a, s_1..s_{ncod-1}) and domain (the rest, b): [(charge, rows,
    cols, offset)] in ascending charge order (mh_bs_coupled_info, host only)."""
    lib = _lib.load()
    nq = len(layout.phys[0])
    L, R = _Sectors(layout.lspace), _Sectors(layout.rspace)
    ph = _charges(layout.phys, nq)
    nc = _lib.c_int()
    _lib.raise_for(lib.mh_bs_coupled_info(L.ref(), R.ref(), layout.d, _hp(ph), layout.nphys, ncod,
                                          _lib.ctypes.byref(nc), None, None, None, None), None,
                   "coupled_info")
    n = nc.value
    q = np.zeros((max(n, 1), nq), dtype=np.int32)
    rows = np.zeros(max(n, 1), dtype=np.int32)
    cols = np.zeros(max(n, 1), dtype=np.int32)
    offs = np.zeros(n + 1, dtype=np.int64)
    _lib.raise_for(lib.mh_bs_coupled_info(L.ref(), R.ref(), layout.d, _hp(ph), layout.nphys, ncod,
                                          _lib.ctypes.byref(nc), _hp(q), _hp(rows), _hp(cols),
                                          _hp(offs)), None, "coupled_info")
    return [(tuple(int(v) for v in q[i]), int(rows[i]), int(cols[i]), int(offs[i]))
            for i in range(n)]


def _coupled_call(name, t_layout, ncod, src, dst):
    from .engine import engine
    lib = _lib.load()
    eng = engine(src.device)
    nq = len(t_layout.phys[0])
    L, R = _Sectors(t_layout.lspace), _Sectors(t_layout.rspace)
    ph = _charges(t_layout.phys, nq)
    h = eng.bind_stream()
    eng.check(getattr(lib, name)(h, L.ref(), R.ref(), t_layout.d, _hp(ph), t_layout.nphys, ncod,
                                 _dp(src), _dp(dst)), name)


def to_coupled(t, ncod):
    """BlockTensor -> {charge: (rows x cols) device matrix}: TensorKit's
    blocks(t)[c] of the tensor with codomain (a, s_1..s_{ncod-1}), domain
    (the other physical legs, b) (mh_bs_to_coupled)."""
    info = coupled_info(t.layout, ncod)
    total = info[-1][3] + info[-1][1] * info[-1][2] if info else 0
    buf = torch.empty(max(total, 1), dtype=torch.float64, device=t.data.device)
    _coupled_call("mh_bs_to_coupled", t.layout, ncod, t.data, buf)
    return {q: buf[o:o + r * c].view(c, r).t() for q, r, c, o in info}


def from_coupled(layout, ncod, mats):
    """{charge: matrix} (TensorKit's blocks, as to_coupled returns them) ->
    BlockTensor of ``layout`` (mh_bs_from_coupled)."""
    info = coupled_info(layout, ncod)
    if sorted(mats) != [q for q, *_ in info]:
        raise _lib.DimensionMismatch(2, "coupled sectors do not match the layout")
    dev = next(iter(mats.values())).device if mats else "cuda"
    parts = []
    for q, r, c, o in info:
        m = mats[q]
        if tuple(m.shape) != (r, c):
            raise _lib.DimensionMismatch(2, f"coupled block {q}: {tuple(m.shape)} vs {(r, c)}")
        parts.append(m.t().contiguous().reshape(-1))
    buf = torch.cat(parts).to(torch.float64) if parts else torch.zeros(1, dtype=torch.float64,
                                                                        device=dev)
    out = torch.empty(max(layout.size, 1), dtype=torch.float64, device=buf.device)
    _coupled_call("mh_bs_from_coupled", layout, ncod, buf, out)
    return BlockTensor(layout, out[:layout.size])
