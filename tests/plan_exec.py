"""Numpy executor of block-sparse plan tables (test infrastructure).

Runs the stage tables a planner produces with the semantics that
include/mpsheff.h states for mh_bs_plan_add_gemm / mh_bs_plan_add_mode, so
the planner can be checked against the dense oracle without a GPU.  It also
asserts every access lies inside its buffer, as the library's own checks do.
"""
import numpy as np


def _operand(buf, off, rows, cols, ld, contig_rows):
    if contig_rows:   # element (r, c) at off + r + c*ld
        last = off + (rows - 1) + (cols - 1) * ld
        assert 0 <= off and last < buf.size
        return np.lib.stride_tricks.as_strided(buf[off:], shape=(rows, cols),
                                               strides=(8, 8 * ld))
    last = off + (rows - 1) * ld + (cols - 1)
    assert 0 <= off and last < buf.size
    return np.lib.stride_tricks.as_strided(buf[off:], shape=(rows, cols), strides=(8 * ld, 8))


def execute(tables, bufs):
    """bufs: list of 6 float64 arrays (None allowed for unused ones); the
    intermediates are allocated here when absent."""
    bufs = list(bufs)
    for i in (4, 5):
        if bufs[i] is None:
            bufs[i] = np.full(tables.sizes[i], np.nan)
    for kind, t in tables.stages:
        if kind == "gemm":
            A, B, C = bufs[t["a_src"]], bufs[t["b_src"]], bufs[t["c_src"]]
            for (M, N, s0, ns, ldc), coff in zip(t["task_i"], t["task_coff"]):
                acc = np.zeros((M, N))
                for s in range(s0, s0 + ns):
                    a_off, b_off = t["seg_off"][s]
                    K, lda, ldb = t["seg_i"][s]
                    a = _operand(A, a_off, M, K, lda, bool(t["amc"]))
                    b = _operand(B, b_off, K, N, ldb, bool(t["bkc"]))
                    acc += a @ b
                c = _operand(C, coff, M, N, ldc, True)
                c[...] = acc
        else:
            I, O = bufs[t["in_src"]], bufs[t["out_src"]]
            for d, (rows, cols, ld, c0, nc) in enumerate(t["dst_i"]):
                acc = np.zeros((rows, cols))
                for c in range(c0, c0 + nc):
                    acc += t["coef"][c] * _operand(I, t["src_off"][c], rows, cols,
                                                   t["src_ld"][c], True)
                _operand(O, t["dst_off"][d], rows, cols, ld, True)[...] = acc
    return bufs
