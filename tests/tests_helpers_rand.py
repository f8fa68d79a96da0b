"""Random graded layouts shared by the block-sparse tests."""
import numpy as np

from finitemptools_amd import blocksparse as bs


def rand_space(rng, charges, lo=1, hi=5):
    return bs.Space([(q, int(rng.integers(lo, hi + 1))) for q in charges])


def reachable(S, phys, nphys=1):
    acc = {}
    for q, d in S.sectors:
        stack = [q]
        for _ in range(nphys):
            stack = [bs.qadd(x, p) for x in stack for p in phys]
        for x in stack:
            acc[x] = 1
    return list(acc)


def rand_layouts(rng, phys):
    """(two-site layout, one-site layout) over random Hubbard-like sectors"""
    Sl = rand_space(rng, [(a, b) for a in range(1, 4) for b in range(1, 4)], 1, 6)
    two = bs.BlockLayout(Sl, phys, 2, rand_space(rng, reachable(Sl, phys, 2), 1, 6))
    one = bs.BlockLayout(Sl, phys, 1, rand_space(rng, reachable(Sl, phys, 1), 1, 6))
    return two, one
