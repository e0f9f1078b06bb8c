"""CPU proof that the sweep kernels' index maps stay in bounds and touch each
amplitude exactly once (or exactly the amplitudes a gate changes).

The functions below restate the address arithmetic of
subquokka_dev_amd/csrc/kernels/sweep.hpp (k_dense, k_diag, k_perm) for every
(tile, lane, register) the launch covers; the assertions are the invariants
(bijection onto [0, 2^n), no overlap), not a second copy of the formulas.
"""
import itertools

import numpy as np
import pytest


def insert_zero(x, pos):
    lo = x & ((1 << pos) - 1)
    return ((x >> pos) << (pos + 1)) | lo


def insert_one(x, pos):
    return insert_zero(x, pos) | (1 << pos)


def dense_addresses(n, targets, vb):
    s = vb + 6
    high = [t for t in targets if t >= s]
    n_tiles = 1 << (n - s - len(high))
    tiles = np.arange(n_tiles, dtype=np.int64)
    base = tiles << s
    for h in high:
        base = insert_zero(base, h)
    out = []
    for lane in range(64):
        b = base + (lane << vb)
        for r in range(1 << len(high)):
            off = b.copy()
            for j, h in enumerate(high):
                if (r >> j) & 1:
                    off = off + (1 << h)
            for e in range(1 << vb):
                out.append(off + e)
    return np.concatenate(out)


@pytest.mark.parametrize("vb", [0, 1])
@pytest.mark.parametrize("n", [8, 11])
def test_dense_map_is_bijection(vb, n):
    for k in (1, 2):
        for targets in itertools.combinations(range(n), k):
            if n < vb + 6 + sum(1 for t in targets if t >= vb + 6):
                continue
            a = dense_addresses(n, list(targets), vb)
            assert a.min() >= 0 and a.max() < (1 << n)
            assert len(a) == 1 << n and len(np.unique(a)) == 1 << n, targets


def diag_visits(n, fixed, vb):
    n_vec = 1 << (n - vb - len(fixed))
    v = np.arange(n_vec, dtype=np.int64)
    i = v << vb
    for f in sorted(fixed):
        i = insert_one(i, f)
    return np.concatenate([i + e for e in range(1 << vb)])


@pytest.mark.parametrize("vb", [0, 1])
def test_diag_map_visits_exactly_the_fixed_one_half(vb):
    n = 10
    for fixed in ([], [vb], [vb + 2], [3, 7], [vb, 9]):
        a = diag_visits(n, fixed, vb)
        assert a.max() < (1 << n) and len(np.unique(a)) == len(a)
        want = [i for i in range(1 << n) if all((i >> f) & 1 for f in fixed)]
        assert sorted(a.tolist()) == want


def perm_moves(n, pairs, vb):
    q = [p for pr in pairs for p in pr]
    srt = sorted(q)
    n_vec = 1 << (n - vb - len(q))
    v = np.arange(n_vec, dtype=np.int64)
    base = v << vb
    for p in srt:
        base = insert_zero(base, p)
    loads, stores = [], []
    for x in range(1 << len(q)):
        bits = [(x >> j) & 1 for j in range(len(q))]
        moves = any(bits[2 * j] != bits[2 * j + 1] for j in range(len(pairs)))
        if not moves:
            continue
        off = base.copy()
        for j, p in enumerate(q):
            if bits[j]:
                off = off + (1 << p)
        for e in range(1 << vb):
            loads.append(off + e)
    return np.concatenate(loads)


@pytest.mark.parametrize("vb", [0, 1])
def test_perm_map_touches_only_moving_amplitudes(vb):
    n = 10
    for pairs in ([(vb, 5)], [(3, 9)], [(vb + 1, 8), (vb + 2, 9)], [(4, 6), (5, 7)]):
        a = perm_moves(n, pairs, vb)
        assert a.max() < (1 << n) and len(np.unique(a)) == len(a)

        def moving(i):
            return any(((i >> p) & 1) != ((i >> r) & 1) for p, r in pairs)
        want = [i for i in range(1 << n) if moving(i)]
        assert sorted(a.tolist()) == want


def tile_addresses(n, q, bs=256):
    """Global offsets touched by k_tile (kernels/tile.hpp) for tile set q."""
    k = len(q)
    lb = bs.bit_length() - 1
    m_per = (1 << k) // bs
    tiles = np.arange(1 << (n - k), dtype=np.int64)
    base = tiles.copy()
    for b in q:
        base = insert_zero(base, b)
    out = []
    for tid in range(bs):
        off_lo = sum(1 << q[i] for i in range(lb) if (tid >> i) & 1)
        for m in range(m_per):
            off_hi = sum(1 << q[i] for i in range(lb, k) if ((m << lb) >> i) & 1)
            out.append(base + off_lo + off_hi)
    return np.concatenate(out)


@pytest.mark.parametrize("q", [list(range(10)), [0, 1, 2, 3, 4, 5, 7, 9, 11, 13], list(range(6)) + [8, 10, 12, 13, 14, 15]])
def test_tile_map_is_bijection(q):
    n = 16
    a = tile_addresses(n, q)
    assert a.min() >= 0 and a.max() < (1 << n)
    assert len(np.unique(a)) == 1 << n


def test_tile_lds_indices_in_range():
    for k in (10, 11, 12, 13):
        d = 1 << k
        for b in range(k):
            p = np.arange(d // 2)
            i0 = insert_zero(p, b)
            assert i0.max() | (1 << b) < d and len(np.unique(np.concatenate([i0, i0 | (1 << b)]))) == d
        for b0 in range(k):
            for b1 in range(b0 + 1, k):
                p = np.arange(d // 4)
                i = insert_zero(insert_zero(p, b0), b1)
                allq = np.concatenate([i, i | (1 << b0), i | (1 << b1), i | (1 << b0) | (1 << b1)])
                assert allq.max() < d and len(np.unique(allq)) == d
