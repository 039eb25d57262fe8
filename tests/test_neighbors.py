"""Neighbour search known answers (test/test_neighbors.cu) on the oracle and
the GPU: grid build, the working set permuted into cell order, findNeighbors,
and the result mapped back to input order and compared with an all-pairs
FP32 brute force."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"


def pos(*xyz):
    return np.array([[x, y, z, 1.0] for x, y, z in xyz], dtype=np.float32)


def brute_force(p, h):
    p = np.asarray(p, np.float32)[:, :3]
    h = np.float32(h)
    d = p[:, None, :] - p[None, :, :]
    d2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    m = d2 <= h * h
    np.fill_diagonal(m, False)
    return [sorted(np.nonzero(row)[0].tolist()) for row in m]


def search(backend, p, h, cell=1.0, maxn=256, mn=(0.0, 0.0, 0.0), mx=(5.0, 5.0, 5.0)):
    """runNeighborSearch (test/test_neighbors.cu:24-112); returns per input
    particle (count, sorted neighbour list in input indices)."""
    n = len(p)
    g = backend.grid(n, mn, mx, cell)
    sk, si, cs, ce = g.build(p)
    ps = p[si]
    nbr, cnt, flag = backend.k.find_neighbors(ps, cs, ce, g.gridSize(), mn, cell, n, h, maxn, n)
    counts = np.zeros(n, np.int64)
    lists = [None] * n
    for s in range(n):
        o = si[s]
        counts[o] = cnt[s]
        stored = min(cnt[s], maxn)
        lists[o] = sorted(int(si[nbr[k * n + s]]) for k in range(stored))
    return counts, lists, flag


def expect_correct(backend, p, h, **kw):
    counts, lists, _ = search(backend, p, h, **kw)
    exp = brute_force(p, h)
    assert [len(e) for e in exp] == list(counts)
    assert lists == exp
    return counts, lists


def test_single_particle(backend):
    counts, _, _ = search(backend, pos((2.0, 2.0, 2.0)), 0.75)
    assert counts[0] == 0


def test_pair_within_and_outside_radius(backend):
    _, lists, _ = search(backend, pos((1.0, 1.0, 1.0), (1.5, 1.0, 1.0)), 0.75)
    assert lists == [[1], [0]]
    counts, _, _ = search(backend, pos((1.0, 1.0, 1.0), (2.0, 1.0, 1.0)), 0.75)
    assert list(counts) == [0, 0]


def test_inclusive_radius_and_one_ulp(backend):
    counts, _, _ = search(backend, pos((1.0, 1.0, 1.0), (1.5, 1.0, 1.0)), 0.5)
    assert list(counts) == [1, 1]
    r = np.float32(0.5)
    below = np.nextafter(r, np.float32(0))
    above = np.nextafter(r, np.float32(np.inf))
    assert list(search(backend, pos((0.0, 1.0, 1.0), (below, 1.0, 1.0)), 0.5)[0]) == [1, 1]
    assert list(search(backend, pos((0.0, 1.0, 1.0), (above, 1.0, 1.0)), 0.5)[0]) == [0, 0]


def test_same_adjacent_and_across_cell_boundaries(backend):
    expect_correct(backend, pos((2.1, 2.2, 2.3), (2.4, 2.4, 2.4), (2.7, 2.5, 2.3)), 0.75)
    assert list(search(backend, pos((0.9, 2.0, 2.0), (1.1, 2.0, 2.0)), 0.3)[0]) == [1, 1]
    assert list(search(backend, pos((0.9999, 2.0, 2.0), (1.0001, 2.0, 2.0)), 0.01)[0]) == [1, 1]


def test_domain_boundaries(backend):
    p = pos((0.01, 0.01, 0.01), (0.11, 0.01, 0.01), (4.89, 4.99, 4.99), (4.99, 4.99, 4.99))
    assert list(search(backend, p, 0.15)[0]) == [1, 1, 1, 1]


def test_line_counts(backend):
    p = pos(*[(0.5 + i * 0.5, 2.0, 2.0) for i in range(8)])
    counts, _ = expect_correct(backend, p, 1.01)
    assert list(counts) == [2, 3, 4, 4, 4, 4, 3, 2]


def lattice(side, spacing):
    return pos(*[(1.0 + x * spacing, 1.0 + y * spacing, 1.0 + z * spacing)
                 for z in range(side) for y in range(side) for x in range(side)])


@pytest.mark.parametrize("side,idx", [(3, (0, 1, 4, 13)), (5, (0, 2, 12, 62))])
def test_lattice_counts_and_symmetry(backend, side, idx):
    counts, lists = expect_correct(backend, lattice(side, 0.5), 0.9, cell=0.5)
    assert [counts[i] for i in idx] == [7, 11, 17, 26]
    for i, lst in enumerate(lists):
        for j in lst:
            assert i in lists[j]


def test_sparse_dense_and_identical(backend):
    sparse = pos((0.1, 0.1, 0.1), (1.1, 3.7, 2.2), (2.5, 0.8, 4.1), (3.9, 2.4, 0.7),
                 (4.8, 4.8, 4.8))
    assert list(search(backend, sparse, 0.2)[0]) == [0] * 5
    dense = pos(*[(2.0 + 0.001 * i, 2.0, 2.0) for i in range(32)])
    assert list(search(backend, dense, 0.5)[0]) == [31] * 32
    same = pos(*[(2.0, 2.0, 2.0)] * 4)
    counts, _ = expect_correct(backend, same, 0.75)
    assert list(counts) == [3, 3, 3, 3]


def test_golden_random_257(backend):
    d = np.load(GOLDEN / "neighbors_257.npz")
    p = np.concatenate([d["positions"], np.ones((257, 1), np.float32)], 1)
    counts, lists, _ = search(backend, p, float(d["radius"]), cell=float(d["cellSize"]))
    m = np.unpackbits(d["adjacency"], axis=1)[:, :257].astype(bool)
    assert lists == [sorted(np.nonzero(r)[0].tolist()) for r in m]
    assert list(counts) == list(m.sum(1))


def test_overflow_still_counts(backend):
    counts, _, flag = search(backend, pos(*[(2.0, 2.0, 2.0)] * 12), 0.75, maxn=8)
    assert list(counts) == [11] * 12 and flag == 1


def test_radius_beyond_integer_range(backend):
    expect_correct(backend, pos((-0.1, 0.0, 0.0), (1.0, 1.0, 1.0), (5.1, 5.1, 5.1)), 1e10)


def test_six_directions_and_diagonal(backend):
    p = pos((2.01, 2.01, 2.01), (1.99, 2.01, 2.01), (2.01, 1.99, 2.01), (2.01, 2.01, 1.99),
            (1.99, 1.99, 1.99), (5.01, 5.01, 5.01), (4.99, 4.99, 4.99), (-0.01, -0.01, -0.01),
            (0.01, 0.01, 0.01))
    _, lists = expect_correct(backend, p, 0.1)
    for i, lst in enumerate(lists):
        for j in lst:
            assert i in lists[j]


def test_slot_major_layout_with_padded_stride(backend):
    """Stride 8 > n = 5, capacity 4 (test/test_neighbors.cu:400-459)."""
    p = pos((3.1, 1.0, 1.0), (0.1, 1.0, 1.0), (3.3, 1.0, 1.0), (0.3, 1.0, 1.0), (0.5, 1.0, 1.0))
    n, h, maxn, stride = 5, 0.45, 4, 8
    g = backend.grid(n, (0, 0, 0), (5, 5, 5), 1.0)
    sk, si, cs, ce = g.build(p)
    ps = p[si]
    nbr, cnt, _ = backend.k.find_neighbors(ps, cs, ce, g.gridSize(), (0, 0, 0), 1.0, n, h, maxn,
                                           stride, neighborsFill=-1, with_flag=False)
    exp = brute_force(ps, h)
    for i in range(n):
        assert cnt[i] == len(exp[i])
        assert sorted(int(nbr[k * stride + i]) for k in range(cnt[i])) == exp[i]
    # Slots past each count keep the fill value.
    for i in range(n):
        for k in range(cnt[i], maxn):
            assert nbr[k * stride + i] == 0xFFFFFFFF
