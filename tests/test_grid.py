"""Spatial grid known answers (test/test_spatial_grid.cu) on the oracle and the
GPU, the golden fixture of the randomized test, and the SpatialGrid API
contract (validation, strong guarantee) of the library."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"
MIN, MAX, CS = (0.0, 0.0, 0.0), (4.0, 3.0, 2.0), 1.0


def pos(*xyz):
    return np.array([[x, y, z, 1.0] for x, y, z in xyz], dtype=np.float32)


def reference_build(p, mn, mx, cs):
    """std::stable_sort-based expectation (test/test_spatial_grid.cu:77-142)."""
    p = np.asarray(p, np.float32)[:, :3]
    dims = np.ceil((np.asarray(mx, np.float64) - np.asarray(mn, np.float64)) / cs).astype(np.int64)
    q = np.floor((p - np.asarray(mn, np.float32)) / np.float32(cs))
    c = np.clip(q.astype(np.int64), 0, dims - 1)
    keys = (c[:, 0] + c[:, 1] * dims[0] + c[:, 2] * dims[0] * dims[1]).astype(np.uint32)
    order = np.argsort(keys, kind="stable").astype(np.uint32)
    sk = keys[order]
    cells = int(np.prod(dims))
    start = np.full(cells, -1, np.int32)
    end = np.full(cells, -1, np.int32)
    for i, k in enumerate(sk):
        if i == 0 or sk[i - 1] != k:
            start[k] = i
        if i == len(sk) - 1 or sk[i + 1] != k:
            end[k] = i + 1
    return sk, order, start, end


def check_equal(actual, expected):
    for a, e in zip(actual, expected):
        assert np.array_equal(a, e)


def test_dimensions_exact_and_ceiled(backend):
    g = backend.grid(10, (-2.0, -4.0, 1.0), (4.0, 2.0, 7.0), 2.0)
    assert tuple(g.gridSize()) == (3, 3, 3) and g.numCells() == 27
    g = backend.grid(10, (-1.0, 2.0, 0.0), (4.1, 7.9, 2.01), 2.0)
    assert tuple(g.gridSize()) == (3, 3, 2) and g.numCells() == 18


@pytest.mark.parametrize("points", [
    [(1.2, 0.4, 1.8)],
    [(0.2, 0.2, 0.2), (1.2, 0.2, 0.2), (2.2, 1.2, 0.2), (3.2, 2.2, 1.2)],
    [(2.1, 1.1, 0.1), (2.4, 1.8, 0.3), (2.9, 1.2, 0.9), (2.5, 1.5, 0.5)],
    [(0.2, 0.2, 0.2), (0.8, 0.7, 0.4), (2.2, 1.2, 0.2), (3.8, 2.8, 1.8), (2.6, 1.7, 0.9),
     (3.1, 2.1, 1.1)],
    [(3.2, 2.2, 1.2), (1.2, 0.2, 0.2), (3.8, 2.8, 1.8), (0.2, 0.2, 0.2), (2.2, 1.2, 0.2)],
])
def test_builds_match_stable_sort_reference(backend, points):
    p = pos(*points)
    g = backend.grid(32, MIN, MAX, CS)
    check_equal(g.build(p), reference_build(p, MIN, MAX, CS))


def test_half_open_ranges_and_empty_cells(backend):
    g = backend.grid(32, MIN, MAX, CS)
    sk, si, cs, ce = g.build(pos((0.1, 0.1, 0.1), (0.2, 0.2, 0.2), (0.3, 0.3, 0.3),
                                 (2.1, 0.1, 0.1), (2.2, 0.2, 0.2)))
    assert (cs[0], ce[0], cs[2], ce[2]) == (0, 3, 3, 5)
    sk, si, cs, ce = g.build(pos((0.1, 0.1, 0.1), (2.1, 1.1, 1.1)))
    others = np.ones(24, bool)
    others[[0, 18]] = False
    assert np.all(cs[others] == -1) and np.all(ce[others] == -1)


@pytest.mark.parametrize("point,key", [
    ((0.0, 0.0, 0.0), 0), ((3.999, 2.999, 1.999), 23), ((-5.0, 1.2, 0.2), 4),
    ((10.0, 1.2, 0.2), 7),
])
def test_known_keys_and_clamping(backend, point, key):
    g = backend.grid(32, MIN, MAX, CS)
    sk, si, _, _ = g.build(pos(point))
    assert sk[0] == key and si[0] == 0


def test_clamps_outside_multiple_axes(backend):
    g = backend.grid(32, MIN, MAX, CS)
    sk, _, _, _ = g.build(pos((-2.0, 8.0, -4.0), (9.0, -3.0, 7.0)))
    assert list(sk) == [8, 15]


def test_first_and_last_cells(backend):
    g = backend.grid(32, MIN, MAX, CS)
    sk, si, cs, ce = g.build(pos((3.5, 2.5, 1.5), (0.5, 0.5, 0.5)))
    assert list(sk) == [0, 23] and list(si) == [1, 0]
    assert (cs[0], ce[0], cs[23], ce[23]) == (0, 1, 1, 2)


def test_clamps_many_into_one_boundary_cell(backend):
    g = backend.grid(32, MIN, MAX, CS)
    sk, _, cs, ce = g.build(pos((-1.0, -1.0, -1.0), (-10.0, -2.0, -3.0), (-0.01, -0.01, -0.01)))
    assert list(sk) == [0, 0, 0] and (cs[0], ce[0]) == (0, 3)


def test_clamps_beyond_integer_range(backend):
    g = backend.grid(32, MIN, MAX, CS)
    sk, si, _, _ = g.build(pos((1e30, 1e30, 1e30), (-1e30, -1e30, -1e30)))
    assert list(sk) == [0, 23] and list(si) == [1, 0]


def test_nan_position_maps_to_cell_zero(backend):
    """static_cast<int>(NaN) is 0 on the reference's GPU path, then clamped."""
    g = backend.grid(32, MIN, MAX, CS)
    sk, _, _, _ = g.build(pos((np.nan, 1.5, 0.5)))
    assert sk[0] == 4


def test_repeated_builds_and_cleared_ranges(backend):
    g = backend.grid(6, MIN, MAX, CS)
    a = pos((0.1, 0.1, 0.1), (1.1, 1.1, 0.1), (3.1, 2.1, 1.1))
    b = pos((2.1, 0.1, 0.1), (2.2, 0.2, 0.2))
    check_equal(g.build(a), reference_build(a, MIN, MAX, CS))
    sk, si, cs, ce = g.build(b)
    check_equal((sk, si, cs, ce), reference_build(b, MIN, MAX, CS))
    assert cs[0] == -1 and cs[5] == -1 and cs[23] == -1 and (cs[2], ce[2]) == (0, 2)


def test_empty_build_clears_ranges(backend):
    g = backend.grid(32, MIN, MAX, CS)
    _, _, cs, _ = g.build(pos((1.2, 0.4, 1.8)))
    assert cs[13] != -1
    sk, si, cs, ce = g.build(np.zeros((0, 4), np.float32))
    assert len(sk) == 0 and np.all(cs == -1) and np.all(ce == -1)


def test_golden_random_8193(backend):
    d = np.load(GOLDEN / "grid_8193.npz")
    p = np.concatenate([d["positions"], np.ones((len(d["positions"]), 1), np.float32)], 1)
    g = backend.grid(len(p), tuple(d["minBounds"]), tuple(d["maxBounds"]), float(d["cellSize"]))
    assert tuple(g.gridSize()) == tuple(d["dims"])
    sk, si, cs, ce = g.build(p)
    assert np.array_equal(sk, d["sortedKeys"])
    assert np.array_equal(si, d["sortedIndices"])
    assert np.array_equal(cs, d["cellStart"]) and np.array_equal(ce, d["cellEnd"])
    assert np.all(np.diff(sk.astype(np.int64)) >= 0)
    assert np.array_equal(sk, d["keys"][si])


def test_count_above_capacity_rejected(backend):
    g = backend.grid(3, MIN, MAX, CS)
    with pytest.raises(backend.E.OutOfRange):
        g.build(pos((0.1, 0.1, 0.1), (0.2, 0.2, 0.2), (0.3, 0.3, 0.3), (0.4, 0.4, 0.4)))


@pytest.mark.parametrize("args,exc", [
    ((0, (0, 0, 0), (1, 1, 1), 1.0), "InvalidArgument"),
    ((1, (0, 0, 0), (1, 1, 1), 0.0), "InvalidArgument"),
    ((1, (0, 0, 0), (1, 1, 1), -0.5), "InvalidArgument"),
    ((1, (np.nan, 0, 0), (1, 1, 1), 1.0), "InvalidArgument"),
    ((1, (0, 0, 0), (1, 1, 1), np.nan), "InvalidArgument"),
    ((1, (0, 0, 0), (0, 2, 2), 1.0), "InvalidArgument"),
    ((1, (0, 0, 0), (2, -1, 2), 1.0), "InvalidArgument"),
    ((1, (0, 0, 0), (2, 2, 0), 1.0), "InvalidArgument"),
    ((1, (0, 0, 0), (65536.0, 65536.0, 2.0), 1.0), "LengthError"),
    ((2 ** 31, (0, 0, 0), (1, 1, 1), 1.0), "LengthError"),
])
def test_initialize_validation(backend, args, exc):
    with pytest.raises(getattr(backend.E, exc)):
        backend.grid(*args)


@pytest.mark.gpu
def test_failed_reinitialization_preserves_configuration(gpu):
    import pbf_amd

    g = pbf_amd.SpatialGrid()
    g.initialize(2, (0, 0, 0), (2, 2, 2), 1.0)
    with pytest.raises(pbf_amd.LengthError):
        g.initialize(100, (0, 0, 0), (1e20, 2, 2), 1.0)
    assert g.maxParticles() == 2 and g.gridSize().x == 2 and g.maxBounds().x == 2.0


@pytest.mark.gpu
def test_null_build_input_and_uninitialized_grid(gpu):
    import pbf_amd

    g = pbf_amd.SpatialGrid()
    with pytest.raises(pbf_amd.LogicError):
        g.build(None, 1)
    g.initialize(2, (0, 0, 0), (2, 2, 2), 1.0)
    with pytest.raises(pbf_amd.InvalidArgument):
        g.build(None, 1)
