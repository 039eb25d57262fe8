"""Regenerates the committed fixtures of tests/golden/.

Inputs: the seeded random particle sets of the reference's randomized grid
and neighbour tests (gen_mt19937.cpp, compiled with g++).
Expected outputs: computed here with numpy, independently of the oracle --
  * grid: FP32 cell coordinates floor((p - min) / cellSize) clamped into the
    grid, keys x + y*gx + z*gx*gy, numpy's stable argsort, and first/last
    occurrence per key for the ranges (the reference's own CPU check,
    test/test_spatial_grid.cu:77-142);
  * neighbours: all-pairs FP32 brute force d2 = (dx*dx + dy*dy) + dz*dz <= h*h
    excluding self (test/test_neighbors.cu:114-136).

Usage: python tests/golden/make_golden.py
"""
from __future__ import annotations

import subprocess
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def generate_inputs(tmp: Path) -> dict[str, np.ndarray]:
    exe = tmp / "gen"
    subprocess.run(["g++", "-O1", "-std=c++17", str(HERE / "gen_mt19937.cpp"), "-o", str(exe)],
                   check=True)
    subprocess.run([str(exe), str(tmp)], check=True)
    return {name: np.fromfile(tmp / f"{name}.f32", dtype=np.float32).reshape(-1, 3)
            for name in ("grid_8193", "neighbors_257")}


def grid_expected(p: np.ndarray, mn, mx, cs):
    p = p.astype(np.float32)
    mn = np.asarray(mn, np.float32)
    dims = np.ceil((np.asarray(mx, np.float64) - np.asarray(mn, np.float64)) / np.float64(cs))
    dims = dims.astype(np.int64)
    q = np.floor((p - mn) / np.float32(cs))          # float32 arithmetic throughout
    cell = np.clip(q.astype(np.int64), 0, dims - 1)
    keys = (cell[:, 0] + cell[:, 1] * dims[0] + cell[:, 2] * dims[0] * dims[1]).astype(np.uint32)
    order = np.argsort(keys, kind="stable").astype(np.uint32)
    sk = keys[order]
    cells = int(np.prod(dims))
    start = np.full(cells, -1, np.int32)
    end = np.full(cells, -1, np.int32)
    uniq, first = np.unique(sk, return_index=True)
    _, last = np.unique(sk[::-1], return_index=True)
    start[uniq] = first
    end[uniq] = len(sk) - last
    return dims, keys, sk, order, start, end


def neighbors_expected(p: np.ndarray, h: float):
    p = p.astype(np.float32)
    h = np.float32(h)
    d = p[:, None, :] - p[None, :, :]
    d2 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
    m = d2 <= h * h
    np.fill_diagonal(m, False)
    return m


def main() -> None:
    with tempfile.TemporaryDirectory() as t:
        inputs = generate_inputs(Path(t))
    g = inputs["grid_8193"]
    mn, mx, cs = (-8.0, -5.0, -3.0), (9.0, 7.0, 6.0), 0.75
    dims, keys, sk, order, start, end = grid_expected(g, mn, mx, cs)
    np.savez_compressed(HERE / "grid_8193.npz", positions=g, minBounds=np.float32(mn),
                        maxBounds=np.float32(mx), cellSize=np.float32(cs), dims=dims, keys=keys,
                        sortedKeys=sk, sortedIndices=order, cellStart=start, cellEnd=end)
    nb = inputs["neighbors_257"]
    m = neighbors_expected(nb, 0.85)
    np.savez_compressed(HERE / "neighbors_257.npz", positions=nb, radius=np.float32(0.85),
                        cellSize=np.float32(0.6), minBounds=np.float32((0, 0, 0)),
                        maxBounds=np.float32((5, 5, 5)), adjacency=np.packbits(m, axis=1))
    print("wrote grid_8193.npz, neighbors_257.npz")


if __name__ == "__main__":
    main()
