"""Frames replayed from captured HIP graphs equal frames launched kernel by
kernel, bit for bit.

A frame permutes the working-set buffers (five grid builds per substep each
swap them with their alternates; the velocity update swaps x and p*), so the
solver keeps one graph per starting buffer assignment: one substep per frame
and three substeps per frame cycle through two assignments, two substeps per
frame through one.  Every replay is checked against the eager solver, with
eager work (copy-out) and a synchronisation between replays; a collider
change between runs must drop the graphs (their kernel arguments are baked).
"""
from __future__ import annotations

import numpy as np
import pytest

from pbf_amd import scenes

pytestmark = pytest.mark.gpu


def _solver(gpu, sc, graphs):
    s = gpu.Solver()
    s.initialize(sc.n, sc.minBounds, sc.maxBounds, sc.params)
    s.setGraphs(graphs)
    s.setParticles(sc.positions, sc.velocities)
    return s


def _diff(x, y):
    bad = np.any(x.view(np.uint32) != y.view(np.uint32), axis=1)
    d = np.abs(x.astype(np.float64) - y.astype(np.float64))
    return f"{int(bad.sum())} of {len(x)} particles, first {np.flatnonzero(bad)[:5]}, max |d| {d.max():.3g}"


def _same(a, b, label):
    pa, pb = a.copyPositionsToHost(), b.copyPositionsToHost()
    va, vb = a.copyVelocitiesToHost(), b.copyVelocitiesToHost()
    assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32)), \
        f"{label}: positions differ in {_diff(pa, pb)}"
    assert np.array_equal(va.view(np.uint32), vb.view(np.uint32)), \
        f"{label}: velocities differ in {_diff(va, vb)}"


@pytest.mark.parametrize("substeps", [1, 2, 3])
def test_graph_frames_equal_eager_frames(gpu, substeps):
    sc = scenes.small_block(16, 16, 16, substeps=substeps)
    eager = _solver(gpu, sc, False)
    eager2 = _solver(gpu, sc, False)
    graphed = _solver(gpu, sc, True)
    for rep in range(4):
        for s in (eager, eager2, graphed):
            s.run(2)
        _same(eager, eager2, f"substeps {substeps}, run(2) #{rep}, two eager solvers")
        _same(eager, graphed, f"substeps {substeps}, run(2) #{rep}")
    for rep in range(3):
        for s in (eager, eager2, graphed):
            s.step()
        _same(eager, eager2, f"substeps {substeps}, step #{rep}, two eager solvers")
        _same(eager, graphed, f"substeps {substeps}, step #{rep}")
    print(f"[substeps {substeps}] 11 frames, graphs equal eager", flush=True)


def test_collider_change_drops_graphs(gpu):
    sc = scenes.small_block(16, 16, 16)
    eager = _solver(gpu, sc, False)
    eager2 = _solver(gpu, sc, False)
    graphed = _solver(gpu, sc, True)
    eager.run(3)
    eager2.run(3)
    graphed.run(3)
    _same(eager, eager2, "two eager solvers")
    _same(eager, graphed, "before the sphere")
    sphere = [((1.2, 0.3, 0.4), 0.25)]
    eager.setSpheres(sphere)
    graphed.setSpheres(sphere)
    eager.run(3)
    graphed.run(3)
    _same(eager, graphed, "with the sphere")
    eager.setParticles(sc.positions, sc.velocities)
    graphed.setParticles(sc.positions, sc.velocities)
    eager.run(2)
    graphed.run(2)
    _same(eager, graphed, "after setParticles")


def test_profiling_runs_eagerly_and_matches(gpu):
    sc = scenes.small_block(12, 10, 8)
    a = _solver(gpu, sc, True)
    b = _solver(gpu, sc, True)
    b.setProfiling(True)
    a.run(3)
    b.run(3)
    _same(a, b, "profiled")
    assert any(ms > 0 for ms, _ in b.kernelTimes().values())
