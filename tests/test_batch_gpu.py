"""Batched scenes: every scene of a PBFBatchSolver equals, bit for bit, a
PBFSolver holding that scene alone (looped-versus-batched equality).

The batch sorts all scenes with one radix sort of offset keys, sweeps their
tiles in one list build (the lane form included: a spray scene puts many
sparse tiles into one 256-particle range) and runs every constraint pass as
one launch, with per-scene grids and containers looked up by slot; the
single-scene solver runs the same per-particle arithmetic in the same order.
"""
from __future__ import annotations

import numpy as np
import pytest

from pbf_amd import scenes

pytestmark = pytest.mark.gpu


def _batch_and_loop(gpu, scs, frames):
    import pbf_amd

    params = scs[0].params
    b = pbf_amd.PBFBatchSolver()
    b.initialize([(s.n, s.minBounds, s.maxBounds) for s in scs], params)
    b.setParticles(np.concatenate([s.positions for s in scs]),
                   np.concatenate([s.velocities for s in scs]))
    b.run(frames)
    bp, bv = b.copyPositionsToHost(), b.copyVelocitiesToHost()
    first = 0
    for k, s in enumerate(scs):
        one = gpu.Solver()
        one.initialize(s.n, s.minBounds, s.maxBounds, params)
        one.setParticles(s.positions, s.velocities)
        one.run(frames)
        p, v = one.copyPositionsToHost(), one.copyVelocitiesToHost()
        sl = slice(first, first + s.n)
        assert np.array_equal(bp[sl].view(np.uint32), p.view(np.uint32)), f"scene {k}: positions"
        assert np.array_equal(bv[sl].view(np.uint32), v.view(np.uint32)), f"scene {k}: velocities"
        first += s.n
    return b


def test_batched_scenes_equal_looped_scenes(gpu):
    from test_parity_gpu import block_with_spray

    scs = [scenes.batch_scene(k, nmin=300, nmax=6000, substeps=1) for k in range(9)]
    scs.insert(4, block_with_spray())
    b = _batch_and_loop(gpu, scs, 3)
    assert b.counts() == (len(scs), sum(s.n for s in scs))
    print(f"{len(scs)} scenes, {sum(s.n for s in scs)} particles, 3 frames: batched == looped",
          flush=True)


def test_one_scene_batch_equals_solver(gpu):
    _batch_and_loop(gpu, [scenes.small_block(16, 12, 10, substeps=2)], 2)


def test_batched_graph_frames_equal_eager(gpu):
    import pbf_amd

    scs = [scenes.batch_scene(k, nmin=500, nmax=3000, substeps=1) for k in range(5)]
    runs = []
    for graphs in (False, True):
        b = pbf_amd.PBFBatchSolver()
        b.initialize([(s.n, s.minBounds, s.maxBounds) for s in scs], scs[0].params)
        b.setGraphs(graphs)
        b.setParticles(np.concatenate([s.positions for s in scs]),
                       np.concatenate([s.velocities for s in scs]))
        for _ in range(3):
            b.run(2)
        runs.append(b.copyPositionsToHost())
    assert np.array_equal(runs[0].view(np.uint32), runs[1].view(np.uint32))


def test_batch_raises_the_overflow_of_one_scene(gpu):
    """A scene that overflows its neighbour capacity alone (the literal
    headline strengths diverge in the first frame, as in the oracle) makes
    the batch raise the same error; the other scenes cannot mask it."""
    import pbf_amd

    bad = scenes.dam_break(1, literal_strengths=True)
    good = scenes.batch_scene(3, nmin=500, nmax=2000, substeps=1)
    params = bad.params
    one = gpu.Solver()
    one.initialize(bad.n, bad.minBounds, bad.maxBounds, params)
    one.setParticles(bad.positions, bad.velocities)
    with pytest.raises(pbf_amd.NeighborOverflow):
        one.step()
    b = pbf_amd.PBFBatchSolver()
    b.initialize([(good.n, good.minBounds, good.maxBounds), (bad.n, bad.minBounds, bad.maxBounds)],
                 params)
    b.setParticles(np.concatenate([good.positions, bad.positions]),
                   np.concatenate([good.velocities, bad.velocities]))
    with pytest.raises(pbf_amd.NeighborOverflow):
        b.step()
