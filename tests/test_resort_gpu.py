"""Grid builds by re-sort (pbf_grid_rebuild, pbf_solver_set_incremental_sort).

A re-sort takes the positions in the order of the previous build's sort,
keeps the slots whose cell key did not change (already in order), sorts the
changed ones per splitter segment and merges them in.  Its outputs must be
pbf_grid_build's bit for bit, so every case here checks against the
std::stable_sort expectation of test/test_spatial_grid.cu:77-142
(test_grid.reference_build):

  * small moves (5% of the particles to a neighbouring cell), ten rebuilds
    in a row, at particle counts around the 1,024-slot tile (1, 1,023,
    1,024, 1,025) and at 300,000;
  * no key changed (the merge copies the kept slots) and every key changed;
  * one segment far above the LDS capacity of 1,024 entries (most
    particles jump above the previous largest key), which sorts that
    segment in global memory;
  * positions not in the previous build's order at all (correct, only
    slower) and a changed particle count (falls back to the full sort);
  * the solver and the batched solver step bitwise the same with the
    re-sort on and off, eagerly and from graphs, with and without skin
    lists;
  * pbf_grid_rebuild validates its arguments like pbf_grid_build;
  * the 1M and 8M benchmark rollouts end bitwise equal both ways.
"""
from __future__ import annotations

import numpy as np
import pytest

from test_grid import reference_build

pytestmark = pytest.mark.gpu

MIN, MAX, CS = (0.0, 0.0, 0.0), (16.0, 8.0, 8.0), 0.25


def _box(n, rng, lo=(0.0, 0.0, 0.0), hi=MAX):
    p = np.zeros((n, 4), np.float32)
    p[:, :3] = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    p[:, 3] = 1.0
    return p


def _check(got, p):
    exp = reference_build(p, MIN, MAX, CS)
    for name, a, e in zip(("sortedKeys", "sortedIndices", "cellStart", "cellEnd"), got, exp):
        assert np.array_equal(a, e), name


def _move(p, rng, frac=0.05, step=CS):
    q = p.copy()
    sel = rng.random(len(q)) < frac
    q[sel, :3] += rng.choice([-step, step], (sel.sum(), 3)).astype(np.float32)
    q[:, :3] = np.clip(q[:, :3], 0.0, np.asarray(MAX, np.float32) - 1e-3)
    return q


@pytest.mark.parametrize("n", [1, 1023, 1024, 1025, 300_000])
def test_rebuild_chain_of_small_moves(gpu, n):
    rng = np.random.default_rng(n)
    g = gpu.grid(n, MIN, MAX, CS)
    p = _box(n, rng)
    sk, si, cs, ce = g.build(p)
    for it in range(10):
        p = _move(p[si], rng)  # the previous build's order, then a small move
        got = g.rebuild(p)
        _check(got, p)
        si = got[1]
    print(f"n={n}: ten rebuilds equal the stable sort", flush=True)


def test_rebuild_without_change_and_with_every_key_changed(gpu):
    rng = np.random.default_rng(7)
    n = 50_000
    g = gpu.grid(n, MIN, MAX, CS)
    p = _box(n, rng)
    _, si, _, _ = g.build(p)
    p = p[si]
    _check(g.rebuild(p), p)  # nothing changed: the identity order
    q = p.copy()
    q[:, 0] = np.float32(MAX[0]) - np.float32(1e-3) - q[:, 0]  # mirror: every key changes
    _check(g.rebuild(q), q)


def test_rebuild_with_a_segment_beyond_lds(gpu):
    # The previous build holds every particle in the lowest quarter of the
    # box; 80% of them then jump to the top, above every previous key, so
    # all those entries fall into the last tile's segment (about 32,000
    # entries against an LDS capacity of 1,024).
    rng = np.random.default_rng(11)
    n = 40_000
    g = gpu.grid(n, MIN, MAX, CS)
    p = _box(n, rng, hi=(16.0, 8.0, 2.0))
    _, si, _, _ = g.build(p)
    p = p[si]
    sel = rng.random(n) < 0.8
    p[sel, 2] += np.float32(5.5)
    _check(g.rebuild(p), p)


def test_rebuild_of_unordered_positions_and_other_counts(gpu):
    rng = np.random.default_rng(5)
    g = gpu.grid(20_000, MIN, MAX, CS)
    p = _box(20_000, rng)
    g.build(p)
    q = _box(20_000, rng)  # unrelated positions: correct, only slower
    _check(g.rebuild(q), q)
    r = _box(12_345, rng)  # another count: the full sort
    _check(g.rebuild(r), r)
    _check(g.rebuild(r), r)  # the same positions again: no key changed


def _solver(sc, resort, graphs, skin=0.0):
    import pbf_amd

    s = pbf_amd.PBFSolver()
    s.initialize(sc.n, sc.minBounds, sc.maxBounds, sc.params)
    s.setParticles(sc.positions, sc.velocities)
    s.setIncrementalSort(resort)
    s.setGraphs(graphs)
    if skin:
        s.setNeighborSkin(skin)
    return s


def _bits(a):
    return a.view(np.uint32)


@pytest.mark.parametrize("graphs", [False, True])
@pytest.mark.parametrize("skin", [0.0, 0.3])
def test_solver_steps_equal_with_resort_on_and_off(gpu, graphs, skin):
    from test_parity_gpu import block_with_spray

    sc = block_with_spray()
    on, off = _solver(sc, True, graphs, skin), _solver(sc, False, graphs, skin)
    for f in range(4):
        on.step()
        off.step()
    assert np.array_equal(_bits(on.copyPositionsToHost()), _bits(off.copyPositionsToHost()))
    assert np.array_equal(_bits(on.copyVelocitiesToHost()), _bits(off.copyVelocitiesToHost()))
    # setParticles in between: the next build sorts in full again
    on.setParticles(sc.positions, sc.velocities)
    off.setParticles(sc.positions, sc.velocities)
    on.run(2)
    off.run(2)
    assert np.array_equal(_bits(on.copyPositionsToHost()), _bits(off.copyPositionsToHost()))
    assert np.array_equal(_bits(on.copyVelocitiesToHost()), _bits(off.copyVelocitiesToHost()))


def test_batch_steps_equal_with_resort_on_and_off(gpu):
    import pbf_amd
    from pbf_amd import scenes

    scs = [scenes.batch_scene(k, nmin=500, nmax=6000) for k in range(10)]
    out = []
    for resort in (True, False):
        b = pbf_amd.PBFBatchSolver()
        b.initialize([(s.n, s.minBounds, s.maxBounds) for s in scs], scs[0].params)
        b.setParticles(np.concatenate([s.positions for s in scs]),
                       np.concatenate([s.velocities for s in scs]))
        b.setIncrementalSort(resort)
        b.setGraphs(True)
        b.run(3)
        out.append((b.copyPositionsToHost(), b.copyVelocitiesToHost()))
    assert np.array_equal(_bits(out[0][0]), _bits(out[1][0]))
    assert np.array_equal(_bits(out[0][1]), _bits(out[1][1]))


def test_rebuild_validates_like_build(gpu):
    """pbf_grid_rebuild raises what pbf_grid_build raises (spatial_grid.cu:
    139-160): before initialize, on null positions, above capacity; a grid
    initialised again (another capacity) forgets the old order."""
    import pbf_amd

    g = pbf_amd.SpatialGrid()
    with pytest.raises(pbf_amd.LogicError):
        g.rebuild(None, 1)
    g.initialize(1000, MIN, MAX, CS)
    with pytest.raises(pbf_amd.InvalidArgument):
        g.rebuild(None, 1)
    rng = np.random.default_rng(3)
    p = _box(1000, rng)
    d = pbf_amd.DeviceArray.from_host(p, 16)
    with pytest.raises(pbf_amd.OutOfRange):
        g.rebuild(d, 1001)
    g.rebuild(d, 1000)
    sk, si, cs, ce = g.read(1000)
    _check((sk, si, cs, ce), p)
    g.initialize(2000, MIN, MAX, CS)  # a new grid: no previous order
    q = _box(2000, rng)
    dq = pbf_amd.DeviceArray.from_host(q, 16)
    g.rebuild(dq, 2000)
    _check(g.read(2000), q)
    r = _move(q[g.read(2000)[1]], rng)
    dr = pbf_amd.DeviceArray.from_host(r, 16)
    g.rebuild(dr, 2000)
    _check(g.read(2000), r)
    g.rebuild(dr, 0)  # an empty build clears the ranges, as build does
    _, _, cs0, ce0 = g.read(0)
    assert np.all(cs0 == -1) and np.all(ce0 == -1)


@pytest.mark.parametrize("config,frames", [(2, 300), (3, 30)])
def test_benchmark_rollouts_equal_with_resort_on_and_off(gpu, config, frames):
    """The 1M dam break over 300 frames and the 8M double dam break over 30
    frames, frames replayed from graphs: the re-sorting builds (splashes
    and fast flow included) end bitwise where the radix-sorted ones do."""
    import pbf_amd
    from pbf_amd import scenes

    sc = scenes.dam_break(config)
    out = []
    for resort in (True, False):
        s = pbf_amd.PBFSolver()
        s.initialize(sc.n, sc.minBounds, sc.maxBounds, sc.params)
        s.setParticles(sc.positions, sc.velocities)
        s.setIncrementalSort(resort)
        s.setGraphs(True)
        s.run(frames)
        out.append((s.copyPositionsToHost(), s.copyVelocitiesToHost()))
        del s
    print(f"{sc.name}: {frames} frames, max |v| "
          f"{np.linalg.norm(out[0][1][:, :3], axis=1).max():.3f}", flush=True)
    assert np.array_equal(_bits(out[0][0]), _bits(out[1][0]))
    assert np.array_equal(_bits(out[0][1]), _bits(out[1][1]))
