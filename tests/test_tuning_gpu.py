"""Launch tunings and the autotuner (PBFSolver.setTuning / autotune).

  * the workgroup size of the neighbour passes and their particles per
    thread do not change any result: states bitwise equal at 64, 128, 192
    and 256 threads and at 1, 2 and 4 particles per thread;
  * autotune leaves the particle state and staleFrames() as they were,
    applies the fastest candidate, and the frames after it equal a solver
    given that tuning directly, bit for bit;
  * invalid tunings raise invalid_argument.
"""
from __future__ import annotations

import numpy as np
import pytest

import pbf_amd
from pbf_amd import scenes
from test_parity_gpu import _solver_for

pytestmark = pytest.mark.gpu


def _state(s):
    return s.copyPositionsToHost(), s.copyVelocitiesToHost()


@pytest.mark.parametrize("scene", ["dam_break", "spray"])
def test_pass_block_and_particles_per_thread_do_not_change_results(gpu, scene):
    from test_parity_gpu import block_with_spray

    sc = scenes.dam_break(1) if scene == "dam_break" else block_with_spray()
    ref = None
    for block, ppt in ((256, 1), (64, 1), (128, 1), (192, 1), (256, 2), (64, 4), (192, 2)):
        s = _solver_for(sc, gpu.Solver)
        s.setTuning(0.0, block, ppt)
        s.run(3)
        st = _state(s)
        if ref is None:
            ref = st
        else:
            assert np.array_equal(st[0].view(np.uint32), ref[0].view(np.uint32)), (block, ppt)
            assert np.array_equal(st[1].view(np.uint32), ref[1].view(np.uint32)), (block, ppt)


def test_autotune_keeps_state_and_applies_the_fastest(gpu):
    sc = scenes.dam_break(1)
    s = _solver_for(sc, gpu.Solver)
    s.run(2)
    before = _state(s)
    stale = s.staleFrames()
    chosen, times = s.autotune(2)
    print("[autotune] " + ", ".join(f"({a:.2f}, {b}, {c}): {ms:.3f} ms" for a, b, c, ms in times),
          flush=True)
    after = _state(s)
    assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])
    assert s.staleFrames() == stale
    assert len(times) == pbf_amd.DEFAULT_TUNING_COUNT
    best = min(times, key=lambda t: t[3])
    assert abs(chosen.neighborSkin - best[0]) < 1e-7 and chosen.passBlock == best[1]
    assert chosen.particlesPerThread == best[2]
    t = s.tuning()
    assert abs(t.neighborSkin - best[0]) < 1e-7 and t.passBlock == best[1]
    assert t.particlesPerThread == best[2]

    twin = _solver_for(sc, gpu.Solver)
    twin.run(2)
    twin.setTuning(chosen.neighborSkin, chosen.passBlock, chosen.particlesPerThread)
    s.run(3)
    twin.run(3)
    a, b = _state(s), _state(twin)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_autotune_explicit_candidates_and_batch(gpu):
    scs = [scenes.batch_scene(k, nmin=500, nmax=3000, substeps=1) for k in range(4)]
    b = pbf_amd.PBFBatchSolver()
    b.initialize([(sc.n, sc.minBounds, sc.maxBounds) for sc in scs], scs[0].params)
    b.setParticles(np.concatenate([x.positions for x in scs]),
                   np.concatenate([x.velocities for x in scs]))
    before = b.copyPositionsToHost()
    chosen, times = b.autotune(1, [(0.0, 256), (0.0, 128, 2)])
    assert len(times) == 2 and (chosen.passBlock, chosen.particlesPerThread) in ((256, 1), (128, 2))
    assert np.array_equal(before, b.copyPositionsToHost())


def test_invalid_tunings_raise(gpu):
    sc = scenes.dam_break(1)
    s = _solver_for(sc, gpu.Solver)
    for skin, block in ((0.1, 96), (0.1, 512), (0.1, 0), (-1.0, 256), (2.0, 256)):
        with pytest.raises(ValueError):
            s.setTuning(skin, block)
    with pytest.raises(ValueError):
        s.setTuning(0.0, 256, 3)
    with pytest.raises(ValueError):
        s.autotune(1, [(0.0, 100)])
    with pytest.raises(ValueError):
        s.autotune(0)
