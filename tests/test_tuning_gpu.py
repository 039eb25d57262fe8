"""Launch tunings and the autotuner (PBFSolver.setTuning / autotune).

  * the workgroup size of the neighbour passes does not change any result:
    states bitwise equal at 64, 128, 192 and 256 threads;
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


def test_pass_block_does_not_change_results(gpu):
    sc = scenes.dam_break(1)
    ref = None
    for block in (256, 64, 128, 192):
        s = _solver_for(sc, gpu.Solver)
        s.setTuning(0.0, block)
        s.run(3)
        st = _state(s)
        if ref is None:
            ref = st
        else:
            assert np.array_equal(st[0], ref[0]) and np.array_equal(st[1], ref[1]), block


def test_autotune_keeps_state_and_applies_the_fastest(gpu):
    sc = scenes.dam_break(1)
    s = _solver_for(sc, gpu.Solver)
    s.run(2)
    before = _state(s)
    stale = s.staleFrames()
    chosen, times = s.autotune(2)
    print("[autotune] " + ", ".join(f"({a:.2f}, {b}): {ms:.3f} ms" for a, b, ms in times),
          flush=True)
    after = _state(s)
    assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])
    assert s.staleFrames() == stale
    best = min(times, key=lambda t: t[2])
    assert abs(chosen.neighborSkin - best[0]) < 1e-7 and chosen.passBlock == best[1]
    t = s.tuning()
    assert abs(t.neighborSkin - best[0]) < 1e-7 and t.passBlock == best[1]

    twin = _solver_for(sc, gpu.Solver)
    twin.run(2)
    twin.setTuning(chosen.neighborSkin, chosen.passBlock)
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
    chosen, times = b.autotune(1, [(0.0, 256), (0.0, 128)])
    assert len(times) == 2 and chosen.passBlock in (128, 256)
    assert np.array_equal(before, b.copyPositionsToHost())


def test_invalid_tunings_raise(gpu):
    sc = scenes.dam_break(1)
    s = _solver_for(sc, gpu.Solver)
    for skin, block in ((0.1, 96), (0.1, 512), (0.1, 0), (-1.0, 256), (2.0, 256)):
        with pytest.raises(ValueError):
            s.setTuning(skin, block)
    with pytest.raises(ValueError):
        s.autotune(1, [(0.0, 100)])
    with pytest.raises(ValueError):
        s.autotune(0)
