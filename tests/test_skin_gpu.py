"""Skin lists (PBFSolver.setNeighborSkin) against the FP64 oracle and the
reference's builds.

With a skin, a substep builds its grid and lists once, around the container
points of the predicted positions, with radius h * (1 + skin); every pass
tests each entry against h.  The bars:
  * one step against the FP64 oracle: the parity bars of test_parity_gpu
    (positions 1e-5 relative per particle, velocities 1e-5 in the norm and
    per particle max(1e-5 |v|, 6 ulp(|x|) / dt)); colliders: the contact bars;
  * a frame whose lists go stale runs again with the reference's builds: with
    a skin so small that every frame goes stale, every frame equals the
    reference-build solver's bit for bit, and run(n) equals n step() calls;
  * overflow and collision non-convergence raise as in the oracle;
  * bitwise determinism, graph replay equal to eager frames, batched scenes
    equal to scenes stepped alone.
"""
from __future__ import annotations

import numpy as np
import pytest

import pbf_amd
from pbf_amd import scenes
from test_parity_gpu import (_compare_state, _compare_with_contacts, _solver_for,
                             block_with_spray)

pytestmark = pytest.mark.gpu


def _skinned(scene, cls, skin):
    s = _solver_for(scene, cls)
    s.setNeighborSkin(skin)
    s.setParticles(scene.positions, scene.velocities)
    return s


def _state(s):
    return s.copyPositionsToHost(), s.copyVelocitiesToHost()


@pytest.mark.parametrize("skin", [0.05, 0.1, 0.3])
def test_one_step_parity_dam_break_config1(gpu, skin):
    from oracle import OracleSolver

    sc = scenes.dam_break(1)
    g = _skinned(sc, gpu.Solver, skin)
    o = _solver_for(sc, lambda: OracleSolver("f64"))
    g.step()
    o.step()
    print(f"[config1 skin {skin}] stale frames {g.staleFrames()}", flush=True)
    _compare_state(*_state(g), *_state(o), sc.params.dt, f"skin {skin} config1 step1")


@pytest.mark.parametrize("skin", [0.1, 0.3])
def test_one_step_parity_block_with_spray(gpu, skin):
    from oracle import OracleSolver

    sc = block_with_spray()
    g = _skinned(sc, gpu.Solver, skin)
    o = _solver_for(sc, lambda: OracleSolver("f64"))
    g.step()
    o.step()
    print(f"[spray skin {skin}] stale frames {g.staleFrames()}", flush=True)
    _compare_state(*_state(g), *_state(o), sc.params.dt, f"skin {skin} block+spray step1")


def test_per_step_parity_colliders(gpu):
    """Spheres, boxes and planes push particles out of the container's
    points: the displacement check measures from the container point, so
    every collider type keeps the lists complete or marks them stale."""
    from oracle import OracleSolver

    sc = scenes.small_block(24, 12, 16)
    sc.positions[:, 1] += 0.30
    sc.spheres = [((0.45, 0.16, 0.4), 0.12)]
    sc.boxes = [((0.9, 0.16, 0.4), (0.12, 0.12, 0.2))]
    sc.planes = [((0.0, 0.03, 0.0), (0.0, 1.0, 0.0))]
    sc.velocities[:, 1] = -1.5
    g = _skinned(sc, gpu.Solver, 0.1)
    o = _solver_for(sc, lambda: OracleSolver("f64"))
    for k in range(8):
        p0, v0 = o.copyPositionsToHost(), o.copyVelocitiesToHost()
        g.setParticles(p0, v0)
        o.setParticles(p0, v0)
        g.step()
        o.step()
        _compare_with_contacts(*_state(g), *_state(o), sc, f"skin colliders step{k + 1}")
    print(f"[skin colliders] stale frames {g.staleFrames()}", flush=True)


def test_every_frame_stale_equals_reference_builds(gpu):
    """A skin of 1e-6 h goes stale in every frame: each frame runs again
    from its checkpoint with the reference's builds, so the state equals a
    solver without skin bit for bit, through run(n) as through step()."""
    sc = scenes.dam_break(1)
    ref = _solver_for(sc, gpu.Solver)
    a = _skinned(sc, gpu.Solver, 1e-6)
    b = _skinned(sc, gpu.Solver, 1e-6)
    ref.run(4)
    a.run(4)
    for _ in range(4):
        b.step()
    assert a.staleFrames() == 4 and b.staleFrames() == 4
    pr, vr = _state(ref)
    for s in (a, b):
        p, v = _state(s)
        assert np.array_equal(p, pr) and np.array_equal(v, vr)


def calm_block(nx=16, ny=12, nz=12, vx=0.0, shift=0.0):
    """A weightless block.  The reference does not clamp the density
    constraint, so the under-dense surface pulls itself in during the
    first two frames (corrections past a skin of 0.3 h: stale frames); then
    it rests, moving far less.  With vx it drifts toward the +x wall,
    shifted `shift` metres from the -x wall, and its impact makes large
    corrections again."""
    sc = scenes.small_block(nx, ny, nz)
    sc.params.gravity = pbf_amd.float3(0.0, 0.0, 0.0)
    sc.positions[:, 0] += shift
    sc.velocities[:, 0] = vx
    return sc


def settled(gpu, sc, frames=3):
    """The scene after `frames` frames with the reference's builds."""
    s = _solver_for(sc, gpu.Solver)
    s.run(frames)
    sc.positions, sc.velocities = _state(s)
    return sc


def test_run_equals_steps_with_stale_frames_between(gpu):
    """run(n) takes up the frames after a stale one exactly as n step()
    calls do: the block drifts in clean frames, then hits the wall."""
    sc = calm_block(vx=3.0, shift=2.1)
    a = _skinned(sc, gpu.Solver, 0.3)
    b = _skinned(sc, gpu.Solver, 0.3)
    a.run(12)
    stale = []
    for _ in range(12):
        b.step()
        stale.append(b.staleFrames())
    print(f"[run vs steps] stale frames after each step: {stale}", flush=True)
    assert a.staleFrames() == b.staleFrames()
    assert 0 < a.staleFrames() < 12, "the workload should mix stale and clean frames"
    pa, va = _state(a)
    pb, vb = _state(b)
    assert np.array_equal(pa, pb) and np.array_equal(va, vb)


def test_calm_block_parity_without_stale_frames(gpu):
    """Clean frames: every pass runs over the skin lists.  Each step starts
    the library (skin 0.3 and the reference's builds) and the FP64 and FP32
    oracles from the same FP64 state.  Positions are held to 1e-5 per
    particle and velocities to the per-particle floor; a resting weightless
    block has velocities near zero, so the velocity norm is held to the
    larger of 1e-5 and twice the FP32 oracle's own distance from FP64."""
    from oracle import OracleSolver

    sc = settled(gpu, calm_block())
    g = _skinned(sc, gpu.Solver, 0.3)
    ref = _solver_for(sc, gpu.Solver)
    o = _solver_for(sc, lambda: OracleSolver("f64"))
    o32 = _solver_for(sc, lambda: OracleSolver("f32"))
    dt = sc.params.dt
    for k in range(3):
        p0, v0 = o.copyPositionsToHost(), o.copyVelocitiesToHost()
        for s in (g, ref, o, o32):
            s.setParticles(p0, v0)
            s.step()
        po, vo = _state(o)
        norm = lambda v: float(np.linalg.norm(v[:, :3] - vo[:, :3]) / np.linalg.norm(vo[:, :3]))
        bar = max(1e-5, 2.0 * norm(o32.copyVelocitiesToHost()))
        pg, vg = _state(g)
        print(f"[calm step{k + 1}] velocity norm error: skin {norm(vg):.3e}, reference builds "
              f"{norm(ref.copyVelocitiesToHost()):.3e}, FP32 oracle "
              f"{norm(o32.copyVelocitiesToHost()):.3e}", flush=True)
        assert norm(vg) <= bar
        rel_p = (np.linalg.norm(pg[:, :3].astype(np.float64) - po[:, :3], axis=1)
                 / np.linalg.norm(po[:, :3].astype(np.float64), axis=1))
        assert rel_p.max() <= 1e-5
        dv = np.linalg.norm(vg[:, :3].astype(np.float64) - vo[:, :3], axis=1)
        ulp_x = np.spacing(np.abs(po[:, :3]).max(axis=1).astype(np.float32)).astype(np.float64)
        vn = np.linalg.norm(vo[:, :3].astype(np.float64), axis=1)
        assert np.all(dv <= np.maximum(1e-5 * vn, 6.0 * ulp_x / dt))
    assert g.staleFrames() == 0


def test_bitwise_determinism(gpu):
    sc = scenes.dam_break(1)
    a = _skinned(sc, gpu.Solver, 0.1)
    b = _skinned(sc, gpu.Solver, 0.1)
    a.run(10)
    b.run(10)
    pa, va = _state(a)
    pb, vb = _state(b)
    assert np.array_equal(pa, pb) and np.array_equal(va, vb)


def test_graph_replay_equals_eager(gpu):
    sc = scenes.dam_break(1)
    a = _skinned(sc, gpu.Solver, 0.1)
    b = _skinned(sc, gpu.Solver, 0.1)
    b.setGraphs(True)
    for k in range(6):
        a.step()
        b.step()
        pa, va = _state(a)
        pb, vb = _state(b)
        assert np.array_equal(pa, pb) and np.array_equal(va, vb), f"frame {k + 1}"


def test_graph_replay_with_stale_frames(gpu):
    sc = scenes.dam_break(1)
    ref = _solver_for(sc, gpu.Solver)
    b = _skinned(sc, gpu.Solver, 1e-6)
    b.setGraphs(True)
    ref.run(3)
    b.run(3)
    assert b.staleFrames() == 3
    pr, vr = _state(ref)
    pb, vb = _state(b)
    assert np.array_equal(pr, pb) and np.array_equal(vr, vb)


def test_overflow_raises_like_the_oracle(gpu):
    from oracle import NeighborOverflow, OracleSolver

    sc = scenes.dam_break(1, literal_strengths=True)
    g = _skinned(sc, gpu.Solver, 0.1)
    o = _solver_for(sc, lambda: OracleSolver("f64"))
    with pytest.raises(NeighborOverflow):
        o.step()
    with pytest.raises(pbf_amd.NeighborOverflow):
        g.step()


def test_nonconvergence_raises_like_the_oracle(gpu):
    from oracle import CollisionNonConvergence, OracleSolver

    sc = scenes.small_block(24, 20, 16)
    sc.boxes = [((1.0, 0.2, 0.4), (0.1, 0.2, 0.2))]
    sc.planes = [((0.0, 0.02, 0.0), (0.0, 1.0, 0.0))]
    g = _skinned(sc, gpu.Solver, 0.1)
    o = _solver_for(sc, lambda: OracleSolver("f64"))
    with pytest.raises(CollisionNonConvergence):
        o.step()
    with pytest.raises(pbf_amd.CollisionNonConvergence):
        g.step()


def test_skin_setter_validates(gpu):
    sc = scenes.dam_break(1)
    s = _solver_for(sc, gpu.Solver)
    for bad in (-0.1, 1.5, float("nan"), float("inf")):
        with pytest.raises(ValueError):
            s.setNeighborSkin(bad)
    assert s.neighborSkin() == 0.0
    s.setNeighborSkin(0.25)
    assert abs(s.neighborSkin() - 0.25) < 1e-7
    s.step()
    s.setNeighborSkin(0.0)
    s.step()


def _batch_vs_alone(gpu, scs, skin, frames):
    batch = pbf_amd.PBFBatchSolver()
    batch.initialize([(sc.n, sc.minBounds, sc.maxBounds) for sc in scs], scs[0].params)
    batch.setNeighborSkin(skin)
    batch.setParticles(np.concatenate([s.positions for s in scs]),
                       np.concatenate([s.velocities for s in scs]))
    batch.run(frames)
    pb, vb = batch.copyPositionsToHost(), batch.copyVelocitiesToHost()
    stale, f = [], 0
    for sc in scs:
        s = _skinned(sc, gpu.Solver, skin)
        s.run(frames)
        stale.append(s.staleFrames())
        p, v = _state(s)
        assert np.array_equal(pb[f:f + sc.n], p) and np.array_equal(vb[f:f + sc.n], v)
        f += sc.n
    return batch.staleFrames(), stale


def test_batched_equals_scenes_alone(gpu):
    """Weightless settled blocks of several sizes and boxes: clean skin
    frames."""
    scs = []
    for k, (a, b, c) in enumerate([(6, 8, 7), (12, 9, 10), (16, 12, 12), (9, 14, 6)]):
        sc = scenes.small_block(a, b, c, seed=0x5EED0100 + k)
        sc.params.gravity = pbf_amd.float3(0.0, 0.0, 0.0)
        scs.append(settled(gpu, sc))
    for sc in scs[1:]:
        sc.params = scs[0].params
    bst, alone = _batch_vs_alone(gpu, scs, 0.3, 3)
    print(f"[batch skin calm] stale frames: batch {bst}, alone {alone}", flush=True)
    assert bst == 0 and sum(alone) == 0


def test_batched_every_frame_stale(gpu):
    """A skin of 1e-6 h: every frame of the batch and of every scene alone
    runs again with the reference's builds."""
    scs = [scenes.batch_scene(k, nmin=300, nmax=4000, substeps=1) for k in range(5)]
    bst, alone = _batch_vs_alone(gpu, scs, 1e-6, 2)
    assert bst == 2 and alone == [2] * len(scs)


def test_full_size_properties_config2(gpu):
    sc = scenes.dam_break(2)
    s = _skinned(sc, gpu.Solver, 0.1)
    s.run(3)
    p, v = _state(s)
    assert np.all(np.isfinite(p)) and np.all(np.isfinite(v))
    r = sc.params.particleRadius
    tol = 4e-6
    assert p[:, 0].min() >= r - tol and p[:, 0].max() <= sc.maxBounds[0] - r + tol
    assert p[:, 1].min() >= r - tol
    assert p[:, 2].min() >= r - tol and p[:, 2].max() <= sc.maxBounds[2] - r + tol
    assert np.array_equal(p[:, 3], sc.positions[:, 3])
    print(f"[config2 skin 0.1] stale frames {s.staleFrames()} of 3", flush=True)
