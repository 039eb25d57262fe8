"""The benchmark workloads at their full sizes on the GPU.

  * the 8,388,608-particle double dam break: one substep against the FP64
    oracle from a state four frames in (both started from the GPU's copied
    state), with the contact-tolerant bars of the one-step parity tests (the
    velocity norm bar raised to the FP32 oracle's own distance from FP64
    where that is above 1e-5); and
    five frames of the benchmark's three-substep frames checked for finite
    state inside the container, the grid of the GPU's own positions
    bit-exact against the oracle's;
  * the 1M dam break at one substep per step: one step against the FP64
    oracle from a state ten steps in, at the same bars;
  * the 256-scene batched workload (about 10.2M particles): one batched
    frame equals, bit for bit, each of a seeded subset of scenes stepped
    alone at its full particle count.
"""
from __future__ import annotations

import time

import numpy as np
import pytest

from pbf_amd import scenes

pytestmark = pytest.mark.gpu


def _in_container(p, sc, tol=4e-6):
    r = sc.params.particleRadius
    assert np.all(np.isfinite(p))
    assert p[:, 0].min() >= r - tol and p[:, 0].max() <= sc.maxBounds[0] - r + tol
    assert p[:, 1].min() >= r - tol
    assert p[:, 2].min() >= r - tol and p[:, 2].max() <= sc.maxBounds[2] - r + tol


def test_config3_one_substep_parity_with_fp64_oracle(gpu):
    from oracle import OracleSolver
    from test_parity_gpu import _compare_with_contacts, _contact_mask, _loose_misses, _solver_for

    sc = scenes.dam_break(3, substeps=1)
    t = time.perf_counter()
    g = _solver_for(sc, gpu.Solver)
    g.run(4)
    sc.positions, sc.velocities = g.copyPositionsToHost(), g.copyVelocitiesToHost()
    g = _solver_for(sc, gpu.Solver)
    g.step()
    pg, vg = g.copyPositionsToHost(), g.copyVelocitiesToHost()
    del g
    print(f"8M: GPU frames done in {time.perf_counter() - t:.1f} s; FP64 oracle step", flush=True)
    out = {}
    for prec in ("f64", "f32"):
        t = time.perf_counter()
        o = _solver_for(sc, lambda: OracleSolver(prec))
        o.step()
        out[prec] = (o.copyPositionsToHost(), o.copyVelocitiesToHost())
        del o
        print(f"8M: {prec} oracle step {time.perf_counter() - t:.1f} s", flush=True)
    po, vo = out["f64"]
    contact = _contact_mask(po, sc) | _contact_mask(pg, sc)

    def norm_rel(v, mask=None):
        d, r = v[:, :3].astype(np.float64) - vo[:, :3], vo[:, :3].astype(np.float64)
        if mask is not None:
            d, r = d[mask], r[mask]
        return np.linalg.norm(d) / np.linalg.norm(r)

    gpu_err, f32_err = norm_rel(vg), norm_rel(out["f32"][1])
    print(f"8M velocity norm error vs FP64: GPU {gpu_err:.3e} (away from contacts "
          f"{norm_rel(vg, ~contact):.3e}); FP32 oracle {f32_err:.3e} "
          f"(away {norm_rel(out['f32'][1], ~contact):.3e})", flush=True)
    # The per-particle bars as in the one-step parity tests; the norm bar is
    # 1e-5 or, where an FP32 evaluation of the reference's own arithmetic
    # (the FP32 oracle, same input) is farther than that from FP64, that
    # distance: measured 4.04e-5 for the FP32 oracle and 1.49e-5 for the GPU
    # (the state is the GPU's own frame 5, so it moves with the GPU's
    # rounding).
    # Contact branches: the FP32 oracle takes the other branch than FP64 at
    # a few contact particles (measured 1 on this state, 2 on an earlier
    # one, where the GPU flipped one of the two), each then off by the whole
    # branch; the GPU may flip as many.
    f32_miss, _ = _loose_misses(*out["f32"], po, vo, sc)
    _compare_with_contacts(pg, vg, po, vo, sc, "double dam break 8M, frame 5",
                           norm_bar=max(1e-5, f32_err), flips_allowed=int(f32_miss.sum()))


def test_config2_one_substep_parity_with_fp64_oracle(gpu):
    """The 1M dam break (the headline workload) at one substep per step, as
    the 8M test: one step from a state ten steps in, against the FP64
    oracle from the same copied state, at the one-step bars
    (contact-tolerant, the velocity norm bar at max(1e-5, the FP32 oracle's
    own distance from FP64))."""
    from oracle import OracleSolver
    from test_parity_gpu import _compare_with_contacts, _loose_misses, _solver_for

    sc = scenes.dam_break(2, substeps=1)
    g = _solver_for(sc, gpu.Solver)
    g.run(10)
    sc.positions, sc.velocities = g.copyPositionsToHost(), g.copyVelocitiesToHost()
    g = _solver_for(sc, gpu.Solver)
    g.step()
    pg, vg = g.copyPositionsToHost(), g.copyVelocitiesToHost()
    del g
    out = {}
    for prec in ("f64", "f32"):
        o = _solver_for(sc, lambda: OracleSolver(prec))
        o.step()
        out[prec] = (o.copyPositionsToHost(), o.copyVelocitiesToHost())
        del o
    po, vo = out["f64"]

    def norm_rel(v):
        d = v[:, :3].astype(np.float64) - vo[:, :3]
        return np.linalg.norm(d) / np.linalg.norm(vo[:, :3].astype(np.float64))

    f32_err = norm_rel(out["f32"][1])
    print(f"1M step 11 velocity norm error vs FP64: GPU {norm_rel(vg):.3e}, "
          f"FP32 oracle {f32_err:.3e}", flush=True)
    f32_miss, _ = _loose_misses(*out["f32"], po, vo, sc)
    _compare_with_contacts(pg, vg, po, vo, sc, "dam break 1M, step 11",
                           norm_bar=max(1e-5, f32_err), flips_allowed=int(f32_miss.sum()))


def test_config3_frames_stay_inside_and_grid_is_bitexact(gpu, oracle64):
    from test_parity_gpu import _solver_for

    sc = scenes.dam_break(3)
    s = _solver_for(sc, gpu.Solver)
    s.run(5)
    p, v = s.copyPositionsToHost(), s.copyVelocitiesToHost()
    del s
    _in_container(p, sc)
    assert np.all(np.isfinite(v))
    assert np.array_equal(p[:, 3], sc.positions[:, 3])
    print(f"8M: 5 frames of {sc.params.substeps} substeps, max |v| "
          f"{np.linalg.norm(v[:, :3], axis=1).max():.3f}", flush=True)
    g = gpu.grid(sc.n, sc.minBounds, sc.maxBounds, sc.params.smoothingRadius)
    sk, si, cst, cen = g.build(p)
    osk, osi, ocs, oce, _, _ = oracle64.grid_build(p, sc.minBounds, sc.maxBounds,
                                                   sc.params.smoothingRadius)
    assert np.array_equal(sk, osk) and np.array_equal(si, osi)
    assert np.array_equal(cst, ocs) and np.array_equal(cen, oce)


def test_config5_batched_frame_equals_scenes_alone(gpu):
    import pbf_amd

    scs = scenes.batch_scenes(256)
    params = scs[0].params
    n = sum(s.n for s in scs)
    b = pbf_amd.PBFBatchSolver()
    b.initialize([(s.n, s.minBounds, s.maxBounds) for s in scs], params)
    b.setParticles(np.concatenate([s.positions for s in scs]),
                   np.concatenate([s.velocities for s in scs]))
    b.step()
    bp, bv = b.copyPositionsToHost(), b.copyVelocitiesToHost()
    del b
    print(f"256 scenes, {n} particles: one batched frame", flush=True)
    first = np.cumsum([0] + [s.n for s in scs])
    rng = np.random.default_rng(0xBA7C)
    picks = sorted(rng.choice(len(scs), 16, replace=False).tolist())
    # the largest and the smallest scene as well
    sizes = [s.n for s in scs]
    picks = sorted(set(picks) | {int(np.argmax(sizes)), int(np.argmin(sizes))})
    for k in picks:
        s = scs[k]
        one = gpu.Solver()
        one.initialize(s.n, s.minBounds, s.maxBounds, params)
        one.setParticles(s.positions, s.velocities)
        one.step()
        sl = slice(first[k], first[k + 1])
        assert np.array_equal(bp[sl].view(np.uint32), one.copyPositionsToHost().view(np.uint32)), \
            f"scene {k} ({s.n} particles): positions"
        assert np.array_equal(bv[sl].view(np.uint32), one.copyVelocitiesToHost().view(np.uint32)), \
            f"scene {k} ({s.n} particles): velocities"
    print(f"scenes {picks} equal alone and batched", flush=True)
