"""The library's adjoint (pbf_adjoint_*) against FP64 central differences.

The differences are taken of a genuinely FP64 forward (adjoint_ref.run_full:
every state in FP64, neighbour pairs rebuilt at every grid build, the
container projection and contact velocities live), so they see the true
derivative of the step; the adjoint holds the forward's lists and discrete
decisions fixed, which is the same derivative wherever no decision sits
within the difference step.  Loss: the mean squared distance of the final
positions to the initial block moved by +0.2 in x.
"""
from __future__ import annotations

import numpy as np
import pytest

import adjoint_ref as A
from pbf_amd import scenes

pytestmark = pytest.mark.gpu


def _case(nx, substeps, seed=5, iterations=4):
    sc = scenes.small_block(nx, nx, nx, substeps=substeps)
    P = sc.params
    P.xsphViscosity = 1e-4
    P.vorticityStrength = 1e-4
    P.solverIterations = iterations
    rng = np.random.default_rng(seed)
    v0 = sc.velocities.copy()
    v0[:, :3] = rng.normal(0.0, 0.1, (sc.n, 3)).astype(np.float32)
    Q = A.Params(dt=P.dt / P.substeps, h=P.smoothingRadius, mass=P.particleMass,
                 rho0=P.restDensity, eps_lambda=P.lambdaEpsilon,
                 gravity=(P.gravity.x, P.gravity.y, P.gravity.z), scorr_k=P.scorrK,
                 scorr_n=P.scorrN, scorr_dq=P.scorrDeltaQ, radius=P.particleRadius,
                 box_min=sc.minBounds, box_max=sc.maxBounds, iterations=iterations)
    return sc, P, Q, v0, rng


@pytest.mark.parametrize("nx,substeps,frames,every,iterations",
                         [(10, 1, 4, 2, 4), (12, 2, 5, 2, 4), (10, 1, 3, 2, 5), (10, 1, 3, 3, 6)])
def test_adjoint_matches_fp64_central_differences(gpu, nx, substeps, frames, every, iterations):
    import pbf_amd

    sc, P, Q, v0, rng = _case(nx, substeps, iterations=iterations)
    x0 = sc.positions
    target = x0[:, :3].astype(np.float64) + np.array([0.2, 0.0, 0.0])

    def loss(x):
        d = x - target
        return np.mean(np.sum(d * d, 1))

    s = gpu.Solver()
    s.initialize(sc.n, sc.minBounds, sc.maxBounds, P)
    s.setParticles(x0, v0)
    adj = pbf_amd.PBFAdjoint(s)
    adj.forward(frames, every)
    xT = s.copyPositionsToHost()
    xbar = np.zeros_like(xT)
    xbar[:, :3] = (2.0 * (xT[:, :3] - target) / sc.n).astype(np.float32)
    _, v0b, cb, eb = adj.backward(xbar, None)
    # backward leaves the solver where forward ended
    assert np.array_equal(s.copyPositionsToHost(), xT)

    x = x0[:, :3].astype(np.float64)
    v = v0[:, :3].astype(np.float64)
    S = frames * substeps
    c, e = P.xsphViscosity, P.vorticityStrength

    def L64(v_, c_=c, e_=e):
        return loss(A.run_full(x, v_, S, Q, c_, e_)[0])

    assert abs(L64(v) - loss(xT[:, :3].astype(np.float64))) < 1e-6 * L64(v)
    fd, ad = [], []
    for k in rng.choice(sc.n, 5, replace=False):
        for d in range(3):
            vp, vm = v.copy(), v.copy()
            vp[k, d] += 1e-4
            vm[k, d] -= 1e-4
            fd.append((L64(vp) - L64(vm)) / 2e-4)
            ad.append(float(v0b[k, d]))
    fd, ad = np.array(fd), np.array(ad)
    err = np.linalg.norm(fd - ad) / np.linalg.norm(fd)
    print(f"[{sc.n} particles, {S} substeps, {iterations} iterations] v0: {len(fd)} coordinates, "
          f"relative error {err:.2e}", flush=True)
    assert err <= 1e-3
    # The parameters move every particle, so a difference step of 1e-3 or
    # 1e-2 of their value already carries some contact decisions across
    # their switch (the loss of the two-substep case is piecewise smooth in
    # c: differences of -5.30, -5.90, -5.265081, -5.265085 at steps of
    # 1e-2 .. 1e-5 relative, against the adjoint's -5.265085); 1e-5 stays
    # inside one smooth piece.
    for name, base, grad in (("xsphViscosity", c, cb), ("vorticityStrength", e, eb)):
        hh = 1e-5 * base
        f = ((L64(v, c_=c + hh) - L64(v, c_=c - hh)) if name == "xsphViscosity"
             else (L64(v, e_=e + hh) - L64(v, e_=e - hh))) / (2 * hh)
        print(f"  d/d{name}: fd {f:+.6e} adjoint {grad:+.6e}", flush=True)
        assert abs(f - grad) <= 1e-3 * abs(f)


# A sphere right of the block, a rounded box inside it and a tilted plane
# under its left part: the block spreads into the sphere, its particles are
# pushed out of the box, and the plane lifts the bottom layer, so every
# collider projects positions and acts on contact velocities within the
# rollout.  Kept apart from each other and from the walls by more than a
# particle diameter.
GPU_COLLIDERS = dict(
    spheres=(((0.55, 0.15, 0.25), 0.08),),
    boxes=(((0.25, 0.3, 0.25), (0.05, 0.05, 0.05)),),
    planes=(((0.0, 0.04, 0.0), (0.05, 1.0, 0.02)),),
)


# Without XSPH (c = 0) the velocities of particles sliding along the
# tilted plane keep a normal component of rounding size, so whether their
# contact acts is decided by rounding: over 3-4 frames the FP32 forward and
# the FP64 restatement then take different branches for a few particles
# (the FP32 oracle does too).  The c = 0 cases therefore run 2 frames (two
# checkpoint segments), on a seed whose FP32 oracle stays on the FP64
# branches (positions within 5e-7).
@pytest.mark.parametrize("c,e,seed,frames,every", [
    (1e-4, 1e-4, 9, 4, 2), (1e-4, 0.0, 9, 4, 2), (0.0, 1e-4, 10, 2, 1), (0.0, 0.0, 10, 2, 1)])
def test_adjoint_with_colliders_matches_fp64_central_differences(gpu, c, e, seed, frames, every):
    """Spheres, rounded boxes and planes in the collision solve and the
    contacts (each projects positions and acts on contact velocities within
    the rollout); XSPH and vorticity confinement each on or off (c = 0 or
    e = 0: the passes are skipped and the derivatives are those at 0, which
    central differences of the FP64 forward see: its terms are linear in c
    and e there)."""
    from dataclasses import replace

    import pbf_amd

    sc, P, Q, v0, rng = _case(10, 1, seed=seed)
    P.xsphViscosity, P.vorticityStrength = c, e
    P.collisionRestitution, P.collisionFriction = 0.2, 0.1
    Q = replace(Q, restitution=0.2, friction=0.1, **GPU_COLLIDERS)
    x0 = sc.positions
    target = x0[:, :3].astype(np.float64) + np.array([0.2, 0.0, 0.0])

    def loss(x):
        d = x - target
        return np.mean(np.sum(d * d, 1))

    s = gpu.Solver()
    s.initialize(sc.n, sc.minBounds, sc.maxBounds, P)
    s.setSpheres(list(GPU_COLLIDERS["spheres"]))
    s.setBoxes(list(GPU_COLLIDERS["boxes"]))
    s.setPlanes(list(GPU_COLLIDERS["planes"]))
    s.setParticles(x0, v0)
    adj = pbf_amd.PBFAdjoint(s)
    adj.forward(frames, every)
    xT = s.copyPositionsToHost()
    xbar = np.zeros_like(xT)
    xbar[:, :3] = (2.0 * (xT[:, :3] - target) / sc.n).astype(np.float32)
    _, v0b, cb, eb = adj.backward(xbar, None)

    x = x0[:, :3].astype(np.float64)
    v = v0[:, :3].astype(np.float64)
    S = frames * P.substeps
    A_ = __import__("adjoint_ref")
    kinds, acted, xx, vv = set(), set(), x, v
    for _ in range(S):
        xx, vv, aux = A_.forward(xx, vv, lambda p: A_.pairs_within(p, Q.h), Q, c, e,
                                 keep_aux=True)
        kinds |= {ev["kind"] for it in aux["it"] for ev in it["events"]}
        acted |= {cd["kind"] for cd in aux["contacts"] if cd["act"].any()}
    assert {"sphere", "box", "plane"} <= kinds and {"sphere", "box", "plane"} <= acted

    def L64(v_, c_=c, e_=e):
        return loss(A_.run_full(x, v_, S, Q, c_, e_)[0])

    base = L64(v)
    assert abs(base - loss(xT[:, :3].astype(np.float64))) < 1e-5 * base
    fd, ad = [], []
    for k in rng.choice(sc.n, 5, replace=False):
        for d in range(3):
            vp, vm = v.copy(), v.copy()
            vp[k, d] += 1e-4
            vm[k, d] -= 1e-4
            fd.append((L64(vp) - L64(vm)) / 2e-4)
            ad.append(float(v0b[k, d]))
    fd, ad = np.array(fd), np.array(ad)
    err = np.linalg.norm(fd - ad) / np.linalg.norm(fd)
    grads = {}
    for name, base_val, grad in (("xsphViscosity", c, cb), ("vorticityStrength", e, eb)):
        hh = 1e-5 * (base_val if base_val > 0 else 1e-4)
        f = ((L64(v, c_=c + hh) - L64(v, c_=c - hh)) if name == "xsphViscosity"
             else (L64(v, e_=e + hh) - L64(v, e_=e - hh))) / (2 * hh)
        grads[name] = (f, grad)
    print(f"[colliders, c {c}, e {e}, {frames} frames] v0: "
          f"{len(fd)} coordinates, relative error {err:.2e}; "
          + "; ".join(f"d/d{k} fd {f:+.6e} adjoint {g:+.6e}" for k, (f, g) in grads.items()),
          flush=True)
    assert err <= 1e-3
    for name, (f, g) in grads.items():
        assert abs(f - g) <= 1e-3 * abs(f), name


def test_torch_rollout_gradients_equal_the_adjoint(gpu):
    """torch.autograd through pbf_amd.autograd.rollout hands back exactly
    the adjoint's gradients for a loss built from torch operations."""
    import torch

    import pbf_amd
    from pbf_amd.autograd import rollout

    sc, P, Q, v0, rng = _case(8, 1)
    target = torch.from_numpy(sc.positions[:, :3].astype(np.float64)) + torch.tensor([0.2, 0.0, 0.0])
    v = torch.from_numpy(v0[:, :3].astype(np.float64)).requires_grad_(True)
    c = torch.tensor(P.xsphViscosity, dtype=torch.float64, requires_grad=True)
    e = torch.tensor(P.vorticityStrength, dtype=torch.float64, requires_grad=True)
    xT = rollout(v, c, e, sc, 3, 2)
    loss = ((xT - target) ** 2).sum(1).mean()
    loss.backward()

    s = gpu.Solver()
    s.initialize(sc.n, sc.minBounds, sc.maxBounds, P)
    s.setParticles(sc.positions, v0)
    adj = pbf_amd.PBFAdjoint(s)
    adj.forward(3, 2)
    x = s.copyPositionsToHost()
    xbar = np.zeros_like(x)
    xbar[:, :3] = (2.0 * (x[:, :3] - target.numpy()) / sc.n).astype(np.float32)
    x0b, v0b, cb, eb = adj.backward(xbar, None)
    # The position and velocity adjoints are sums in a fixed order (each slot
    # pulls its pair terms; no atomics), so they are bitwise repeatable; the
    # two parameter adjoints add per-wave FP64 partials atomically, equal to
    # FP64 rounding.
    assert np.array_equal(v.grad.numpy(), v0b[:, :3].astype(np.float64))
    assert abs(c.grad.item() - cb) <= 1e-12 * abs(cb)
    assert abs(e.grad.item() - eb) <= 1e-12 * abs(eb)
    x0b2, v0b2, cb2, eb2 = adj.backward(xbar, None)
    assert np.array_equal(x0b.view(np.uint32), x0b2.view(np.uint32))
    assert np.array_equal(v0b.view(np.uint32), v0b2.view(np.uint32))
    assert abs(cb2 - cb) <= 1e-12 * abs(cb) and abs(eb2 - eb) <= 1e-12 * abs(eb)


def test_taped_forward_equals_plain_run(gpu):
    """A plain run moves only p* at the grid builds inside the constraint
    loop and the other records once per substep; the adjoint's taped forward
    (one segment: every substep taped) moves every record at every build.
    Both must give the same state, bit for bit, on a dam-break block with
    every effect on and several substeps."""
    import pbf_amd

    sc = scenes.dam_break(1, substeps=2)
    P = sc.params
    runs = []
    for taped in (False, True):
        s = gpu.Solver()
        s.initialize(sc.n, sc.minBounds, sc.maxBounds, P)
        s.setParticles(sc.positions, sc.velocities)
        if taped:
            pbf_amd.PBFAdjoint(s).forward(3, 3)
        else:
            s.run(3)
        runs.append((s.copyPositionsToHost(), s.copyVelocitiesToHost()))
    (xa, va), (xb, vb) = runs
    assert np.array_equal(xa.view(np.uint32), xb.view(np.uint32))
    assert np.array_equal(va.view(np.uint32), vb.view(np.uint32))


def _differentiable():
    import pbf_amd

    sc, target = scenes.differentiable_block()
    s = pbf_amd.PBFSolver()
    s.initialize(sc.n, sc.minBounds, sc.maxBounds, sc.params)
    s.setParticles(sc.positions, sc.velocities)
    return sc, target, s


def _loss_bar(x, target, n):
    xbar = np.zeros_like(x)
    xbar[:, :3] = (2.0 * (x[:, :3] - target) / n).astype(np.float32)
    return xbar


def test_differentiable_workload_matches_fp64_differences(gpu):
    """The 65,536-particle differentiable workload over 5 frames with
    checkpoints every 2 (backward replays two segments from their
    checkpoints): the gradients with respect to 32 seeded initial-velocity
    coordinates, the XSPH viscosity and the vorticity strength against FP64
    central differences of the same rollout (tests/golden/adjoint_fd_64k_f5.npz,
    written by tests/golden/make_adjoint_fd.py)."""
    from pathlib import Path

    import pbf_amd

    fx = np.load(Path(__file__).resolve().parent / "golden" / "adjoint_fd_64k_f5.npz")
    frames = int(fx["frames"])
    sc, target, s = _differentiable()
    adj = pbf_amd.PBFAdjoint(s)
    adj.forward(frames, 2)
    x = s.copyPositionsToHost()
    loss = float(np.mean(np.sum((x[:, :3] - target) ** 2, 1)))
    _, v0b, cb, eb = adj.backward(_loss_bar(x, target, sc.n), None)
    ad = v0b[fx["particles"], fx["axes"]].astype(np.float64)
    err = np.linalg.norm(ad - fx["fd_v"]) / np.linalg.norm(fx["fd_v"])
    worst = np.argsort(-np.abs(ad - fx["fd_v"]))[:4]
    print("  worst coordinates (adjoint, fd, fd at the finer step): " + "; ".join(
        f"{ad[k]:+.5e} {fx['fd_v'][k]:+.5e} {fx['fd_v_fine'][k]:+.5e}" for k in worst),
        flush=True)
    print(f"[65536 particles, {frames} frames] loss {loss:.9e} (FP64 {float(fx['loss']):.9e}); "
          f"32 velocity coordinates: relative error {err:.2e}; d/dc {cb:+.6e} "
          f"(fd {float(fx['fd_c']):+.6e}); d/de {eb:+.6e} (fd {float(fx['fd_e']):+.6e})",
          flush=True)
    assert abs(loss - float(fx["loss"])) <= 1e-5 * float(fx["loss"])
    assert err <= 1e-3
    assert abs(cb - float(fx["fd_c"])) <= 1e-3 * abs(float(fx["fd_c"]))
    assert abs(eb - float(fx["fd_e"])) <= 1e-3 * abs(float(fx["fd_e"]))


def test_differentiable_workload_60_frames_properties(gpu):
    """The full 60-frame rollout with checkpoints every 10 (five segments
    replayed).  Its gradient leaves FP32's range (it grows by orders of
    magnitude per frame past 10 frames), so the checks are properties: the
    taped forward ends where a plain run ends, bit for bit; a zero loss
    gradient gives zero gradients; backward is bitwise repeatable and
    leaves the solver where forward ended."""
    import pbf_amd

    sc, target, plain = _differentiable()
    plain.run(60)
    xp, vp = plain.copyPositionsToHost(), plain.copyVelocitiesToHost()
    del plain
    _, _, s = _differentiable()
    adj = pbf_amd.PBFAdjoint(s)
    adj.forward(60, 10)
    x, v = s.copyPositionsToHost(), s.copyVelocitiesToHost()
    assert np.array_equal(x.view(np.uint32), xp.view(np.uint32))
    assert np.array_equal(v.view(np.uint32), vp.view(np.uint32))
    x0b, v0b, cb, eb = adj.backward(np.zeros_like(x), None)
    assert not np.any(x0b) and not np.any(v0b) and cb == 0.0 and eb == 0.0
    xbar = _loss_bar(x, target, sc.n)
    g1 = adj.backward(xbar, None)
    g2 = adj.backward(xbar, None)
    assert np.array_equal(g1[0].view(np.uint32), g2[0].view(np.uint32))
    assert np.array_equal(g1[1].view(np.uint32), g2[1].view(np.uint32))
    assert np.array_equal(s.copyPositionsToHost().view(np.uint32), x.view(np.uint32))
    print(f"65536 particles, 60 frames: |dL/dv0| {np.linalg.norm(g1[1][:, :3]):.3e}, "
          f"dL/dc {g1[2]:.3e}", flush=True)
