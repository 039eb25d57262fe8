"""PBFSolver known answers (test/test_pbf_solver.cu) on the oracle and the GPU,
and the solver's API contract (validation, exception types, empty sets)."""
from __future__ import annotations

import math

import numpy as np
import pytest

from backends import pair_density, self_density, valid_params

MIN, MAX = (-2.0, -2.0, -2.0), (12.0, 12.0, 12.0)


def f4(*rows):
    return np.array(rows, dtype=np.float32)


def near(a, b, tol):
    return np.all(np.abs(np.asarray(a, np.float32) - np.asarray(b, np.float32)) <= tol)


def make(backend, capacity, params, mn=MIN, mx=MAX):
    s = backend.Solver()
    s.initialize(capacity, mn, mx, params)
    return s


def collision_params(**kw):
    """collisionParams() (test/test_pbf_solver.cu:85-97)."""
    p = valid_params(dt=0.1, particleRadius=0.25, smoothingRadius=0.5, solverIterations=4,
                     substeps=1, gravity=(0.0, 0.0, 0.0), collisionRestitution=0.0,
                     collisionFriction=0.0)
    p.restDensity = self_density(p.particleMass, p.smoothingRadius)
    for k, v in kw.items():
        if k == "gravity":
            from pbf_amd import float3
            v = float3(*v)
        setattr(p, k, v)
    return p


def integrate_reference(x, v, dt, g, substeps):
    x = np.array(x, np.float32)
    v = np.array(v, np.float32)
    sdt = np.float32(dt) / np.float32(substeps)
    g = np.asarray(g, np.float32)
    for _ in range(substeps):
        v[:3] += g * sdt
        x[:3] += v[:3] * sdt
    return x, v


def test_single_particle_ballistic(backend):
    p = valid_params(dt=0.2, substeps=4, gravity=(1.0, -10.0, 2.0))
    s = make(backend, 1, p)
    x0, v0 = f4([1, 5, 2, 17]), f4([2, 3, -1, 23])
    s.setParticles(x0, v0)
    s.step()
    ex, ev = integrate_reference(x0[0], v0[0], 0.2, (1.0, -10.0, 2.0), 4)
    assert near(s.copyPositionsToHost()[0], ex, 1e-5) and near(s.copyVelocitiesToHost()[0], ev, 1e-5)


def test_pair_separates_symmetrically(backend):
    p = valid_params(dt=0.1, gravity=(0.0, 0.0, 0.0), solverIterations=3, substeps=1)
    p.restDensity = self_density(p.particleMass, p.smoothingRadius)
    s = make(backend, 2, p)
    x0 = f4([4.75, 5, 5, 3], [5.25, 5, 5, 4])
    v0 = f4([0, 0, 0, 7], [0, 0, 0, 8])
    s.setParticles(x0, v0)
    s.step()
    x, v = s.copyPositionsToHost(), s.copyVelocitiesToHost()
    assert x[0, 0] < 4.75 and x[1, 0] > 5.25 and abs(x[0, 0] + x[1, 0] - 10.0) <= 1e-5
    assert np.all(np.abs(x[:, 1:3] - 5.0) <= 1e-6) and np.array_equal(x[:, 3], x0[:, 3])
    assert v[0, 0] < 0 and v[1, 0] > 0 and abs(v[0, 0] + v[1, 0]) <= 1e-4
    assert np.all(np.abs(v[:, 1:3]) <= 1e-6) and np.array_equal(v[:, 3], v0[:, 3])


def test_reordering_preserves_api_order(backend):
    p = valid_params(gravity=(0.0, 0.0, 0.0))
    p.restDensity = self_density(p.particleMass, p.smoothingRadius)
    s = make(backend, 5, p)
    x0 = f4([9, 5, 5, 11], [7, 5, 5, 12], [5, 5, 5, 13], [3, 5, 5, 14], [1, 5, 5, 15])
    v0 = f4([0.1, 0, 0, 21], [0.2, 0, 0, 22], [0.3, 0, 0, 23], [0.4, 0, 0, 24], [0.5, 0, 0, 25])
    s.setParticles(x0, v0)
    s.step()
    ex = x0.copy()
    ex[:, 0] = x0[:, 0] + np.float32(0.1) * v0[:, 0]
    assert near(s.copyPositionsToHost(), ex, 1e-5) and near(s.copyVelocitiesToHost(), v0, 1e-5)


def test_run_frame_count_independent_of_iterations(backend):
    p = valid_params(dt=0.1, gravity=(0.0, 0.0, 0.0), solverIterations=3, substeps=1)
    s = make(backend, 1, p)
    x0, v0 = f4([1, 2, 3, 5]), f4([2, -1, 0.5, 6])
    s.setParticles(x0, v0)
    s.run(2)
    ex = x0.copy()
    ex[0, :3] += v0[0, :3] * np.float32(0.2)
    assert near(s.copyPositionsToHost(), ex, 1e-5) and near(s.copyVelocitiesToHost(), v0, 1e-5)


def test_run_without_count_is_one_frame(backend):
    p = valid_params(dt=0.1, solverIterations=4, gravity=(0.0, 0.0, 0.0))
    s = make(backend, 1, p)
    s.setParticles(f4([1, 2, 3, 4]), f4([2, -1, 0.5, 5]))
    s.run()
    assert near(s.copyPositionsToHost(), f4([1.2, 1.9, 3.05, 4]), 1e-5)
    assert near(s.copyVelocitiesToHost(), f4([2, -1, 0.5, 5]), 1e-5)


def test_many_particles_across_blocks(backend):
    n, width, spacing = 1025, 11, 0.75
    p = valid_params(dt=0.02, smoothingRadius=0.5, particleMass=0.25, lambdaEpsilon=0.1,
                     solverIterations=2, substeps=2, gravity=(0.5, -9.8, -0.25))
    p.restDensity = self_density(p.particleMass, p.smoothingRadius)
    s = make(backend, n + 100, p)
    i = np.arange(n)
    x, y, z = i % width, (i // width) % width, i // (width * width)
    pos = np.stack([0.25 + spacing * x, 0.25 + spacing * y, 0.25 + spacing * z, 1.0 + i % 5], 1)
    vel = np.stack([0.1 + 1e-5 * x, 0.2 + 1e-5 * y, -0.3 + 1e-5 * z, 10.0 + i % 7], 1)
    pos = pos.astype(np.float32)
    vel = vel.astype(np.float32)
    s.setParticles(pos, vel)
    s.step()
    out_p, out_v = s.copyPositionsToHost(), s.copyVelocitiesToHost()
    for k in range(n):
        ex, ev = integrate_reference(pos[k], vel[k], 0.02, (0.5, -9.8, -0.25), 2)
        assert near(out_p[k], ex, 2e-4) and near(out_v[k], ev, 2e-4)
    assert np.all(np.isfinite(out_p)) and np.all(np.isfinite(out_v))


def test_container_restitution_and_friction(backend):
    p = collision_params(particleRadius=0.5, collisionRestitution=0.5, collisionFriction=0.25)
    s = make(backend, 1, p)
    s.setParticles(f4([-1.4, 5, 5, 7]), f4([-2, 4, 0, 8]))
    s.step()
    assert near(s.copyPositionsToHost(), f4([-1.5, 5.4, 5, 7]), 1e-5)
    assert near(s.copyVelocitiesToHost(), f4([1, 3, 0, 8]), 1e-5)


def test_restitution_uses_incoming_velocity_at_contact(backend):
    p = collision_params(particleRadius=0.5, collisionRestitution=0.5)
    s = make(backend, 1, p)
    s.setParticles(f4([-1.5, 5, 5, 1]), f4([-2, 0, 0, 2]))
    s.step()
    assert abs(s.copyPositionsToHost()[0, 0] + 1.5) <= 1e-6
    assert abs(s.copyVelocitiesToHost()[0, 0] - 1.0) <= 1e-6


def test_existing_inelastic_contact_applies_friction(backend):
    p = collision_params(gravity=(0.0, -9.8, 0.0), collisionFriction=0.25)
    s = make(backend, 1, p)
    s.setParticles(f4([5, -1.75, 5, 1]), f4([2, 0, 0, 0]))
    s.step()
    v = s.copyVelocitiesToHost()
    assert abs(v[0, 0] - 1.5) <= 1e-5 and abs(v[0, 1]) <= 1e-5


def test_neighbor_overflow_raises(backend):
    p = valid_params(gravity=(0.0, 0.0, 0.0), particleRadius=0.01)
    s = make(backend, 258, p)
    s.setParticles(np.tile(f4([5, 5, 5, 1]), (258, 1)), np.zeros((258, 4), np.float32))
    with pytest.raises(backend.E.NeighborOverflow):
        s.step()


@pytest.mark.parametrize("kind", ["sphere", "box", "plane"])
def test_forwards_each_collider(backend, kind):
    s = make(backend, 1, collision_params())
    if kind == "sphere":
        s.setSpheres([((3.0, 5.0, 5.0), 0.5)])
        x0, v0, ex = f4([4, 5, 5, 1]), f4([-4, 0, 0, 2]), f4([3.75, 5, 5, 1])
    elif kind == "box":
        s.setBoxes([((3.0, 5.0, 5.0), (0.5, 0.5, 0.5))])
        x0, v0, ex = f4([4, 5, 5, 3]), f4([-4, 0, 0, 4]), f4([3.75, 5, 5, 3])
    else:
        s.setPlanes([((0.0, 2.0, 0.0), (0.0, 2.0, 0.0))])
        x0, v0, ex = f4([5, 2.4, 5, 5]), f4([0, -3, 0, 6]), f4([5, 2.25, 5, 5])
    s.setParticles(x0, v0)
    s.step()
    ev = v0.copy()
    ev[0, :3] = 0.0
    assert near(s.copyPositionsToHost(), ex, 1e-5) and near(s.copyVelocitiesToHost(), ev, 1e-5)


def test_cleared_sphere_no_longer_acts(backend):
    s = make(backend, 1, collision_params())
    s.setSpheres([((3.0, 5.0, 5.0), 0.5)])
    s.setSpheres([])
    s.setParticles(f4([4, 5, 5, 1]), f4([-4, 0, 0, 2]))
    s.step()
    assert near(s.copyPositionsToHost(), f4([3.6, 5, 5, 1]), 1e-5)
    assert near(s.copyVelocitiesToHost(), f4([-4, 0, 0, 2]), 1e-5)


def test_combined_colliders_one_step(backend):
    s = make(backend, 3, collision_params())
    s.setSpheres([((2.0, 6.0, 2.0), 0.5)])
    s.setBoxes([((6.0, 6.0, 2.0), (0.5, 0.5, 0.5))])
    s.setPlanes([((0.0, 2.0, 0.0), (0.0, 1.0, 0.0))])
    s.setParticles(f4([3, 6, 2, 1], [7, 6, 2, 2], [10, 2.4, 2, 3]),
                   f4([-4, 0, 0, 4], [-4, 0, 0, 5], [0, -3, 0, 6]))
    s.step()
    assert near(s.copyPositionsToHost(), f4([2.75, 6, 2, 1], [6.75, 6, 2, 2], [10, 2.25, 2, 3]),
                1e-5)
    assert near(s.copyVelocitiesToHost(), f4([0, 0, 0, 4], [0, 0, 0, 5], [0, 0, 0, 6]), 1e-5)


def test_xsph_uses_post_solve_neighbors(backend):
    p = valid_params(dt=0.1, gravity=(0.0, 0.0, 0.0), solverIterations=1, substeps=1,
                     particleRadius=0.25, smoothingRadius=1.0, xsphViscosity=1.0)
    p.restDensity = self_density(p.particleMass, 1.0) + pair_density(p.particleMass, 0.8, 1.0)
    s = make(backend, 2, p)
    s.setSpheres([((5.0, 5.0, 5.0), 0.5)])
    s.setParticles(f4([4.6, 5, 5, 1], [5.4, 5, 5, 2]), f4([0, 0, 0, 3], [0, 0, 0, 4]))
    s.step()
    x, v = s.copyPositionsToHost(), s.copyVelocitiesToHost()
    assert abs(x[0, 0] - 4.25) <= 1e-5 and abs(x[1, 0] - 5.75) <= 1e-5
    assert abs(v[0, 0] + 3.5) <= 1e-4 and abs(v[1, 0] - 3.5) <= 1e-4
    assert v[0, 3] == 3.0 and v[1, 3] == 4.0


def test_fixed_scenario_repeats_exactly(backend):
    p = collision_params(dt=0.01, substeps=2, solverIterations=3, xsphViscosity=0.03,
                         vorticityStrength=0.02, scorrK=0.001, scorrN=4, scorrDeltaQ=0.15,
                         gravity=(0.0, -1.0, 0.0))
    x0 = f4([4.7, 6, 5, 1], [5.1, 6, 5, 2], [4.9, 6.35, 5, 3], [5.3, 6.35, 5, 4])
    v0 = f4([0.1, 0, 0, 5], [-0.1, 0, 0, 6], [0, 0.1, 0, 7], [0, -0.1, 0, 8])
    a, b = make(backend, 4, p), make(backend, 4, p)
    a.setParticles(x0, v0)
    b.setParticles(x0, v0)
    a.run(10)
    b.run(10)
    assert np.all(np.isfinite(a.copyPositionsToHost()))
    assert np.array_equal(a.copyPositionsToHost(), b.copyPositionsToHost())
    assert np.array_equal(a.copyVelocitiesToHost(), b.copyVelocitiesToHost())


def test_stability_dense_fluid_with_sphere_and_box(backend):
    """1,000 particles x 1,200 frames with a sphere and a box
    (test/test_correctness.cu:183-267): finite, inside every constraint,
    contacts made, settled mean height < 1, RMS speed < 1."""
    side, frames = 10, 1200
    n = side ** 3
    p = valid_params(dt=1.0 / 60.0, substeps=2, solverIterations=4, restDensity=1000.0,
                     particleMass=float(np.float32(1000) * np.float32(0.12) ** 3),
                     particleRadius=0.05, smoothingRadius=0.24, lambdaEpsilon=100.0,
                     gravity=(0.0, -9.81, 0.0), collisionFriction=0.03, scorrK=0.0001, scorrN=4,
                     scorrDeltaQ=0.072, xsphViscosity=0.00001, vorticityStrength=0.00001)
    sphere = ((0.8, 0.45, 1.0), 0.22)
    box = ((1.9, 0.35, 1.0), (0.22, 0.35, 0.4))
    s = make(backend, n, p, (0, 0, 0), (3, 3, 2))
    s.setSpheres([sphere])
    s.setBoxes([box])
    g = np.arange(side)
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([0.35 + x.ravel() * 0.12, 1.1 + y.ravel() * 0.12, 0.4 + z.ravel() * 0.12,
                    np.ones(n)], 1).astype(np.float32)
    vel = np.tile(f4([1, 0, 0.15, 0]), (n, 1))
    s.setParticles(pos, vel)
    tol = 4e-6
    contacts_s = contacts_b = 0
    for frame in range(frames):
        s.step()
        q, v = s.copyPositionsToHost().astype(np.float64), s.copyVelocitiesToHost()
        assert np.all(np.isfinite(q)) and np.all(np.isfinite(v))
        assert q[:, 0].min() >= 0.05 - tol and q[:, 0].max() <= 3 - 0.05 + tol
        assert q[:, 1].min() >= 0.05 - tol and q[:, 1].max() < 20
        assert q[:, 2].min() >= 0.05 - tol and q[:, 2].max() <= 2 - 0.05 + tol
        ds = np.linalg.norm(q[:, :3] - np.asarray(sphere[0]), axis=1)
        assert ds.min() >= sphere[1] + 0.05 - tol
        contacts_s += int(np.sum(ds < sphere[1] + 0.05 + 1e-4))
        dd = np.maximum(np.abs(q[:, :3] - np.asarray(box[0])) - np.asarray(box[1]), 0.0)
        db = np.linalg.norm(dd, axis=1)
        assert db.min() >= 0.05 - tol
        contacts_b += int(np.sum(db < 0.05 + 1e-4))
        assert np.max(np.sum(v[:, :3].astype(np.float64) ** 2, 1)) < 50.0 ** 2
    assert contacts_s > 0 and contacts_b > 0
    assert q[:, 1].mean() < 1.0
    assert math.sqrt(np.mean(np.sum(v[:, :3].astype(np.float64) ** 2, 1))) < 1.0


# ---------------------------------------------------------------- API contract
@pytest.mark.gpu
def test_calls_before_initialize(gpu):
    import pbf_amd

    s = pbf_amd.PBFSolver()
    x = f4([0, 0, 0, 1])
    with pytest.raises(pbf_amd.LogicError):
        s.setParticles(x, x, 1)
    with pytest.raises(pbf_amd.LogicError):
        s.step()
    with pytest.raises(pbf_amd.LogicError):
        s.run()
    with pytest.raises(pbf_amd.OutOfRange):
        s.copyPositionsToHost(np.zeros((1, 4), np.float32), 1)


@pytest.mark.gpu
def test_invalid_capacity_and_bounds(gpu):
    import pbf_amd

    p = valid_params()
    with pytest.raises(pbf_amd.InvalidArgument):
        pbf_amd.PBFSolver().initialize(0, (0, 0, 0), (1, 1, 1), p)
    with pytest.raises(pbf_amd.LengthError):
        pbf_amd.PBFSolver().initialize(2 ** 31, (0, 0, 0), (1, 1, 1), p)
    with pytest.raises(pbf_amd.InvalidArgument):
        pbf_amd.PBFSolver().initialize(8, (1, 0, 0), (1, 2, 2), p)
    with pytest.raises(pbf_amd.InvalidArgument):
        pbf_amd.PBFSolver().initialize(8, (0, 0, 0), (1, 2, 1), p)


@pytest.mark.gpu
@pytest.mark.parametrize("field,value", [
    ("dt", 0.0), ("dt", float("nan")), ("restDensity", -1.0), ("restDensity", float("inf")),
    ("particleMass", 0.0), ("particleMass", float("nan")), ("particleRadius", 0.0),
    ("particleRadius", float("inf")), ("collisionRestitution", -0.01),
    ("collisionRestitution", 1.01), ("collisionRestitution", float("nan")),
    ("collisionFriction", -0.01), ("collisionFriction", 1.01), ("collisionFriction", float("inf")),
    ("smoothingRadius", -1.0), ("smoothingRadius", float("inf")), ("lambdaEpsilon", 0.0),
    ("lambdaEpsilon", -0.01), ("lambdaEpsilon", float("nan")), ("solverIterations", 0),
    ("solverIterations", -1), ("substeps", 0), ("substeps", -1), ("vorticityStrength", -0.01),
    ("vorticityStrength", float("inf")), ("xsphViscosity", -1.0), ("scorrK", -1.0),
])
def test_invalid_parameters(gpu, field, value):
    import pbf_amd

    p = valid_params(**{field: value})
    with pytest.raises(pbf_amd.InvalidArgument):
        pbf_amd.PBFSolver().initialize(8, (0, 0, 0), (2, 2, 2), p)


@pytest.mark.gpu
def test_invalid_gravity_scorr_and_substep(gpu):
    import pbf_amd

    for g in ((float("inf"), -9.8, 0.0), (0.0, float("nan"), 0.0), (0.0, -9.8, -float("inf"))):
        with pytest.raises(pbf_amd.InvalidArgument):
            pbf_amd.PBFSolver().initialize(8, (0, 0, 0), (2, 2, 2), valid_params(gravity=g))
    for kw in (dict(scorrK=0.1, scorrN=0), dict(scorrK=0.1, scorrN=4, scorrDeltaQ=1.0),
               dict(scorrK=0.1, scorrN=4, scorrDeltaQ=-0.1)):
        with pytest.raises(pbf_amd.InvalidArgument):
            pbf_amd.PBFSolver().initialize(8, (0, 0, 0), (2, 2, 2), valid_params(**kw))
    tiny = valid_params(dt=float(np.finfo(np.float32).smallest_subnormal), substeps=2)
    with pytest.raises(pbf_amd.InvalidArgument):
        pbf_amd.PBFSolver().initialize(8, (0, 0, 0), (2, 2, 2), tiny)


@pytest.mark.gpu
def test_particle_input_validation_and_strong_guarantee(gpu):
    import pbf_amd

    s = pbf_amd.PBFSolver()
    s.initialize(2, MIN, MAX, valid_params())
    x, v = f4([1, 2, 3, 1]), f4([0, 0, 0, 7])
    with pytest.raises(pbf_amd.OutOfRange):
        s.setParticles(np.tile(x, (3, 1)), np.tile(v, (3, 1)), 3)
    with pytest.raises(pbf_amd.InvalidArgument):
        s.setParticles(None, v, 1)
    with pytest.raises(pbf_amd.InvalidArgument):
        s.setParticles(x, None, 1)
    s.setParticles(f4([2, 3, 4, 1]), f4([0, 0, 0, 0]), 1)
    bad = f4([np.nan, 6, 7, 1])
    with pytest.raises(pbf_amd.InvalidArgument):
        s.setParticles(bad, v, 1)
    with pytest.raises(pbf_amd.InvalidArgument):
        s.setParticles(x, f4([0, 0, np.inf, 0]), 1)
    assert np.array_equal(s.copyPositionsToHost(np.zeros((1, 4), np.float32), 1), f4([2, 3, 4, 1]))


@pytest.mark.gpu
def test_output_requests_and_empty_sets(gpu):
    import pbf_amd

    s = pbf_amd.PBFSolver()
    s.initialize(4, MIN, MAX, valid_params())
    s.setParticles(f4([1, 2, 3, 1], [2, 3, 4, 2]), f4([0, 0, 0, 3], [1, 1, 1, 4]), 2)
    out = np.zeros((3, 4), np.float32)
    for n in (1, 3):
        with pytest.raises(pbf_amd.OutOfRange):
            s.copyPositionsToHost(out, n)
        with pytest.raises(pbf_amd.OutOfRange):
            s.copyVelocitiesToHost(out, n)
    with pytest.raises(pbf_amd.InvalidArgument):
        pbf_amd._check(pbf_amd.lib().pbf_solver_copy_positions_to_host(s._h, None, 2))
    s.setParticles(None, None, 0)
    s.step()
    s.run()
    assert s.copyPositionsToHost(None, 0) is None


@pytest.mark.gpu
def test_upload_download_and_smaller_replacement(gpu):
    import pbf_amd

    s = pbf_amd.PBFSolver()
    s.initialize(8, MIN, MAX, valid_params())
    x = f4([1, 2, 3, 4], [-1, -2, -3, -4], [5.5, 6.5, 7.5, 8.5])
    v = f4([0.1, 0.2, 0.3, 9], [-0.1, -0.2, -0.3, 10], [1, 2, 3, 11])
    s.setParticles(x, v)
    assert np.array_equal(s.copyPositionsToHost(), x) and np.array_equal(s.copyVelocitiesToHost(), v)
    s.setParticles(np.tile(f4([1, 1, 1, 1]), (6, 1)), np.tile(f4([2, 2, 2, 2]), (6, 1)))
    x2, v2 = f4([3, 4, 5, 6], [7, 8, 9, 10]), f4([-1, -2, -3, 11], [-4, -5, -6, 12])
    s.setParticles(x2, v2)
    assert np.array_equal(s.copyPositionsToHost(), x2) and np.array_equal(s.copyVelocitiesToHost(), v2)


def test_batch_validation():
    """The batch validates like PBFSolver::initialize, per scene, before it
    touches the device."""
    import pbf_amd

    b = pbf_amd.PBFBatchSolver()
    p = valid_params()
    with pytest.raises(pbf_amd.InvalidArgument):
        b.initialize([], p)
    with pytest.raises(pbf_amd.InvalidArgument):
        b.initialize([(10, MIN, MAX), (0, MIN, MAX)], p)
    with pytest.raises(pbf_amd.InvalidArgument):
        b.initialize([(10, MIN, MAX), (10, (0, 0, 0), (1, -1, 1))], p)
    with pytest.raises(pbf_amd.InvalidArgument):
        b.initialize([(10, MIN, MAX)], valid_params(dt=-1.0))
    with pytest.raises(pbf_amd.InvalidArgument):
        b.initialize([(10, MIN, MAX), (10, (0, 0, 0), (0.01, 1, 1))], p)
    with pytest.raises(pbf_amd.LogicError):
        b.setParticles(np.zeros((1, 4), np.float32), np.zeros((1, 4), np.float32))
    with pytest.raises(pbf_amd.LogicError):
        b.step()
