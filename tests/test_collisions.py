"""Collision known answers (test/test_collision_kernels.cu,
test/test_collision_system.cu) on the oracle and the GPU, plus the
CollisionSystem API contract of the library."""
from __future__ import annotations

import numpy as np
import pytest

TOL = 1e-5
LARGE = ((-10.0, -10.0, -10.0), (10.0, 10.0, 10.0))
UNIT = ((0.0, 0.0, 0.0), (10.0, 10.0, 10.0))


def f4(*rows):
    return np.array(rows, dtype=np.float32)


def near(a, b, tol=TOL):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    return np.all(np.abs(a[..., :3] - b[..., :3]) <= tol) and np.array_equal(a[..., 3], b[..., 3])


# ---------------------------------------------------------------- kernels
def test_container_clamps_five_walls_top_open(backend):
    out, flag = backend.k.solve_container(f4([-2, -3, -4, 1], [12, 15, 14, 2]), UNIT, 1.0)
    assert near(out[0], [1, 1, 1, 1]) and near(out[1], [9, 15, 9, 2]) and flag == 1


def test_container_leaves_interior_and_boundary(backend):
    p = f4([5, 5, 5, 3], [1, 1, 9, 4])
    out, flag = backend.k.solve_container(p, UNIT, 1.0)
    assert near(out, p) and flag == 0


def test_container_leaves_elements_past_count(backend):
    p = np.tile(f4([-1, -1, -1, 5]), (258, 1))
    p[-1] = (-7, -8, -9, 6)
    out, _ = backend.k.solve_container(p, UNIT, 0.5, particleCount=257)
    assert near(out[256], [0.5, 0.5, 0.5, 5]) and near(out[-1], p[-1])


def test_sphere_projects_and_handles_center(backend):
    out, flag = backend.k.solve_spheres(f4([2, 2, 3, 1], [1, 2, 3, 2]), [((1, 2, 3), 2.0)], 0.5)
    assert near(out[0], [3.5, 2, 3, 1]) and near(out[1], [3.5, 2, 3, 2]) and flag == 1


def test_sphere_leaves_contact_and_outside(backend):
    p = f4([1.5, 0, 0, 3], [2, 0, 0, 4])
    out, flag = backend.k.solve_spheres(p, [((0, 0, 0), 1.0)], 0.5)
    assert near(out, p) and flag == 0


def test_spheres_in_array_order(backend):
    out, _ = backend.k.solve_spheres(f4([0, 0, 0, 7]), [((0, 0, 0), 1.0), ((2.5, 0, 0), 1.0)],
                                     0.5)
    assert near(out[0], [1, 0, 0, 7])


def test_empty_collider_arrays(backend):
    p = f4([1, 2, 3, 4])
    assert near(backend.k.solve_spheres(p, [], 0.5)[0], p)
    assert near(backend.k.solve_boxes(p, [], 0.5)[0], p)
    assert near(backend.k.solve_planes(p, [], 0.5)[0], p)


def test_box_face_edge_corner(backend):
    box = [((0, 0, 0), (1, 1, 1))]
    out, _ = backend.k.solve_boxes(f4([1.2, 0, 0, 1], [1.2, 1.2, 0, 2], [1.2, 1.2, 1.2, 3]), box,
                                   0.5)
    e2 = 1 + 0.5 / np.sqrt(2.0)
    e3 = 1 + 0.5 / np.sqrt(3.0)
    assert near(out[0], [1.5, 0, 0, 1])
    assert abs(out[1, 0] - e2) <= TOL and abs(out[1, 1] - e2) <= TOL and out[1, 2] == 0.0
    assert np.all(np.abs(out[2, :3] - e3) <= TOL)


def test_box_interior_to_nearest_face(backend):
    box = [((0, 0, 0), (1, 1, 1))]
    out, _ = backend.k.solve_boxes(f4([0.8, 0, 0, 1], [-0.8, 0, 0, 2], [0, 0.8, 0, 3],
                                      [0, 0, -0.8, 4]), box, 0.5)
    assert near(out, f4([1.5, 0, 0, 1], [-1.5, 0, 0, 2], [0, 1.5, 0, 3], [0, 0, -1.5, 4]))


def test_box_center_tie_uses_positive_x(backend):
    out, _ = backend.k.solve_boxes(f4([0, 0, 0, 9]), [((0, 0, 0), (1, 1, 1))], 0.5)
    assert near(out[0], [1.5, 0, 0, 9])


def test_box_leaves_contact_and_outside(backend):
    p = f4([1.5, 0, 0, 1], [2, 2, 0, 2])
    out, flag = backend.k.solve_boxes(p, [((0, 0, 0), (1, 1, 1))], 0.5)
    assert near(out, p) and flag == 0


def test_plane_normalizes_and_projects(backend):
    out, _ = backend.k.solve_planes(f4([2, 1.2, 3, 1], [2, -4, 3, 2]), [((0, 1, 0), (0, 2, 0))],
                                    0.5)
    assert near(out, f4([2, 1.5, 3, 1], [2, 1.5, 3, 2]))


def test_plane_leaves_contact_and_positive_side(backend):
    p = f4([0, 1.5, 0, 3], [0, 2, 0, 4])
    out, flag = backend.k.solve_planes(p, [((0, 1, 0), (0, 1, 0))], 0.5)
    assert near(out, p) and flag == 0


def test_planes_ignore_zero_normal_in_order(backend):
    planes = [((0, 0, 0), (0, 0, 0)), ((1, 0, 0), (1, 0, 0)), ((0, 2, 0), (0, 1, 0))]
    out, _ = backend.k.solve_planes(f4([0, 0, 3, 5]), planes, 0.5)
    assert near(out[0], [1.5, 2.5, 3, 5])


def test_velocities_every_container_wall(backend):
    p = f4([0.5, 5, 5, 1], [9.5, 5, 5, 2], [5, 0.5, 5, 3], [5, 5, 0.5, 4], [5, 5, 9.5, 5],
           [5, 20, 5, 6])
    v = f4([-2, 4, 0, 11], [2, 4, 0, 12], [4, -2, 0, 13], [4, 0, -2, 14], [4, 0, 2, 15],
           [1, 2, 3, 16])
    _, out = backend.k.resolve_velocities(p, v, UNIT, particleRadius=0.5, restitution=0.5,
                                          friction=0.25)
    exp = f4([1, 3, 0, 11], [-1, 3, 0, 12], [3, 1, 0, 13], [3, 0, 1, 14], [3, 0, -1, 15],
             [1, 2, 3, 16])
    assert near(out, exp)


def test_velocities_interior_and_separating(backend):
    v = f4([2, 3, 4, 5], [-2, 3, 4, 6])
    _, out = backend.k.resolve_velocities(f4([0.5, 5, 5, 1], [5, 5, 5, 2]), v, UNIT)
    assert near(out, v)


def test_velocity_contact_tolerance(backend):
    _, out = backend.k.resolve_velocities(f4([0.50005, 5, 5, 1]), f4([-2, 0, 0, 2]), UNIT)
    assert near(out[0], [0, 0, 0, 2])


def test_velocity_tolerance_every_collider(backend):
    _, s = backend.k.resolve_velocities(f4([1.50005, 0, 0, 3]), f4([-2, 0, 0, 1]), LARGE,
                                        spheres=[((0, 0, 0), 1.0)])
    _, b = backend.k.resolve_velocities(f4([1.50005, 0, 0, 4]), f4([-2, 0, 0, 1]), LARGE,
                                        boxes=[((0, 0, 0), (1, 1, 1))])
    _, pl = backend.k.resolve_velocities(f4([0, 0.50005, 0, 5]), f4([0, -2, 0, 2]), LARGE,
                                         planes=[((0, 0, 0), (0, 1, 0))])
    assert abs(s[0, 0]) <= TOL and abs(b[0, 0]) <= TOL and abs(pl[0, 1]) <= TOL


def test_sphere_restitution_friction_direction(backend):
    p = f4([1.5, 0, 0, 1], [1.5, 0, 0, 2], [2, 0, 0, 3])
    v = f4([-2, 4, 0, 4], [2, 4, 0, 5], [-2, 4, 0, 6])
    _, out = backend.k.resolve_velocities(p, v, LARGE, spheres=[((0, 0, 0), 1.0)],
                                          restitution=0.5, friction=0.25)
    assert near(out[0], [1, 3, 0, 4]) and near(out[1], v[1]) and near(out[2], v[2])


def test_full_restitution_and_friction(backend):
    _, out = backend.k.resolve_velocities(f4([1.5, 0, 0, 1]), f4([-2, 4, 0, 2]), LARGE,
                                          spheres=[((0, 0, 0), 1.0)], restitution=1.0,
                                          friction=1.0)
    assert near(out[0], [2, 0, 0, 2])


def test_sphere_center_degeneracy(backend):
    v = f4([-2, 3, 4, 5])
    _, out = backend.k.resolve_velocities(f4([0, 0, 0, 1]), v, LARGE, spheres=[((0, 0, 0), 1.0)])
    assert near(out, v)


def test_rounded_box_edge_and_interior(backend):
    e = np.float32(1.0 + 0.5 / np.sqrt(2.0))
    p = f4([e, e, 0, 1], [0.8, 0, 0, 2], [3, 0, 0, 3])
    v = f4([-2, -2, 2, 4], [-2, 4, 0, 5], [-2, 4, 0, 6])
    _, out = backend.k.resolve_velocities(p, v, LARGE, boxes=[((0, 0, 0), (1, 1, 1))],
                                          friction=0.5)
    assert near(out[0], [0, 0, 1, 4]) and near(out[1], [0, 2, 0, 5]) and near(out[2], v[2])


def test_non_unit_plane_and_zero_normal(backend):
    planes = [((0, 0, 0), (0, 0, 0)), ((0, 1, 0), (0, 2, 0))]
    _, out = backend.k.resolve_velocities(f4([0, 1.5, 0, 1]), f4([4, -2, 0, 2]), LARGE,
                                          planes=planes, restitution=0.5, friction=0.25)
    assert near(out[0], [3, 1, 0, 2])


def test_plane_separating_and_distant(backend):
    v = f4([1, 2, 3, 4], [1, -2, 3, 5])
    _, out = backend.k.resolve_velocities(f4([0, 0.5, 0, 6], [0, 1, 0, 7]), v, LARGE,
                                          planes=[((0, 0, 0), (0, 1, 0))])
    assert near(out, v)


def test_velocities_past_count(backend):
    p = np.tile(f4([-9.5, 0, 0, 1]), (258, 1))
    v = np.tile(f4([-2, 0, 0, 2]), (258, 1))
    p[-1] = (-9.5, 3, 4, 7)
    v[-1] = (-8, 5, 6, 9)
    pout, out = backend.k.resolve_velocities(p, v, LARGE, particleCount=257)
    assert near(pout, p) and near(out[256], [0, 0, 0, 2]) and near(out[-1], v[-1])


# ---------------------------------------------------------------- system
def test_rounded_contacts_converge_despite_rounding(backend):
    cs = backend.collision(LARGE)
    cs.setSpheres([((1.3, 2.1, -0.7), 0.7)])
    for i in range(1, 65):
        out = cs.solve(f4([np.float32(1.3) + np.float32(i) * np.float32(0.001), 2.2, -0.65, 1]),
                       0.1)
        d = out[0, :3].astype(np.float64) - np.array([np.float32(1.3), np.float32(2.1),
                                                       np.float32(-0.7)], np.float64)
        assert abs(np.linalg.norm(d) - 0.8) <= 2e-6
    cs.clearSpheres()
    cs.setBoxes([((0, 0, 0), (1, 1, 1))])
    cs.solve(f4([1.01, 1.03, 0, 1]), 0.1)
    cs.clearBoxes()
    cs.setPlanes([((0, 0, 0), (1, 2, 3))])
    cs.solve(f4([0.01, 0.02, 0.03, 1]), 0.1)


def test_floor_stops_normal_velocity_damps_tangent(backend):
    cs = backend.collision(UNIT)
    p = cs.solve(f4([2, 0.2, 2, 1]), 0.5)
    v = cs.resolveVelocities(p, f4([1, -2, 0, 0]), 0.5, 0.0, 0.1)
    assert abs(p[0, 1] - 0.5) <= TOL and abs(v[0, 0] - 0.9) <= TOL and abs(v[0, 1]) <= TOL


def test_sphere_impenetrable(backend):
    cs = backend.collision(LARGE)
    cs.setSpheres([((0, 0, 0), 1.0)])
    p = cs.solve(f4([0.2, 0, 0, 1]), 0.5)
    v = cs.resolveVelocities(p, f4([-3, 0, 0, 0]), 0.5, 0.0, 0.0)
    assert abs(p[0, 0] - 1.5) <= TOL and abs(v[0, 0]) <= TOL


def test_box_edge_rounded_contact(backend):
    cs = backend.collision(LARGE)
    cs.setBoxes([((0, 0, 0), (1, 1, 1))])
    p = cs.solve(f4([1.1, 1.1, 0, 1]), 0.5)
    assert abs(np.hypot(p[0, 0] - 1.0, p[0, 1] - 1.0) - 0.5) <= TOL


def test_plane_stops_normal_velocity(backend):
    cs = backend.collision(LARGE)
    cs.setPlanes([((0, 0, 0), (0, 1, 0))])
    p = cs.solve(f4([0, 0.1, 0, 1]), 0.5)
    v = cs.resolveVelocities(p, f4([0, -2, 0, 0]), 0.5, 0.0, 0.0)
    assert abs(p[0, 1] - 0.5) <= TOL and abs(v[0, 1]) <= TOL


def test_non_convergent_constraints_raise(backend):
    cs = backend.collision(((-2.0, -2.0, -2.0), (2.0, 2.0, 2.0)))
    cs.setSpheres([((-1.5, 0, 0), 0.5)])
    with pytest.raises(backend.E.CollisionNonConvergence):
        cs.solve(f4([-1.6, 0, 0, 1]), 0.25)


def test_multiple_collider_types(backend):
    cs = backend.collision(LARGE)
    cs.setSpheres([((-5, 5, 0), 1.0)])
    cs.setBoxes([((0, 5, 0), (1, 1, 1))])
    cs.setPlanes([((0, 3, 0), (0, 1, 0))])
    p = cs.solve(f4([-5.2, 5, 0, 1], [1.2, 5, 0, 2], [5, 2.8, 0, 3]), 0.5)
    v = cs.resolveVelocities(p, f4([2, 0, 0, 4], [-2, 0, 0, 5], [0, -2, 0, 6]), 0.5, 0.0, 0.0)
    assert abs(p[0, 0] + 6.5) <= TOL and abs(p[1, 0] - 1.5) <= TOL and abs(p[2, 1] - 3.5) <= TOL
    assert abs(v[0, 0]) <= TOL and abs(v[1, 0]) <= TOL and abs(v[2, 1]) <= TOL


def test_container_update_and_cleared_colliders(backend):
    cs = backend.collision(UNIT)
    assert abs(cs.solve(f4([-1, 5, 5, 7]), 0.5)[0, 0] - 0.5) <= TOL
    cs2 = backend.collision(((-5.0, -5.0, -5.0), (5.0, 5.0, 5.0)))
    assert abs(cs2.solve(f4([-6, 0, 0, 8]), 0.5)[0, 0] + 4.5) <= TOL
    cs = backend.collision(LARGE)
    cs.setSpheres([((0, 0, 0), 1.0)])
    cs.setBoxes([((3, 0, 0), (1, 1, 1))])
    cs.setPlanes([((0, 3, 0), (0, 1, 0))])
    cs.setSpheres([])
    cs.clearBoxes()
    cs.setPlanes([])
    p = f4([0, 0, 0, 4])
    assert np.array_equal(cs.solve(p, 0.5), p)


# ---------------------------------------------------------------- API contract
@pytest.mark.gpu
def test_collision_system_validation(gpu):
    import pbf_amd

    cs = pbf_amd.CollisionSystem(LARGE)
    d = pbf_amd.DeviceArray(16)
    nan, inf = float("nan"), float("inf")
    with pytest.raises(pbf_amd.InvalidArgument):
        cs.solve(None, 1, 0.5)
    for r in (0.0, -0.5, nan):
        with pytest.raises(pbf_amd.InvalidArgument):
            cs.solve(d, 1, r)
    cs.solve(d, 0, 0.5)
    narrow = pbf_amd.CollisionSystem(((0, 0, 0), (0.5, 1, 1)))
    with pytest.raises(pbf_amd.InvalidArgument):
        narrow.solve(d, 1, 0.3)
    with pytest.raises(pbf_amd.InvalidArgument):
        narrow.resolveVelocities(d, d, 1, 0.3, 0.0, 0.0)
    for args in ((None, d, 1, 0.5, 0.0, 0.0), (d, None, 1, 0.5, 0.0, 0.0),
                 (d, d, 1, 0.0, 0.0, 0.0), (d, d, 1, nan, 0.0, 0.0),
                 (d, d, 1, 0.5, -0.1, 0.0), (d, d, 1, 0.5, 1.1, 0.0), (d, d, 1, 0.5, nan, 0.0),
                 (d, d, 1, 0.5, 0.0, -0.1), (d, d, 1, 0.5, 0.0, 1.1), (d, d, 1, 0.5, 0.0, nan)):
        with pytest.raises(pbf_amd.InvalidArgument):
            cs.resolveVelocities(*args)
    cs.resolveVelocities(d, d, 0, 0.5, 0.0, 0.0)
    with pytest.raises(pbf_amd.InvalidArgument):
        pbf_amd.CollisionSystem(((0, 0, 0), (0, 1, 1)))
    with pytest.raises(pbf_amd.InvalidArgument):
        pbf_amd.CollisionSystem(((0, 0, 0), (1, nan, 1)))
    for s in (((0, 0, 0), 0.0), ((0, 0, 0), -1.0), ((0, 0, 0), nan), ((0, 0, 0), inf),
              ((nan, 0, 0), 1.0)):
        with pytest.raises(pbf_amd.InvalidArgument):
            cs.setSpheres([s])
    for he in ((0, 1, 1), (1, -1, 1), (1, 1, nan)):
        with pytest.raises(pbf_amd.InvalidArgument):
            cs.setBoxes([((0, 0, 0), he)])
    with pytest.raises(pbf_amd.InvalidArgument):
        cs.setBoxes([((0, nan, 0), (1, 1, 1))])
    for pl in (((0, 0, 0), (0, 0, 0)), ((0, 0, 0), (inf, 0, 0)), ((0, 0, nan), (0, 1, 0))):
        with pytest.raises(pbf_amd.InvalidArgument):
            cs.setPlanes([pl])


@pytest.mark.gpu
def test_collision_system_use_before_container_and_launch_limits(gpu):
    import pbf_amd

    d = pbf_amd.DeviceArray(16)
    cs = pbf_amd.CollisionSystem()
    with pytest.raises(pbf_amd.LogicError):
        cs.solve(d, 1, 0.5)
    with pytest.raises(pbf_amd.LogicError):
        cs.resolveVelocities(d, d, 1, 0.5, 0.0, 0.0)
    cs.setContainer(LARGE)
    big = 2 ** 64 - 1
    with pytest.raises(pbf_amd.LengthError):
        cs.solve(d, big, 0.1)
    with pytest.raises(pbf_amd.LengthError):
        cs.resolveVelocities(d, d, big, 0.1, 0.0, 0.0)


@pytest.mark.gpu
def test_failed_setter_keeps_previous_colliders(gpu):
    import pbf_amd

    cs = pbf_amd.CollisionSystem(LARGE)
    cs.setSpheres([((0, 0, 0), 1.0)])
    with pytest.raises(pbf_amd.InvalidArgument):
        cs.setSpheres([((5, 0, 0), -1.0)])
    d = pbf_amd.DeviceArray.from_host(f4([0, 0, 0, 1]))
    cs.solve(d, 1, 0.5)
    assert abs(d.to_numpy(np.float32, (1, 4))[0, 0] - 1.5) <= TOL
    cs.setSpheres([((5, 0, 0), 1.0)])
    d2 = pbf_amd.DeviceArray.from_host(f4([0, 0, 0, 1]))
    cs.solve(d2, 1, 0.5)
    assert np.array_equal(d2.to_numpy(np.float32, (1, 4)), f4([0, 0, 0, 1]))
