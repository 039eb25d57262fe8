"""Known answers of the particle kernels on hand-built neighbour lists
(test/test_pbf_solver_kernels.cu:133-767, test/test_integration.cu), run on
the oracle and on the GPU.  Lists are slot-major: nbr[k * stride + i]."""
from __future__ import annotations

import numpy as np

PI = np.float32(3.14159265358979323846)


def f4(*rows):
    return np.array(rows, dtype=np.float32)


def test_integrate_1025_particles(backend):
    n = 1025
    i = np.arange(n, dtype=np.float32)
    x = np.stack([1 + 0.01 * i, 2 - 0.02 * i, 3 + 0.03 * i, np.ones(n, np.float32)], 1)
    v = np.stack([-2 + 0.005 * i, 0.5 - 0.003 * i, -1 + 0.002 * i, np.zeros(n, np.float32)], 1)
    x = x.astype(np.float32)
    v = v.astype(np.float32)
    dt, g = np.float32(0.1), (0.0, -10.0, 0.0)
    p, vout = backend.k.integrate(x, v, dt, g)
    ev = v[:, :3] + np.asarray(g, np.float32) * dt
    ep = x[:, :3] + ev * dt
    assert np.all(np.abs(vout[:, :3] - ev) <= 1e-5)
    assert np.all(np.abs(p[:, :3] - ep) <= 1e-5)
    assert np.array_equal(p[:, 3], x[:, 3])


def test_lambda_across_partial_blocks(backend):
    n, maxn = 513, 3
    p = np.tile(f4([5, 5, 5, 9]), (n, 1))
    p[0] = (0, 0, 0, 7)
    p[1] = (0.5, 0.5, 0.25, -3)
    p[2] = (0, 0.6, 0.8, 2)
    p[3] = (0, 0, 1.25, -8)
    nbr = np.zeros(n * maxn, np.uint32)
    nbr[0], nbr[n], nbr[2 * n] = 1, 2, 3
    cnt = np.zeros(n, np.int32)
    cnt[0] = maxn
    mass, rho0, eps = np.float32(2.0), np.float32(3.0), np.float32(0.01)
    lam = backend.k.compute_lambda(p, nbr, cnt, maxn, n, 1.0, mass, rho0, eps, n)
    selfw = np.float32(315.0) / (np.float32(64.0) * PI)
    nw = selfw * (np.float32(1.0) - np.float32(0.5625)) ** 3
    pair = mass * (selfw + nw)
    selfd = mass * selfw
    grad = np.float32(45.0) / PI * np.float32(0.25) ** 2
    gsum = 2 * (mass / rho0 * grad) ** 2
    assert abs(lam[0] - (-(pair / rho0 - 1) / (gsum + eps))) <= 1e-5
    assert abs(lam[1] - (-(selfd / rho0 - 1) / eps)) <= 1e-5
    assert abs(lam[-1] - (-(selfd / rho0 - 1) / eps)) <= 1e-5


def test_delta_position_and_w_zero(backend):
    n, maxn = 257, 3
    p = np.tile(f4([4, 4, 4, 1]), (n, 1))
    p[0] = (0, 0, 0, 3)
    p[1] = (0.25, 0.5, 0.5, 5)
    p[2] = (0, 0.6, 0.8, 7)
    p[3] = (0, 0, 1.25, -1)
    nbr = np.zeros(n * maxn, np.uint32)
    nbr[0], nbr[n], nbr[2 * n] = 1, 2, 3
    cnt = np.zeros(n, np.int32)
    cnt[0] = maxn
    lam = np.zeros(n, np.float32)
    lam[:3] = (2.0, -0.5, 3.0)
    d = backend.k.compute_delta_position(p, nbr, cnt, maxn, lam, n, 1.0, 2.0, 4.0, 0.0, 0, 0.0,
                                         n)
    gm = np.float32(45.0) / PI * np.float32(0.25) ** 2
    scale = np.float32(2.0) / np.float32(4.0) * (lam[0] + lam[1])
    ex = scale * gm / 3
    assert abs(d[0, 0] - ex) <= 1e-5 and abs(d[0, 1] - 2 * ex) <= 1e-5
    assert abs(d[0, 2] - 2 * ex) <= 1e-5
    assert d[0, 3] == 0.0 and np.all(d[1, :3] == 0.0) and d[-1, 3] == 0.0


def test_apply_delta_updates_only_xyz(backend):
    n = 513
    i = np.arange(n, dtype=np.float32)
    p = np.stack([i, -i, 2 * i, 10 + i], 1).astype(np.float32)
    d = np.stack([-0.25 * i, 0.5 * i, -0.75 * i, -100 - i], 1).astype(np.float32)
    out = backend.k.apply_delta_position(p, d)
    assert np.array_equal(out[:, :3], p[:, :3] + d[:, :3])
    assert np.array_equal(out[:, 3], p[:, 3])


def test_update_velocity_and_position(backend):
    n = 513
    i = np.arange(n, dtype=np.float32)
    x = np.stack([0.5 * i, -2 * i, i + 3, -10 - i], 1).astype(np.float32)
    p = np.stack([x[:, 0] + 0.125 * (i + 1), x[:, 1] - 0.25 * (i + 2), x[:, 2] + 0.5 * (i + 3),
                  100 + i], 1).astype(np.float32)
    v = np.stack([-7 - i, 8 + i, 9 - i, 50 + i], 1).astype(np.float32)
    inv = np.float32(8.0)
    xo, po, vo = backend.k.update_velocity_and_position(x, p, v, n, inv)
    assert np.array_equal(xo, p)
    assert np.array_equal(po, p)
    assert np.array_equal(vo[:, :3], (p[:, :3] - x[:, :3]) * inv)
    assert np.array_equal(vo[:, 3], v[:, 3])


def test_update_velocity_leaves_guard_element(backend):
    n = 257
    x = np.tile(f4([-1, -2, -3, -4]), (n + 1, 1))
    p = np.tile(f4([5, 6, 7, 8]), (n + 1, 1))
    v = np.tile(f4([9, 10, 11, 12]), (n + 1, 1))
    xo, _, vo = backend.k.update_velocity_and_position(x, p, v, n, 2.0)
    assert np.array_equal(vo[:n, :3], np.tile(f4([12, 16, 20]), (n, 1)))
    assert np.array_equal(xo[n], x[n]) and np.array_equal(vo[n], v[n])


def test_xsph_disabled_leaves_velocities(backend):
    p = f4([0, 0, 0, 1], [0.25, 0, 0, 2])
    v = f4([1, -2, 3, 4], [-4, 5, -6, 7])
    out, vin = backend.k.apply_xsph_viscosity(p, [1, 0, 0, 0], [1, 1], 2, v, 2, 1.0, 0.0, 2)
    assert np.array_equal(out, v) and np.array_equal(vin, v)


def test_xsph_no_neighbors_and_equal_velocities(backend):
    p = f4([0, 0, 0, 1], [0.25, 0, 0, 2], [0.5, 0, 0, 3])
    v = f4([2, -1, 0.5, 4], [2, -1, 0.5, 5], [-3, 1, 7, 6])
    out, _ = backend.k.apply_xsph_viscosity(p, [1, 0, 0, 0, 0, 0], [1, 1, 0], 2, v, 3, 1.0,
                                            0.25, 3)
    assert np.array_equal(out, v)


def test_xsph_poly6_weighted_jacobi_correction(backend):
    p = f4([0, 0, 0, 1], [0.5, 0, 0, 2], [2, 0, 0, 3])
    v = f4([1, 2, -1, 4], [-3, 4, 5, 5], [10, -2, 1, 6])
    # Particle 2 is listed for particle 0 but lies outside h.
    out, _ = backend.k.apply_xsph_viscosity(p, [1, 0, 0, 2, 0, 0], [2, 0, 0], 2, v, 3, 1.0,
                                            0.15, 3)
    w = np.float32(315.0) / (np.float32(64.0) * PI) * (np.float32(1.0) - np.float32(0.25)) ** 3
    exp = v[0, :3] + np.float32(0.15) * (v[1, :3] - v[0, :3]) * w
    assert np.all(np.abs(out[0, :3] - exp) <= 1e-6)
    assert out[0, 3] == v[0, 3] and out[1, 0] == v[1, 0] and out[2, 0] == v[2, 0]


def test_vorticity_disabled_and_no_neighbors(backend):
    p = f4([0, 0, 0, 1], [0.5, 0, 0, 2])
    v = f4([1, -2, 3, 4], [-4, 5, -6, 7])
    nbr, cnt = [1, 0], [1, 0]
    w = backend.k.compute_vorticity(p, v, nbr, cnt, 1, 2, 1.0, 2)
    assert np.all(w[1, :3] == 0.0)
    out, _ = backend.k.apply_vorticity_confinement(p, nbr, cnt, 1, w, v, 2, 1.0, 0.1, 0.0, 2)
    assert np.array_equal(out, v)
    out2, _ = backend.k.apply_vorticity_confinement(p, nbr, cnt, 1, w, v, 2, 1.0, 0.1, 2.0, 2)
    assert np.array_equal(out2[1], v[1])


def test_vorticity_uniform_velocity_zero_eta(backend):
    p = f4([0, 0, 0, 1], [0.5, 0, 0, 2], [0, 0.5, 0, 3])
    v = np.tile(f4([2, -1, 0.5, 9]), (3, 1))
    nbr, cnt = [1, 0, 0, 2, 2, 1], [2, 2, 2]
    w = backend.k.compute_vorticity(p, v, nbr, cnt, 2, 3, 1.0, 3)
    out, _ = backend.k.apply_vorticity_confinement(p, nbr, cnt, 2, w, v, 3, 1.0, 0.1, 3.0, 3)
    assert np.all(w[:, :3] == 0.0)
    assert np.all(np.isfinite(out)) and np.array_equal(out[:, :3], v[:, :3])


def _spiky_ref(d, h):
    r = np.sqrt(np.float32(d @ d))
    if r <= 0 or r > h:
        return np.zeros(3, np.float32)
    mag = np.float32(-45.0) / (PI * np.float32(h) ** 6) * (np.float32(h) - r) ** 2
    return d / r * mag


def test_vorticity_matches_cpu_reference_and_curl_sign(backend):
    h, dt, strength = np.float32(1.0), np.float32(0.1), np.float32(0.25)
    p = f4([0, 0, 0, 1], [0.5, 0, 0, 2], [0, 0.5, 0, 3])
    v = f4([0, 0, 0, 4], [0, 0.5, 0, 5], [-0.5, 0, 0, 6])  # v = (-y, x, 0)
    nbr = np.array([1, 0, 0, 2, 2, 1], np.uint32)
    cnt = np.array([2, 2, 2], np.int32)
    n = 3
    ew = np.zeros((n, 3), np.float32)
    for i in range(n):
        for k in range(cnt[i]):
            j = nbr[k * n + i]
            g = _spiky_ref(p[i, :3] - p[j, :3], h)
            ew[i] += np.cross(g, v[j, :3] - v[i, :3])
    ev = v.copy()
    for i in range(n):
        eta = np.zeros(3, np.float32)
        for k in range(cnt[i]):
            j = nbr[k * n + i]
            eta += (np.linalg.norm(ew[j]) - np.linalg.norm(ew[i])) * _spiky_ref(p[i, :3] - p[j, :3], h)
        el = np.linalg.norm(eta)
        nrm = eta / el if el > 1e-6 else np.zeros(3, np.float32)
        ev[i, :3] += dt * strength * np.cross(nrm, ew[i])
    w = backend.k.compute_vorticity(p, v, nbr, cnt, 2, n, h, n)
    out, vin = backend.k.apply_vorticity_confinement(p, nbr, cnt, 2, w, v, n, h, dt, strength, n)
    assert np.all(np.abs(w[:, :3] - ew) <= 1e-5) and np.all(w[:, 3] == 0.0)
    assert np.all(np.abs(out[:, :3] - ev[:, :3]) <= 1e-5)
    assert np.array_equal(out[:, 3], v[:, 3]) and np.array_equal(vin, v)
    assert w[0, 2] > 0.0  # v = (-y, x, 0) has positive z curl


def test_invalid_neighbor_counts_poison_outputs(backend):
    """A count outside [0, maxNeighbors] NaN-poisons xyz and keeps w
    (src/pbf/pbf_solver_kernels.cu:41-45, :124-131, :225-233, :281-284, :335-338)."""
    p = f4([0, 0, 0, 1], [0.5, 0, 0, 2])
    v = f4([1, 2, 3, 4], [5, 6, 7, 8])
    nbr = [1, 0, 0, 0]
    cnt = [3, -1]  # maxNeighbors = 2
    lam = backend.k.compute_lambda(p, nbr, cnt, 2, 2, 1.0, 1.0, 1.0, 0.1, 2)
    assert np.all(np.isnan(lam))
    d = backend.k.compute_delta_position(p, nbr, cnt, 2, [0.1, 0.2], 2, 1.0, 1.0, 1.0, 0, 0, 0, 2)
    assert np.all(np.isnan(d[:, :3])) and np.all(d[:, 3] == 0.0)
    xs, _ = backend.k.apply_xsph_viscosity(p, nbr, cnt, 2, v, 2, 1.0, 0.1, 2)
    assert np.all(np.isnan(xs[:, :3])) and np.array_equal(xs[:, 3], v[:, 3])
    w = backend.k.compute_vorticity(p, v, nbr, cnt, 2, 2, 1.0, 2)
    assert np.all(np.isnan(w[:, :3])) and np.all(w[:, 3] == 0.0)
    c, _ = backend.k.apply_vorticity_confinement(p, nbr, cnt, 2, w, v, 2, 1.0, 0.1, 0.5, 2)
    assert np.all(np.isnan(c[:, :3])) and np.array_equal(c[:, 3], v[:, 3])
