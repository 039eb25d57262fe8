"""Host-array wrappers of the library's kernel launchers.

Each function mirrors one free kernel of the reference (include/pbf/*.cuh)
with the same argument meaning: it uploads numpy arrays, launches the gfx950
kernel through the C ABI on the default stream, waits, and returns the
outputs as numpy arrays.  Arrays may be longer than `particleCount`; the
elements past it are passed through, as in the reference's guard-element
tests.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import (BoxCollider, Container, DeviceArray, PlaneCollider, SphereCollider, _box, _check,
               _f3, _plane, _sphere, as_float4, float3, int3, lib, synchronize)


def _dev(a, dtype, min_bytes=0):
    arr = np.ascontiguousarray(np.asarray(a, dtype=dtype))
    return DeviceArray.from_host(arr, min_bytes), arr


def _f4dev(a):
    arr = as_float4(a)
    return DeviceArray.from_host(arr), arr


def _out_f4(n):
    return DeviceArray(max(16, 16 * n))


def poly6(displacements, smoothingRadius):
    d = np.ascontiguousarray(np.asarray(displacements, dtype=np.float32).reshape(-1, 3))
    dd = DeviceArray.from_host(d)
    out = DeviceArray(max(4, 4 * len(d)))
    _check(lib().pbf_k_poly6(dd.ptr, len(d), float(smoothingRadius), out.ptr, None))
    synchronize()
    return out.to_numpy(np.float32, (len(d),))


def spiky_gradient(displacements, smoothingRadius):
    d = np.ascontiguousarray(np.asarray(displacements, dtype=np.float32).reshape(-1, 3))
    dd = DeviceArray.from_host(d)
    out = DeviceArray(max(12, 12 * len(d)))
    _check(lib().pbf_k_spiky_gradient(dd.ptr, len(d), float(smoothingRadius), out.ptr, None))
    synchronize()
    return out.to_numpy(np.float32, (len(d), 3))


def artificial_pressure(distanceSquared, smoothingRadius, scorrK, scorrN, scorrDeltaQ):
    d = np.ascontiguousarray(np.asarray(distanceSquared, dtype=np.float32).reshape(-1))
    dd = DeviceArray.from_host(d)
    out = DeviceArray(max(4, 4 * len(d)))
    _check(lib().pbf_k_artificial_pressure(dd.ptr, len(d), float(smoothingRadius), float(scorrK),
                                           int(scorrN), float(scorrDeltaQ), out.ptr, None))
    synchronize()
    return out.to_numpy(np.float32, (len(d),))


def integrate(position, velocity, dt, gravity, particleCount=None):
    dx, x = _f4dev(position)
    dv, v = _f4dev(velocity)
    n = len(x) if particleCount is None else int(particleCount)
    dp = _out_f4(len(x))
    _check(lib().pbf_k_integrate(dx.ptr, dp.ptr, dv.ptr, n, float(dt), _f3(gravity), None))
    synchronize()
    return dp.to_numpy(np.float32, x.shape), dv.to_numpy(np.float32, v.shape)


def find_neighbors(predicted, cellStart, cellEnd, gridSize, minBounds, cellSize, particleCount,
                   smoothingRadius, maxNeighbors, particleStride, neighborsFill=None,
                   with_flag=True):
    """Returns (neighbors flat uint32[maxNeighbors * stride], counts, overflowFlag)."""
    dp, p = _f4dev(predicted)
    n = int(particleCount)
    if isinstance(cellStart, int):
        cs_ptr, ce_ptr = cellStart, cellEnd
        keep = []
    else:
        dcs, _ = _dev(cellStart, np.int32)
        dce, _ = _dev(cellEnd, np.int32)
        cs_ptr, ce_ptr = dcs.ptr, dce.ptr
        keep = [dcs, dce]
    slots = int(maxNeighbors) * int(particleStride)
    dn = DeviceArray(max(4, 4 * slots))
    if neighborsFill is not None:
        dn.fill_bytes(neighborsFill)
    dc = DeviceArray(max(4, 4 * n))
    df = DeviceArray(4)
    df.fill_bytes(0)
    g = gridSize if isinstance(gridSize, int3) else int3(*gridSize)
    _check(lib().pbf_k_find_neighbors(dp.ptr, cs_ptr, ce_ptr, g, _f3(minBounds), float(cellSize),
                                      n, float(smoothingRadius), dn.ptr, dc.ptr,
                                      int(maxNeighbors), int(particleStride),
                                      df.ptr if with_flag else None, None))
    synchronize()
    del keep
    return (dn.to_numpy(np.uint32, (slots,)), dc.to_numpy(np.int32, (n,)),
            int(df.to_numpy(np.int32, (1,))[0]))


def compute_lambda(predicted, neighbors, counts, maxNeighbors, particleCount, smoothingRadius,
                   particleMass, restDensity, epsilon, stride):
    dp, p = _f4dev(predicted)
    dn, _ = _dev(neighbors, np.uint32)
    dc, _ = _dev(counts, np.int32)
    n = int(particleCount)
    dl = DeviceArray(max(4, 4 * n))
    _check(lib().pbf_k_compute_lambda(dp.ptr, dn.ptr, dc.ptr, int(maxNeighbors), n,
                                      float(smoothingRadius), float(particleMass),
                                      float(restDensity), float(epsilon), dl.ptr, int(stride),
                                      None))
    synchronize()
    return dl.to_numpy(np.float32, (n,))


def compute_delta_position(predicted, neighbors, counts, maxNeighbors, lambdas, particleCount,
                           smoothingRadius, particleMass, restDensity, scorrK, scorrN,
                           scorrDeltaQ, stride):
    dp, p = _f4dev(predicted)
    dn, _ = _dev(neighbors, np.uint32)
    dc, _ = _dev(counts, np.int32)
    dl, _ = _dev(lambdas, np.float32)
    n = int(particleCount)
    dd = _out_f4(n)
    _check(lib().pbf_k_compute_delta_position(dp.ptr, dn.ptr, dc.ptr, int(maxNeighbors), dl.ptr,
                                              n, float(smoothingRadius), float(particleMass),
                                              float(restDensity), float(scorrK), int(scorrN),
                                              float(scorrDeltaQ), dd.ptr, int(stride), None))
    synchronize()
    return dd.to_numpy(np.float32, (n, 4))


def apply_delta_position(predicted, deltas, particleCount=None):
    dp, p = _f4dev(predicted)
    dd, _ = _f4dev(deltas)
    n = len(p) if particleCount is None else int(particleCount)
    _check(lib().pbf_k_apply_delta_position(dp.ptr, dd.ptr, n, None))
    synchronize()
    return dp.to_numpy(np.float32, p.shape)


def update_velocity_and_position(positions, predicted, velocities, particleCount, inverseDt):
    dx, x = _f4dev(positions)
    dp, _ = _f4dev(predicted)
    dv, v = _f4dev(velocities)
    _check(lib().pbf_k_update_velocity_and_position(dx.ptr, dp.ptr, dv.ptr, int(particleCount),
                                                    float(inverseDt), None))
    synchronize()
    return dx.to_numpy(np.float32, x.shape), dp.to_numpy(np.float32, x.shape), \
        dv.to_numpy(np.float32, v.shape)


def apply_xsph_viscosity(predicted, neighbors, counts, maxNeighbors, inputVelocities,
                         particleCount, smoothingRadius, xsphViscosity, stride):
    dp, _ = _f4dev(predicted)
    dn, _ = _dev(neighbors, np.uint32)
    dc, _ = _dev(counts, np.int32)
    dvi, vi = _f4dev(inputVelocities)
    n = int(particleCount)
    dvo = _out_f4(n)
    _check(lib().pbf_k_apply_xsph_viscosity(dp.ptr, dn.ptr, dc.ptr, int(maxNeighbors), dvi.ptr,
                                            dvo.ptr, n, float(smoothingRadius),
                                            float(xsphViscosity), int(stride), None))
    synchronize()
    return dvo.to_numpy(np.float32, (n, 4)), dvi.to_numpy(np.float32, vi.shape)


def compute_vorticity(positions, velocities, neighbors, counts, maxNeighbors, particleCount,
                      smoothingRadius, stride):
    dp, _ = _f4dev(positions)
    dv, _ = _f4dev(velocities)
    dn, _ = _dev(neighbors, np.uint32)
    dc, _ = _dev(counts, np.int32)
    n = int(particleCount)
    dw = _out_f4(n)
    _check(lib().pbf_k_compute_vorticity(dp.ptr, dv.ptr, dn.ptr, dc.ptr, int(maxNeighbors), n,
                                         float(smoothingRadius), dw.ptr, int(stride), None))
    synchronize()
    return dw.to_numpy(np.float32, (n, 4))


def apply_vorticity_confinement(positions, neighbors, counts, maxNeighbors, vorticity,
                                inputVelocities, particleCount, smoothingRadius, dt,
                                vorticityStrength, stride):
    dp, _ = _f4dev(positions)
    dn, _ = _dev(neighbors, np.uint32)
    dc, _ = _dev(counts, np.int32)
    dw, _ = _f4dev(vorticity)
    dvi, vi = _f4dev(inputVelocities)
    n = int(particleCount)
    dvo = _out_f4(n)
    _check(lib().pbf_k_apply_vorticity_confinement(dp.ptr, dn.ptr, dc.ptr, int(maxNeighbors),
                                                   dw.ptr, dvi.ptr, dvo.ptr, n,
                                                   float(smoothingRadius), float(dt),
                                                   float(vorticityStrength), int(stride), None))
    synchronize()
    return dvo.to_numpy(np.float32, (n, 4)), dvi.to_numpy(np.float32, vi.shape)


def gather_solver_state(sortedIndices, positions, predicted, velocities, incoming, stableIds):
    di, idx = _dev(sortedIndices, np.uint32)
    n = len(idx)
    arrays = [_f4dev(a)[0] for a in (positions, predicted, velocities, incoming)]
    dids, _ = _dev(stableIds, np.uint32)
    outs = [_out_f4(n) for _ in range(4)]
    oid = DeviceArray(max(4, 4 * n))
    _check(lib().pbf_k_gather_solver_state(di.ptr, *[a.ptr for a in arrays], dids.ptr,
                                           *[o.ptr for o in outs], oid.ptr, n, None))
    synchronize()
    return [o.to_numpy(np.float32, (n, 4)) for o in outs] + [oid.to_numpy(np.uint32, (n,))]


def initialize_stable_particle_ids(particleCount):
    n = int(particleCount)
    d = DeviceArray(max(4, 4 * n))
    _check(lib().pbf_k_initialize_stable_particle_ids(d.ptr, n, None))
    synchronize()
    return d.to_numpy(np.uint32, (n,))


def scatter_float4_by_stable_id(values, stableIds):
    dv, v = _f4dev(values)
    di, _ = _dev(stableIds, np.uint32)
    n = len(v)
    out = _out_f4(n)
    _check(lib().pbf_k_scatter_float4_by_stable_id(dv.ptr, di.ptr, out.ptr, n, None))
    synchronize()
    return out.to_numpy(np.float32, (n, 4))


def _colliders(items, conv, ctype):
    items = [conv(x) for x in items]
    if not items:
        return DeviceArray(16), 0
    arr = (ctype * len(items))(*items)
    d = DeviceArray(C.sizeof(arr))
    _check(lib().pbf_memcpy_htod(d.ptr, C.addressof(arr), C.sizeof(arr)))
    return d, len(items)


def _container(c):
    if not isinstance(c, Container):
        mn, mx = c
        c = Container(_f3(mn), _f3(mx))
    d = DeviceArray(C.sizeof(c))
    _check(lib().pbf_memcpy_htod(d.ptr, C.addressof(c), C.sizeof(c)))
    return d


def _solve(fn, positions, extra, particleRadius, particleCount):
    dp, p = _f4dev(positions)
    n = len(p) if particleCount is None else int(particleCount)
    flag = DeviceArray(4)
    flag.fill_bytes(0)
    _check(fn(dp.ptr, n, *extra, float(particleRadius), flag.ptr, None))
    synchronize()
    return dp.to_numpy(np.float32, p.shape), int(flag.to_numpy(np.int32, (1,))[0])


def solve_container(positions, container, particleRadius, particleCount=None):
    dc = _container(container)
    return _solve(lib().pbf_k_solve_container, positions, [dc.ptr], particleRadius, particleCount)


def solve_spheres(positions, spheres, particleRadius, particleCount=None):
    d, k = _colliders(spheres, _sphere, SphereCollider)
    return _solve(lib().pbf_k_solve_spheres, positions, [d.ptr, k], particleRadius, particleCount)


def solve_boxes(positions, boxes, particleRadius, particleCount=None):
    d, k = _colliders(boxes, _box, BoxCollider)
    return _solve(lib().pbf_k_solve_boxes, positions, [d.ptr, k], particleRadius, particleCount)


def solve_planes(positions, planes, particleRadius, particleCount=None):
    d, k = _colliders(planes, _plane, PlaneCollider)
    return _solve(lib().pbf_k_solve_planes, positions, [d.ptr, k], particleRadius, particleCount)


def resolve_velocities(positions, velocities, container, spheres=(), boxes=(), planes=(),
                       particleRadius=0.5, restitution=0.0, friction=0.0, particleCount=None,
                       incoming=None):
    dp, p = _f4dev(positions)
    dv, v = _f4dev(velocities)
    din = dv if incoming is None else _f4dev(incoming)[0]
    n = len(p) if particleCount is None else int(particleCount)
    dc = _container(container)
    ds, ns = _colliders(spheres, _sphere, SphereCollider)
    db, nb = _colliders(boxes, _box, BoxCollider)
    dl, nl = _colliders(planes, _plane, PlaneCollider)
    _check(lib().pbf_k_resolve_velocities(dp.ptr, din.ptr, dv.ptr, n, dc.ptr, ds.ptr, ns, db.ptr,
                                          nb, dl.ptr, nl, float(particleRadius),
                                          float(restitution), float(friction), None))
    synchronize()
    return dp.to_numpy(np.float32, p.shape), dv.to_numpy(np.float32, v.shape)
