"""pbf_grid_build_static (the grid build that allocates nothing) against
pbf_grid_build through the C ABI: bit-exact sorted keys, sorted indices,
cellStart and cellEnd on two calls in a row, the second with a smaller
input (the static scratch keeps the first call's contents), plus the size
and argument checks that run before any device call."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MIN, MAX, CS = (0.0, 0.0, 0.0), (12.8, 6.4, 6.4), 0.1  # 128 x 64 x 64 cells


def _positions(seed, n):
    rng = np.random.default_rng(seed)
    p = np.empty((n, 4), np.float32)
    # Mostly inside the grid, some outside every face (clamped keys).
    p[:, 0] = rng.uniform(-0.5, 13.3, n)
    p[:, 1] = rng.uniform(-0.5, 6.9, n)
    p[:, 2] = rng.uniform(-0.5, 6.9, n)
    p[:, 3] = 1.0
    return p


def _static_build(d_pos, n, cells, stream=None):
    import pbf_amd

    sk = pbf_amd.DeviceArray(max(4 * n, 4))
    si = pbf_amd.DeviceArray(max(4 * n, 4))
    cs = pbf_amd.DeviceArray(4 * cells)
    ce = pbf_amd.DeviceArray(4 * cells)
    st = pbf_amd.lib().pbf_grid_build_static(
        d_pos.ptr, n, pbf_amd.make_float3(*MIN), pbf_amd.make_float3(*MAX), C.c_float(CS),
        sk.ptr, si.ptr, cs.ptr, ce.ptr, cells, stream)
    pbf_amd._check(st)
    pbf_amd.synchronize()
    return (sk.to_numpy(np.uint32, (n,)), si.to_numpy(np.uint32, (n,)),
            cs.to_numpy(np.int32, (cells,)), ce.to_numpy(np.int32, (cells,)))


def test_static_build_equals_grid_build_on_two_calls(gpu):
    import pbf_amd

    g = pbf_amd.SpatialGrid()
    g.initialize(1_048_576, MIN, MAX, CS)
    cells = g.numCells()
    assert cells == 128 * 64 * 64
    for seed, n in ((11, 900_001), (12, 70_001)):
        p = _positions(seed, n)
        d = pbf_amd.DeviceArray.from_host(p)
        print(f"static grid build: seed {seed}, {n} particles, {cells} cells", flush=True)
        g.build(d, n)
        expect = g.read(n)
        got = _static_build(d, n, cells)
        for name, a, b in zip(("sortedKeys", "sortedIndices", "cellStart", "cellEnd"), got,
                              expect):
            assert np.array_equal(a, b), f"{name} differs (seed {seed}, n {n})"


def test_static_build_on_a_stream_and_at_the_maximum(gpu):
    import pbf_amd

    n = 1_048_576  # PBF_GRID_BUILD_STATIC_MAX_PARTICLES
    g = pbf_amd.SpatialGrid()
    g.initialize(n, MIN, MAX, CS)
    cells = g.numCells()
    p = _positions(13, n)
    d = pbf_amd.DeviceArray.from_host(p)
    print(f"static grid build at the maximum: {n} particles", flush=True)
    g.build(d, n)
    expect = g.read(n)
    s = C.c_void_p()
    pbf_amd._check(pbf_amd.lib().pbf_stream_create(C.byref(s)))
    try:
        got = _static_build(d, n, cells, s.value)
    finally:
        pbf_amd.lib().pbf_stream_destroy(s.value)
    for a, b in zip(got, expect):
        assert np.array_equal(a, b)


def test_static_build_rejects_sizes_and_arguments(gpu):
    import pbf_amd

    L = pbf_amd.lib()
    mn, mx = pbf_amd.make_float3(*MIN), pbf_amd.make_float3(*MAX)
    cells = 128 * 64 * 64
    big = 1_048_577
    assert L.pbf_grid_build_static(None, big, mn, mx, C.c_float(CS), None, None, None, None,
                                   cells, None) == 4  # PBF_ERR_LENGTH
    assert b"PBF_GRID_BUILD_STATIC_MAX_PARTICLES" in L.pbf_last_error()
    assert L.pbf_grid_build_static(None, 10, mn, mx, C.c_float(0.0), None, None, None, None,
                                   cells, None) == 1  # cell size
    d = pbf_amd.DeviceArray(16 * 10)
    cs = pbf_amd.DeviceArray(4 * cells)
    assert L.pbf_grid_build_static(d.ptr, 10, mn, mx, C.c_float(CS), d.ptr, d.ptr, cs.ptr,
                                   cs.ptr, cells - 1, None) == 3  # cell capacity
    # An empty build clears the ranges, as pbf_grid_build does.
    cs.fill_bytes(0)
    ce = pbf_amd.DeviceArray(4 * cells)
    ce.fill_bytes(0)
    pbf_amd._check(L.pbf_grid_build_static(None, 0, mn, mx, C.c_float(CS), None, None, cs.ptr,
                                           ce.ptr, cells, None))
    pbf_amd.synchronize()
    assert np.all(cs.to_numpy(np.int32, (cells,)) == -1)
    assert np.all(ce.to_numpy(np.int32, (cells,)) == -1)
