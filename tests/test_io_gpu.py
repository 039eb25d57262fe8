"""Upload and copy-out through the pinned stage: multi-chunk round trips
(8 MiB chunks of 524288 records), the strong guarantee when a non-finite
value sits in a later chunk, and copy-out straight into pinned memory."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 1_200_001  # three chunks, the last one partial
MIN, MAX = (0.0, 0.0, 0.0), (8.0, 8.0, 8.0)


def _params():
    from pbf_amd import scenes

    return scenes.dam_break_params()


def _records(seed, n, lo, hi, w):
    rng = np.random.default_rng(seed)
    r = rng.uniform(lo, hi, (n, 4)).astype(np.float32)
    r[:, 3] = w
    return r


def test_multichunk_round_trip_and_strong_guarantee(gpu):
    import pbf_amd

    s = pbf_amd.PBFSolver()
    s.initialize(N, MIN, MAX, _params())
    x = _records(1, N, 0.5, 7.5, 1.0)
    v = _records(2, N, -1.0, 1.0, 0.0)
    s.setParticles(x, v)
    assert np.array_equal(s.copyPositionsToHost(), x)
    assert np.array_equal(s.copyVelocitiesToHost(), v)
    print(f"{N} records up and down through the stage", flush=True)

    x2, v2 = x.copy(), v.copy()
    v2[N - 3, 1] = np.nan  # in the third chunk of the second array
    with pytest.raises(pbf_amd.InvalidArgument):
        s.setParticles(x2, v2)
    x2[N - 2, 0] = np.inf  # in the third chunk of the first array
    with pytest.raises(pbf_amd.InvalidArgument):
        s.setParticles(x2, v)
    assert np.array_equal(s.copyPositionsToHost(), x), "positions kept after a rejected upload"
    assert np.array_equal(s.copyVelocitiesToHost(), v), "velocities kept after a rejected upload"


def test_pinned_upload_and_copy_out(gpu):
    import pbf_amd

    n = 700_000
    s = pbf_amd.PBFSolver()
    s.initialize(n, MIN, MAX, _params())
    x = pbf_amd.pinned_empty((n, 4))
    v = pbf_amd.pinned_empty((n, 4))
    x[:] = _records(3, n, 0.5, 7.5, 2.0)
    v[:] = _records(4, n, -1.0, 1.0, 3.0)
    s.setParticles(x, v)
    out = pbf_amd.pinned_empty((n, 4))
    s.copyPositionsToHost(out, n)
    assert np.array_equal(out, x)
    s.copyVelocitiesToHost(out, n)
    assert np.array_equal(out, v)
    bad = pbf_amd.pinned_empty((n, 4))
    bad[:] = x
    bad[n // 2, 2] = np.nan
    with pytest.raises(pbf_amd.InvalidArgument):
        s.setParticles(bad, v)
    assert np.array_equal(s.copyPositionsToHost(), x), "kept after a rejected pinned upload"


def test_host_array_calls_reject_device_memory(gpu):
    """setParticles and the host copy-outs read or write their arrays on the
    host, so a device pointer is an invalid argument (the device-array calls
    take those), and the state is left as it was."""
    import pbf_amd

    n = 1000
    s = pbf_amd.PBFSolver()
    s.initialize(n, MIN, MAX, _params())
    x = _records(5, n, 0.5, 7.5, 1.0)
    v = _records(6, n, -1.0, 1.0, 0.0)
    s.setParticles(x, v)
    d = pbf_amd.DeviceArray.from_host(x)
    L = pbf_amd.lib()
    assert L.pbf_solver_set_particles(s._h, d.ptr, v.ctypes.data, n) == 1
    assert L.pbf_solver_set_particles(s._h, x.ctypes.data, d.ptr, n) == 1
    assert L.pbf_solver_copy_positions_to_host(s._h, d.ptr, n) == 1
    assert L.pbf_solver_copy_velocities_to_host(s._h, d.ptr, n) == 1
    assert np.array_equal(s.copyPositionsToHost(), x)
    assert np.array_equal(s.copyVelocitiesToHost(), v)
    s.setParticlesDevice(d, pbf_amd.DeviceArray.from_host(v), n)
    out = pbf_amd.DeviceArray(16 * n)
    s.copyVelocitiesToDevice(out, n)
    assert np.array_equal(out.to_numpy(np.float32, (n, 4)), v)
