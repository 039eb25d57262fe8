"""Python binding of libpbf_amd.so, the gfx950 Position Based Fluids library.

The classes mirror the reference C++ API of DusanCojic/Fluid-Simulator
(include/pbf/pbf_solver.hpp, spatial_grid.cuh, collision_system.hpp): same
member names, argument meaning and error behaviour.  C++ exception types map
onto Python exceptions:

    std::invalid_argument -> InvalidArgument (ValueError)
    std::logic_error      -> LogicError (RuntimeError)
    std::out_of_range     -> OutOfRange (IndexError)
    std::length_error     -> LengthError (ValueError)
    std::overflow_error   -> NeighborOverflow (OverflowError)
    std::runtime_error    -> CollisionNonConvergence / DeviceError (RuntimeError)

Float4 arrays are numpy float32 arrays of shape (n, 4).  Everything runs
through the HIP library; there is no CPU path, and importing fails loudly
when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

__all__ = [
    "lib", "LIB_PATH", "float3", "int3", "SimulationParams", "Container", "SphereCollider",
    "BoxCollider", "PlaneCollider", "make_float3", "Tuning", "DeviceArray", "PBFSolver",
    "SpatialGrid", "PBFBatchSolver", "PBFAdjoint", "pinned_empty",
    "CollisionSystem", "kernels", "PbfError", "InvalidArgument", "LogicError", "OutOfRange",
    "LengthError", "NeighborOverflow", "CollisionNonConvergence", "DeviceError",
    "STATUS_NEIGHBOR_OVERFLOW", "STATUS_COLLISION_NONCONVERGENCE", "STATUS_LIST_STALE",
]

LIB_PATH = Path(__file__).resolve().parent / "libpbf_amd.so"

STATUS_NEIGHBOR_OVERFLOW = 0x1
STATUS_COLLISION_NONCONVERGENCE = 0x2
STATUS_LIST_STALE = 0x4


class PbfError(Exception):
    """Base class of the library's errors."""


class InvalidArgument(PbfError, ValueError):
    pass


class LogicError(PbfError, RuntimeError):
    pass


class OutOfRange(PbfError, IndexError):
    pass


class LengthError(PbfError, ValueError):
    pass


class NeighborOverflow(PbfError, OverflowError):
    pass


class CollisionNonConvergence(PbfError, RuntimeError):
    pass


class DeviceError(PbfError, RuntimeError):
    pass


_ERRORS = {1: InvalidArgument, 2: LogicError, 3: OutOfRange, 4: LengthError,
           5: NeighborOverflow, 6: CollisionNonConvergence, 7: DeviceError}


# --------------------------------------------------------------------------
# Plain data types (layouts of pbf_amd.h)
# --------------------------------------------------------------------------
class float3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]

    def __iter__(self):
        return iter((self.x, self.y, self.z))

    def __repr__(self):
        return f"float3({self.x}, {self.y}, {self.z})"


class int3(C.Structure):
    _fields_ = [("x", C.c_int), ("y", C.c_int), ("z", C.c_int)]

    def __iter__(self):
        return iter((self.x, self.y, self.z))


def make_float3(x, y, z) -> float3:
    return float3(float(x), float(y), float(z))


def _f3(v) -> float3:
    if isinstance(v, float3):
        return v
    x, y, z = v
    return float3(float(x), float(y), float(z))


class SimulationParams(C.Structure):
    """SimulationParams (include/pbf/simulation_params.hpp:6-29), 72 bytes."""
    _fields_ = [("dt", C.c_float), ("restDensity", C.c_float), ("particleMass", C.c_float),
                ("particleRadius", C.c_float), ("collisionRestitution", C.c_float),
                ("collisionFriction", C.c_float), ("smoothingRadius", C.c_float),
                ("lambdaEpsilon", C.c_float), ("solverIterations", C.c_int),
                ("substeps", C.c_int), ("gravity", float3), ("scorrK", C.c_float),
                ("scorrN", C.c_int), ("scorrDeltaQ", C.c_float), ("xsphViscosity", C.c_float),
                ("vorticityStrength", C.c_float)]


class Container(C.Structure):
    _fields_ = [("min", float3), ("max", float3)]


class SphereCollider(C.Structure):
    _fields_ = [("center", float3), ("radius", C.c_float)]


class BoxCollider(C.Structure):
    _fields_ = [("center", float3), ("halfExtents", float3)]


class PlaneCollider(C.Structure):
    _fields_ = [("point", float3), ("normal", float3)]


assert C.sizeof(SimulationParams) == 72
assert C.sizeof(Container) == 24 and C.sizeof(SphereCollider) == 16
assert C.sizeof(BoxCollider) == 24 and C.sizeof(PlaneCollider) == 24


def _sphere(s):
    if isinstance(s, SphereCollider):
        return s
    c, r = s
    return SphereCollider(_f3(c), float(r))


def _box(b):
    if isinstance(b, BoxCollider):
        return b
    c, h = b
    return BoxCollider(_f3(c), _f3(h))


def _plane(p):
    if isinstance(p, PlaneCollider):
        return p
    pt, nrm = p
    return PlaneCollider(_f3(pt), _f3(nrm))


# --------------------------------------------------------------------------
# Library loading and signatures
# --------------------------------------------------------------------------
_vp = C.c_void_p
_sz = C.c_size_t
_f = C.c_float
_i = C.c_int
_st = C.c_int

_SIGNATURES = {
    "pbf_last_error": (C.c_char_p, []),
    "pbf_version": (C.c_char_p, []),
    "pbf_device_alloc": (_st, [C.POINTER(_vp), _sz]),
    "pbf_device_free": (_st, [_vp]),
    "pbf_host_alloc": (_st, [C.POINTER(_vp), _sz]),
    "pbf_host_free": (_st, [_vp]),
    "pbf_memcpy_htod": (_st, [_vp, _vp, _sz]),
    "pbf_memcpy_dtoh": (_st, [_vp, _vp, _sz]),
    "pbf_memcpy_dtod": (_st, [_vp, _vp, _sz]),
    "pbf_memset": (_st, [_vp, _i, _sz]),
    "pbf_device_synchronize": (_st, []),
    "pbf_stream_create": (_st, [C.POINTER(_vp)]),
    "pbf_stream_destroy": (_st, [_vp]),
    "pbf_stream_synchronize": (_st, [_vp]),
    "pbf_k_integrate": (_st, [_vp, _vp, _vp, _i, _f, float3, _vp]),
    "pbf_k_find_neighbors": (_st, [_vp, _vp, _vp, int3, float3, _f, _sz, _f, _vp, _vp, _i, _sz,
                                   _vp, _vp]),
    "pbf_k_compute_lambda": (_st, [_vp, _vp, _vp, _i, _sz, _f, _f, _f, _f, _vp, _sz, _vp]),
    "pbf_k_compute_delta_position": (_st, [_vp, _vp, _vp, _i, _vp, _sz, _f, _f, _f, _f, _i, _f,
                                           _vp, _sz, _vp]),
    "pbf_k_apply_delta_position": (_st, [_vp, _vp, _sz, _vp]),
    "pbf_k_update_velocity_and_position": (_st, [_vp, _vp, _vp, _sz, _f, _vp]),
    "pbf_k_apply_xsph_viscosity": (_st, [_vp, _vp, _vp, _i, _vp, _vp, _sz, _f, _f, _sz, _vp]),
    "pbf_k_compute_vorticity": (_st, [_vp, _vp, _vp, _vp, _i, _sz, _f, _vp, _sz, _vp]),
    "pbf_k_apply_vorticity_confinement": (_st, [_vp, _vp, _vp, _i, _vp, _vp, _vp, _sz, _f, _f, _f,
                                                _sz, _vp]),
    "pbf_k_gather_solver_state": (_st, [_vp] * 11 + [_sz, _vp]),
    "pbf_k_initialize_stable_particle_ids": (_st, [_vp, _sz, _vp]),
    "pbf_k_scatter_float4_by_stable_id": (_st, [_vp, _vp, _vp, _sz, _vp]),
    "pbf_k_solve_container": (_st, [_vp, _sz, _vp, _f, _vp, _vp]),
    "pbf_k_solve_spheres": (_st, [_vp, _sz, _vp, _sz, _f, _vp, _vp]),
    "pbf_k_solve_boxes": (_st, [_vp, _sz, _vp, _sz, _f, _vp, _vp]),
    "pbf_k_solve_planes": (_st, [_vp, _sz, _vp, _sz, _f, _vp, _vp]),
    "pbf_k_resolve_velocities": (_st, [_vp, _vp, _vp, _sz, _vp, _vp, _sz, _vp, _sz, _vp, _sz,
                                       _f, _f, _f, _vp]),
    "pbf_k_poly6": (_st, [_vp, _sz, _f, _vp, _vp]),
    "pbf_k_spiky_gradient": (_st, [_vp, _sz, _f, _vp, _vp]),
    "pbf_k_artificial_pressure": (_st, [_vp, _sz, _f, _f, _i, _f, _vp, _vp]),
    "pbf_grid_create": (_st, [C.POINTER(_vp)]),
    "pbf_grid_destroy": (_st, [_vp]),
    "pbf_grid_initialize": (_st, [_vp, _sz, float3, float3, _f]),
    "pbf_grid_build": (_st, [_vp, _vp, _sz]),
    "pbf_grid_rebuild": (_st, [_vp, _vp, _sz]),
    "pbf_grid_build_static": (_st, [_vp, _sz, float3, float3, _f, _vp, _vp, _vp, _vp, _sz, _vp]),
    "pbf_grid_sorted_keys": (_vp, [_vp]),
    "pbf_grid_sorted_indices": (_vp, [_vp]),
    "pbf_grid_cell_start": (_vp, [_vp]),
    "pbf_grid_cell_end": (_vp, [_vp]),
    "pbf_grid_info": (_st, [_vp, C.POINTER(int3), C.POINTER(_sz), C.POINTER(_sz),
                            C.POINTER(_f), C.POINTER(float3), C.POINTER(float3)]),
    "pbf_collision_create": (_st, [C.POINTER(_vp)]),
    "pbf_collision_destroy": (_st, [_vp]),
    "pbf_collision_set_container": (_st, [_vp, Container]),
    "pbf_collision_set_spheres": (_st, [_vp, _vp, _sz]),
    "pbf_collision_clear_spheres": (_st, [_vp]),
    "pbf_collision_set_boxes": (_st, [_vp, _vp, _sz]),
    "pbf_collision_clear_boxes": (_st, [_vp]),
    "pbf_collision_set_planes": (_st, [_vp, _vp, _sz]),
    "pbf_collision_clear_planes": (_st, [_vp]),
    "pbf_collision_solve": (_st, [_vp, _vp, _sz, _f]),
    "pbf_collision_resolve_velocities": (_st, [_vp, _vp, _vp, _vp, _sz, _f, _f, _f]),
    "pbf_solver_create": (_st, [C.POINTER(_vp)]),
    "pbf_solver_destroy": (_st, [_vp]),
    "pbf_solver_initialize": (_st, [_vp, _sz, float3, float3, C.POINTER(SimulationParams)]),
    "pbf_solver_set_particles": (_st, [_vp, _vp, _vp, _sz]),
    "pbf_solver_set_particles_device": (_st, [_vp, _vp, _vp, _sz]),
    "pbf_solver_set_spheres": (_st, [_vp, _vp, _sz]),
    "pbf_solver_set_boxes": (_st, [_vp, _vp, _sz]),
    "pbf_solver_set_planes": (_st, [_vp, _vp, _sz]),
    "pbf_solver_step": (_st, [_vp]),
    "pbf_solver_run_one": (_st, [_vp]),
    "pbf_solver_run": (_st, [_vp, _sz]),
    "pbf_solver_copy_positions_to_host": (_st, [_vp, _vp, _sz]),
    "pbf_solver_copy_velocities_to_host": (_st, [_vp, _vp, _sz]),
    "pbf_solver_copy_positions_to_device": (_st, [_vp, _vp, _sz]),
    "pbf_solver_copy_velocities_to_device": (_st, [_vp, _vp, _sz]),
    "pbf_solver_counts": (_st, [_vp, C.POINTER(_sz), C.POINTER(_sz)]),
    "pbf_solver_stream": (_vp, [_vp]),
    "pbf_solver_set_profiling": (_st, [_vp, _i]),
    "pbf_solver_set_graphs": (_st, [_vp, _i]),
    "pbf_solver_set_incremental_sort": (_st, [_vp, _i]),
    "pbf_solver_set_fused_xsph": (_st, [_vp, _i]),
    "pbf_solver_set_neighbor_skin": (_st, [_vp, C.c_float]),
    "pbf_solver_neighbor_skin": (_st, [_vp, C.POINTER(C.c_float), C.POINTER(_sz)]),
    "pbf_solver_set_tuning": (_st, [_vp, _vp]),
    "pbf_solver_get_tuning": (_st, [_vp, _vp]),
    "pbf_solver_autotune": (_st, [_vp, _sz, _vp, _sz, _vp, _vp]),
    "pbf_solver_kernel_times": (_st, [_vp, C.c_char_p, _sz, C.POINTER(C.c_double),
                                      C.POINTER(C.c_uint64), _sz, C.POINTER(_sz)]),
    "pbf_solver_neighbor_lists": (_st, [_vp, _vp, _vp, _vp, _sz, _i]),
    "pbf_adjoint_create": (_st, [_vp, C.POINTER(_vp)]),
    "pbf_adjoint_destroy": (_st, [_vp]),
    "pbf_adjoint_forward": (_st, [_vp, _sz, _sz]),
    "pbf_adjoint_backward": (_st, [_vp, _vp, _vp, _sz, _vp, _vp, C.POINTER(C.c_double),
                                   C.POINTER(C.c_double)]),
    "pbf_batch_create": (_st, [C.POINTER(_vp)]),
    "pbf_batch_destroy": (_st, [_vp]),
    "pbf_batch_initialize": (_st, [_vp, _vp, _sz, _vp]),
    "pbf_batch_set_particles": (_st, [_vp, _vp, _vp, _sz]),
    "pbf_batch_step": (_st, [_vp]),
    "pbf_batch_run": (_st, [_vp, _sz]),
    "pbf_batch_copy_positions_to_host": (_st, [_vp, _vp, _sz]),
    "pbf_batch_copy_velocities_to_host": (_st, [_vp, _vp, _sz]),
    "pbf_batch_counts": (_st, [_vp, C.POINTER(_sz), C.POINTER(_sz)]),
    "pbf_batch_set_graphs": (_st, [_vp, _i]),
    "pbf_batch_set_incremental_sort": (_st, [_vp, _i]),
    "pbf_batch_set_fused_xsph": (_st, [_vp, _i]),
    "pbf_batch_set_neighbor_skin": (_st, [_vp, C.c_float]),
    "pbf_batch_stale_frames": (_st, [_vp, C.POINTER(_sz)]),
    "pbf_batch_set_tuning": (_st, [_vp, _vp]),
    "pbf_batch_autotune": (_st, [_vp, _sz, _vp, _sz, _vp, _vp]),
    "pbf_batch_set_profiling": (_st, [_vp, _i]),
    "pbf_batch_kernel_times": (_st, [_vp, C.c_char_p, _sz, C.POINTER(C.c_double),
                                     C.POINTER(C.c_uint64), _sz, C.POINTER(_sz)]),
}

_LIB = None


def lib() -> C.CDLL:
    """The loaded libpbf_amd.so; raises if the HIP library was not built."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build the HIP library first "
                "(python fluid-simulator_amd/build.py or __graft_entry__.build())")
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = handle
    return _LIB


def declared_symbols() -> list[str]:
    return sorted(_SIGNATURES)


def _check(status: int) -> None:
    if status != 0:
        msg = lib().pbf_last_error()
        raise _ERRORS.get(status, DeviceError)(msg.decode() if msg else f"status {status}")


def _ptr(a):
    """Device pointer of a DeviceArray / int, or host pointer of a numpy array."""
    if a is None:
        return None
    if isinstance(a, DeviceArray):
        return a.ptr
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    if isinstance(a, int):
        return a
    raise TypeError(f"cannot take a pointer of {type(a)}")


def as_float4(a, n=None) -> np.ndarray:
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    if arr.ndim == 1:
        arr = arr.reshape(-1, 4)
    if arr.shape[-1] != 4:
        raise InvalidArgument("float4 arrays must have shape (n, 4)")
    return arr


# --------------------------------------------------------------------------
# Pinned host memory
# --------------------------------------------------------------------------
class Tuning(C.Structure):
    """pbf_tuning: the skin (cell-size multiplier 1 + neighborSkin), the
    workgroup size of the neighbour passes and their particles per thread."""
    _fields_ = [("neighborSkin", C.c_float), ("passBlock", C.c_int),
                ("particlesPerThread", C.c_int)]

    def __repr__(self):
        return (f"Tuning(neighborSkin={self.neighborSkin:.3g}, passBlock={self.passBlock}, "
                f"particlesPerThread={self.particlesPerThread})")


DEFAULT_TUNING_COUNT = 24
DEFAULT_TUNINGS = [(s, b, p) for s in (0.0, 0.1, 0.2, 0.3) for b in (64, 128, 256) for p in (1, 2)]


def _tuning(t) -> Tuning:
    """A Tuning from (skin, block) or (skin, block, particles per thread)."""
    return Tuning(float(t[0]), int(t[1]), int(t[2]) if len(t) > 2 else 1)


def _autotune(fn, handle, frames, candidates):
    if candidates:
        arr = (Tuning * len(candidates))(*[_tuning(c) for c in candidates])
        count, ptr = len(candidates), C.cast(arr, _vp)
    else:
        count, ptr = 0, None
    ms = (C.c_double * max(count, DEFAULT_TUNING_COUNT))()
    chosen = Tuning()
    _check(fn(handle, frames, ptr, count, C.cast(ms, _vp), C.cast(C.pointer(chosen), _vp)))
    tried = [_tuning(c) for c in (candidates or DEFAULT_TUNINGS)]
    return chosen, [(t.neighborSkin, t.passBlock, t.particlesPerThread, ms[k])
                    for k, t in enumerate(tried)]


class _PinnedBlock:
    def __init__(self, nbytes: int):
        p = _vp()
        _check(lib().pbf_host_alloc(C.byref(p), max(int(nbytes), 1)))
        self.ptr = p.value

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().pbf_host_free(self.ptr)
            self.ptr = None


def pinned_empty(shape, dtype=np.float32) -> np.ndarray:
    """A numpy array in pinned (page-locked) host memory: the solver's
    uploads from and copy-outs into it run as one DMA each.  The allocation
    is freed with the last array that views it."""
    dtype = np.dtype(dtype)
    count = int(np.prod(shape))
    nbytes = max(count * dtype.itemsize, 1)
    block = _PinnedBlock(nbytes)
    buf = (C.c_char * nbytes).from_address(block.ptr)
    buf._block = block  # the buffer, and every array over it, keeps the block alive
    return np.frombuffer(buf, dtype=dtype, count=count).reshape(shape)

# --------------------------------------------------------------------------
# Device memory
# --------------------------------------------------------------------------
class DeviceArray:
    """An owned device allocation (CudaBuffer<T>, include/pbf/cuda_buffer.hpp)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = _vp()
        _check(lib().pbf_device_alloc(C.byref(p), self.nbytes))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a: np.ndarray, min_bytes: int = 0) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(max(a.nbytes, min_bytes, 4))
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise OutOfRange("Source is larger than device buffer")
        _check(lib().pbf_memcpy_htod(self.ptr, a.ctypes.data, a.nbytes))

    def download(self, out: np.ndarray) -> np.ndarray:
        if not out.flags.c_contiguous:
            raise InvalidArgument("destination must be contiguous")
        if out.nbytes > self.nbytes:
            raise OutOfRange("Copy exceeds device buffer size")
        _check(lib().pbf_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def to_numpy(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        return self.download(out)

    def fill_bytes(self, value: int) -> None:
        _check(lib().pbf_memset(self.ptr, int(value), self.nbytes))

    def free(self) -> None:
        if getattr(self, "ptr", None):
            lib().pbf_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize() -> None:
    _check(lib().pbf_device_synchronize())


def read_device(ptr: int, dtype, shape) -> np.ndarray:
    out = np.empty(shape, dtype=dtype)
    if out.nbytes:
        _check(lib().pbf_memcpy_dtoh(out.ctypes.data, ptr, out.nbytes))
    return out


# --------------------------------------------------------------------------
# SpatialGrid
# --------------------------------------------------------------------------
class SpatialGrid:
    """SpatialGrid (include/pbf/spatial_grid.cuh:11-62)."""

    def __init__(self):
        h = _vp()
        _check(lib().pbf_grid_create(C.byref(h)))
        self._h = h.value

    def __del__(self):
        if getattr(self, "_h", None):
            lib().pbf_grid_destroy(self._h)
            self._h = None

    def initialize(self, maxParticles, minBounds, maxBounds, cellSize):
        _check(lib().pbf_grid_initialize(self._h, int(maxParticles), _f3(minBounds),
                                         _f3(maxBounds), float(cellSize)))

    def build(self, predictedPositions, particleCount):
        _check(lib().pbf_grid_build(self._h, _ptr(predictedPositions), int(particleCount)))

    def rebuild(self, predictedPositions, particleCount):
        """build's results, re-sorted from the last build (fast when the
        positions are its particles in its sorted order, moved a little)."""
        _check(lib().pbf_grid_rebuild(self._h, _ptr(predictedPositions), int(particleCount)))

    def _info(self):
        g, nc, mp, cs, mn, mx = int3(), _sz(), _sz(), _f(), float3(), float3()
        _check(lib().pbf_grid_info(self._h, C.byref(g), C.byref(nc), C.byref(mp), C.byref(cs),
                                   C.byref(mn), C.byref(mx)))
        return g, nc.value, mp.value, cs.value, mn, mx

    def sortedKeys(self):
        return lib().pbf_grid_sorted_keys(self._h)

    def sortedIndices(self):
        return lib().pbf_grid_sorted_indices(self._h)

    def cellStart(self):
        return lib().pbf_grid_cell_start(self._h)

    def cellEnd(self):
        return lib().pbf_grid_cell_end(self._h)

    def gridSize(self):
        return self._info()[0]

    def numCells(self):
        return self._info()[1]

    def maxParticles(self):
        return self._info()[2]

    def cellSize(self):
        return self._info()[3]

    def minBounds(self):
        return self._info()[4]

    def maxBounds(self):
        return self._info()[5]

    def read(self, n):
        """(sortedKeys, sortedIndices, cellStart, cellEnd) as numpy arrays."""
        cells = self.numCells()
        return (read_device(self.sortedKeys(), np.uint32, (n,)),
                read_device(self.sortedIndices(), np.uint32, (n,)),
                read_device(self.cellStart(), np.int32, (cells,)),
                read_device(self.cellEnd(), np.int32, (cells,)))


# --------------------------------------------------------------------------
# CollisionSystem
# --------------------------------------------------------------------------
def _struct_array(items, conv, ctype):
    items = [conv(x) for x in items]
    arr = (ctype * max(1, len(items)))(*items) if items else None
    return arr, len(items)


class CollisionSystem:
    """CollisionSystem (include/pbf/collision_system.hpp:11-57)."""

    def __init__(self, container=None):
        h = _vp()
        _check(lib().pbf_collision_create(C.byref(h)))
        self._h = h.value
        if container is not None:
            self.setContainer(container)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().pbf_collision_destroy(self._h)
            self._h = None

    def setContainer(self, container):
        if not isinstance(container, Container):
            mn, mx = container
            container = Container(_f3(mn), _f3(mx))
        _check(lib().pbf_collision_set_container(self._h, container))

    def setSpheres(self, spheres):
        arr, n = _struct_array(spheres, _sphere, SphereCollider)
        _check(lib().pbf_collision_set_spheres(self._h, arr, n))

    def clearSpheres(self):
        _check(lib().pbf_collision_clear_spheres(self._h))

    def setBoxes(self, boxes):
        arr, n = _struct_array(boxes, _box, BoxCollider)
        _check(lib().pbf_collision_set_boxes(self._h, arr, n))

    def clearBoxes(self):
        _check(lib().pbf_collision_clear_boxes(self._h))

    def setPlanes(self, planes):
        arr, n = _struct_array(planes, _plane, PlaneCollider)
        _check(lib().pbf_collision_set_planes(self._h, arr, n))

    def clearPlanes(self):
        _check(lib().pbf_collision_clear_planes(self._h))

    def solve(self, predictedPositions, particleCount, particleRadius):
        _check(lib().pbf_collision_solve(self._h, _ptr(predictedPositions), int(particleCount),
                                         float(particleRadius)))

    def resolveVelocities(self, positions, *args):
        """resolveVelocities(positions, velocities, n, r, e, f) or
        resolveVelocities(positions, incoming, velocities, n, r, e, f)."""
        if len(args) == 5:
            velocities, n, r, e, f = args
            incoming = velocities
        else:
            incoming, velocities, n, r, e, f = args
        _check(lib().pbf_collision_resolve_velocities(self._h, _ptr(positions), _ptr(incoming),
                                                      _ptr(velocities), int(n), float(r),
                                                      float(e), float(f)))


# --------------------------------------------------------------------------
# PBFSolver
# --------------------------------------------------------------------------
class PBFSolver:
    """PBFSolver (include/pbf/pbf_solver.hpp:13-70)."""

    def __init__(self):
        h = _vp()
        _check(lib().pbf_solver_create(C.byref(h)))
        self._h = h.value

    def __del__(self):
        if getattr(self, "_h", None):
            lib().pbf_solver_destroy(self._h)
            self._h = None

    def initialize(self, maxParticles, minBounds, maxBounds, params: SimulationParams):
        _check(lib().pbf_solver_initialize(self._h, int(maxParticles), _f3(minBounds),
                                           _f3(maxBounds), C.byref(params)))

    def setParticles(self, positions, velocities, particleCount=None):
        pos = None if positions is None else as_float4(positions)
        vel = None if velocities is None else as_float4(velocities)
        n = particleCount if particleCount is not None else (0 if pos is None else len(pos))
        _check(lib().pbf_solver_set_particles(self._h, _ptr(pos), _ptr(vel), int(n)))

    def setParticlesDevice(self, positions, velocities, particleCount):
        _check(lib().pbf_solver_set_particles_device(self._h, _ptr(positions), _ptr(velocities),
                                                     int(particleCount)))

    def setSpheres(self, spheres):
        arr, n = _struct_array(spheres, _sphere, SphereCollider)
        _check(lib().pbf_solver_set_spheres(self._h, arr, n))

    def setBoxes(self, boxes):
        arr, n = _struct_array(boxes, _box, BoxCollider)
        _check(lib().pbf_solver_set_boxes(self._h, arr, n))

    def setPlanes(self, planes):
        arr, n = _struct_array(planes, _plane, PlaneCollider)
        _check(lib().pbf_solver_set_planes(self._h, arr, n))

    def step(self):
        _check(lib().pbf_solver_step(self._h))

    def run(self, frameCount=None):
        if frameCount is None:
            _check(lib().pbf_solver_run_one(self._h))
        else:
            _check(lib().pbf_solver_run(self._h, int(frameCount)))

    def _copy(self, fn, out, particleCount):
        n = int(particleCount) if particleCount is not None else self.particleCount()
        if out is None and n != 0 and particleCount is not None and particleCount < 0:
            raise InvalidArgument("negative count")
        if out is None:
            out_arr = np.zeros((n, 4), dtype=np.float32) if n else None
        else:
            out_arr = out
        _check(fn(self._h, _ptr(out_arr), n))
        return out_arr

    def copyPositionsToHost(self, out=None, particleCount=None):
        return self._copy(lib().pbf_solver_copy_positions_to_host, out, particleCount)

    def copyVelocitiesToHost(self, out=None, particleCount=None):
        return self._copy(lib().pbf_solver_copy_velocities_to_host, out, particleCount)

    def copyPositionsToDevice(self, out: "DeviceArray", particleCount):
        _check(lib().pbf_solver_copy_positions_to_device(self._h, _ptr(out), int(particleCount)))

    def copyVelocitiesToDevice(self, out: "DeviceArray", particleCount):
        _check(lib().pbf_solver_copy_velocities_to_device(self._h, _ptr(out), int(particleCount)))

    def particleCount(self) -> int:
        n, m = _sz(), _sz()
        _check(lib().pbf_solver_counts(self._h, C.byref(n), C.byref(m)))
        return n.value

    def maxParticles(self) -> int:
        n, m = _sz(), _sz()
        _check(lib().pbf_solver_counts(self._h, C.byref(n), C.byref(m)))
        return m.value

    def stream(self) -> int:
        return lib().pbf_solver_stream(self._h) or 0

    def setGraphs(self, enabled: bool):
        """Replay frames from captured HIP graphs, or launch eagerly (default)."""
        _check(lib().pbf_solver_set_graphs(self._h, 1 if enabled else 0))

    def setIncrementalSort(self, enabled: bool):
        """Grid builds by re-sort from the previous build (default) or by
        full radix sort; the results are bitwise the same."""
        _check(lib().pbf_solver_set_incremental_sort(self._h, 1 if enabled else 0))

    def setFusedXsph(self, enabled: bool):
        """XSPH inside the post-solve grid build (default) or as a pass of its
        own; the results are bitwise the same."""
        _check(lib().pbf_solver_set_fused_xsph(self._h, 1 if enabled else 0))

    def setNeighborSkin(self, skin: float):
        """Skin lists: one grid build per substep, radius h * (1 + skin); 0 runs
        the reference's build before every iteration and after the solve."""
        _check(lib().pbf_solver_set_neighbor_skin(self._h, float(skin)))

    def neighborSkin(self) -> float:
        sk, st = C.c_float(), _sz()
        _check(lib().pbf_solver_neighbor_skin(self._h, C.byref(sk), C.byref(st)))
        return sk.value

    def staleFrames(self) -> int:
        """Frames run again with the reference's builds (stale skin lists)."""
        sk, st = C.c_float(), _sz()
        _check(lib().pbf_solver_neighbor_skin(self._h, C.byref(sk), C.byref(st)))
        return st.value

    def setTuning(self, neighborSkin: float, passBlock: int = 256, particlesPerThread: int = 1):
        t = Tuning(float(neighborSkin), int(passBlock), int(particlesPerThread))
        _check(lib().pbf_solver_set_tuning(self._h, C.cast(C.pointer(t), _vp)))

    def tuning(self) -> Tuning:
        t = Tuning()
        _check(lib().pbf_solver_get_tuning(self._h, C.cast(C.pointer(t), _vp)))
        return t

    def autotune(self, frames: int, candidates=None):
        """Times `frames` frames of each (skin, passBlock[, particlesPerThread])
        candidate (default set when None) from the current state and keeps
        the fastest.  Returns (chosen Tuning, [(skin, passBlock,
        particlesPerThread, ms per frame), ...])."""
        return _autotune(lib().pbf_solver_autotune, self._h, frames, candidates)

    def setProfiling(self, enabled: bool):
        _check(lib().pbf_solver_set_profiling(self._h, 1 if enabled else 0))

    def kernelTimes(self) -> dict:
        cap = 64
        names = C.create_string_buffer(8192)
        ms = (C.c_double * cap)()
        launches = (C.c_uint64 * cap)()
        count = _sz()
        _check(lib().pbf_solver_kernel_times(self._h, names, 8192, ms, launches, cap,
                                             C.byref(count)))
        keys = names.value.decode().split("\n")[:count.value]
        return {k: (ms[i], launches[i]) for i, k in enumerate(keys)}

    def neighborLists(self, maxNeighbors: int = 256):
        """(positions, counts, neighbours) of the last grid build: working-set
        order, neighbours slot-major as (maxNeighbors, n)."""
        n = self.particleCount()
        pos = np.zeros((n, 4), np.float32)
        cnt = np.zeros(n, np.int32)
        nbr = np.zeros((maxNeighbors, n), np.uint32)
        _check(lib().pbf_solver_neighbor_lists(self._h, pos.ctypes.data, cnt.ctypes.data,
                                               nbr.ctypes.data, n, int(maxNeighbors)))
        return pos, cnt, nbr


from . import kernels  # noqa: E402  (device kernel wrappers need the classes above)


class _batch_scene(C.Structure):
    _fields_ = [("particleCount", C.c_size_t), ("minBounds", float3), ("maxBounds", float3)]


class PBFBatchSolver:
    """Many independent scenes advanced together (pbf_batch_*): each scene
    keeps its own grid and container, the parameters are shared, and every
    stage of the step is one launch for all scenes.  Each scene's result is
    bitwise that of a PBFSolver holding it alone."""

    def __init__(self):
        h = _vp()
        _check(lib().pbf_batch_create(C.byref(h)))
        self._h = h.value

    def __del__(self):
        if getattr(self, "_h", None):
            lib().pbf_batch_destroy(self._h)
            self._h = None

    def initialize(self, scenes, params: SimulationParams):
        """scenes: (particleCount, minBounds, maxBounds) triples."""
        arr = (_batch_scene * max(len(scenes), 1))()
        for k, (n, lo, hi) in enumerate(scenes):
            arr[k] = _batch_scene(int(n), _f3(lo), _f3(hi))
        _check(lib().pbf_batch_initialize(self._h, C.cast(arr, _vp), len(scenes),
                                          C.cast(C.byref(params), _vp)))

    def setParticles(self, positions, velocities, particleCount=None):
        pos = None if positions is None else as_float4(positions)
        vel = None if velocities is None else as_float4(velocities)
        n = particleCount if particleCount is not None else (0 if pos is None else len(pos))
        _check(lib().pbf_batch_set_particles(self._h, _ptr(pos), _ptr(vel), int(n)))

    def step(self):
        _check(lib().pbf_batch_step(self._h))

    def run(self, frames: int):
        _check(lib().pbf_batch_run(self._h, int(frames)))

    def counts(self):
        s, n = _sz(), _sz()
        _check(lib().pbf_batch_counts(self._h, C.byref(s), C.byref(n)))
        return s.value, n.value

    def _copy(self, fn, out):
        n = self.counts()[1]
        out = np.zeros((n, 4), dtype=np.float32) if out is None else out
        _check(fn(self._h, _ptr(out), n))
        return out

    def copyPositionsToHost(self, out=None):
        return self._copy(lib().pbf_batch_copy_positions_to_host, out)

    def copyVelocitiesToHost(self, out=None):
        return self._copy(lib().pbf_batch_copy_velocities_to_host, out)

    def setGraphs(self, enabled: bool):
        _check(lib().pbf_batch_set_graphs(self._h, 1 if enabled else 0))

    def setIncrementalSort(self, enabled: bool):
        _check(lib().pbf_batch_set_incremental_sort(self._h, 1 if enabled else 0))

    def setFusedXsph(self, enabled: bool):
        _check(lib().pbf_batch_set_fused_xsph(self._h, 1 if enabled else 0))

    def setNeighborSkin(self, skin: float):
        _check(lib().pbf_batch_set_neighbor_skin(self._h, float(skin)))

    def staleFrames(self) -> int:
        st = _sz()
        _check(lib().pbf_batch_stale_frames(self._h, C.byref(st)))
        return st.value

    def setTuning(self, neighborSkin: float, passBlock: int = 256, particlesPerThread: int = 1):
        t = Tuning(float(neighborSkin), int(passBlock), int(particlesPerThread))
        _check(lib().pbf_batch_set_tuning(self._h, C.cast(C.pointer(t), _vp)))

    def autotune(self, frames: int, candidates=None):
        return _autotune(lib().pbf_batch_autotune, self._h, frames, candidates)

    def setProfiling(self, enabled: bool):
        _check(lib().pbf_batch_set_profiling(self._h, 1 if enabled else 0))

    def kernelTimes(self) -> dict:
        cap = 64
        names = C.create_string_buffer(8192)
        ms = (C.c_double * cap)()
        launches = (C.c_uint64 * cap)()
        count = _sz()
        _check(lib().pbf_batch_kernel_times(self._h, names, 8192, ms, launches, cap,
                                            C.byref(count)))
        keys = names.value.decode().split("\n")[:count.value]
        return {k: (ms[i], launches[i]) for i, k in enumerate(keys)}


class PBFAdjoint:
    """Reverse-mode derivatives of a PBFSolver's frames (pbf_adjoint_*).

    forward(frames, every) runs the solver and keeps checkpoints;
    backward(xbar, vbar) returns (x0bar, v0bar, viscosity_bar,
    vorticity_bar) for dL/dx, dL/dv of the final state given in API order.
    Neighbour lists and the forward's discrete decisions are held fixed."""

    def __init__(self, solver: "PBFSolver"):
        self._solver = solver  # keeps the solver alive
        h = _vp()
        _check(lib().pbf_adjoint_create(solver._h, C.byref(h)))
        self._h = h.value

    def __del__(self):
        if getattr(self, "_h", None):
            lib().pbf_adjoint_destroy(self._h)
            self._h = None

    def forward(self, frames: int, checkpoint_every: int = 10):
        _check(lib().pbf_adjoint_forward(self._h, int(frames), int(checkpoint_every)))

    def backward(self, xbar=None, vbar=None):
        n = self._solver.particleCount()
        xb = None if xbar is None else as_float4(xbar)
        vb = None if vbar is None else as_float4(vbar)
        x0b = np.zeros((n, 4), np.float32)
        v0b = np.zeros((n, 4), np.float32)
        cb, eb = C.c_double(), C.c_double()
        _check(lib().pbf_adjoint_backward(self._h, _ptr(xb), _ptr(vb), n, _ptr(x0b), _ptr(v0b),
                                          C.byref(cb), C.byref(eb)))
        return x0b, v0b, cb.value, eb.value

