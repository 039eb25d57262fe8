"""ctypes binding of the CPU oracle (pbf_oracle.c).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker and as the timed CPU baseline.

`Oracle("f64")` is the parity oracle (FP64 arithmetic, FP32 storage points);
`Oracle("f32")` is the FP32 OpenMP baseline.  The Python surface mirrors the
library binding (pbf_amd) so one test body runs against both.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


class OracleError(Exception):
    pass


class InvalidArgument(OracleError, ValueError):
    pass


class LogicError(OracleError, RuntimeError):
    pass


class OutOfRange(OracleError, IndexError):
    pass


class LengthError(OracleError, ValueError):
    pass


class NeighborOverflow(OracleError, OverflowError):
    pass


class CollisionNonConvergence(OracleError, RuntimeError):
    pass


_ERRORS = {1: InvalidArgument, 2: LogicError, 3: OutOfRange, 4: LengthError,
           5: NeighborOverflow, 6: CollisionNonConvergence}


def _check(st):
    if st != 0:
        raise _ERRORS.get(st, OracleError)(f"oracle status {st}")


class f3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class i3(C.Structure):
    _fields_ = [("x", C.c_int), ("y", C.c_int), ("z", C.c_int)]


class Params(C.Structure):
    _fields_ = [("dt", C.c_float), ("restDensity", C.c_float), ("particleMass", C.c_float),
                ("particleRadius", C.c_float), ("collisionRestitution", C.c_float),
                ("collisionFriction", C.c_float), ("smoothingRadius", C.c_float),
                ("lambdaEpsilon", C.c_float), ("solverIterations", C.c_int),
                ("substeps", C.c_int), ("gravity", f3), ("scorrK", C.c_float),
                ("scorrN", C.c_int), ("scorrDeltaQ", C.c_float), ("xsphViscosity", C.c_float),
                ("vorticityStrength", C.c_float)]


class ContainerS(C.Structure):
    _fields_ = [("min", f3), ("max", f3)]


class SphereS(C.Structure):
    _fields_ = [("center", f3), ("radius", C.c_float)]


class BoxS(C.Structure):
    _fields_ = [("center", f3), ("half", f3)]


class PlaneS(C.Structure):
    _fields_ = [("point", f3), ("normal", f3)]


def _v3(v) -> f3:
    if hasattr(v, "x"):
        return f3(v.x, v.y, v.z)
    x, y, z = v
    return f3(float(x), float(y), float(z))


def _params(p) -> Params:
    if isinstance(p, Params):
        return p
    q = Params()
    for name, _ in Params._fields_:
        val = getattr(p, name)
        setattr(q, name, _v3(val) if name == "gravity" else val)
    return q


def _arr(items, conv, ctype):
    items = [conv(x) for x in items]
    return ((ctype * len(items))(*items) if items else None), len(items)


def _sphere(s):
    if hasattr(s, "radius"):
        return SphereS(_v3(s.center), s.radius)
    c, r = s
    return SphereS(_v3(c), float(r))


def _box(b):
    if hasattr(b, "halfExtents"):
        return BoxS(_v3(b.center), _v3(b.halfExtents))
    c, h = b
    return BoxS(_v3(c), _v3(h))


def _plane(p):
    if hasattr(p, "normal"):
        return PlaneS(_v3(p.point), _v3(p.normal))
    pt, nrm = p
    return PlaneS(_v3(pt), _v3(nrm))


def _container(c) -> ContainerS:
    if hasattr(c, "min"):
        return ContainerS(_v3(c.min), _v3(c.max))
    mn, mx = c
    return ContainerS(_v3(mn), _v3(mx))


def f4(a) -> np.ndarray:
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    return arr.reshape(-1, 4)


def _p(a):
    return None if a is None else a.ctypes.data


_vp = C.c_void_p
_sz = C.c_size_t
_f = C.c_float
_i = C.c_int

_SIGS = {
    "orc_precision": (C.c_char_p, []),
    "orc_poly6": (_f, [f3, _f]),
    "orc_spiky": (None, [f3, _f, C.POINTER(f3)]),
    "orc_scorr": (_f, [_f, _f, _f, _i, _f]),
    "orc_grid_dims": (_i, [_sz, f3, f3, _f, C.POINTER(i3), C.POINTER(_sz)]),
    "orc_cell_keys": (None, [_vp, _sz, f3, _f, i3, _vp]),
    "orc_grid_build": (None, [_vp, _sz, f3, _f, i3, _sz, _vp, _vp, _vp, _vp]),
    "orc_find_neighbors": (_i, [_vp, _vp, _vp, i3, f3, _f, _sz, _f, _vp, _vp, _i, _sz]),
    "orc_integrate": (None, [_vp, _vp, _vp, _sz, _f, f3]),
    "orc_lambda": (None, [_vp, _vp, _vp, _i, _sz, _f, _f, _f, _f, _vp, _sz]),
    "orc_delta": (None, [_vp, _vp, _vp, _i, _vp, _sz, _f, _f, _f, _f, _i, _f, _vp, _sz]),
    "orc_apply_delta": (None, [_vp, _vp, _sz]),
    "orc_update_velocity": (None, [_vp, _vp, _vp, _sz, _f]),
    "orc_xsph": (None, [_vp, _vp, _vp, _i, _vp, _vp, _sz, _f, _f, _sz]),
    "orc_vorticity": (None, [_vp, _vp, _vp, _vp, _i, _sz, _f, _vp, _sz]),
    "orc_confine": (None, [_vp, _vp, _vp, _i, _vp, _vp, _vp, _sz, _f, _f, _f, _sz]),
    "orc_solve_container": (None, [_vp, _sz, C.POINTER(ContainerS), _f, C.POINTER(_i)]),
    "orc_solve_spheres": (None, [_vp, _sz, _vp, _sz, _f, C.POINTER(_i)]),
    "orc_solve_boxes": (None, [_vp, _sz, _vp, _sz, _f, C.POINTER(_i)]),
    "orc_solve_planes": (None, [_vp, _sz, _vp, _sz, _f, C.POINTER(_i)]),
    "orc_collision_solve": (_i, [_vp, _sz, C.POINTER(ContainerS), _vp, _sz, _vp, _sz, _vp, _sz,
                                 _f]),
    "orc_resolve_velocities": (None, [_vp, _vp, _vp, _sz, C.POINTER(ContainerS), _vp, _sz, _vp,
                                      _sz, _vp, _sz, _f, _f, _f]),
    "orc_solver_create": (_i, [C.POINTER(_vp), _sz, f3, f3, C.POINTER(Params)]),
    "orc_solver_destroy": (None, [_vp]),
    "orc_solver_set_particles": (_i, [_vp, _vp, _vp, _sz]),
    "orc_solver_set_colliders": (_i, [_vp, _vp, _sz, _vp, _sz, _vp, _sz]),
    "orc_solver_step": (_i, [_vp]),
    "orc_solver_run": (_i, [_vp, _sz]),
    "orc_solver_get": (_i, [_vp, _vp, _vp, _sz]),
    "orc_solver_working_set": (_i, [_vp, _vp, _vp, _sz]),
}

_LIBS: dict[str, C.CDLL] = {}


def build() -> None:
    """Compiles liboracle_f64.so and liboracle_f32.so (make in oracle/)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def load(precision: str = "f64", path: str | None = None) -> C.CDLL:
    """The oracle library of `precision` ("f64", "f32"); with `path`, the
    library at that path (a host-native build of the FP32 form, say) is
    loaded under that name instead."""
    if precision not in _LIBS:
        if path is None:
            path = HERE / f"liboracle_{precision}.so"
            if not path.exists():
                build()
        lib = C.CDLL(str(path))
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIBS[precision] = lib
    return _LIBS[precision]


class Oracle:
    """Kernel-level functions with the signatures of pbf_amd.kernels."""

    InvalidArgument = InvalidArgument
    LogicError = LogicError
    OutOfRange = OutOfRange
    LengthError = LengthError
    NeighborOverflow = NeighborOverflow
    CollisionNonConvergence = CollisionNonConvergence

    def __init__(self, precision: str = "f64"):
        self.precision = precision
        self.lib = load(precision)

    # SPH helpers
    def poly6(self, displacements, smoothingRadius):
        d = np.asarray(displacements, dtype=np.float32).reshape(-1, 3)
        return np.array([self.lib.orc_poly6(f3(*map(float, r)), smoothingRadius) for r in d],
                        dtype=np.float32)

    def spiky_gradient(self, displacements, smoothingRadius):
        d = np.asarray(displacements, dtype=np.float32).reshape(-1, 3)
        out = np.zeros((len(d), 3), dtype=np.float32)
        g = f3()
        for k, r in enumerate(d):
            self.lib.orc_spiky(f3(*map(float, r)), smoothingRadius, C.byref(g))
            out[k] = (g.x, g.y, g.z)
        return out

    def artificial_pressure(self, distanceSquared, smoothingRadius, scorrK, scorrN, scorrDeltaQ):
        d = np.asarray(distanceSquared, dtype=np.float32).reshape(-1)
        return np.array([self.lib.orc_scorr(float(x), smoothingRadius, scorrK, int(scorrN),
                                            scorrDeltaQ) for x in d], dtype=np.float32)

    # grid
    def grid_dims(self, maxParticles, minBounds, maxBounds, cellSize):
        g, cells = i3(), _sz()
        _check(self.lib.orc_grid_dims(int(maxParticles), _v3(minBounds), _v3(maxBounds),
                                      float(cellSize), C.byref(g), C.byref(cells)))
        return (g.x, g.y, g.z), cells.value

    def cell_keys(self, positions, minBounds, cellSize, gridSize):
        p = f4(positions)
        keys = np.zeros(len(p), dtype=np.uint32)
        self.lib.orc_cell_keys(_p(p), len(p), _v3(minBounds), float(cellSize), i3(*gridSize),
                               _p(keys))
        return keys

    def grid_build(self, positions, minBounds, maxBounds, cellSize):
        p = f4(positions)
        dims, cells = self.grid_dims(max(1, len(p)), minBounds, maxBounds, cellSize)
        n = len(p)
        sk = np.zeros(n, dtype=np.uint32)
        si = np.zeros(n, dtype=np.uint32)
        cs = np.zeros(cells, dtype=np.int32)
        ce = np.zeros(cells, dtype=np.int32)
        self.lib.orc_grid_build(_p(p), n, _v3(minBounds), float(cellSize), i3(*dims), cells,
                                _p(sk), _p(si), _p(cs), _p(ce))
        return sk, si, cs, ce, dims, cells

    def find_neighbors(self, predicted, cellStart, cellEnd, gridSize, minBounds, cellSize,
                       particleCount, smoothingRadius, maxNeighbors, particleStride,
                       neighborsFill=None, with_flag=True):
        p = f4(predicted)
        cs = np.ascontiguousarray(cellStart, dtype=np.int32)
        ce = np.ascontiguousarray(cellEnd, dtype=np.int32)
        n = int(particleCount)
        slots = int(maxNeighbors) * int(particleStride)
        nbr = np.zeros(max(1, slots), dtype=np.uint32)
        if neighborsFill is not None:
            nbr[:] = np.uint32(0xFFFFFFFF if neighborsFill == -1 else neighborsFill)
        cnt = np.zeros(max(1, n), dtype=np.int32)
        ov = self.lib.orc_find_neighbors(_p(p), _p(cs), _p(ce), i3(*gridSize), _v3(minBounds),
                                         float(cellSize), n, float(smoothingRadius), _p(nbr),
                                         _p(cnt), int(maxNeighbors), int(particleStride))
        return nbr[:slots], cnt[:n], int(ov) if with_flag else 0

    # particle kernels
    def integrate(self, position, velocity, dt, gravity, particleCount=None):
        x = f4(position).copy()
        v = f4(velocity).copy()
        n = len(x) if particleCount is None else int(particleCount)
        p = np.zeros_like(x)
        self.lib.orc_integrate(_p(x), _p(p), _p(v), n, float(dt), _v3(gravity))
        return p, v

    def compute_lambda(self, predicted, neighbors, counts, maxNeighbors, particleCount,
                       smoothingRadius, particleMass, restDensity, epsilon, stride):
        p = f4(predicted)
        nb = np.ascontiguousarray(neighbors, dtype=np.uint32)
        c = np.ascontiguousarray(counts, dtype=np.int32)
        n = int(particleCount)
        lam = np.zeros(max(1, n), dtype=np.float32)
        self.lib.orc_lambda(_p(p), _p(nb), _p(c), int(maxNeighbors), n, float(smoothingRadius),
                            float(particleMass), float(restDensity), float(epsilon), _p(lam),
                            int(stride))
        return lam[:n]

    def compute_delta_position(self, predicted, neighbors, counts, maxNeighbors, lambdas,
                               particleCount, smoothingRadius, particleMass, restDensity, scorrK,
                               scorrN, scorrDeltaQ, stride):
        p = f4(predicted)
        nb = np.ascontiguousarray(neighbors, dtype=np.uint32)
        c = np.ascontiguousarray(counts, dtype=np.int32)
        lam = np.ascontiguousarray(lambdas, dtype=np.float32)
        n = int(particleCount)
        d = np.zeros((max(1, n), 4), dtype=np.float32)
        self.lib.orc_delta(_p(p), _p(nb), _p(c), int(maxNeighbors), _p(lam), n,
                           float(smoothingRadius), float(particleMass), float(restDensity),
                           float(scorrK), int(scorrN), float(scorrDeltaQ), _p(d), int(stride))
        return d[:n]

    def apply_delta_position(self, predicted, deltas, particleCount=None):
        p = f4(predicted).copy()
        d = f4(deltas)
        n = len(p) if particleCount is None else int(particleCount)
        self.lib.orc_apply_delta(_p(p), _p(d), n)
        return p

    def update_velocity_and_position(self, positions, predicted, velocities, particleCount,
                                     inverseDt):
        x = f4(positions).copy()
        p = f4(predicted).copy()
        v = f4(velocities).copy()
        self.lib.orc_update_velocity(_p(x), _p(p), _p(v), int(particleCount), float(inverseDt))
        return x, p, v

    def apply_xsph_viscosity(self, predicted, neighbors, counts, maxNeighbors, inputVelocities,
                             particleCount, smoothingRadius, xsphViscosity, stride):
        p = f4(predicted)
        nb = np.ascontiguousarray(neighbors, dtype=np.uint32)
        c = np.ascontiguousarray(counts, dtype=np.int32)
        vi = f4(inputVelocities).copy()
        n = int(particleCount)
        vo = np.zeros((max(1, n), 4), dtype=np.float32)
        self.lib.orc_xsph(_p(p), _p(nb), _p(c), int(maxNeighbors), _p(vi), _p(vo), n,
                          float(smoothingRadius), float(xsphViscosity), int(stride))
        return vo[:n], vi

    def compute_vorticity(self, positions, velocities, neighbors, counts, maxNeighbors,
                          particleCount, smoothingRadius, stride):
        p = f4(positions)
        v = f4(velocities)
        nb = np.ascontiguousarray(neighbors, dtype=np.uint32)
        c = np.ascontiguousarray(counts, dtype=np.int32)
        n = int(particleCount)
        w = np.zeros((max(1, n), 4), dtype=np.float32)
        self.lib.orc_vorticity(_p(p), _p(v), _p(nb), _p(c), int(maxNeighbors), n,
                               float(smoothingRadius), _p(w), int(stride))
        return w[:n]

    def apply_vorticity_confinement(self, positions, neighbors, counts, maxNeighbors, vorticity,
                                    inputVelocities, particleCount, smoothingRadius, dt,
                                    vorticityStrength, stride):
        p = f4(positions)
        nb = np.ascontiguousarray(neighbors, dtype=np.uint32)
        c = np.ascontiguousarray(counts, dtype=np.int32)
        w = f4(vorticity)
        vi = f4(inputVelocities).copy()
        n = int(particleCount)
        vo = np.zeros((max(1, n), 4), dtype=np.float32)
        self.lib.orc_confine(_p(p), _p(nb), _p(c), int(maxNeighbors), _p(w), _p(vi), _p(vo), n,
                             float(smoothingRadius), float(dt), float(vorticityStrength),
                             int(stride))
        return vo[:n], vi

    # collisions
    def _solve(self, name, positions, extra, particleRadius, particleCount):
        p = f4(positions).copy()
        n = len(p) if particleCount is None else int(particleCount)
        flag = C.c_int(0)
        getattr(self.lib, name)(_p(p), n, *extra, float(particleRadius), C.byref(flag))
        return p, flag.value

    def solve_container(self, positions, container, particleRadius, particleCount=None):
        c = _container(container)
        return self._solve("orc_solve_container", positions, [C.byref(c)], particleRadius,
                           particleCount)

    def solve_spheres(self, positions, spheres, particleRadius, particleCount=None):
        a, k = _arr(spheres, _sphere, SphereS)
        return self._solve("orc_solve_spheres", positions, [a, k], particleRadius, particleCount)

    def solve_boxes(self, positions, boxes, particleRadius, particleCount=None):
        a, k = _arr(boxes, _box, BoxS)
        return self._solve("orc_solve_boxes", positions, [a, k], particleRadius, particleCount)

    def solve_planes(self, positions, planes, particleRadius, particleCount=None):
        a, k = _arr(planes, _plane, PlaneS)
        return self._solve("orc_solve_planes", positions, [a, k], particleRadius, particleCount)

    def collision_solve(self, positions, container, spheres=(), boxes=(), planes=(),
                        particleRadius=0.5):
        p = f4(positions).copy()
        c = _container(container)
        s, ns = _arr(spheres, _sphere, SphereS)
        b, nb = _arr(boxes, _box, BoxS)
        pl, npl = _arr(planes, _plane, PlaneS)
        st = self.lib.orc_collision_solve(_p(p), len(p), C.byref(c), s, ns, b, nb, pl, npl,
                                          float(particleRadius))
        return p, st

    def resolve_velocities(self, positions, velocities, container, spheres=(), boxes=(),
                           planes=(), particleRadius=0.5, restitution=0.0, friction=0.0,
                           particleCount=None, incoming=None):
        x = f4(positions).copy()
        v = f4(velocities).copy()
        vin = v.copy() if incoming is None else f4(incoming)
        n = len(x) if particleCount is None else int(particleCount)
        c = _container(container)
        s, ns = _arr(spheres, _sphere, SphereS)
        b, nb = _arr(boxes, _box, BoxS)
        pl, npl = _arr(planes, _plane, PlaneS)
        if incoming is None:
            # The in-place overload reads the incoming velocity from the same buffer.
            self.lib.orc_resolve_velocities(_p(x), _p(v), _p(v), n, C.byref(c), s, ns, b, nb, pl,
                                            npl, float(particleRadius), float(restitution),
                                            float(friction))
        else:
            self.lib.orc_resolve_velocities(_p(x), _p(vin), _p(v), n, C.byref(c), s, ns, b, nb,
                                            pl, npl, float(particleRadius), float(restitution),
                                            float(friction))
        return x, v


class OracleSolver:
    """PBFSolver restated on the CPU (same member names as pbf_amd.PBFSolver)."""

    def __init__(self, precision: str = "f64"):
        self.lib = load(precision)
        self._h = None
        self._spheres, self._boxes, self._planes = [], [], []
        self._count = 0

    def __del__(self):
        if getattr(self, "_h", None):
            self.lib.orc_solver_destroy(self._h)
            self._h = None

    def initialize(self, maxParticles, minBounds, maxBounds, params):
        h = _vp()
        prm = _params(params)
        _check(self.lib.orc_solver_create(C.byref(h), int(maxParticles), _v3(minBounds),
                                          _v3(maxBounds), C.byref(prm)))
        if self._h:
            self.lib.orc_solver_destroy(self._h)
        self._h = h.value
        self._count = 0
        self._apply_colliders()

    def _need(self):
        if not self._h:
            raise LogicError("PBFSolver must be initialized")

    def setParticles(self, positions, velocities, particleCount=None):
        self._need()
        p = None if positions is None else f4(positions)
        v = None if velocities is None else f4(velocities)
        n = particleCount if particleCount is not None else (0 if p is None else len(p))
        _check(self.lib.orc_solver_set_particles(self._h, _p(p), _p(v), int(n)))
        self._count = int(n)

    def _apply_colliders(self):
        if not self._h:
            return
        s, ns = _arr(self._spheres, _sphere, SphereS)
        b, nb = _arr(self._boxes, _box, BoxS)
        pl, npl = _arr(self._planes, _plane, PlaneS)
        self.lib.orc_solver_set_colliders(self._h, s, ns, b, nb, pl, npl)

    def setSpheres(self, spheres):
        self._spheres = list(spheres)
        self._apply_colliders()

    def setBoxes(self, boxes):
        self._boxes = list(boxes)
        self._apply_colliders()

    def setPlanes(self, planes):
        self._planes = list(planes)
        self._apply_colliders()

    def step(self):
        self._need()
        _check(self.lib.orc_solver_step(self._h))

    def run(self, frameCount=None):
        self._need()
        _check(self.lib.orc_solver_run(self._h, 1 if frameCount is None else int(frameCount)))

    def particleCount(self):
        return self._count

    def copyPositionsToHost(self, out=None, particleCount=None):
        n = self._count if particleCount is None else int(particleCount)
        pos = np.zeros((n, 4), dtype=np.float32) if out is None else out
        _check(self.lib.orc_solver_get(self._h, _p(pos), None, n))
        return pos

    def copyVelocitiesToHost(self, out=None, particleCount=None):
        n = self._count if particleCount is None else int(particleCount)
        vel = np.zeros((n, 4), dtype=np.float32) if out is None else out
        _check(self.lib.orc_solver_get(self._h, None, _p(vel), n))
        return vel

    def workingSet(self):
        n = self._count
        p = np.zeros((n, 4), dtype=np.float32)
        ids = np.zeros(n, dtype=np.uint32)
        _check(self.lib.orc_solver_working_set(self._h, _p(p), _p(ids), n))
        return p, ids


def threads() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
