"""The two backends the tests run against, behind one Python surface.

OracleBackend: the CPU oracle (oracle/pbf_oracle.c, FP64 build) -- the checker.
GpuBackend:    libpbf_amd.so through its C ABI -- the product.

Both expose `k` (kernel functions with the argument meaning of the
reference's free kernels), `Solver()` (PBFSolver surface), `grid(...)` and
`collision(...)` adapters that take and return host numpy arrays, and `E`
(the error classes).
"""
from __future__ import annotations

import numpy as np

import pbf_amd
from pbf_amd import SimulationParams, float3


def f4(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32)).reshape(-1, 4)


class _OracleGrid:
    def __init__(self, k, maxParticles, minBounds, maxBounds, cellSize):
        self.k = k
        self.dims, self.cells = k.grid_dims(maxParticles, minBounds, maxBounds, cellSize)
        self.maxParticles = maxParticles
        self.minBounds, self.maxBounds, self.cellSize = minBounds, maxBounds, cellSize

    def gridSize(self):
        return self.dims

    def numCells(self):
        return self.cells

    def build(self, positions):
        from oracle import OutOfRange

        p = f4(positions)
        if len(p) > self.maxParticles:
            raise OutOfRange("Particle count exceeds maximum number of particles")
        if len(p) == 0:
            return (np.zeros(0, np.uint32), np.zeros(0, np.uint32),
                    np.full(self.cells, -1, np.int32), np.full(self.cells, -1, np.int32))
        sk, si, cs, ce, _, _ = self.k.grid_build(p, self.minBounds, self.maxBounds, self.cellSize)
        return sk, si, cs, ce


class _GpuGrid:
    def __init__(self, maxParticles, minBounds, maxBounds, cellSize):
        self.g = pbf_amd.SpatialGrid()
        self.g.initialize(maxParticles, minBounds, maxBounds, cellSize)

    def gridSize(self):
        g = self.g.gridSize()
        return (g.x, g.y, g.z)

    def numCells(self):
        return self.g.numCells()

    def build(self, positions):
        p = f4(positions)
        d = pbf_amd.DeviceArray.from_host(p, 16)
        self.g.build(d, len(p))
        return self.g.read(len(p))

    def rebuild(self, positions):
        p = f4(positions)
        d = pbf_amd.DeviceArray.from_host(p, 16)
        self.g.rebuild(d, len(p))
        return self.g.read(len(p))


class _OracleCollision:
    """CollisionSystem numerics through the oracle (validation is the library's)."""

    def __init__(self, k, container):
        self.k = k
        self.container = container
        self.spheres, self.boxes, self.planes = [], [], []

    def setSpheres(self, s):
        self.spheres = list(s)

    def setBoxes(self, b):
        self.boxes = list(b)

    def setPlanes(self, p):
        self.planes = list(p)

    def clearSpheres(self):
        self.spheres = []

    def clearBoxes(self):
        self.boxes = []

    def clearPlanes(self):
        self.planes = []

    def solve(self, positions, particleRadius):
        from oracle import CollisionNonConvergence

        p, st = self.k.collision_solve(positions, self.container, self.spheres, self.boxes,
                                       self.planes, particleRadius)
        if st == 6:
            raise CollisionNonConvergence("Collision constraints did not converge")
        return p

    def resolveVelocities(self, positions, velocities, particleRadius, restitution, friction,
                          incoming=None):
        _, v = self.k.resolve_velocities(positions, velocities, self.container, self.spheres,
                                         self.boxes, self.planes, particleRadius, restitution,
                                         friction, incoming=incoming)
        return v


class _GpuCollision:
    def __init__(self, container):
        self.c = pbf_amd.CollisionSystem(container)

    def __getattr__(self, name):
        return getattr(self.c, name)

    def solve(self, positions, particleRadius):
        p = f4(positions)
        d = pbf_amd.DeviceArray.from_host(p, 16)
        self.c.solve(d, len(p), particleRadius)
        return d.to_numpy(np.float32, p.shape)

    def resolveVelocities(self, positions, velocities, particleRadius, restitution, friction,
                          incoming=None):
        p = f4(positions)
        v = f4(velocities)
        dp = pbf_amd.DeviceArray.from_host(p, 16)
        dv = pbf_amd.DeviceArray.from_host(v, 16)
        if incoming is None:
            self.c.resolveVelocities(dp, dv, len(p), particleRadius, restitution, friction)
        else:
            di = pbf_amd.DeviceArray.from_host(f4(incoming), 16)
            self.c.resolveVelocities(dp, di, dv, len(p), particleRadius, restitution, friction)
        return dv.to_numpy(np.float32, v.shape)


class OracleBackend:
    name = "oracle"

    def __init__(self):
        import oracle

        self.E = oracle
        self.k = oracle.Oracle("f64")

    def Solver(self):
        import oracle

        return oracle.OracleSolver("f64")

    def grid(self, maxParticles, minBounds, maxBounds, cellSize):
        return _OracleGrid(self.k, maxParticles, minBounds, maxBounds, cellSize)

    def collision(self, container):
        return _OracleCollision(self.k, container)


class GpuBackend:
    name = "gpu"

    def __init__(self):
        from pbf_amd import kernels

        pbf_amd.lib()
        self.E = pbf_amd
        self.k = kernels

    def Solver(self):
        return pbf_amd.PBFSolver()

    def grid(self, maxParticles, minBounds, maxBounds, cellSize):
        return _GpuGrid(maxParticles, minBounds, maxBounds, cellSize)

    def collision(self, container):
        return _GpuCollision(container)


PI = np.float32(3.14159265358979323846)


def valid_params(**overrides) -> SimulationParams:
    """validParams() of test/test_pbf_solver.cu:17-36."""
    p = SimulationParams()
    p.dt = 0.1
    p.restDensity = float(np.float32(315.0) / (np.float32(64.0) * PI))
    p.particleMass = 1.0
    p.particleRadius = 1.0
    p.collisionRestitution = 0.0
    p.collisionFriction = 0.03
    p.smoothingRadius = 1.0
    p.lambdaEpsilon = 0.01
    p.solverIterations = 1
    p.substeps = 1
    p.gravity = float3(0.0, -9.8, 0.0)
    p.scorrK = 0.0
    p.scorrN = 0
    p.scorrDeltaQ = 0.0
    p.xsphViscosity = 0.0
    p.vorticityStrength = 0.0
    for k, v in overrides.items():
        if k == "gravity":
            v = float3(*v)
        setattr(p, k, v)
    return p


def self_density(mass, h):
    """selfDensity (test/test_pbf_solver.cu:38-41)."""
    return float(np.float32(mass) * np.float32(315.0) /
                 (np.float32(64.0) * PI * np.float32(h) ** np.float32(3.0)))


def pair_density(mass, distance, h):
    """pairDensityContribution (test/test_pbf_solver.cu:43-50)."""
    h = np.float32(h)
    diff = h * h - np.float32(distance) * np.float32(distance)
    return float(np.float32(mass) * np.float32(315.0) / (np.float32(64.0) * PI * h ** np.float32(9.0))
                 * diff * diff * diff)
