"""Seeded dam-break workloads (synthetic inputs, generated on the host).

A scene is a block of particles on a cubic lattice at spacing s = h/2 with
per-coordinate jitter drawn from SplitMix64, inside an open-top box that is
four block-widths long, two block-heights tall and one block deep.  The
physical parameters are the same for every scene.

    dam_break(1): 32 x 32 x 32 = 32,768 particles, box 6.4 x 3.2 x 1.6
    dam_break(2): 128 x 64 x 128 = 1,048,576 particles, box 25.6 x 12.8 x 6.4
    dam_break(3): two 128 x 256 x 128 blocks against opposite walls,
                  8,388,608 particles, box 25.6 x 25.6 x 6.4

The lattice density at s = h/2 with unit mass is 8078.2013, so the particle
mass 6378 / 8078.2013 = 0.789532 starts the block at the rest density 6378.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import SimulationParams, float3

H = 0.1
SPACING = 0.5 * H
JITTER = 0.001
GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int, start: int = 0) -> np.ndarray:
    """Outputs start..start+count-1 of SplitMix64 seeded with `seed`."""
    with np.errstate(over="ignore"):
        n = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + n * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        return z ^ (z >> np.uint64(31))


def uniform_jitter(seed: int, count: int, start: int = 0) -> np.ndarray:
    """Uniform values in [-JITTER, JITTER): top 24 bits of each draw."""
    u = (splitmix64(seed, count, start) >> np.uint64(40)).astype(np.float64) * (1.0 / 16777216.0)
    return (-JITTER + 2.0 * JITTER * u)


def dam_break_params(substeps: int = 1, iterations: int = 4,
                     literal_strengths: bool = False) -> SimulationParams:
    """Parameters of the dam-break scenes.

    The artificial-pressure, XSPH and vorticity strengths are those of the
    reference's own benchmark (benchmark/pbf_benchmark.cu:63-83: k = 1e-4,
    c = 1e-5, epsilon = 1e-5).  The headline workload's literal strengths
    (k = 0.1, c = 0.01, epsilon = 1e-4; `literal_strengths=True`) make the
    reference's constraint iteration diverge within the first steps (the
    neighbour count passes 256 and the step throws overflow_error), so they
    only serve the tests of that error path.
    """
    p = SimulationParams()
    p.dt = 1.0 / 60.0
    p.restDensity = 6378.0
    p.particleMass = 6378.0 / 8078.2013
    p.particleRadius = 0.025
    p.collisionRestitution = 0.0
    p.collisionFriction = 0.0
    p.smoothingRadius = H
    p.lambdaEpsilon = 141.2
    p.solverIterations = iterations
    p.substeps = substeps
    p.gravity = float3(0.0, -9.8, 0.0)
    p.scorrN = 4
    p.scorrDeltaQ = 0.2 * H
    if literal_strengths:
        p.scorrK = 0.1
        p.xsphViscosity = 0.01
        p.vorticityStrength = 1e-4
    else:
        p.scorrK = 1e-4
        p.xsphViscosity = 1e-5
        p.vorticityStrength = 1e-5
    return p


@dataclass
class Scene:
    name: str
    positions: np.ndarray          # (n, 4) float32, w = 1
    velocities: np.ndarray         # (n, 4) float32
    minBounds: tuple
    maxBounds: tuple
    params: SimulationParams
    spheres: list = field(default_factory=list)
    boxes: list = field(default_factory=list)
    planes: list = field(default_factory=list)

    @property
    def n(self) -> int:
        return len(self.positions)


def lattice_block(nx: int, ny: int, nz: int, seed: int, first: int = 0,
                  mirror_x_at: float | None = None) -> np.ndarray:
    """Block of nx*ny*nz particles, fill order x fastest, then z, then y."""
    count = nx * ny * nz
    k = np.arange(count, dtype=np.int64)
    ix = k % nx
    iz = (k // nx) % nz
    iy = k // (nx * nz)
    j = uniform_jitter(seed, 3 * count, 3 * first).reshape(count, 3)
    x = SPACING / 2 + ix * SPACING
    if mirror_x_at is not None:
        x = mirror_x_at - x
    out = np.empty((count, 4), dtype=np.float32)
    out[:, 0] = (x + j[:, 0]).astype(np.float32)
    out[:, 1] = (SPACING / 2 + iy * SPACING + j[:, 1]).astype(np.float32)
    out[:, 2] = (SPACING / 2 + iz * SPACING + j[:, 2]).astype(np.float32)
    out[:, 3] = 1.0
    return out


# Substeps per frame of each scene (dt = 1/60 s per frame).  The taller
# blocks of scenes 2 and 3 reach the far wall faster than h per 1/60 s, where
# the 4-iteration constraint solve no longer holds the density and the
# neighbour count passes 256 (the reference throws overflow_error at frame
# ~30 with one substep); two (three) substeps per frame keep them stable.
DEFAULT_SUBSTEPS = {1: 1, 2: 2, 3: 3}


def dam_break(config: int, substeps: int | None = None, iterations: int = 4,
              literal_strengths: bool = False) -> Scene:
    seed = 0x5EED0000 + config
    if substeps is None:
        substeps = DEFAULT_SUBSTEPS[config]
    params = dam_break_params(substeps, iterations, literal_strengths)
    if config == 1:
        nx = ny = nz = 32
        pos = lattice_block(nx, ny, nz, seed)
        box = (4 * nx * SPACING, 2 * ny * SPACING, nz * SPACING)
    elif config == 2:
        # 1,048,576 particles: a 128 x 64 x 128 block (half the height of a
        # 128^3 cube) in the 25.6 x 12.8 x 6.4 box of the 128^3 construction.
        nx, ny, nz = 128, 64, 128
        pos = lattice_block(nx, ny, nz, seed)
        box = (4 * nx * SPACING, 4 * ny * SPACING, nz * SPACING)
    elif config == 3:
        nx, ny, nz = 128, 256, 128
        box = (4 * nx * SPACING, 2 * ny * SPACING, nz * SPACING)
        a = lattice_block(nx, ny, nz, seed)
        b = lattice_block(nx, ny, nz, seed, first=nx * ny * nz, mirror_x_at=box[0])
        pos = np.concatenate([a, b])
    else:
        raise ValueError(f"unknown dam-break config {config}")
    vel = np.zeros_like(pos)
    vel[:, 3] = 0.0
    return Scene(f"dam_break_{config}", pos, vel, (0.0, 0.0, 0.0),
                 tuple(float(np.float32(v)) for v in box), params)


def small_block(nx: int, ny: int, nz: int, seed: int = 0x5EED00FF, box_scale=(4, 2, 1),
                substeps: int = 1, iterations: int = 4) -> Scene:
    """A smaller dam break of the same construction (parity tests)."""
    params = dam_break_params(substeps, iterations)
    pos = lattice_block(nx, ny, nz, seed)
    box = (box_scale[0] * nx * SPACING, box_scale[1] * ny * SPACING, box_scale[2] * nz * SPACING)
    vel = np.zeros_like(pos)
    return Scene(f"block_{nx}x{ny}x{nz}", pos, vel, (0.0, 0.0, 0.0),
                 tuple(float(np.float32(v)) for v in box), params)


def batch_scene(index: int, nmin: int = 2048, nmax: int = 200_000, substeps: int = 2,
                iterations: int = 4) -> Scene:
    """Scene `index` of the batched workload: a dam-break block of
    n = round(exp(U(ln nmin, ln nmax))) particles with aspect ratios
    a1, a2 ~ U[0.5, 2] (nz = round(cbrt(n / (a1 a2))), nx = round(a1 nz),
    ny = ceil(n / (nx nz)), the fill truncated to n), in its own open-top box
    of the dam-break construction.  Seed 0xBA7C0000 + index: its first three
    SplitMix64 outputs draw n, a1 and a2, the rest the jitter.  Two substeps
    per frame, as the 1M dam break (the tallest blocks are about 3 m)."""
    seed = 0xBA7C0000 + index
    u = (splitmix64(seed, 3) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    n = int(round(np.exp(np.log(nmin) + u[0] * (np.log(nmax) - np.log(nmin)))))
    a1, a2 = 0.5 + 1.5 * u[1], 0.5 + 1.5 * u[2]
    nz = max(1, int(round(np.cbrt(n / (a1 * a2)))))
    nx = max(1, int(round(a1 * nz)))
    ny = -(-n // (nx * nz))
    pos = lattice_block(nx, ny, nz, seed, first=1)[:n]
    box = (4 * nx * SPACING, 2 * ny * SPACING, nz * SPACING)
    vel = np.zeros_like(pos)
    params = dam_break_params(substeps, iterations)
    return Scene(f"batch_{index}", pos, vel, (0.0, 0.0, 0.0),
                 tuple(float(np.float32(v)) for v in box), params)


def batch_scenes(count: int = 256, **kw) -> list:
    """The batched workload: scenes 0 .. count-1 (about 43,000 particles
    each on average, 11.06M for 256)."""
    return [batch_scene(k, **kw) for k in range(count)]


def differentiable_block(substeps: int = 1, iterations: int = 4) -> tuple:
    """The differentiable workload: 65,536 particles (a 40 x 41 x 40
    lattice truncated in fill order) in the dam-break box of a 40 x 41 x 40
    block, initial velocities N(0, 0.1^2) per coordinate from seed 0xD1FF,
    and the loss target: the initial block moved by +0.2 in x.  Returns
    (scene with those velocities, target positions (n, 3)).

    The gradient of that loss with respect to the viscosity is -4.22 after
    5 frames and -23.3 after 10 (the adjoint and an FP64 restatement agree
    to 4 digits), then grows by orders of magnitude per frame as the block
    collapses (about 3.5e6 after 20 frames in FP64): past about 20 frames it
    leaves FP32's range.  Runs of 60 frames measure the adjoint's cost, not
    a usable gradient."""
    nx, ny, nz = 40, 41, 40
    n = 65_536
    seed = 0xD1FF
    pos = lattice_block(nx, ny, nz, seed)[:n]
    rng = np.random.default_rng(seed)
    vel = np.zeros_like(pos)
    vel[:, :3] = rng.normal(0.0, 0.1, (n, 3)).astype(np.float32)
    box = (4 * nx * SPACING, 2 * ny * SPACING, nz * SPACING)
    params = dam_break_params(substeps, iterations)
    sc = Scene("differentiable_64k", pos, vel, (0.0, 0.0, 0.0),
               tuple(float(np.float32(v)) for v in box), params)
    target = pos[:, :3].astype(np.float64) + np.array([0.2, 0.0, 0.0])
    return sc, target

