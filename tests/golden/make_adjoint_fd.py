"""Writes adjoint_fd_64k_f<frames>.npz: FP64 central differences of the
differentiable workload's loss, the fixture tests/test_adjoint_gpu.py checks the library's
adjoint against at full size.

Workload: pbf_amd.scenes.differentiable_block() (65,536 particles, initial
velocities N(0, 0.1^2) from seed 0xD1FF, one substep per 1/60 s frame),
--frames frames (5 for the committed fixture; the loss's sensitivity grows by
orders of magnitude per frame).  Loss: mean over particles of
|x_T - target|^2, target = the initial block moved by +0.2 in x.  The forward is adjoint_ref.run_full: every
state in FP64, neighbour pairs rebuilt at every grid build, container
projection and contact velocities live (the function central differences
see; tests/test_adjoint.py pins its adjoint and its agreement with the FP64
oracle's step).

Differences: 32 velocity coordinates (particles and axes drawn from
numpy's default_rng(0xD1FF + 32)) at steps of 1e-4 and 2.5e-5 (their
agreement bounds the truncation error), and the XSPH viscosity and
vorticity strength at a step of 1e-5 of their value.  About seven minutes
on eight cores at 5 frames (133 forwards of 5 substeps).

Usage: python tests/golden/make_adjoint_fd.py [--frames 5] [--workers 8]
"""
from __future__ import annotations

import argparse
import os
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
for p in (REPO / "fluid-simulator_amd", REPO / "tests"):
    sys.path.insert(0, str(p))

V_STEPS = (1e-4, 2.5e-5)
PARAM_STEP = 1e-5  # relative
COORDS = 32


def _setup():
    import numpy as np

    import adjoint_ref as A
    from pbf_amd import scenes

    sc, target = scenes.differentiable_block()
    P = sc.params
    Q = A.Params(dt=P.dt / P.substeps, h=P.smoothingRadius, mass=P.particleMass,
                 rho0=P.restDensity, eps_lambda=P.lambdaEpsilon,
                 gravity=(P.gravity.x, P.gravity.y, P.gravity.z), scorr_k=P.scorrK,
                 scorr_n=P.scorrN, scorr_dq=P.scorrDeltaQ, radius=P.particleRadius,
                 box_min=sc.minBounds, box_max=sc.maxBounds,
                 restitution=P.collisionRestitution, friction=P.collisionFriction,
                 iterations=P.solverIterations)
    x = sc.positions[:, :3].astype(np.float64)
    v = sc.velocities[:, :3].astype(np.float64)
    return sc, target, Q, x, v


def coordinates(n):
    import numpy as np

    rng = np.random.default_rng(0xD1FF + 32)
    return rng.choice(n, COORDS, replace=False), rng.integers(0, 3, COORDS)


def _loss(job):
    """job = (frames, particle or -1, axis or parameter name, signed step)."""
    import numpy as np

    import adjoint_ref as A

    sc, target, Q, x, v = _setup()
    frames, k, axis, step = job
    c, e = float(sc.params.xsphViscosity), float(sc.params.vorticityStrength)
    if k >= 0:
        v = v.copy()
        v[k, axis] += step
    elif axis == "c":
        c += step
    elif axis == "e":
        e += step
    xT = A.run_full(x, v, frames * sc.params.substeps, Q, c, e)[0]
    d = xT - target
    return float(np.mean(np.sum(d * d, 1)))


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--frames", type=int, default=5)
    ap.add_argument("--workers", type=int, default=8)
    a = ap.parse_args()
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    import multiprocessing as mp

    import numpy as np

    F = a.frames
    sc, _, _, _, _ = _setup()
    parts, axes = coordinates(sc.n)
    c, e = float(sc.params.xsphViscosity), float(sc.params.vorticityStrength)
    jobs = [(F, -1, "base", 0.0)]
    for hs in V_STEPS:
        for k, ax in zip(parts, axes):
            jobs += [(F, int(k), int(ax), hs), (F, int(k), int(ax), -hs)]
    jobs += [(F, -1, "c", PARAM_STEP * c), (F, -1, "c", -PARAM_STEP * c),
             (F, -1, "e", PARAM_STEP * e), (F, -1, "e", -PARAM_STEP * e)]
    with mp.Pool(a.workers) as pool:
        vals = pool.map(_loss, jobs, chunksize=1)
    base = vals[0]
    fd_v = []
    for s_i, hs in enumerate(V_STEPS):
        pv = np.array(vals[1 + 2 * COORDS * s_i:1 + 2 * COORDS * (s_i + 1)]).reshape(COORDS, 2)
        fd_v.append((pv[:, 0] - pv[:, 1]) / (2 * hs))
    q = vals[1 + 2 * COORDS * len(V_STEPS):]
    fd_c = (q[0] - q[1]) / (2 * PARAM_STEP * c)
    fd_e = (q[2] - q[3]) / (2 * PARAM_STEP * e)
    out = HERE / f"adjoint_fd_64k_f{F}.npz"
    np.savez(out, particles=parts, axes=axes, fd_v=fd_v[0], fd_v_fine=fd_v[1], fd_c=fd_c,
             fd_e=fd_e, loss=base, frames=F, v_steps=np.array(V_STEPS), param_step=PARAM_STEP)
    agree = np.linalg.norm(fd_v[0] - fd_v[1]) / np.linalg.norm(fd_v[0])
    print(f"wrote {out}: loss {base:.9e}, |fd_v| {np.linalg.norm(fd_v[0]):.4e} "
          f"(steps agree to {agree:.2e}), dL/dc {fd_c:.6e}, dL/de {fd_e:.6e}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
