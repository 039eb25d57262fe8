"""Throughput of the PBF step on one MI355X.

Workload: the seeded dam break at 1,048,576 particles (128 x 64 x 128 block,
h = 0.1, rest density 6378, 4 constraint iterations, 1/60 s frames of 2
substeps, XSPH, vorticity confinement, s_corr, container collisions); state
resident in HBM before the timed region.  The throughput unit is the
particle-step = one particle advanced by one PBF substep; --steps counts
PBFSolver::step() calls (frames).

Prints one JSON line:
  value          particle-steps/s over exactly --steps timed steps
  roofline       dominant kernel's algorithmic bytes / measured launch time
                 (HIP events on the solver's stream) against 8 TB/s
  cpu_baseline   the FP32 OpenMP restatement of the reference dataflow on the
                 host cores, on a bounded sample of the same workload
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "fluid-simulator_amd"))

HBM_PEAK_GBS = 8000.0
KMEAN = 29  # mean neighbours within h of the workload at rest

# Algorithmic HBM bytes per particle for one launch of each step kernel:
# reads and writes of particle records (16 B), keys/ids/counts (4 B) and
# neighbour-list slots (4 B each, K of them); neighbour records gathered
# through the caches are not counted.
ALG_BYTES = {
    "predict_key": 32 + 48 + 4,
    "keys": 16 + 4,
    "gather": 4 + 4 * 16 + 4 + 4 * 16 + 4,
    "find_neighbors": 16 + 4 + 4 * KMEAN,
    "lambda": 16 + 4 + 4 + 4 * KMEAN,
    "delta_apply_collide": 16 + 4 + 16 + 4 + 4 * KMEAN,
    "update_xsph": 16 + 16 + 16 + 16 + 4 + 4 * KMEAN,
    "update_velocity": 16 + 16 + 16 + 16,
    "vorticity": 16 + 16 + 16 + 4 + 4 * KMEAN,
    "confine_resolve": 16 + 16 + 16 + 16 + 16 + 4 + 4 * KMEAN,
    "resolve_velocities": 16 + 16 + 16 + 16,
}


def alg_bytes(kernel: str, n: int, key_bits: int, cells: int) -> float | None:
    if kernel == "radix_sort":
        return (16.0 * ((key_bits + 7) // 8) + 4.0 * ((key_bits + 7) // 8)) * n
    if kernel == "cell_ranges":
        return 4.0 * n + 8.0 * cells
    b = ALG_BYTES.get(kernel)
    return None if b is None else float(b) * n


def cpu_baseline(scene, seconds_budget: float = 25.0) -> dict:
    sys.path.insert(0, str(REPO / "oracle"))
    from oracle import OracleSolver, threads

    s = OracleSolver("f32")
    s.initialize(scene.n, scene.minBounds, scene.maxBounds, scene.params)
    s.setParticles(scene.positions, scene.velocities)
    t = time.perf_counter()
    s.step()  # warm-up
    first = time.perf_counter() - t
    reps = max(1, min(5, int(seconds_budget / max(first, 1e-3)) - 1))
    times = []
    for _ in range(reps):
        t = time.perf_counter()
        s.step()
        times.append(time.perf_counter() - t)
    med = statistics.median(times)
    sub = scene.params.substeps
    return {
        "value": scene.n * sub / med,
        "unit": "particle-steps/s",
        "cores": threads(),
        "kind": "port",
        "sample": f"{scene.name}: {scene.n} particles, 1 warm-up frame then median of {reps} "
                  f"timed frames ({sub} substeps each) of the FP32 OpenMP restatement "
                  f"(5 grid builds per substep, slot-major neighbour lists, one loop per "
                  f"reference kernel)",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", type=int, default=2, choices=[1, 2, 3])
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    a = ap.parse_args()

    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is a design note, not a run mode)"}))
        return 2

    import pbf_amd
    from pbf_amd import scenes

    scene = scenes.dam_break(a.config)
    n = scene.n
    solver = pbf_amd.PBFSolver()
    solver.initialize(n, scene.minBounds, scene.maxBounds, scene.params)
    solver.setParticles(scene.positions, scene.velocities)

    # One timed "step" is one PBFSolver::step() call, i.e. one frame of
    # `substeps` PBF substeps; throughput counts particle-substeps (the
    # reference benchmark's particle-substeps/s, pbf_benchmark.cu:201-211).
    substeps = scene.params.substeps
    solver.run(a.warmup)
    pbf_amd.synchronize()
    t0 = time.perf_counter()
    solver.run(a.steps)
    pbf_amd.synchronize()
    elapsed = time.perf_counter() - t0
    value = n * a.steps * substeps / elapsed

    # Per-kernel launch times (HIP events on the solver's stream), after the
    # timed region so the events do not perturb it.
    solver.setProfiling(True)
    for _ in range(a.profile_steps):
        solver.step()
    times = solver.kernelTimes()
    solver.setProfiling(False)

    grid = pbf_amd.SpatialGrid()
    grid.initialize(1, scene.minBounds, scene.maxBounds, scene.params.smoothingRadius)
    cells = grid.numCells()
    key_bits = max(0, (cells - 1).bit_length())
    kernel_rows = {}
    for name, (ms, launches) in times.items():
        avg = ms / max(launches, 1)
        b = alg_bytes(name, n, key_bits, cells)
        kernel_rows[name] = {"ms_total": ms, "launches": launches, "avg_ms": avg,
                             "gbs": (b / (avg * 1e-3) / 1e9) if (b and avg > 0) else None}
    dominant = max(kernel_rows, key=lambda k: kernel_rows[k]["ms_total"])
    drow = kernel_rows[dominant]
    achieved = drow["gbs"]
    roofline = {
        "bound": "hbm",
        "kernel": dominant,
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
        "traffic": None,
        "avg_launch_ms": drow["avg_ms"],
        "alg_bytes_per_launch": alg_bytes(dominant, n, key_bits, cells),
    }

    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(scene)

    out = {
        "metric": "particle-steps/s @1M (4 iters) + % HBM roofline",
        "step_unit": "one PBF substep of one particle (integrate, 4 constraint iterations, "
                     "XSPH, vorticity, contacts); a timed step() call is one frame of "
                     f"{substeps} substeps",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{scene.name}: {n} particles, 4 iterations, "
                               f"{substeps} substeps per 1/60 s frame, "
                               "XSPH + vorticity + s_corr + container",
                   "particles": n, "iterations": scene.params.solverIterations,
                   "substeps": substeps, "parallelism": "single-gpu"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernel_rows,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
