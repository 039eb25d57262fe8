"""Throughput of the PBF step on one MI355X.

Workload: the seeded dam break at 1,048,576 particles (128 x 64 x 128 block,
h = 0.1, rest density 6378, 4 constraint iterations, 1/60 s frames of 2
substeps, XSPH, vorticity confinement, s_corr, container collisions); state
resident in HBM before the timed region.  The throughput unit is the
particle-step = one particle advanced by one PBF substep; --steps counts
PBFSolver::step() calls (frames).

After the warm-up frames the library's autotuner times every launch
tuning (the skin of the neighbour lists, i.e. the grid's cell-size
multiplier, by the neighbour passes' workgroup size) on a few frames from
the warmed-up state, restores that state and keeps the fastest.  Two passes
over the same frames from the same seeded state and tuning (the step is
bitwise deterministic): the first is timed end to end for `value`; the
second records HIP events around every launch on the solver's stream, which
gives each kernel's average launch duration over exactly the timed frames.

Prints one JSON line:
  value          particle-steps/s over exactly --steps timed steps
  tuning         the autotuner's candidates, their milliseconds per frame,
                 the choice, and the stale frames of the timed pass
  roofline       dominant kernel (the neighbour-list build): algorithmic
                 HBM bytes per launch (bytes per particle of its phase x
                 particles) / its average launch duration, against 8 TB/s;
                 `traffic` is the measured HBM bytes per launch from the
                 committed rocprofv3 counter summary, when present
  valu           the same kernel against the unpacked FP32 vector peak
                 (flop model per neighbour and per candidate test)
  step_roofline  the whole step's algorithmic bytes per particle-step x
                 value, against 8 TB/s (this dataflow: five grid builds and
                 five list builds per substep, as the reference)
  b_alg_roofline the same with the compulsory-byte model of one build per
                 step and no stored lists (356 + 56 I + 16 ceil(b/8) + 8 C/N)
  grid_sort      the same frames with every grid build radix-sorted in full
                 instead of re-sorted from the previous build (bitwise the
                 same results): ms per frame and sort microseconds per build
                 of both (--no-sort-comparison skips it)
  double_dam_break_8m  the 8,388,608-particle double dam break (3 substeps
                 per frame) with the same tuning: particle-steps/s over 10
                 frames after 3 warm-up frames, and its rooflines (--no-8m
                 skips it)
  batched        the batched workload (256 seeded scenes of 2,048-200,000
                 particles, about 10M in all): one PBFBatchSolver (frames
                 replayed from HIP graphs, and eager) against a loop of
                 PBFSolver.step() over one solver per scene (eager, and with
                 graphs), the speedups, and the share of a batched frame its
                 kernels occupy (--no-batched skips it)
  adjoint        the differentiable workload (65,536 particles, 60 frames,
                 checkpoints every 10, and one 60-frame segment taped by
                 forward): the adjoint's passes against a plain run of the
                 same frames (--no-adjoint skips it)
  cpu_baseline   the FP32 OpenMP restatement of the reference dataflow on
                 the host cores (built on the machine that runs it with
                 -march=native, OMP_PROC_BIND=close, the CPU model and thread
                 count recorded), on a bounded sample of the same workload;
                 `samples` adds the 32K dam break (every step of its 100) and
                 the 8M double dam break (substeps after warm-up)

`value` is the median over --reps runs of exactly --steps timed frames, each
run started from the same warmed-up state (the reference benchmark takes the
median of 5 runs, benchmark/pbf_benchmark.cu:158-211); `runs_ms_per_step`
lists every run.  `ms_per_substep` = `ms_per_step` / substeps per frame.
`headline_8m_step_roofline` repeats the 8M double dam break's compulsory-byte
fraction at the top level: the step fraction the 8M target is quoted on.

`--copy-calibration BYTES` instead copies BYTES of device memory once with
the library's device-to-device copy: the known-size copy whose HBM counters
scale the solver kernels' counters.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import statistics
import sys
import time
from pathlib import Path

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "fluid-simulator_amd"))

HBM_PEAK_GBS = 8000.0
# MI355X FP32 vector FMA peak without packed-FP32 instructions (which this
# build disables): 256 CUs x 4 SIMD x 16 lanes x 2 flop x 2.4 GHz.
VALU_PEAK_TFLOPS = 78.6

KMEAN, CAND = 29, 218  # neighbours and candidates per particle at rest
LIST = 4 + 4 * KMEAN    # a neighbour list: its count and its entries
# Algorithmic HBM bytes per particle of each step phase: float4 records,
# 4-byte keys / ids / lambdas / list entries read or written once, neighbour
# records gathered on chip, plus the list the step stores and reads.
PHASE_BYTES = {
    "predict_key": 72,                 # read x, v; write p*, v_in, key, id
    "keys": 20,                        # read p*; write key
    "gather": 104,                     # read id + 3 records; write 3 records + id
    # A build inside the constraint loop moves p* and composes the record
    # permutation (read idx, p*, perm; write p*, perm; the sorted key for the
    # ranges): 48.  The substep's last build moves every record through the
    # composed permutation: 148.  Per launch, averaged over a substep's five
    # builds at 4 iterations: (4 * 48 + 148) / 5 = 68.
    "ranges_gather": 68,
    "build_lists": 20 + LIST,          # read p*, key; write the list
    # The post-solve build with XSPH summed as the list is written: also
    # read x, write v (the list is not read back).
    "build_lists_xsph": 52 + LIST,
    "lambda": 20 + LIST,               # read p*, the list; write lambda
    "delta_apply_collide": 40 + LIST,  # read p*, lambda, the list; write p' and its key
    "update_xsph": 48 + LIST,          # read x, p*, the list; write v
    "update_velocity": 48,
    "vorticity": 48 + LIST,            # read p*, v, the list; write omega
    "confine_resolve": 80 + LIST,      # read 4 records, the list; write v
    "resolve_velocities": 64,
}
# The two kinds of build permutation (bytes per particle, as above).
RANGES_GATHER_DEFERRED, RANGES_GATHER_FULL = 48, 148
# A build's re-sort from the previous build (k_resort_*): keys and previous
# keys read (8), keys read again by the merge (4), sorted keys and slots
# written (8); per changed key its segment (written and read, 8) and its
# (key, slot) entry (written and read, 16).
RESORT_BYTES, RESORT_CHANGED_BYTES = 20, 24
# Keys that change per build on the 1M dam break at frame 60 (FP32 oracle):
# 20.8% at a substep's first build, 4.3-4.6% at the other four, 7.7% over
# the five.
RESORT_CHANGED = 0.077
RESORT_PER_PARTICLE = RESORT_BYTES + RESORT_CHANGED_BYTES * RESORT_CHANGED
# Flop model: per neighbour of each pass (the FP32 operations of the
# reference's expression trees) and 8 per candidate radius test.
FLOPS_PER_NEIGHBOUR = {"lambda": 38, "delta_apply_collide": 37, "update_xsph": 22,
                       "vorticity": 34, "confine_resolve": 32, "build_lists": 0,
                       "build_lists_xsph": 22}
TESTING = {"build_lists", "build_lists_xsph"}  # passes that run the radius test
TRAFFIC_SUMMARY = REPO / "profiles" / "build_lists_dam_break_1M_traffic.md"
TRAFFIC_KERNEL = "build_lists"
COUNTER_SUMMARY = REPO / "profiles" / "list_passes_dam_break_1M_counters.md"
CUS, CLOCK_HZ = 256, 2.4e9  # MI355X compute units and peak engine clock


def alg_bytes(kernel: str, n: int, key_bits: int, cells: int,
              kmean: float | None = None) -> float | None:
    """Algorithmic bytes of one launch.  With kmean (the mean neighbour
    count measured over the timed frames), the list term 4 K of the list
    passes uses it instead of the count at rest."""
    if kernel == "radix_sort":
        return 16.0 * ((key_bits + 7) // 8) * n
    if kernel == "resort":
        return RESORT_PER_PARTICLE * n
    if kernel == "cell_ranges":
        return 4.0 * n + 8.0 * cells
    if kernel == "fill_cells":
        return 4.0 * cells
    b = PHASE_BYTES.get(kernel)
    if b is None:
        return None
    if kmean is not None and kernel in FLOPS_PER_NEIGHBOUR:
        b += 4.0 * (kmean - KMEAN)
    return float(b) * n


def step_bytes_per_particle(iterations: int, key_bits: int, cells: int, n: int,
                            resort: bool = True, fused_xsph: bool = True) -> float:
    """Algorithmic bytes of one substep of this dataflow per particle: the
    reference's five grid builds (one per constraint iteration plus the
    post-solve rebuild), each a sort (a re-sort from the previous build, or
    with resort=False the radix sort), the ranges, the permute and the
    neighbour lists; the keys come from the prediction and the Delta-p
    passes.  With fused_xsph (the solver's default) the post-solve build
    sums XSPH as it writes the list: x read and v written (32) in place of
    the separate pass."""
    builds = iterations + 1
    sort = RESORT_PER_PARTICLE if resort else 16 * ((key_bits + 7) // 8)
    build = sort + 4 * cells / n + PHASE_BYTES["build_lists"]
    permute = RANGES_GATHER_DEFERRED * iterations + RANGES_GATHER_FULL
    return (PHASE_BYTES["predict_key"] + builds * build + permute
            + iterations * (PHASE_BYTES["lambda"] + PHASE_BYTES["delta_apply_collide"])
            + (32 if fused_xsph else PHASE_BYTES["update_xsph"]) + PHASE_BYTES["vorticity"]
            + PHASE_BYTES["confine_resolve"])


def measured_valu_insts() -> float | None:
    """SQ_INSTS_VALU per launch of the list build from the committed
    counter summary."""
    if not COUNTER_SUMMARY.exists():
        return None
    m = re.search(r"valu_insts_per_launch k_build_lists:\s*([0-9.eE+]+)",
                  COUNTER_SUMMARY.read_text())
    return float(m.group(1)) if m else None


def measured_traffic() -> float | None:
    """HBM bytes per launch of the list build from the committed counter
    summary (FETCH_SIZE and WRITE_SIZE scaled by the copy calibration)."""
    if not TRAFFIC_SUMMARY.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", TRAFFIC_SUMMARY.read_text())
    return float(m.group(1)) if m else None


NATIVE_FLAGS = ["-O3", "-march=native", "-fPIC", "-shared", "-std=c11", "-fopenmp",
                "-ffp-contract=off", "-fno-fast-math", "-DORC_REAL=float"]


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _native_oracle() -> tuple:
    """The FP32 OpenMP oracle compiled for this host with -march=native (in a
    temporary directory), or the prebuilt x86-64-v3 library when gcc fails."""
    import subprocess
    import tempfile

    out = Path(tempfile.mkdtemp(prefix="pbf_cpu_")) / "liboracle_f32_native.so"
    src = REPO / "oracle" / "pbf_oracle.c"
    r = subprocess.run(["gcc", *NATIVE_FLAGS, str(src), "-o", str(out), "-lm"],
                       capture_output=True, text=True)
    if r.returncode == 0:
        return str(out), "gcc " + " ".join(NATIVE_FLAGS)
    return None, "prebuilt liboracle_f32.so (-march=x86-64-v3): native build failed"


def _time_oracle(OracleSolver, scene, warmup: int, timed: int) -> tuple:
    s = OracleSolver("f32_native")
    s.initialize(scene.n, scene.minBounds, scene.maxBounds, scene.params)
    s.setParticles(scene.positions, scene.velocities)
    for _ in range(warmup):
        s.step()
    times = []
    for _ in range(timed):
        t = time.perf_counter()
        s.step()
        times.append(time.perf_counter() - t)
    return times


def cpu_samples() -> dict:
    """Child-process entry (--cpu-samples): the FP32 OpenMP restatement of
    the reference dataflow (5 grid builds per substep, slot-major neighbour
    lists, one loop per reference kernel) on the three dam breaks."""
    sys.path.insert(0, str(REPO / "oracle"))
    import oracle
    from oracle import OracleSolver, threads
    from pbf_amd import scenes

    path, build = _native_oracle()
    oracle.load("f32_native", path or str(REPO / "oracle" / "liboracle_f32.so"))
    cores = threads()
    common = {"unit": "particle-steps/s", "cores": cores, "kind": "port",
              "cpu_model": cpu_model(), "build": build,
              "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
              "omp_places": os.environ.get("OMP_PLACES")}
    out = {}
    # 1M (the measured workload): one warm-up frame, median of 5 frames.
    sc = scenes.dam_break(2)
    t = _time_oracle(OracleSolver, sc, 1, 5)
    out["dam_break_2"] = dict(common, value=sc.n * sc.params.substeps / statistics.median(t),
                              sample=f"{sc.name}: {sc.n} particles, 1 warm-up frame then the "
                                     f"median of 5 timed frames of {sc.params.substeps} "
                                     "substeps", seconds=sum(t))
    # 32K: every step of the workload's 100.
    sc = scenes.dam_break(1)
    t = _time_oracle(OracleSolver, sc, 0, 100)
    out["dam_break_1"] = dict(common, value=sc.n * sc.params.substeps * 100 / sum(t),
                              sample=f"{sc.name}: {sc.n} particles, all 100 steps timed",
                              seconds=sum(t))
    # 8M: substeps of the same per-substep work (one substep per frame),
    # two warm-up substeps then the median of 5.
    sc = scenes.dam_break(3, substeps=1)
    t = _time_oracle(OracleSolver, sc, 2, 5)
    out["dam_break_3"] = dict(common, value=sc.n / statistics.median(t),
                              sample=f"{sc.name}: {sc.n} particles at one substep per frame "
                                     "(the per-substep work of its 3-substep frames), 2 "
                                     "warm-up substeps then the median of 5 timed",
                              seconds=sum(t))
    return out


def cpu_baseline() -> dict:
    """Runs cpu_samples in a child process with OMP_PROC_BIND=close and
    OMP_PLACES=cores set before its OpenMP runtime starts (and its memory,
    about 10 GB for the 8M lists, returned when it exits)."""
    import subprocess

    env = dict(os.environ, OMP_PROC_BIND="close", OMP_PLACES="cores")
    r = subprocess.run([sys.executable, str(Path(__file__).resolve()), "--cpu-samples"],
                       capture_output=True, text=True, env=env)
    if r.returncode != 0:
        return {"error": r.stderr[-2000:]}
    samples = json.loads(r.stdout.strip().splitlines()[-1])
    head = dict(samples["dam_break_2"])
    head["samples"] = {k: v for k, v in samples.items() if k != "dam_break_2"}
    return head


def copy_calibration(nbytes: int) -> int:
    import ctypes as C

    import pbf_amd

    L = pbf_amd.lib()
    src, dst = C.c_void_p(), C.c_void_p()
    pbf_amd._check(L.pbf_device_alloc(C.byref(src), nbytes))
    pbf_amd._check(L.pbf_device_alloc(C.byref(dst), nbytes))
    pbf_amd._check(L.pbf_memset(src, 1, nbytes))
    pbf_amd._check(L.pbf_device_synchronize())
    pbf_amd._check(L.pbf_memcpy_dtod(dst, src, nbytes))
    pbf_amd._check(L.pbf_device_synchronize())
    pbf_amd._check(L.pbf_device_free(src))
    pbf_amd._check(L.pbf_device_free(dst))
    print(json.dumps({"copy_calibration_bytes": nbytes}), flush=True)
    return 0


def run_frames(pbf_amd, scene, warmup: int, steps: int, profile: bool, tuning=None,
               tune_frames: int = 0, reps: int = 1, count_every: int = 0, resort: bool = True):
    """Warm-up frames, then (tune_frames > 0) the library's autotuner from
    the warmed-up state (it restores that state), or the given tuning, then
    `reps` runs of `steps` timed frames, each started from the warmed-up
    state (copied out after the warm-up and set again before the run, so
    every run computes the same frames).  With count_every > 0 the last run
    also reads the neighbour counts every count_every frames (outside its
    timing: that run is then not timed).  Returns (seconds per run, kernel
    times, tuning report, mean neighbour count or None).  resort=False sorts
    every grid build in full (the results are bitwise the same)."""
    solver = pbf_amd.PBFSolver()
    solver.initialize(scene.n, scene.minBounds, scene.maxBounds, scene.params)
    solver.setParticles(scene.positions, scene.velocities)
    solver.setIncrementalSort(resort)
    solver.run(warmup)
    pbf_amd.synchronize()
    x0, v0 = solver.copyPositionsToHost(), solver.copyVelocitiesToHost()
    report = None
    if tune_frames > 0:
        chosen, table = solver.autotune(tune_frames)
        report = {"chosen": {"neighbor_skin": chosen.neighborSkin, "pass_block": chosen.passBlock,
                             "particles_per_thread": chosen.particlesPerThread},
                  "frames_per_candidate": tune_frames,
                  "candidates": [{"neighbor_skin": a, "pass_block": b, "particles_per_thread": c,
                                  "ms_per_frame": ms} for a, b, c, ms in table]}
    elif tuning is not None:
        solver.setTuning(tuning["neighbor_skin"], tuning["pass_block"],
                         tuning.get("particles_per_thread", 1))
    stale0 = solver.staleFrames()
    runs = []
    for _ in range(reps):
        solver.setParticles(x0, v0)
        if profile:
            solver.setProfiling(True)
        pbf_amd.synchronize()
        t0 = time.perf_counter()
        solver.run(steps)
        pbf_amd.synchronize()
        runs.append(time.perf_counter() - t0)
    times = solver.kernelTimes() if profile else {}
    if report is not None:
        report["stale_frames_timed"] = solver.staleFrames() - stale0
    kmean = None
    if count_every > 0:
        solver.setProfiling(False)
        solver.setParticles(x0, v0)
        counts = []
        for f in range(0, steps, count_every):
            solver.run(min(count_every, steps - f))
            counts.append(float(solver.neighborLists(0)[1].mean()))
        kmean = statistics.fmean(counts)
    return runs, times, report, kmean


def sort_comparison(pbf_amd, scene, warmup: int, steps: int, tuning, resort_s: float,
                    resort_rows: dict) -> dict:
    """The same frames with every grid build radix-sorted in full (the
    results are bitwise the same): milliseconds per frame of both, and the
    sort's microseconds per build (events around the launches)."""
    runs, _, _, _ = run_frames(pbf_amd, scene, warmup, steps, profile=False, tuning=tuning,
                               resort=False)
    _, times, _, _ = run_frames(pbf_amd, scene, warmup, steps, profile=True, tuning=tuning,
                                resort=False)
    radix = times.get("radix_sort", (0.0, 0))
    re = resort_rows.get("resort", {})
    return {"resort_ms_per_step": resort_s * 1e3 / steps, "radix_ms_per_step": runs[0] * 1e3 / steps,
            "resort_us_per_build": re.get("avg_ms", 0.0) * 1e3 if re else None,
            "radix_us_per_build": radix[0] / max(radix[1], 1) * 1e3,
            "builds": radix[1], "resort_bytes_per_particle": RESORT_PER_PARTICLE}


def batched_vs_looped(pbf_amd, scenes, count: int, frames: int, warmup: int = 2,
                      looped: bool = True) -> dict:
    """The batched workload (`count` seeded scenes of 2,048-200,000
    particles, shared parameters): one PBFBatchSolver against looping
    PBFSolver.step() over one solver per scene, each frame.  Every scene's
    result is bitwise the same either way (tests/test_batch_gpu.py,
    tests/test_full_size_gpu.py).  Both sides run with frames replayed from
    HIP graphs and launched eagerly; `speedup` is batched (graphs) over the
    eager loop, the plain single-scene path.  `kernel_busy` is the share of
    a batched frame its kernels occupy (events around every launch, summed):
    the launch gaps a persistent per-substep kernel could remove are at most
    the rest."""
    import numpy as np

    scs = scenes.batch_scenes(count)
    n = sum(s.n for s in scs)
    sub = scs[0].params.substeps
    b = pbf_amd.PBFBatchSolver()
    b.initialize([(s.n, s.minBounds, s.maxBounds) for s in scs], scs[0].params)
    b.setParticles(np.concatenate([s.positions for s in scs]),
                   np.concatenate([s.velocities for s in scs]))
    b.run(warmup)
    pbf_amd.synchronize()
    chosen, _ = b.autotune(1)
    tb = {}
    for graphs in (True, False):
        b.setGraphs(graphs)
        b.run(1)  # graph capture out of the timing
        pbf_amd.synchronize()
        t0 = time.perf_counter()
        b.run(frames)
        pbf_amd.synchronize()
        tb[graphs] = (time.perf_counter() - t0) / frames
    b.setProfiling(True)
    b.run(frames)
    pbf_amd.synchronize()
    busy = sum(ms for ms, _ in b.kernelTimes().values()) / 1e3 / frames
    stale = b.staleFrames()
    del b
    if not looped:
        return {"scenes": count, "particles": n, "substeps": sub, "frames": frames,
                "batched_value": n * sub / tb[True], "unit": "particle-steps/s",
                "batched_ms_per_step": tb[True] * 1e3,
                "batched_eager_ms_per_step": tb[False] * 1e3, "kernel_busy": busy / tb[False]}
    solvers = []
    for s in scs:
        o = pbf_amd.PBFSolver()
        o.initialize(s.n, s.minBounds, s.maxBounds, s.params)
        o.setTuning(chosen.neighborSkin, chosen.passBlock, chosen.particlesPerThread)
        o.setParticles(s.positions, s.velocities)
        o.run(warmup)
        solvers.append(o)
    tl = {}
    for graphs in (False, True):
        for o in solvers:
            o.setGraphs(graphs)
            o.step()  # capture
        pbf_amd.synchronize()
        t0 = time.perf_counter()
        for _ in range(frames):
            for o in solvers:
                o.step()
        tl[graphs] = (time.perf_counter() - t0) / frames
    return {"scenes": count, "particles": n, "substeps": sub, "frames": frames,
            "tuning": {"neighbor_skin": chosen.neighborSkin, "pass_block": chosen.passBlock,
                       "particles_per_thread": chosen.particlesPerThread},
            "stale_frames": stale,
            "batched_value": n * sub / tb[True], "looped_value": n * sub / tl[False],
            "unit": "particle-steps/s",
            "batched_ms_per_step": tb[True] * 1e3, "batched_eager_ms_per_step": tb[False] * 1e3,
            "looped_ms_per_step": tl[False] * 1e3, "looped_graphs_ms_per_step": tl[True] * 1e3,
            "speedup": tl[False] / tb[True], "speedup_over_looped_graphs": tl[True] / tb[True],
            "kernel_busy": busy / tb[False]}


def double_dam_break(pbf_amd, scenes, tuning, steps: int = 100, warmup: int = 3) -> dict:
    """The 8,388,608-particle double dam break (two 128 x 256 x 128 blocks
    at opposite walls, 3 substeps per frame) with the tuning chosen on the
    1M workload: particle-steps/s over the workload's 100 frames after
    `warmup`, and the compulsory-byte roofline of that rate."""
    sc = scenes.dam_break(3)
    runs, _, _, _ = run_frames(pbf_amd, sc, warmup, steps, profile=False, tuning=tuning)
    elapsed = runs[0]
    sub = sc.params.substeps
    value = sc.n * steps * sub / elapsed
    grid = pbf_amd.SpatialGrid()
    grid.initialize(1, sc.minBounds, sc.maxBounds, sc.params.smoothingRadius)
    cells = grid.numCells()
    bits = max(0, (cells - 1).bit_length())
    b_alg = 356 + 56 * sc.params.solverIterations + 16 * -(-bits // 8) + 8 * cells / sc.n
    return {"workload": f"{sc.name}: {sc.n} particles, 4 iterations, {sub} substeps per "
                        f"1/60 s frame, frames {warmup + 1}-{warmup + steps}",
            "value": value, "unit": "particle-steps/s", "ms_per_step": elapsed * 1e3 / steps,
            "ms_per_substep": elapsed * 1e3 / steps / sub,
            "b_alg_roofline": {"bytes_per_particle_step": b_alg,
                               "achieved": b_alg * value / 1e9, "peak": HBM_PEAK_GBS,
                               "unit": "GB/s", "frac": b_alg * value / 1e9 / HBM_PEAK_GBS},
            "step_roofline": {"bytes_per_particle_step":
                              step_bytes_per_particle(sc.params.solverIterations, bits, cells, sc.n),
                              "frac": step_bytes_per_particle(sc.params.solverIterations, bits,
                                                              cells, sc.n) * value / 1e9
                              / HBM_PEAK_GBS}}


def adjoint_cost(pbf_amd, scenes, frames: int = 60, every: int = 10) -> dict:
    """The differentiable workload (65,536 particles, 60 frames, checkpoints
    every 10): forward() against backward() (the taped replay of every
    checkpoint segment plus the adjoint kernels), and a plain run of the
    same frames."""
    import numpy as np

    sc, target = scenes.differentiable_block()
    s = pbf_amd.PBFSolver()
    s.initialize(sc.n, sc.minBounds, sc.maxBounds, sc.params)
    s.setParticles(sc.positions, sc.velocities)
    s.run(2)  # first-use costs out of the timings
    s.setParticles(sc.positions, sc.velocities)
    t0 = time.perf_counter()
    s.run(frames)
    pbf_amd.synchronize()
    plain = time.perf_counter() - t0
    s.setParticles(sc.positions, sc.velocities)
    adj = pbf_amd.PBFAdjoint(s)
    t0 = time.perf_counter()
    adj.forward(frames, every)
    pbf_amd.synchronize()
    fwd = time.perf_counter() - t0
    x = s.copyPositionsToHost()
    xbar = np.zeros_like(x)
    xbar[:, :3] = (2.0 * (x[:, :3] - target) / sc.n).astype(np.float32)
    t0 = time.perf_counter()
    adj.backward(xbar, None)
    bwd = time.perf_counter() - t0
    # One segment: forward() tapes every substep as it runs them, backward()
    # replays nothing.
    s.setParticles(sc.positions, sc.velocities)
    one = pbf_amd.PBFAdjoint(s)
    t0 = time.perf_counter()
    one.forward(frames, frames)
    pbf_amd.synchronize()
    fwd1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    one.backward(xbar, None)
    bwd1 = time.perf_counter() - t0
    return {"workload": f"{sc.name}: {sc.n} particles, {frames} frames of "
                        f"{sc.params.substeps} substep(s), checkpoints every {every}",
            "run_ms": plain * 1e3, "forward_ms": fwd * 1e3, "backward_ms": bwd * 1e3,
            "backward_over_run": bwd / plain,
            "one_segment": {"forward_ms": fwd1 * 1e3, "backward_ms": bwd1 * 1e3,
                            "forward_plus_backward_over_run": (fwd1 + bwd1) / plain}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", type=int, default=2, choices=[1, 2, 3])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile-pass", action="store_true")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched-versus-looped measurement (256 scenes)")
    ap.add_argument("--batch-scenes", type=int, default=256)
    ap.add_argument("--no-looped", action="store_true",
                    help="batched leg without the loop of per-scene solvers")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the adjoint cost measurement (65,536 particles)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES")
    ap.add_argument("--no-8m", action="store_true",
                    help="skip the 8M double dam break measured after the 1M workload")
    ap.add_argument("--no-sort-comparison", action="store_true",
                    help="skip the frames with every grid build radix-sorted")
    ap.add_argument("--no-autotune", action="store_true",
                    help="time the default tuning (the reference's builds, 256-thread passes)")
    ap.add_argument("--tune-frames", type=int, default=3,
                    help="frames the autotuner times per candidate")
    ap.add_argument("--reps", type=int, default=5,
                    help="timed runs of --steps frames; value is their median")
    ap.add_argument("--cpu-samples", action="store_true", help=argparse.SUPPRESS)
    a = ap.parse_args()

    if a.cpu_samples:  # child process of cpu_baseline(); touches no GPU
        print(json.dumps(cpu_samples()), flush=True)
        return 0

    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is a design note, not a run mode)"}))
        return 2

    import pbf_amd
    from pbf_amd import scenes

    if a.copy_calibration:
        return copy_calibration(a.copy_calibration)

    scene = scenes.dam_break(a.config)
    n = scene.n
    substeps = scene.params.substeps

    # One timed "step" is one PBFSolver::step() call, i.e. one frame of
    # `substeps` PBF substeps; throughput counts particle-substeps (the
    # reference benchmark's particle-substeps/s, pbf_benchmark.cu:201-211).
    runs, _, tuning, _ = run_frames(pbf_amd, scene, a.warmup, a.steps, profile=False,
                                    tune_frames=0 if a.no_autotune else a.tune_frames,
                                    reps=max(1, a.reps))
    elapsed = statistics.median(runs)
    value = n * a.steps * substeps / elapsed

    grid = pbf_amd.SpatialGrid()
    grid.initialize(1, scene.minBounds, scene.maxBounds, scene.params.smoothingRadius)
    cells = grid.numCells()
    key_bits = max(0, (cells - 1).bit_length())

    kernel_rows = {}
    roofline = valu = None
    kmean = None
    if not a.no_profile_pass:
        _, times, _, kmean = run_frames(pbf_amd, scene, a.warmup, a.steps, profile=True,
                                        tuning=tuning["chosen"] if tuning else None,
                                        count_every=10)
        for name, (ms, launches) in times.items():
            avg = ms / max(launches, 1)
            b = alg_bytes(name, n, key_bits, cells, kmean)
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
            "traffic": measured_traffic() if dominant == TRAFFIC_KERNEL else None,
            "avg_launch_ms": drow["avg_ms"],
            "alg_bytes_per_launch": alg_bytes(dominant, n, key_bits, cells, kmean),
            "launches": drow["launches"],
            "mean_neighbours": kmean,
        }
        if dominant in FLOPS_PER_NEIGHBOUR:
            flops = n * ((kmean or KMEAN) * FLOPS_PER_NEIGHBOUR[dominant]
                         + (CAND * 8 if dominant in TESTING else 0))
            tf = flops / (drow["avg_ms"] * 1e-3) / 1e12
            valu = {"kernel": dominant, "achieved": tf, "peak": VALU_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": tf / VALU_PEAK_TFLOPS,
                    "alg_flops_per_launch": flops}
            insts = measured_valu_insts() if dominant == TRAFFIC_KERNEL else None
            if insts is not None:
                # The list build's bound: VALU issue, one wave-instruction
                # per CU per cycle over every CU at the peak clock.
                issue_ms = insts / (CUS * CLOCK_HZ) * 1e3
                valu["issue"] = {"insts_per_launch": insts, "issue_ms": issue_ms,
                                 "frac": issue_ms / drow["avg_ms"]}

    sort_cmp = None
    if not a.no_profile_pass and not a.no_sort_comparison:
        sort_cmp = sort_comparison(pbf_amd, scene, a.warmup, a.steps,
                                   tuning["chosen"] if tuning else None, elapsed, kernel_rows)

    bstep = step_bytes_per_particle(scene.params.solverIterations, key_bits, cells, n)
    step_roof = {"bytes_per_particle_step": bstep,
                 "achieved": bstep * value / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                 "frac": bstep * value / 1e9 / HBM_PEAK_GBS}

    # The survey's algorithmic byte model of a step (one grid build, no
    # materialised lists): B_alg = 356 + 56 I + 16 ceil(b / 8) + 8 C / N.
    b_alg = (356 + 56 * scene.params.solverIterations + 16 * -(-key_bits // 8) + 8 * cells / n)
    b_alg_roof = {"bytes_per_particle_step": b_alg, "achieved": b_alg * value / 1e9,
                  "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": b_alg * value / 1e9 / HBM_PEAK_GBS}

    headline_8m = None
    if a.config == 2 and not a.no_8m:
        try:
            headline_8m = double_dam_break(pbf_amd, scenes, tuning["chosen"] if tuning else None)
        except pbf_amd.PbfError as e:  # reported in the line rather than losing it
            headline_8m = {"error": f"{type(e).__name__}: {e}"}

    batched = None
    if not a.no_batched:
        batched = batched_vs_looped(pbf_amd, scenes, a.batch_scenes, frames=3,
                                    looped=not a.no_looped)

    adjoint = None if a.no_adjoint else adjoint_cost(pbf_amd, scenes)

    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline()

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
        "ms_per_substep": elapsed * 1e3 / a.steps / substeps,
        "runs_ms_per_step": [r * 1e3 / a.steps for r in runs],
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
        "tuning": tuning,
        "roofline": roofline,
        "valu": valu,
        "step_roofline": step_roof,
        "b_alg_roofline": b_alg_roof,
        "headline_8m_step_roofline": (dict(headline_8m["b_alg_roofline"],
                                           workload=headline_8m["workload"],
                                           value=headline_8m["value"])
                                      if headline_8m and "value" in headline_8m else None),
        "double_dam_break_8m": headline_8m,
        "grid_sort": sort_cmp,
        "batched": batched,
        "adjoint": adjoint,
        "cpu_baseline": cpu,
        "kernels": kernel_rows,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
