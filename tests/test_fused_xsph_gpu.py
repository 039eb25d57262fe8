"""XSPH fused into the post-solve grid build (pbf_solver_set_fused_xsph):
the fused build adds each neighbour's applyXsphViscosity term
(src/pbf/pbf_solver_kernels.cu:213-264) as it appends the neighbour to the
list, with the separate pass's arithmetic and in list order, so the step's
results are bitwise those of the separate pass.  Covered: the staged waves,
the per-lane sweep of sparse regions (the spray scene), rows swept from
global memory (the dense block), graph replay, a batch of scenes, and the
adjoint's tape of the XSPH output.
"""
from __future__ import annotations

import numpy as np
import pytest

from pbf_amd import scenes

pytestmark = pytest.mark.gpu


def _bits(a):
    return a.view(np.uint32)


def _solver(sc, fused, graphs=False):
    import pbf_amd

    s = pbf_amd.PBFSolver()
    s.initialize(sc.n, sc.minBounds, sc.maxBounds, sc.params)
    s.setParticles(sc.positions, sc.velocities)
    s.setFusedXsph(fused)
    s.setGraphs(graphs)
    return s


def _same(a, b):
    assert np.array_equal(_bits(a.copyPositionsToHost()), _bits(b.copyPositionsToHost()))
    assert np.array_equal(_bits(a.copyVelocitiesToHost()), _bits(b.copyVelocitiesToHost()))


@pytest.mark.parametrize("graphs", [False, True])
def test_fused_xsph_equals_separate_pass(gpu, graphs):
    from test_parity_gpu import block_with_spray, dense_block

    for sc in (block_with_spray(), dense_block(), scenes.dam_break(1)):
        fused, separate = _solver(sc, True, graphs), _solver(sc, False, graphs)
        for _ in range(3):
            fused.step()
            separate.step()
        _same(fused, separate)
        v = fused.copyVelocitiesToHost()
        assert np.isfinite(v).all()
        print(f"{sc.name}: {sc.n} particles, 3 frames, fused == separate (graphs {graphs})",
              flush=True)


def test_fused_xsph_with_two_substeps_and_no_vorticity(gpu):
    sc = scenes.small_block(20, 16, 12, substeps=2)
    sc.params.vorticityStrength = 0.0
    fused, separate = _solver(sc, True), _solver(sc, False)
    fused.run(4)
    separate.run(4)
    _same(fused, separate)


def test_fused_xsph_batch_equals_separate_pass(gpu):
    import pbf_amd

    scs = [scenes.batch_scene(k, nmin=500, nmax=6000) for k in range(8)]
    out = []
    for fused in (True, False):
        b = pbf_amd.PBFBatchSolver()
        b.initialize([(s.n, s.minBounds, s.maxBounds) for s in scs], scs[0].params)
        b.setParticles(np.concatenate([s.positions for s in scs]),
                       np.concatenate([s.velocities for s in scs]))
        b.setFusedXsph(fused)
        b.run(3)
        out.append((b.copyPositionsToHost(), b.copyVelocitiesToHost()))
    assert np.array_equal(_bits(out[0][0]), _bits(out[1][0]))
    assert np.array_equal(_bits(out[0][1]), _bits(out[1][1]))


def test_adjoint_tape_of_fused_xsph(gpu):
    """The fused build writes the XSPH output into the adjoint's tape: the
    gradients equal those of a forward that runs the separate pass."""
    import pbf_amd

    sc = scenes.small_block(12, 10, 10)
    xbar = np.zeros((sc.n, 4), np.float32)
    xbar[:, :3] = np.random.default_rng(11).standard_normal((sc.n, 3)).astype(np.float32)
    out = []
    for fused in (True, False):
        s = _solver(sc, fused)
        adj = pbf_amd.PBFAdjoint(s)
        adj.forward(4, 2)
        out.append(adj.backward(xbar, None))
    (x0a, v0a, ca, ea), (x0b, v0b, cb, eb) = out
    assert np.array_equal(_bits(x0a), _bits(x0b))
    assert np.array_equal(_bits(v0a), _bits(v0b))
    assert ca == cb and ea == eb
