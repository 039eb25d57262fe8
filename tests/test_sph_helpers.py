"""Closed-form checks of the SPH kernels and s_corr (test/test_sph_helpers.cu:58-171,
test/test_pbf_solver_kernels.cu:77-131), run on the oracle and on the GPU."""
from __future__ import annotations

import math

import numpy as np
import pytest

PI = np.float32(3.14159265358979323846)


def test_poly6_zero_outside_and_at_radius(backend):
    k = backend.k
    assert k.poly6([[3.0, 4.0, 0.0]], 4.0)[0] == 0.0
    assert k.poly6([[0.0, 0.0, 2.0]], 2.0)[0] == 0.0


def test_poly6_inside_radius(backend):
    h = np.float32(2.0)
    expected = np.float32(315.0) / (np.float32(64.0) * PI * h ** np.float32(9.0)) * \
        (h * h - np.float32(1.0)) ** np.float32(3.0)
    assert abs(backend.k.poly6([[1.0, 0.0, 0.0]], 2.0)[0] - expected) <= 1e-6


def test_poly6_same_distance_same_weight(backend):
    w = backend.k.poly6([[1.0, 2.0, 2.0], [-2.0, -1.0, 2.0]], 4.0)
    assert w[0] == w[1]


def test_spiky_zero_at_origin_outside_and_at_radius(backend):
    k = backend.k
    assert np.all(k.spiky_gradient([[0.0, 0.0, 0.0]], 2.0) == 0.0)
    assert np.all(k.spiky_gradient([[3.0, 4.0, 0.0]], 4.0) == 0.0)
    assert np.all(k.spiky_gradient([[0.0, 0.0, 2.0]], 2.0) == 0.0)


def test_spiky_inside_radius(backend):
    h = np.float32(4.0)
    mag = np.float32(-45.0) / (PI * h ** np.float32(6.0)) * (h - np.float32(3.0)) ** 2
    g = backend.k.spiky_gradient([[1.0, 2.0, 2.0]], 4.0)[0]
    assert abs(g[0] - mag / 3) <= 1e-7
    assert abs(g[1] - 2 * mag / 3) <= 1e-7
    assert abs(g[2] - 2 * mag / 3) <= 1e-7


def test_spiky_antisymmetric(backend):
    g = backend.k.spiky_gradient([[1.0, 2.0, 2.0], [-1.0, -2.0, -2.0]], 4.0)
    assert np.array_equal(g[0], -g[1])


@pytest.mark.parametrize("h", [1e-5, 1e5])
def test_poly6_extreme_h_stays_representable(backend, h):
    h32 = float(np.float32(h))
    expected = 315.0 / (64.0 * math.pi * h32 ** 3)
    w = backend.k.poly6([[0.0, 0.0, 0.0], [h32, 0.0, 0.0]], h32)
    assert np.isfinite(w[0]) and abs(w[0] - expected) <= expected * 3e-6
    assert w[1] == 0.0


def test_spiky_extreme_h_stays_representable(backend):
    h = float(np.float32(1e-7))
    expected = -45.0 / (4.0 * math.pi * h ** 4)
    g = backend.k.spiky_gradient([[float(np.float32(h) / np.float32(2)), 0.0, 0.0]], h)[0]
    assert np.isfinite(g[0]) and abs(g[0] - expected) <= -expected * 3e-6
    assert g[1] == 0.0 and g[2] == 0.0


def test_artificial_pressure_paper_formula(backend):
    dq = np.float32(0.3)
    d2 = [dq * dq, 0.25, 1.0, 1.44]
    r = backend.k.artificial_pressure(d2, 1.0, 0.001, 4, 0.3)
    ratio = ((np.float32(1.0) - np.float32(0.25)) / (np.float32(1.0) - dq * dq)) ** np.float32(3.0)
    expected = -np.float32(0.001) * ratio ** np.float32(4.0)
    assert abs(r[0] - (-0.001)) <= 1e-7
    assert abs(r[1] - expected) <= 1e-7
    assert r[2] == 0.0 and r[3] == 0.0


def test_artificial_pressure_disabled(backend):
    assert backend.k.artificial_pressure([0.04], 1.0, 0.0, 4, 0.3)[0] == 0.0
    assert backend.k.artificial_pressure([0.04], 1.0, 0.001, 0, 0.3)[0] == 0.0
    # A reference distance at or beyond h also disables it.
    assert backend.k.artificial_pressure([0.04], 1.0, 0.001, 4, 1.0)[0] == 0.0


def test_artificial_pressure_other_exponents(backend):
    # n = 1 and n = 7 take the generic integer-power path.
    for n in (1, 2, 7):
        r = backend.k.artificial_pressure([0.16], 1.0, 0.01, n, 0.2)[0]
        ratio = ((1.0 - 0.16) / (1.0 - 0.04)) ** 3
        assert abs(r - (-0.01 * ratio ** n)) <= 1e-8
