"""Pins the oracle's Biot-Savart sums (src/VelocityKernel.cpp:30-108).

Checks are the reference's own known-answer tests
(tests/unit/test_VelocityKernel.cpp, tests/integration/test_BiotSavart.cpp,
tests/unit/test_VortexBlobs.cpp:55-83, tests/regression/test_CylinderFlow.cpp:73-96),
closed forms, invariants and a long-double brute force.
"""
import math

import numpy as np
import pytest

import oracle
from conftest import config2_blobs

RP2 = 1.0 / (2.0 * math.pi)


def bs(parts):
    p = np.array(parts, dtype=float).reshape(-1, 4)
    return oracle.bs_self(p[:, 0], p[:, 1], p[:, 2], p[:, 3])


def test_empty():  # test_VelocityKernel.cpp:47-56
    u, w = oracle.bs_self([], [], [], [])
    assert u.size == 0 and w.size == 0


def test_single_particle_zero():  # test_VelocityKernel.cpp:58-71
    u, w = bs([[0.0, 0.0, 1.0, 0.1]])
    assert u[0] == 0.0 and w[0] == 0.0


def test_two_particles_same_direction():  # test_VelocityKernel.cpp:24-45
    u, w = bs([[0.0, 0.0, 1.0, 0.1], [1.0, 0.0, -1.0, 0.1]])
    assert w[0] != 0.0 and w[1] != 0.0 and w[0] * w[1] > 0.0


def test_induced_at_target():  # test_VelocityKernel.cpp:73-97
    u, w = oracle.bs_targets([0.0], [0.0], [1.0], [0.1], [1.0], [0.0])
    assert abs(u[0]) < 1e-10
    assert abs(w[0] - 1.0 / (2.0 * math.pi)) < 0.01


@pytest.mark.parametrize("sign", [-1.0, 1.0])
def test_pair_analytical(sign):
    # test_VelocityKernel.cpp:99-125 / test_BiotSavart.cpp:15-86: d=2, sigma=0.01
    d, G = 2.0, 1.0
    u, w = bs([[-d / 2, 0.0, G, 0.01], [d / 2, 0.0, sign * G, 0.01]])
    ew = G / (2 * math.pi * d)
    if sign < 0:
        assert abs(w[0] - ew) < 0.01 * ew and abs(w[1] - ew) < 0.01 * ew
    else:
        assert abs(w[0] + ew) < 0.01 * ew and abs(w[1] - ew) < 0.01 * ew
    assert abs(u[0]) < 1e-10 and abs(u[1]) < 1e-10
    # far field (r2 = 4 >= 10 sigma^2): exactly G/(2 pi d) up to rounding
    assert abs(abs(w[1]) - ew) <= 4e-16 * ew


def test_pair_d1_within_10pct():  # test_VortexBlobs.cpp:55-83
    u, w = bs([[-0.5, 0.0, 1.0, 0.1], [0.5, 0.0, -1.0, 0.1]])
    ew = 1.0 / (2 * math.pi)
    assert abs(w[0] - ew) < 0.1 * ew and abs(w[1] - ew) < 0.1 * ew


def test_regularised_finite():  # test_BiotSavart.cpp:89-114
    u, w = bs([[-0.0005, 0.0, 1.0, 0.1], [0.0005, 0.0, -1.0, 0.1]])
    assert np.all(np.isfinite(u)) and np.all(np.isfinite(w))
    # inside the core K -> 1/sigma^2: w ~ G d / (2 pi sigma^2)
    ew = 1.0 * 0.001 / (2 * math.pi * 0.01) * (-math.expm1(-1e-6 / 0.01) / (1e-6 / 0.01))
    assert abs(w[0] - ew) < 1e-9 * abs(ew)


def test_leapfrog_symmetry():  # test_BiotSavart.cpp:117-161
    u, w = bs([[-1, 0.5, 1, 0.05], [-1, -0.5, -1, 0.05], [1, 0.5, 1, 0.05], [1, -0.5, -1, 0.05]])
    assert abs(u[0] - u[2]) < 1e-10 and abs(u[1] - u[3]) < 1e-10
    assert np.all(np.isfinite(u)) and np.all(np.isfinite(w))


def test_ring_circulation():  # test_BiotSavart.cpp:164-191
    n = 20
    th = 2 * np.pi * np.arange(n) / n
    u, w = oracle.bs_self(np.cos(th), np.sin(th), np.full(n, 1.0 / n), np.full(n, 0.05))
    assert np.all(np.isfinite(u)) and np.all(np.isfinite(w))
    # equal ring rotates rigidly: |v| equal for all, tangential
    sp = np.hypot(u, w)
    assert np.ptp(sp) < 1e-13 * sp.max()
    assert np.max(np.abs(u * np.cos(th) + w * np.sin(th))) < 1e-13 * sp.max()


def test_closed_form_far_field():
    # any pair with r2 >= 10 sigma^2 gives K = 1/r2 exactly
    rng = np.random.default_rng(3)
    for _ in range(20):
        x = rng.uniform(-2, 2, 2)
        z = rng.uniform(-2, 2, 2)
        c = rng.normal(size=2)
        s = np.full(2, 0.01)
        u, w = oracle.bs_self(x, z, c, s)
        dx, dz = x[0] - x[1], z[0] - z[1]
        r2 = dx * dx + dz * dz
        assert np.isclose(u[0], -dz * c[1] / r2 * RP2, rtol=1e-15, atol=0)
        assert np.isclose(w[0], dx * c[1] / r2 * RP2, rtol=1e-15, atol=0)
        assert np.isclose(u[1], dz * c[0] / r2 * RP2, rtol=1e-15, atol=0)


def test_near_branch_threshold():
    # r2 just below 10 sigma^2 uses the Gaussian core, just above uses 1/r2:
    # the kernel jumps by the factor (1 - e^-10) across the threshold.
    s = 0.1
    d_in = math.sqrt(10.0) * s * (1 - 1e-9)
    d_out = math.sqrt(10.0) * s * (1 + 1e-9)
    _, w_in = oracle.bs_targets([0.0], [0.0], [1.0], [s], [d_in], [0.0])
    _, w_out = oracle.bs_targets([0.0], [0.0], [1.0], [s], [d_out], [0.0])
    ratio = (w_in[0] * d_in) / (w_out[0] * d_out)
    assert abs(ratio - (1 - math.exp(-10.0))) < 1e-8


def test_translation_rotation_invariance():
    x, z, c, s = config2_blobs(300, seed=5)
    u, w = oracle.bs_self(x, z, c, s)
    # translation by an exactly representable shift
    u2, w2 = oracle.bs_self(x + 0.5, z - 0.25, c, s)
    assert np.max(np.abs(u2 - u)) < 1e-10 * np.max(np.abs(u))
    # rotation by 90 degrees: (x, z) -> (-z, x) maps (u, w) -> (-w, u)
    u3, w3 = oracle.bs_self(-z, x, c, s)
    assert np.max(np.abs(u3 + w)) < 1e-12 * np.max(np.abs(w))
    assert np.max(np.abs(w3 - u)) < 1e-12 * np.max(np.abs(u))


def test_antisymmetry_zero_net_force():
    # equal cores: pairwise terms cancel, so sum_i circ_i * v_i = 0
    x, z, c, s = config2_blobs(500, seed=6)
    u, w = oracle.bs_self(x, z, c, s)
    scale = np.sum(np.abs(c * u)) + np.sum(np.abs(c * w))
    assert abs(np.sum(c * u)) < 1e-12 * scale
    assert abs(np.sum(c * w)) < 1e-12 * scale


def test_long_double_brute_force():
    # dense near field: 64 blobs in a 0.2 box with sigma 0.05
    rng = np.random.default_rng(7)
    n = 64
    x = rng.uniform(0, 0.2, n)
    z = rng.uniform(0, 0.2, n)
    c = rng.normal(size=n)
    s = rng.uniform(0.03, 0.07, n)
    u, w = oracle.bs_self(x, z, c, s)
    ul, wl = oracle.bs_self_ld(x, z, c, s)
    scale = max(np.max(np.abs(ul)), np.max(np.abs(wl)))
    assert np.max(np.abs(u - ul)) < 1e-12 * scale
    assert np.max(np.abs(w - wl)) < 1e-12 * scale


def test_targets_equals_self_without_coincidence():
    x, z, c, s = config2_blobs(400, seed=8)
    u, w = oracle.bs_self(x, z, c, s)
    ut, wt = oracle.bs_targets(x, z, c, s, x, z)
    # same loop order; self pair skipped by the r2 <= 1e-15 guard
    assert np.array_equal(u, ut) and np.array_equal(w, wt)


def test_coincident_blobs_nan():
    # two distinct blobs at one point: 0/0 (reference behaviour)
    u, w = bs([[0.3, 0.3, 1.0, 0.1], [0.3, 0.3, 1.0, 0.1]])
    assert np.isnan(u[0]) and np.isnan(w[1])


def test_circulation_unchanged():  # test_CylinderFlow.cpp:73-96
    rng = oracle.Rng(123)
    vals = rng.fill(150).reshape(50, 3)
    x = vals[:, 0] * 10 - 5
    z = vals[:, 1] * 10 - 5
    c = vals[:, 2] * 2 - 1
    c0 = c.copy()
    oracle.bs_self(x, z, c, np.full(50, 0.1))
    assert np.array_equal(c, c0)
