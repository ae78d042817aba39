"""The device random stream (Philox4x32-10 keyed by seed, step and blob
index), which drives every batched, served and device-stream run.

It is pinned through the oracle: the oracle's Philox reproduces the
generator authors' known-answer vectors (tests/test_oracle_random_remesh.py),
and the GPU's random-walk and release offsets equal the oracle's device-stream
mode draw for draw.  On top, the GPU increments have the moments the
reference's regression test asks of the random walk
(tests/regression/test_CylinderFlow.cpp:26-71: each component N(0, 2 nu dt)),
and increments of different components, steps and seeds are uncorrelated.
"""
import numpy as np
import pytest

import oracle
from conftest import rel_err
from dvmpp_amd import sim as dsim

pytestmark = pytest.mark.gpu

NU, DT = 1e-3, 0.005
N = 200_000


def _cloud(n):
    zero = np.zeros(n)
    return {"x": zero.copy(), "z": zero.copy(), "circ": zero.copy(),
            "sigma": np.full(n, 0.01), "u": zero.copy(), "w": zero.copy(), "uvs": zero.copy(),
            "wvs": zero.copy(), "id": np.arange(n, dtype=np.uint32)}


def _gpu_increments(ctx, seed, steps_before=0):
    """Random-walk increments of N blobs started at the origin, at step
    `steps_before` of a device-stream simulation (the diffusion substep does
    not look at the body)."""
    px, pz = oracle.cylinder_points(16)
    s = dsim.Sim(ctx, px, pz, dsim.sim_params(seed=seed, rng_mode=dsim.RNG_DEVICE, nu=NU, dt=DT))
    try:
        s.run(steps_before)
        step = s.step_index()
        s.set_state(_cloud(N))
        s.substep(2)  # diffusion
        st = s.get_state()
        return st["x"], st["z"], step
    finally:
        s.close()


def test_diffusion_equals_oracle_device_stream(gpu_ctx):
    for seed, before in ((9, 0), (10, 3)):
        dx, dz, step = _gpu_increments(gpu_ctx, seed, before)
        px, pz = oracle.cylinder_points(16)
        osim = oracle.Sim(px, pz, oracle.make_params(seed=seed, rng_mode=1, nu=NU, dt=DT))
        osim.set_state(_cloud(N))
        osim.set_clock(step, 0)
        osim.diffuse()
        o = osim.get_state()
        # same draws; the device's log, sqrt, cos and sin round within an ulp or two of libm's
        assert rel_err(dx, o["x"]) <= 1e-13
        assert rel_err(dz, o["z"]) <= 1e-13


def test_release_equals_oracle_device_stream(gpu_ctx):
    px, pz = oracle.cylinder_points(64)
    gam = 2.0 * np.sin(np.linspace(0, 2 * np.pi, 64, endpoint=False) + 0.3)
    g = dsim.Sim(gpu_ctx, px, pz, dsim.sim_params(seed=21, rng_mode=dsim.RNG_DEVICE, nu=NU, dt=DT))
    o = oracle.Sim(px, pz, oracle.make_params(seed=21, rng_mode=1, nu=NU, dt=DT))
    try:
        g.run(2)
        o.set_clock(g.step_index(), 0)
        g.set_state(_cloud(10))
        o.set_state(_cloud(10))
        g.set_gamma(gam)
        o.set_gamma(gam)
        g.substep(3)  # release + append
        o.release()
        a, b = o.get_state(), g.get_state()
        assert a["x"].size == b["x"].size > 64 + 10
        assert np.array_equal(a["id"], b["id"]) and np.array_equal(a["circ"], b["circ"])
        assert rel_err(b["x"], a["x"]) <= 1e-13 and rel_err(b["z"], a["z"]) <= 1e-13
    finally:
        g.close()


def test_random_walk_moments_and_independence(gpu_ctx):
    var = 2 * NU * DT
    (ax, az, _), (bx, bz, _), (cx, _, _) = (_gpu_increments(gpu_ctx, 9, 0),
                                             _gpu_increments(gpu_ctx, 9, 1),
                                             _gpu_increments(gpu_ctx, 10, 0))
    for d in (ax, az, bx, bz, cx):
        assert abs(d.mean()) < 5 * np.sqrt(var / N)
        assert abs(d.var() / var - 1) < 5 * np.sqrt(2.0 / N)
        # Gaussian tails: P(|d| > 3 sd) = 0.27 %
        frac3 = np.mean(np.abs(d) > 3 * np.sqrt(var))
        assert abs(frac3 - 0.0027) < 5 * np.sqrt(0.0027 / N)
    lim = 5 / np.sqrt(N)
    assert abs(np.corrcoef(ax, az)[0, 1]) < lim      # components
    assert abs(np.corrcoef(ax, bx)[0, 1]) < lim      # consecutive steps
    assert abs(np.corrcoef(ax, cx)[0, 1]) < lim      # seeds (slots)
    assert abs(np.corrcoef(ax[:-1], ax[1:])[0, 1]) < lim  # neighbouring blobs
