"""Pins the oracle's RNG (src/Random.cpp:3-18), remesh specification (the
reference's BackgroundMesh source is missing; kernels and API from
tests/unit/test_BackgroundMesh.cpp) and full-step bookkeeping
(src/DVMBase.cpp:86-133)."""
import math

import numpy as np
import pytest

import oracle


# ---------------------------------------------------------------- Random

def test_mt19937_known_answer():
    # C++ standard [rand.predef]: the 10000th output of a default-constructed
    # mt19937 (seed 5489) is 4123659995
    r = oracle.Rng(5489)
    for _ in range(9999):
        r.raw()
    assert r.raw() == 4123659995


def test_canonical_from_two_words():
    # libstdc++ generate_canonical<double,53> on a 32-bit engine draws two words:
    # R = (w1 + w2 * 2^32) / 2^64 (random.tcc:3348-3380)
    a = oracle.Rng(42)
    b = oracle.Rng(42)
    for _ in range(1000):
        w1 = b.raw()
        w2 = b.raw()
        expect = float(np.float64(w1) + np.float64(w2) * 4294967296.0) / 18446744073709551616.0
        assert a.rand() == expect


def test_rand_range_and_seeding():  # test_Random.cpp:10-78
    r = oracle.Rng(42)
    v = r.fill(1000)
    assert np.all(v >= 0.0) and np.all(v < 1.0)
    assert np.array_equal(oracle.Rng(42).fill(100), oracle.Rng(42).fill(100))
    assert not np.array_equal(oracle.Rng(42).fill(100), oracle.Rng(123).fill(100))
    assert len(set(oracle.Rng(42).fill(100))) > 90
    oracle.Rng(0).rand()


def test_diffusion_statistics():  # test_CylinderFlow.cpp:26-71
    p = oracle.make_params(nu=0.001, dt=0.01, seed=42)
    x, z = oracle.cylinder_points(8, radius=0.01)
    sim = oracle.Sim(x, z, p)
    n = 1000
    st = {k: np.zeros(n) for k in oracle.STATE_FIELDS}
    st["x"] = np.full(n, 5.0)
    st["z"] = np.full(n, 5.0)
    st["circ"] = np.ones(n)
    st["sigma"] = np.full(n, 0.1)
    st["id"] = np.arange(n, dtype=np.uint32)
    sim.set_state(st)
    sim.diffuse()
    out = sim.get_state()
    dx = out["x"] - 5.0
    dz = out["z"] - 5.0
    spread = math.sqrt(4 * 0.001 * 0.01)
    assert abs(dx.mean()) < 0.1 and abs(dz.mean()) < 0.1
    for d in (dx, dz):
        assert 0.5 * spread < d.std(ddof=1) < 2.0 * spread
    # Box-Muller: each component is N(0, 2 nu dt); variance within 10%
    assert abs(np.var(dx) / (2 * 0.001 * 0.01) - 1.0) < 0.1
    assert sim.rng_draws() == 2 * n


# ---------------------------------------------------------------- remesh kernels

def W(k, r):
    return oracle.remesh_weight(k, r)


def test_m4prime_values():  # test_BackgroundMesh.cpp:53-78
    assert W(0, 0.0) == 1.0
    assert abs(W(0, 0.5) - 0.5625) < 1e-10
    assert W(0, 2.0) == 0.0 and W(0, 3.0) == 0.0
    assert W(0, -0.7) == W(0, 0.7)


@pytest.mark.parametrize("k,support", [(0, 2), (2, 3)])
def test_partition_of_unity_and_moments(k, support):
    # test_BackgroundMesh.cpp:207-263 (M4' partition of unity in 1-D and 2-D);
    # M4' also reproduces the first and second moments, M6 (B-spline) the first
    for xi in (0.0, 0.3, 0.5, 0.77):
        nodes = np.arange(-support + 1, support + 1)
        w = np.array([W(k, xi - i) for i in nodes])
        assert abs(w.sum() - 1.0) < 1e-14
        assert abs(np.sum(w * nodes) - xi) < 1e-14
        if k == 0:
            assert abs(np.sum(w * nodes ** 2) - xi ** 2) < 1e-14
    w2 = sum(W(k, 0.3 - i) * W(k, 0.7 - j) for i in range(-2, 4) for j in range(-2, 4))
    assert abs(w2 - 1.0) < 1e-10


def test_m6_is_quintic_bspline():
    # M6(0) = 66/120, M6(1) = 26/120, M6(2) = 1/120, M6(3) = 0
    assert abs(W(2, 0.0) - 66 / 120) < 1e-15
    assert abs(W(2, 1.0) - 26 / 120) < 1e-15
    assert abs(W(2, 2.0) - 1 / 120) < 1e-15
    assert W(2, 3.0) == 0.0


def test_gaussian_truncated():  # test_BackgroundMesh.cpp:86-102
    assert W(1, 3.0) == 0.0
    assert W(1, 0.0) > W(1, 1.0)
    assert abs(W(1, 0.0) - 1.0 / math.sqrt(2 * math.pi)) < 1e-16


# ---------------------------------------------------------------- remesh spec

def remesh_params(kernel=0, **kw):
    d = dict(remesh_frequency=1, remesh_kernel=kernel, remesh_xmin=-1.0, remesh_xmax=2.0,
             remesh_zmin=-1.0, remesh_zmax=2.0, remesh_dx=0.25, remesh_dz=0.25,
             remesh_sigma=0.05, remesh_circ_threshold=0.0)
    d.update(kw)
    return oracle.make_params(**d)


@pytest.mark.parametrize("kernel", [0, 1, 2])
def test_remesh_conserves_moments(kernel):
    # test_BackgroundMesh.cpp:167-204 (circulation conserved)
    x = np.array([0.25, 0.75, 0.25, 0.75, 0.4])
    z = np.array([0.25, 0.25, 0.75, 0.75, 0.61])
    c = np.array([0.5, 0.3, -0.2, 0.4, 0.13])
    s = np.full(5, 0.1)
    p = remesh_params(kernel)
    ox, oz, oc, osg = oracle.remesh(p, x, z, c, s)
    assert abs(oc.sum() - c.sum()) < 1e-14
    # first moments (impulse) conserved by all three (Gaussian: normalised,
    # symmetric stencil gives it only approximately)
    tol = 1e-14 if kernel != 1 else 2e-3
    assert abs(np.sum(oc * ox) - np.sum(c * x)) < tol
    assert abs(np.sum(oc * oz) - np.sum(c * z)) < tol
    if kernel == 0:
        assert abs(np.sum(oc * ox ** 2) - np.sum(c * x ** 2)) < 1e-14
    assert np.all(osg == 0.05)


def test_remesh_single_blob_on_node():
    # a blob sitting exactly on node (i, k) maps to that node alone with M4'
    p = remesh_params(0)
    ox, oz, oc, _ = oracle.remesh(p, [0.5], [0.25], [2.0], [0.1])
    assert ox.size == 1 and ox[0] == 0.5 and oz[0] == 0.25 and oc[0] == 2.0


def test_remesh_order_threshold_and_kept():
    p = remesh_params(0, remesh_circ_threshold=1e-3)
    # blob 1 is outside the mesh (stencil leaves it) and is kept as is
    x = np.array([0.3, 5.0, 0.6])
    z = np.array([0.3, 5.0, 0.2])
    c = np.array([1.0, 7.0, -0.5])
    s = np.array([0.1, 0.2, 0.3])
    ox, oz, oc, osg = oracle.remesh(p, x, z, c, s)
    assert ox[-1] == 5.0 and oc[-1] == 7.0 and osg[-1] == 0.2
    grid = slice(0, ox.size - 1)
    lin = np.round((oz[grid] + 1.0) / 0.25) * 13 + np.round((ox[grid] + 1.0) / 0.25)
    assert np.all(np.diff(lin) > 0)  # linear node order k*nx + i
    assert np.all(np.abs(oc[grid]) > 1e-3)


# ---------------------------------------------------------------- full step

def test_step_rng_bookkeeping_and_growth():
    p = oracle.make_params(seed=0)
    x, z = oracle.cylinder_points(64)
    sim = oracle.Sim(x, z, p)
    total = 0
    for k in range(5):
        n0 = sim.size()
        d0 = sim.rng_draws()
        assert sim.step() == 0
        n1 = sim.size()
        # diffusion draws 2 per existing blob, release one per nascent blob
        assert sim.rng_draws() - d0 == 2 * n0 + (n1 - n0)
        assert n1 - n0 >= 64  # at least one blob per panel (src/VortexSheet.cpp:287)
        total = n1
    assert sim.step_index() == 5 and total == sim.size()
    cd = sim.loads()[2]
    assert np.isfinite(cd)


def test_rk3_weights_integrate_constant_velocity():
    # the low-storage RK3 set a = {0, -17/32, -32/27}, b = {1/4, 8/9, 3/4}
    # (src/DVMBase.cpp:184-185) advances a constant velocity by exactly v dt:
    # 1/4 + 8/9 * 15/32 + 3/4 * 12/27 = 1.  A zero-circulation blob 1000 radii
    # away feels only the free stream (sheet influence ~1e-7, i.e. 1e-9 per step).
    x, z = oracle.cylinder_points(32)
    for scheme in (0, 1):
        sim = oracle.Sim(x, z, oracle.make_params(scheme=scheme, ux=0.7, uz=-0.3, dt=0.01))
        st = {k: np.zeros(1) for k in oracle.STATE_FIELDS}
        st["x"][0], st["z"][0], st["sigma"][0] = 500.0, 300.0, 0.05
        st["id"] = np.zeros(1, dtype=np.uint32)
        sim.set_state(st)
        sim.convect()
        out = sim.get_state()
        assert abs(out["x"][0] - 500.0 - 0.007) < 1e-8
        assert abs(out["z"][0] - 300.0 + 0.003) < 1e-8


def test_sim_set_clock_continues_the_run():
    """A state taken at step k with its step index and random-stream position
    continues bit for bit as the run it came from (the lock-step tests
    re-synchronise the oracle to a GPU run this way)."""
    px, pz = oracle.cylinder_points(32)
    p = oracle.make_params(seed=3)
    a = oracle.Sim(px, pz, p)
    for _ in range(5):
        a.step()
    st, draws, k = a.get_state(), a.rng_draws(), a.step_index()
    b = oracle.Sim(px, pz, p)
    b.set_state(st)
    b.set_clock(k, draws)
    for _ in range(2):
        a.step()
        b.step()
    sa, sb = a.get_state(), b.get_state()
    assert all(np.array_equal(sa[f], sb[f]) for f in sa)
    assert a.rng_draws() == b.rng_draws() and a.step_index() == b.step_index() == k + 2
    assert a.loads() == b.loads()
    # a wrong stream position changes the diffusion
    c = oracle.Sim(px, pz, p)
    c.set_state(st)
    c.set_clock(k, draws + 1)
    c.step()
    a2 = oracle.Sim(px, pz, p)
    a2.set_state(st)
    a2.set_clock(k, draws)
    a2.step()
    assert not np.array_equal(c.get_state()["x"], a2.get_state()["x"])


# Philox4x32-10 known-answer vectors of the generator's authors (Random123,
# kat_vectors: philox4x32_10 for the zero, all-ones and pi-digit inputs).
PHILOX_KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,out", PHILOX_KAT)
def test_philox_known_answers(ctr, key, out):
    assert tuple(int(v) for v in oracle.philox4x32(ctr, key)) == out


def test_u01_bounds_and_bits():
    # (m + 0.5) 2^-53 over the 53 bits taken: a > 5 high bits, b > 6 high bits
    assert oracle.u01(0, 0) == 0.5 * 2.0 ** -53
    assert oracle.u01(0xFFFFFFFF, 0xFFFFFFFF) == 1.0 - 0.5 * 2.0 ** -53
    assert oracle.u01(0x20, 0) == (2.0 ** 26 + 0.5) * 2.0 ** -53
    assert oracle.u01(0, 0x40) == 1.5 * 2.0 ** -53
    assert oracle.u01(0x1F, 0x3F) == oracle.u01(0, 0)  # dropped low bits


def test_device_stream_diffusion_statistics():
    """The oracle's device-stream random walk: each component N(0, 2 nu dt),
    components and steps uncorrelated (the device mode's reference moments,
    src/VortexBlobs.cpp:251-265 on Philox draws)."""
    n = 200_000
    px, pz = oracle.cylinder_points(16)
    p = oracle.make_params(rng_mode=1, seed=9, nu=1e-3, dt=0.005)
    sim = oracle.Sim(px, pz, p)
    zero = np.zeros(n)
    st = {"x": np.full(n, 50.0), "z": zero.copy(), "circ": zero.copy(), "sigma": np.full(n, 0.01),
          "u": zero.copy(), "w": zero.copy(), "uvs": zero.copy(), "wvs": zero.copy(),
          "id": np.arange(n, dtype=np.uint32)}
    steps = []
    for k in (1, 2):
        sim.set_state(st)
        sim.set_clock(k, 0)
        sim.diffuse()
        s = sim.get_state()
        steps.append((s["x"] - 50.0, s["z"]))
    var = 2 * 1e-3 * 0.005
    for dx, dz in steps:
        for d in (dx, dz):
            assert abs(d.mean()) < 5 * np.sqrt(var / n)
            assert abs(d.var() / var - 1) < 5 * np.sqrt(2.0 / n)
        assert abs(np.corrcoef(dx, dz)[0, 1]) < 5 / np.sqrt(n)
    assert abs(np.corrcoef(steps[0][0], steps[1][0])[0, 1]) < 5 / np.sqrt(n)
