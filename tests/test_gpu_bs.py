"""GPU parity of the Biot-Savart sums (dvm_bs_self / dvm_bs_targets) against
the oracle (src/VelocityKernel.cpp:30-108).

Tolerances: FP64 form <= 1e-10 relative (normwise max), the FP64 mode's
parity bar; FP32-compensated form <= 1e-5 relative (normwise max and p99 of
per-target relative error), the chosen form's accuracy target.
"""
import math

import numpy as np
import pytest

import dvmpp_amd as dvm
import oracle
from conftest import config2_blobs, rel_err, to_dev

pytestmark = pytest.mark.gpu

F64_TOL = 1e-10
F32C_TOL = 1e-5


def gpu_self(ctx, x, z, c, s, prec):
    X, Z, C, S = to_dev(x, z, c, s)
    u, w = dvm.bs_self(ctx, X, Z, C, S, prec=prec)
    return u.cpu().numpy(), w.cpu().numpy()


def gpu_targets(ctx, sx, sz, sc, ss, tx, tz, prec, guard=True):
    SX, SZ, SC, SS, TX, TZ = to_dev(sx, sz, sc, ss, tx, tz)
    u, w = dvm.bs_targets(ctx, SX, SZ, SC, SS, TX, TZ, prec=prec, guard=guard)
    return u.cpu().numpy(), w.cpu().numpy()


@pytest.mark.parametrize("prec", [dvm.F64, dvm.F32C])
def test_reference_known_answers(gpu_ctx, prec):
    tol = 1e-14 if prec == dvm.F64 else 1e-6
    # single blob: exactly zero (test_VelocityKernel.cpp:58-71)
    u, w = gpu_self(gpu_ctx, [0.0], [0.0], [1.0], [0.1], prec)
    assert u[0] == 0.0 and w[0] == 0.0
    # pair at d = 2, sigma = 0.01 (test_VelocityKernel.cpp:99-125)
    u, w = gpu_self(gpu_ctx, [-1.0, 1.0], [0.0, 0.0], [1.0, -1.0], [0.01, 0.01], prec)
    ew = 1.0 / (2 * math.pi * 2.0)
    assert abs(w[0] - ew) < tol * ew and abs(w[1] - ew) < tol * ew
    assert abs(u[0]) < 1e-10 and abs(u[1]) < 1e-10
    # induced at (1, 0) (test_VelocityKernel.cpp:73-97)
    u, w = gpu_targets(gpu_ctx, [0.0], [0.0], [1.0], [0.1], [1.0], [0.0], prec)
    assert abs(u[0]) < 1e-10 and abs(w[0] - 1 / (2 * math.pi)) < 0.01
    # leapfrog symmetry (test_BiotSavart.cpp:117-161)
    u, w = gpu_self(gpu_ctx, [-1, -1, 1, 1], [0.5, -0.5, 0.5, -0.5], [1, -1, 1, -1],
                    [0.05] * 4, prec)
    assert abs(u[0] - u[2]) < 1e-10 and abs(u[1] - u[3]) < 1e-10


@pytest.mark.parametrize("n", [2, 3, 255, 256, 257, 1000, 4096])
def test_f64_self_matches_oracle(gpu_ctx, n):
    x, z, c, s = config2_blobs(n, seed=n)
    ug, wg = gpu_self(gpu_ctx, x, z, c, s, dvm.F64)
    uo, wo = oracle.bs_self(x, z, c, s)
    assert rel_err(ug, uo) <= F64_TOL and rel_err(wg, wo) <= F64_TOL


def test_f64_self_dense_near_field(gpu_ctx):
    # every pair inside the Gaussian core branch for most of the set
    rng = np.random.default_rng(9)
    n = 3000
    x = rng.uniform(0, 0.3, n)
    z = rng.uniform(0, 0.3, n)
    c = rng.normal(size=n) * 1e-3
    s = rng.uniform(0.02, 0.06, n)
    ug, wg = gpu_self(gpu_ctx, x, z, c, s, dvm.F64)
    uo, wo = oracle.bs_self(x, z, c, s)
    assert rel_err(ug, uo) <= F64_TOL and rel_err(wg, wo) <= F64_TOL


def test_f64_targets_probes_and_guard(gpu_ctx):
    x, z, c, s = config2_blobs(5000, seed=3)
    tx = np.linspace(-1, 15, 1024)
    tz = np.zeros(1024)
    tx[7], tz[7] = x[11], z[11]  # coincident target: the pair is skipped
    ug, wg = gpu_targets(gpu_ctx, x, z, c, s, tx, tz, dvm.F64)
    uo, wo = oracle.bs_targets(x, z, c, s, tx, tz)
    assert np.all(np.isfinite(ug))
    assert rel_err(ug, uo) <= F64_TOL and rel_err(wg, wo) <= F64_TOL


def test_f64_targets_unguarded_collocation(gpu_ctx):
    # the collocation sum of solvevortexsheet has no guard (src/VortexSheet.cpp:181-207)
    x, z, c, s = config2_blobs(20000, seed=4)
    g = oracle.Sheet(*oracle.cylinder_points(256), oracle.make_params()).geometry()
    ug, wg = gpu_targets(gpu_ctx, x, z, c, s, g["xc"], g["zc"], dvm.F64, guard=False)
    uo, wo = oracle.bs_targets(x, z, c, s, g["xc"], g["zc"])
    assert rel_err(ug, uo) <= F64_TOL and rel_err(wg, wo) <= F64_TOL
    # a target on a source: unguarded -> NaN like the reference
    ug, _ = gpu_targets(gpu_ctx, x[:10], z[:10], c[:10], s[:10], x[:1], z[:1], dvm.F64,
                        guard=False)
    assert np.isnan(ug[0])


def test_f64_coincident_blobs_nan(gpu_ctx):
    u, w = gpu_self(gpu_ctx, [0.3, 0.3], [0.3, 0.3], [1.0, 1.0], [0.1, 0.1], dvm.F64)
    assert np.isnan(u[0]) and np.isnan(w[1])


def test_empty_inputs(gpu_ctx):
    import torch
    e = torch.empty(0, dtype=torch.float64, device="cuda")
    u, w = dvm.bs_self(gpu_ctx, e, e, e, e)
    assert u.numel() == 0
    one = torch.ones(3, dtype=torch.float64, device="cuda")
    u, w = dvm.bs_targets(gpu_ctx, e, e, e, e, one, one)
    assert torch.all(u == 0) and torch.all(w == 0)


def _f32c_errors(ug, wg, uo, wo):
    v = np.hypot(uo, wo)
    d = np.hypot(ug - uo, wg - wo)
    per = d / np.maximum(v, 1e-300)
    return float(d.max() / v.max()), float(np.percentile(per, 99))


@pytest.mark.parametrize("n", [1000, 20000])
def test_f32c_self_accuracy(gpu_ctx, n):
    x, z, c, s = config2_blobs(n, seed=n + 1)
    ug, wg = gpu_self(gpu_ctx, x, z, c, s, dvm.F32C)
    uo, wo = oracle.bs_self(x, z, c, s)
    norm_max, p99 = _f32c_errors(ug, wg, uo, wo)
    assert norm_max <= F32C_TOL and p99 <= F32C_TOL, (norm_max, p99)


def test_f32c_targets_accuracy(gpu_ctx):
    x, z, c, s = config2_blobs(20000, seed=12)
    tx = np.linspace(-1, 15, 1024)
    tz = np.zeros(1024)
    ug, wg = gpu_targets(gpu_ctx, x, z, c, s, tx, tz, dvm.F32C)
    uo, wo = oracle.bs_targets(x, z, c, s, tx, tz)
    norm_max, p99 = _f32c_errors(ug, wg, uo, wo)
    assert norm_max <= F32C_TOL and p99 <= F32C_TOL, (norm_max, p99)


def _mixed_blobs(n, seed):
    """Uniform background plus a dense patch with wider cores: tile pairs of
    both the near (regularised) and the far (1/r^2) kind; n is not a multiple
    of any tile size, so the partial last tile is exercised."""
    x, z, c, s = config2_blobs(n, seed=seed)
    rng = np.random.default_rng(seed + 1)
    k = n // 4
    x[:k] = rng.uniform(3.0, 3.4, k)
    z[:k] = rng.uniform(-0.2, 0.2, k)
    s[:k] = rng.uniform(0.01, 0.03, k)
    return x, z, c, s


def _with_variant(ctx, variant, fn):
    ctx.set_f32c_variant(*variant)
    try:
        return fn()
    finally:
        ctx.set_f32c_variant()


@pytest.mark.parametrize("variant", dvm.F32C_VARIANTS)
def test_f32c_every_variant(gpu_ctx, variant):
    # Each compiled launch variant (tile, block, sources per reciprocal) at the
    # smallest and largest occupancy target meets the accuracy bound for the
    # self sum and both target sums.
    x, z, c, s = _mixed_blobs(6001, seed=31)
    tx = np.linspace(-1, 15, 777)
    tz = np.linspace(-0.5, 0.5, 777)
    uo, wo = oracle.bs_self(x, z, c, s)
    pu, pw = oracle.bs_targets(x, z, c, s, tx, tz)
    for occ in (dvm.F32C_OCCUPANCY[0], dvm.F32C_OCCUPANCY[-1]):
        v = tuple(variant) + (occ,)
        ug, wg = _with_variant(gpu_ctx, v, lambda: gpu_self(gpu_ctx, x, z, c, s, dvm.F32C))
        norm_max, p99 = _f32c_errors(ug, wg, uo, wo)
        assert norm_max <= F32C_TOL and p99 <= F32C_TOL, (v, norm_max, p99)
        for guard in (True, False):
            tg = _with_variant(gpu_ctx, v, lambda: gpu_targets(gpu_ctx, x, z, c, s, tx, tz,
                                                               dvm.F32C, guard=guard))
            assert gpu_ctx.f32c_variant() == v
            norm_max, p99 = _f32c_errors(tg[0], tg[1], pu, pw)
            assert norm_max <= F32C_TOL and p99 <= F32C_TOL, (v, guard, norm_max, p99)


@pytest.mark.parametrize("tile", [128, 256])
def test_f32c_matrix_form_large(gpu_ctx, tile):
    # 65,536 blobs: most tile pairs are well separated and go through the
    # matrix-core far field; FP32C against the FP64 form on every target,
    # self sum and both target sums.
    x, z, c, s = _mixed_blobs(65537, seed=77)
    X, Z, C, S = to_dev(x, z, c, s)
    tx = np.linspace(-1, 15, 1024)
    tz = np.full(1024, 0.1)
    TX, TZ = to_dev(tx, tz)
    u64, w64 = (t.cpu().numpy() for t in dvm.bs_self(gpu_ctx, X, Z, C, S, prec=dvm.F64))
    p64 = [t.cpu().numpy() for t in dvm.bs_targets(gpu_ctx, X, Z, C, S, TX, TZ, prec=dvm.F64)]
    v = (tile, 32, 0, 8)
    u32, w32 = _with_variant(gpu_ctx, v, lambda: [t.cpu().numpy() for t in dvm.bs_self(
        gpu_ctx, X, Z, C, S, prec=dvm.F32C)])
    norm_max, p99 = _f32c_errors(u32, w32, u64, w64)
    assert norm_max <= F32C_TOL and p99 <= F32C_TOL, (norm_max, p99)
    p32 = _with_variant(gpu_ctx, v, lambda: [t.cpu().numpy() for t in dvm.bs_targets(
        gpu_ctx, X, Z, C, S, TX, TZ, prec=dvm.F32C)])
    norm_max, p99 = _f32c_errors(p32[0], p32[1], p64[0], p64[1])
    assert norm_max <= F32C_TOL and p99 <= F32C_TOL, (norm_max, p99)


@pytest.mark.parametrize("share", [1, 2, 4])
def test_f32c_shared_reciprocal_wide_domain(gpu_ctx, share):
    # Separations up to ~1.7e4: r^2 products of 4 leave the range where one
    # reciprocal may be shared (1e-34..1e34), so those tile pairs take the
    # per-pair loop while the closer ones share.
    rng = np.random.default_rng(5)
    n = 1000
    x = rng.uniform(0.0, 1.2e4, n)
    z = rng.uniform(0.0, 1.2e4, n)
    c = rng.normal(0.0, 1.0, n)
    s = np.full(n, 5.0)
    uo, wo = oracle.bs_self(x, z, c, s)
    ug, wg = _with_variant(gpu_ctx, (128, 64, share, 4),
                           lambda: gpu_self(gpu_ctx, x, z, c, s, dvm.F32C))
    norm_max, p99 = _f32c_errors(ug, wg, uo, wo)
    assert norm_max <= F32C_TOL and p99 <= F32C_TOL, (norm_max, p99)


def test_f32c_measured_tuning(gpu_ctx):
    # The first call of a size bucket times every variant and keeps the
    # fastest; later calls in the bucket reuse it, and its results are the
    # forced variant's, bit for bit.
    x, z, c, s = _mixed_blobs(20001, seed=44)
    gpu_ctx.set_tuning(dvm.TUNE_MEASURE)
    try:
        u1, w1 = gpu_self(gpu_ctx, x, z, c, s, dvm.F32C)
        v = gpu_ctx.f32c_variant()
        assert v[:3] in [tuple(t) for t in dvm.F32C_VARIANTS] and v[3] in dvm.F32C_OCCUPANCY
        u2, w2 = gpu_self(gpu_ctx, x, z, c, s, dvm.F32C)
        assert gpu_ctx.f32c_variant() == v
    finally:
        gpu_ctx.set_tuning(dvm.TUNE_TABLE)
    u3, w3 = _with_variant(gpu_ctx, v, lambda: gpu_self(gpu_ctx, x, z, c, s, dvm.F32C))
    for a, b in ((u1, u2), (w1, w2), (u1, u3), (w1, w3)):
        assert np.array_equal(a, b)
    uo, wo = oracle.bs_self(x, z, c, s)
    norm_max, p99 = _f32c_errors(u1, w1, uo, wo)
    assert norm_max <= F32C_TOL and p99 <= F32C_TOL, (norm_max, p99)


@pytest.mark.parametrize("prec", [dvm.F32C, dvm.F64])
def test_probe_sums_rerun_bit_for_bit(gpu_ctx, prec):
    # 1,024 probes over 65,536 sources: hundreds of source chunks whose
    # partials the few-target reduce adds in a fixed order (64 targets x 16
    # chunk groups a workgroup), so a second call gives the same bits; the
    # sums still match the oracle's
    x, z, c, s = config2_blobs(65536, seed=31)
    tx = np.linspace(-1, 15, 1024)
    tz = np.zeros(1024)
    u1, w1 = gpu_targets(gpu_ctx, x, z, c, s, tx, tz, prec)
    u2, w2 = gpu_targets(gpu_ctx, x, z, c, s, tx, tz, prec)
    assert np.array_equal(u1, u2) and np.array_equal(w1, w2)
    uo, wo = oracle.bs_targets(x, z, c, s, tx, tz)
    if prec == dvm.F64:
        assert rel_err(u1, uo) <= F64_TOL and rel_err(w1, wo) <= F64_TOL
    else:
        norm_max, p99 = _f32c_errors(u1, w1, uo, wo)
        assert norm_max <= F32C_TOL and p99 <= F32C_TOL, (norm_max, p99)


def test_f32c_variant_rejected(gpu_ctx):
    with pytest.raises(dvm.DvmError) as e:
        gpu_ctx.set_f32c_variant(512, 64, 2, 4)
    assert e.value.status == 1
    with pytest.raises(dvm.DvmError):
        gpu_ctx.set_tuning(5)


def test_full_size_isolated_sum(gpu_ctx):
    # N = 262,144 (the isolated-sum workload): FP64 against the oracle on 256
    # sampled targets (computeInducedAt at a blob's own position skips only the
    # self pair), and FP32C against the GPU FP64 result on all targets.
    n = 262144
    x, z, c, s = config2_blobs(n, seed=2024)
    X, Z, C, S = to_dev(x, z, c, s)
    u64, w64 = dvm.bs_self(gpu_ctx, X, Z, C, S, prec=dvm.F64)
    u32, w32 = dvm.bs_self(gpu_ctx, X, Z, C, S, prec=dvm.F32C)
    u64, w64, u32, w32 = (t.cpu().numpy() for t in (u64, w64, u32, w32))
    idx = np.random.default_rng(0).choice(n, 256, replace=False)
    uo, wo = oracle.bs_targets(x, z, c, s, x[idx], z[idx])
    assert rel_err(u64[idx], uo) <= F64_TOL and rel_err(w64[idx], wo) <= F64_TOL
    norm_max, p99 = _f32c_errors(u32, w32, u64, w64)
    assert norm_max <= F32C_TOL and p99 <= F32C_TOL, (norm_max, p99)


def test_host_aos_entry_points(gpu_ctx):
    import ctypes
    lib = dvm.load()
    x, z, c, s = config2_blobs(777, seed=5)
    parts = np.ascontiguousarray(np.stack([x, z, c, s], axis=1))
    res = np.empty((777, 2))
    st = lib.dvm_bs_self_host(gpu_ctx.h, dvm.F64, 777, parts.ctypes.data_as(ctypes.c_void_p),
                              res.ctypes.data_as(ctypes.c_void_p))
    assert st == 0
    ug, wg = gpu_self(gpu_ctx, x, z, c, s, dvm.F64)
    assert np.array_equal(res[:, 0], ug) and np.array_equal(res[:, 1], wg)
    tx = np.linspace(0, 1, 50)
    tz = np.full(50, 0.5)
    res2 = np.empty((50, 2))
    st = lib.dvm_bs_targets_host(gpu_ctx.h, dvm.F64, 777, parts.ctypes.data_as(ctypes.c_void_p),
                                 50, tx.ctypes.data_as(ctypes.c_void_p),
                                 tz.ctypes.data_as(ctypes.c_void_p),
                                 res2.ctypes.data_as(ctypes.c_void_p))
    assert st == 0
    uo, wo = oracle.bs_targets(x, z, c, s, tx, tz)
    assert rel_err(res2[:, 0], uo) <= F64_TOL and rel_err(res2[:, 1], wo) <= F64_TOL


def test_invalid_precision_rejected(gpu_ctx):
    x, z, c, s = to_dev(*config2_blobs(10))
    with pytest.raises(dvm.DvmError) as e:
        dvm.bs_self(gpu_ctx, x, z, c, s, prec=7)
    assert e.value.status == 1
