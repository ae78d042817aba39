"""DVM_FMM blob-on-blob sums (fmm.hip): the reference's FMMKernel
(src/VelocityKernel.hpp:84-110), whose body there is the direct sum
(src/VelocityKernel.cpp:113-147).  Checked against the oracle's direct sum
(src/VelocityKernel.cpp:30-67): the near field is the same regularised
kernel, so the difference is the far-field truncation, which must fall with
the expansion order as theta^p predicts; tests/unit/test_VelocityKernel.cpp:
141-165's case (N = 100, defaults, 1e-6) is reproduced."""
import numpy as np
import pytest
import torch

import dvmpp_amd as dvm
import oracle
from conftest import rel_err, to_dev
from dvmpp_amd import sim as dsim

pytestmark = pytest.mark.gpu


def _wake(n, seed, sigma=0.01):
    rng = np.random.default_rng(seed)
    # a dense core and a sparse tail, like a wake
    k = n // 2
    x = np.concatenate([rng.normal(0.5, 0.3, k), rng.uniform(-1.0, 6.0, n - k)])
    z = np.concatenate([rng.normal(0.0, 0.2, k), rng.uniform(-1.5, 1.5, n - k)])
    c = rng.normal(0, 1e-3, n)
    s = np.full(n, sigma)
    return x, z, c, s


def _fmm(ctx, x, z, c, s, order, theta):
    ctx.set_fmm(order, theta)
    print("fmm n", x.size, "order", order, "theta", theta, flush=True)
    u, w = dvm.bs_self(ctx, *to_dev(x, z, c, s), prec=dvm.FMM)
    return u.cpu().numpy(), w.cpu().numpy(), ctx.fmm_last_level()


def test_fmm_reruns_bit_for_bit(gpu_ctx):
    # every sum of the fast form has a fixed order (no atomics in any sum), so
    # a second call on the same blobs gives the same bits; a dense cluster
    # (20,000 blobs in a 0.05 square) gives targets with more than 2,048 near
    # sources, summed by a wavefront each
    x, z, c, s = _wake(60000, seed=21)
    rng = np.random.default_rng(22)
    x = np.concatenate([x, rng.uniform(0.0, 0.05, 20000)])
    z = np.concatenate([z, rng.uniform(0.0, 0.05, 20000)])
    c = np.concatenate([c, rng.normal(0.0, 1e-3, 20000)])
    s = np.full(x.size, 0.01)
    u1, w1, lv = _fmm(gpu_ctx, x, z, c, s, 12, 0.5)
    u2, w2, _ = _fmm(gpu_ctx, x, z, c, s, 12, 0.5)
    assert lv >= 2
    assert np.array_equal(u1, u2) and np.array_equal(w1, w2)
    uo, wo = oracle.bs_self(x, z, c, s)
    assert rel_err(u1, uo) <= 2e-6 and rel_err(w1, wo) <= 2e-6


@pytest.mark.parametrize("order,theta,tol", [(10, 0.5, 2e-6), (20, 0.5, 1e-11), (32, 1.0, 1e-8),
                                             (8, 0.7, 1e-3)])
def test_fmm_matches_direct(gpu_ctx, order, theta, tol):
    x, z, c, s = _wake(40000, 3)
    ud, wd = oracle.bs_self(x, z, c, s)
    uf, wf, level = _fmm(gpu_ctx, x, z, c, s, order, theta)
    assert level >= 4  # a real tree, not the direct fallback
    assert rel_err(uf, ud) <= tol and rel_err(wf, wd) <= tol, (rel_err(uf, ud), rel_err(wf, wd))


def test_fmm_error_falls_with_order(gpu_ctx):
    x, z, c, s = _wake(20000, 4)
    ud, _ = oracle.bs_self(x, z, c, s)
    errs = [rel_err(_fmm(gpu_ctx, x, z, c, s, p, 0.5)[0], ud) for p in (4, 8, 12, 16)]
    assert all(b < a for a, b in zip(errs, errs[1:])), errs
    assert errs[-1] < 1e-8


def test_fmm_reference_small_case(gpu_ctx):
    # tests/unit/test_VelocityKernel.cpp:141-165 with the default parameters
    rng = np.random.default_rng(42)
    x, z = rng.uniform(-10, 10, 100), rng.uniform(-10, 10, 100)
    c, s = rng.uniform(-1, 1, 100), np.full(100, 0.1)
    ud, wd = oracle.bs_self(x, z, c, s)
    uf, wf, _ = _fmm(gpu_ctx, x, z, c, s, 10, 0.5)
    assert np.max(np.abs(uf - ud)) <= 1e-6 and np.max(np.abs(wf - wd)) <= 1e-6


def test_fmm_declines_to_direct(gpu_ctx):
    # non-finite input and tiny sets take the FP64 direct sum, bit for bit
    x, z, c, s = _wake(3000, 5)
    x[17] = np.nan
    X, Z, C, S = to_dev(x, z, c, s)
    gpu_ctx.set_fmm(10, 0.5)
    uf, wf = dvm.bs_self(gpu_ctx, X, Z, C, S, prec=dvm.FMM)
    assert gpu_ctx.fmm_last_level() == 0
    ud, wd = dvm.bs_self(gpu_ctx, X, Z, C, S, prec=dvm.F64)
    assert torch.equal(torch.isnan(uf), torch.isnan(ud))
    assert torch.equal(torch.nan_to_num(uf), torch.nan_to_num(ud))
    with pytest.raises(dvm.DvmError):
        gpu_ctx.set_fmm(0, 0.5)
    with pytest.raises(dvm.DvmError):
        gpu_ctx.set_fmm(10, 1.5)
    gpu_ctx.set_fmm(10, 0.5)


def test_fmm_large_core_keeps_regularised_pairs_near(gpu_ctx):
    # cores wide compared with the spacing: the tree stays shallow enough that
    # every pair inside sqrt(10) sigma is summed directly
    x, z, c, s = _wake(20000, 6, sigma=0.2)
    ud, wd = oracle.bs_self(x, z, c, s)
    uf, wf, level = _fmm(gpu_ctx, x, z, c, s, 20, 0.5)
    assert level >= 2
    assert rel_err(uf, ud) <= 1e-10 and rel_err(wf, wd) <= 1e-10


def test_fmm_simulation_step(gpu_ctx):
    # one step of the time loop with FMM sums against the FP64 form from the
    # same state: positions move by dt * (velocity error)
    px, pz = oracle.cylinder_points(128)
    a = dsim.Sim(gpu_ctx, px, pz, dsim.sim_params(seed=2, prec=dvm.F64, rng_mode=dsim.RNG_DEVICE))
    a.run(60)
    st = a.get_state()
    b = dsim.Sim(gpu_ctx, px, pz, dsim.sim_params(seed=2, prec=dvm.FMM, rng_mode=dsim.RNG_DEVICE))
    b.set_state(st)
    b.set_gamma(a.gamma())
    gpu_ctx.set_fmm(20, 0.5)
    for sim in (a, b):
        sim.substep(0)  # increment_step + convect
    assert gpu_ctx.fmm_last_level() >= 2  # b's sum was a tree
    sa, sb = a.get_state(), b.get_state()
    # the collocation sums came from the same tree (FMM targets)
    assert rel_err(b.gamma(), a.gamma()) <= 1e-9
    assert np.max(np.abs(sa["x"] - sb["x"])) <= 1e-9 and np.max(np.abs(sa["z"] - sb["z"])) <= 1e-9
    assert rel_err(sb["u"], sa["u"]) <= 1e-8
