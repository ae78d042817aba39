"""Batched ensemble (dvm_batch_*): each simulation's result is the same alone
and in a mix (bit-for-bit: per-item sum orders are fixed), the graph replay
equals eager launches, a batch slot follows the oracle's step in its
device-stream mode (FP64: to 1e-10) and the single-simulation engine with the
same device random stream (FP32C: to the FP32C tolerance), and a failing
simulation stops alone."""
import numpy as np
import pytest

import dvmpp_amd as dvm
import oracle
from conftest import rel_err
from dvmpp_amd import bodies
from dvmpp_amd import sim as dsim
from dvmpp_amd.batch import Batch, batch_config

pytestmark = pytest.mark.gpu

CASES = [  # body, sheet kw, seed, steps
    (bodies.cylinder(64), dict(nu=1e-3, dt=0.005), 11, 12),
    (bodies.ellipse(48, aspect=2.5, aoa_deg=15), dict(nu=5e-4, dt=0.004, ux=0.8), 12, 9),
    (bodies.rounded_square(40), dict(nu=2e-3, dt=0.006), 13, 15),
]


def run_batch(ctx, cases, prec=dvm.F32C, scheme=dsim.EULER, crossing=dsim.REFLECT, graph=True,
              extra_steps=0):
    cfg = batch_config(slots=8, slot_capacity=1 << 15, max_panels=128, scheme=scheme,
                       crossing=crossing, prec=prec, use_graph=graph)
    b = Batch(ctx, cfg)
    slots = [b.admit(px, pz, dsim.sheet_params(**kw), seed, steps)
             for (px, pz), kw, seed, steps in cases]
    b.step(max(c[3] for c in cases) + extra_steps)
    out = [b.state(s) for s in slots]
    infos = [b.info(s) for s in slots]
    done = b.poll()
    b.close()
    return out, infos, done


def test_alone_equals_in_mix(gpu_ctx):
    mix, infos, done = run_batch(gpu_ctx, CASES)
    assert sorted(s for s, _ in done) == [0, 1, 2] and all(st == 0 for _, st in done)
    for k, case in enumerate(CASES):
        assert infos[k]["step"] == case[3] and infos[k]["active"] == 0
        alone, _, _ = run_batch(gpu_ctx, [case])
        for f in ("x", "z", "circ", "sigma", "u", "w"):
            assert np.array_equal(mix[k][f], alone[0][f]), (k, f)


def test_graph_equals_eager(gpu_ctx):
    a, _, _ = run_batch(gpu_ctx, CASES, graph=True)
    b, _, _ = run_batch(gpu_ctx, CASES, graph=False)
    for k in range(len(CASES)):
        for f in ("x", "z", "circ", "u", "w"):
            assert np.array_equal(a[k][f], b[k][f]), (k, f)


@pytest.mark.parametrize("prec,scheme,crossing", [(dvm.F64, dsim.EULER, dsim.REFLECT),
                                                  (dvm.F64, dsim.RK3, dsim.DELETE),
                                                  (dvm.F32C, dsim.EULER, dsim.REFLECT)])
def test_slot_matches_single_simulation(gpu_ctx, prec, scheme, crossing):
    (px, pz), kw, seed, steps = CASES[0]
    out, infos, _ = run_batch(gpu_ctx, [CASES[0]], prec=prec, scheme=scheme, crossing=crossing)
    sp = dsim.sim_params(scheme=scheme, crossing=crossing, seed=seed, prec=prec,
                         rng_mode=dsim.RNG_DEVICE, **kw)
    s = dsim.Sim(gpu_ctx, px, pz, sp)
    s.run(steps)
    ref = s.get_state()
    assert ref["x"].size == out[0]["x"].size
    tol = 1e-10 if prec == dvm.F64 else 1e-6
    scale = np.max(np.abs(ref["x"]))
    assert np.max(np.abs(ref["x"] - out[0]["x"])) <= tol * scale
    assert np.max(np.abs(ref["z"] - out[0]["z"])) <= tol * scale
    assert np.array_equal(ref["sigma"], out[0]["sigma"])
    lo = s.loads()
    assert abs(lo[2] - infos[0]["loads"][2]) <= max(1e-6, 1e3 * tol) * max(1.0, abs(lo[2]))
    s.close()


@pytest.mark.parametrize("case", [CASES[2], (bodies.ellipse(300, aspect=3.0, aoa_deg=20),
                                             dict(nu=5e-4, dt=0.004), 21, 10)])
def test_f32c_panel_groups_match_single_simulation(gpu_ctx, case):
    # The FP32C slot takes the sheet's velocity by 16-panel groups (40
    # panels: two full groups and one of 8; 300: 18 and one of 12); the
    # single simulation by the body-wide series and panel sums.
    (px, pz), kw, seed, steps = case
    out, infos, _ = run_batch_wide(gpu_ctx, [case])
    sp = dsim.sim_params(seed=seed, prec=dvm.F32C, rng_mode=dsim.RNG_DEVICE, **kw)
    s = dsim.Sim(gpu_ctx, px, pz, sp)
    s.run(steps)
    ref = s.get_state()
    assert ref["x"].size == out[0]["x"].size
    scale = np.max(np.abs(ref["x"]))
    assert np.max(np.abs(ref["x"] - out[0]["x"])) <= 1e-6 * scale
    assert np.max(np.abs(ref["z"] - out[0]["z"])) <= 1e-6 * scale
    assert abs(s.loads()[2] - infos[0]["loads"][2]) <= 1e-6 * max(1.0, abs(s.loads()[2]))
    s.close()


def run_batch_wide(ctx, cases):
    cfg = batch_config(slots=2, slot_capacity=1 << 15, max_panels=512, prec=dvm.F32C)
    b = Batch(ctx, cfg)
    slots = [b.admit(px, pz, dsim.sheet_params(**kw), seed, steps)
             for (px, pz), kw, seed, steps in cases]
    b.step(max(c[3] for c in cases))
    out = [b.state(s) for s in slots]
    infos = [b.info(s) for s in slots]
    done = b.poll()
    b.close()
    return out, infos, done


@pytest.mark.parametrize("prec,tol", [(dvm.F64, 1e-10), (dvm.F32C, 1e-6)])
def test_batch_remesh_matches_single_simulation(gpu_ctx, prec, tol):
    # FP32C: the remeshed slot's next stage re-sorts its blobs for the pair
    # sum and the panel-group sheet velocity
    rp = dsim.remesh_params(frequency=4, kernel=dsim.M4PRIME, xmin=-2.0, xmax=4.0, zmin=-2.0,
                            zmax=2.0, dx=0.05, dz=0.05, sigma=0.06, circ_threshold=1e-9)
    (px, pz), kw, seed, _ = CASES[0]
    cfg = batch_config(slots=2, slot_capacity=1 << 15, max_panels=128, prec=prec, remesh=rp)
    b = Batch(gpu_ctx, cfg)
    slot = b.admit(px, pz, dsim.sheet_params(**kw), seed, 10)
    b.step(10)
    out = b.state(slot)
    b.close()
    s = dsim.Sim(gpu_ctx, px, pz, dsim.sim_params(seed=seed, prec=prec, rng_mode=dsim.RNG_DEVICE,
                                                   remesh=rp, **kw))
    s.run(10)
    ref = s.get_state()
    assert ref["x"].size == out["x"].size
    assert np.max(np.abs(ref["x"] - out["x"])) <= tol * np.max(np.abs(ref["x"]))
    assert np.max(np.abs(ref["z"] - out["z"])) <= tol * np.max(np.abs(ref["z"]))
    assert np.max(np.abs(ref["circ"] - out["circ"])) <= tol * np.max(np.abs(ref["circ"]))
    assert np.array_equal(ref["sigma"], out["sigma"])
    s.close()


def test_failure_stops_one_slot(gpu_ctx):
    cases = list(CASES[:2]) + [(bodies.cylinder(32), dict(ux=0.0, uz=0.0), 5, 10)]
    out, infos, done = run_batch(gpu_ctx, cases)
    status = dict(done)
    assert status[2] == 3  # DVM_E_NAN: zero free stream -> NaN release counts
    assert status[0] == 0 and status[1] == 0
    alone, _, _ = run_batch(gpu_ctx, CASES[:2])
    for k in range(2):
        assert np.array_equal(out[k]["x"], alone[k]["x"])


def test_slot_reuse_and_capacity_error(gpu_ctx):
    cfg = batch_config(slots=2, slot_capacity=1024, max_panels=64)
    b = Batch(gpu_ctx, cfg)
    s0 = b.admit(*bodies.cylinder(64), dsim.sheet_params(), 1, 30)  # overflows 1024 blobs
    s1 = b.admit(*bodies.cylinder(16), dsim.sheet_params(), 2, 3)
    with pytest.raises(dvm.DvmError):
        b.admit(*bodies.cylinder(16), dsim.sheet_params(), 3, 3)  # no free slot
    b.step(30)
    status = dict(b.poll())
    assert status[s1] == 0 and status[s0] == 2  # DVM_E_SIZE
    s2 = b.admit(*bodies.cylinder(16), dsim.sheet_params(), 4, 2)
    assert s2 in (s0, s1)
    b.step(2)
    assert dict(b.poll())[s2] == 0
    b.close()


@pytest.mark.parametrize("scheme,crossing", [(dsim.EULER, dsim.REFLECT), (dsim.RK3, dsim.DELETE)])
def test_fp64_slots_match_oracle(gpu_ctx, scheme, crossing):
    """FP64 slots of one batch against the oracle's time step
    (src/DVMBase.cpp:86-133) with the same device random stream (Philox
    keyed by seed, step and blob index; the oracle's Philox is pinned by the
    generator's known-answer vectors): 8 steps free-running from rest, blob
    counts and core sizes equal, positions, circulations and Cd, Cl to 1e-10."""
    cases = [(bodies.cylinder(128), dict(nu=1e-3, dt=0.005), 31),
             (bodies.ellipse(96, aspect=2.0, aoa_deg=10), dict(nu=5e-4, dt=0.004, ux=0.8), 32),
             (bodies.rounded_square(80), dict(nu=2e-3, dt=0.006), 33)]
    steps = 8
    cfg = batch_config(slots=4, slot_capacity=1 << 14, max_panels=128, scheme=scheme,
                       crossing=crossing, prec=dvm.F64)
    b = Batch(gpu_ctx, cfg)
    slots = [b.admit(px, pz, dsim.sheet_params(**kw), seed, steps) for (px, pz), kw, seed in cases]
    b.step(steps)
    for slot, ((px, pz), kw, seed) in zip(slots, cases):
        got, info = b.state(slot), b.info(slot)
        assert info["status"] == 0 and info["step"] == steps
        sp = dsim.sheet_params(**kw)
        op = oracle.make_params(rho=sp.rho, nu=sp.nu, dt=sp.dt, ux=sp.ux, uz=sp.uz,
                                cutoff_exp=sp.cutoff_exp, max_npv=sp.max_npv, scheme=scheme,
                                crossing=crossing, seed=seed, rng_mode=1)
        osim = oracle.Sim(px, pz, op)
        for _ in range(steps):
            assert osim.step() == 0
        ref = osim.get_state()
        assert ref["x"].size == got["x"].size, (slot, ref["x"].size, got["x"].size)
        assert np.array_equal(ref["sigma"], got["sigma"])
        assert rel_err(got["circ"], ref["circ"]) <= 1e-10
        scale = max(1.0, float(np.max(np.abs(ref["x"]))))
        assert np.max(np.abs(got["x"] - ref["x"])) <= 1e-10 * scale
        assert np.max(np.abs(got["z"] - ref["z"])) <= 1e-10 * scale
        for f in ("u", "w"):
            assert rel_err(got[f], ref[f]) <= 1e-9, (slot, f)
        lo = osim.loads()
        for k in (2, 3):
            assert abs(info["loads"][k] - lo[k]) <= 1e-10 * max(1.0, abs(lo[k])), (slot, k)
    b.close()


def test_set_state_one_slot_matches_single_simulation(gpu_ctx):
    """A one-slot batch started from a saved wake (dvm_batch_set_state) steps
    as the single simulation started from the same state: FP32C RK3 steps
    with remesh, device random stream (positions to 1e-6, Cd to 1e-6)."""
    from dvmpp_amd.workloads import rk3_cylinder, wake_snapshot
    px, pz, sp, sig = rk3_cylinder()
    rp = dsim.remesh_params(frequency=2, kernel=dsim.M4PRIME, xmin=-5.0, xmax=25.0, zmin=-10.0,
                            zmax=10.0, dx=0.01, dz=0.01, sigma=sig, circ_threshold=1e-10)
    snap = wake_snapshot(n_target=40_000, sigma=sig)
    b = Batch(gpu_ctx, batch_config(slots=1, slot_capacity=1 << 17, max_panels=512,
                                    scheme=dsim.RK3, prec=dvm.F32C, remesh=rp))
    slot = b.admit(px, pz, sp, 7, 100)
    b.set_state(slot, snap["x"], snap["z"], snap["circ"], snap["sigma"])
    assert b.info(slot)["n"] == snap["x"].size
    b.step(3)  # a remesh at step 2
    got, info = b.state(slot), b.info(slot)
    b.close()
    s = dsim.Sim(gpu_ctx, px, pz, dsim.SimParams(sp, dsim.RK3, dsim.REFLECT, 7, dvm.F32C,
                                                 dsim.RNG_DEVICE, rp))
    s.set_state(snap)
    s.run(3)
    ref, lo = s.get_state(), s.loads()
    s.close()
    assert ref["x"].size == got["x"].size
    scale = np.max(np.abs(ref["x"]))
    assert np.max(np.abs(ref["x"] - got["x"])) <= 1e-6 * scale
    assert np.max(np.abs(ref["z"] - got["z"])) <= 1e-6 * scale
    assert np.array_equal(ref["sigma"], got["sigma"])
    assert abs(lo[2] - info["loads"][2]) <= 1e-6 and abs(lo[3] - info["loads"][3]) <= 1e-6


def _spread12(v):
    v = v.astype(np.uint32) & np.uint32(0xFFFF)
    for sh, m in ((8, 0x00FF00FF), (4, 0x0F0F0F0F), (2, 0x33333333), (1, 0x55555555)):
        v = (v | (v << np.uint32(sh))) & np.uint32(m)
    return v


def _morton_order(x, z):
    """Stable order of 24-bit Morton keys over the set's bounding box (12 bits
    per axis, truncated; the pair sum's tile order)."""
    x0, z0 = x.min(), z.min()
    ext = max(max(x.max() - x0, z.max() - z0), 1e-300)
    scale = 4095.0 / ext
    qx = np.clip((x - x0) * scale, 0.0, 4095.0).astype(np.uint32)
    qz = np.clip((z - z0) * scale, 0.0, 4095.0).astype(np.uint32)
    keys = _spread12(qx) | (_spread12(qz) << np.uint32(1))
    return np.argsort(keys, kind="stable").astype(np.uint32)


def test_sort_order_is_stable_morton_order(gpu_ctx):
    """The batch's radix sort (several workgroups per slot, counts on the
    device) gives each slot the stable Morton order: slots of 70,001 blobs
    (18 sort tiles), 5,000 (a partial second tile), 100 and one left empty;
    positions on a coarse lattice so that many keys tie."""
    rng = np.random.default_rng(5)
    cfg = batch_config(slots=4, slot_capacity=1 << 17, max_panels=64, prec=dvm.F32C)
    b = Batch(gpu_ctx, cfg)
    (px, pz), kw = bodies.cylinder(32), dict(nu=1e-3, dt=0.005)
    sizes = [70_001, 5_000, 100]
    sets = []
    for n in sizes:
        slot = b.admit(px, pz, dsim.sheet_params(**kw), 3, 5)
        x = 2.0 + np.round(rng.uniform(0, 6, n) * 40) / 40  # outside the unit body
        z = np.round(rng.uniform(-2, 2, n) * 40) / 40
        c = rng.normal(0, 1e-4, n)
        b.set_state(slot, x, z, c, np.full(n, 0.02))
        sets.append((slot, x, z))
    b.admit(px, pz, dsim.sheet_params(**kw), 4, 5)  # its sheet alone: no blobs yet
    b.step(1)  # the sort runs on the positions set above
    for slot, x, z in sets:
        got = b.sort_order(slot, x.size)
        assert np.array_equal(got, _morton_order(x, z)), slot
    b.close()


def _fmm_batch(ctx, with_small, fmm_min, steps=1):
    """A cylinder slot started from a 40K-blob wake (Euler, FP32C), beside two
    small bodies when with_small; fmm_min > 0 gives slots of that many blobs
    the fast multipole sum (order 16)."""
    from dvmpp_amd.workloads import rk3_cylinder, wake_snapshot
    px, pz, sp, sig = rk3_cylinder()
    snap = wake_snapshot(n_target=40_000, sigma=sig)
    b = Batch(ctx, batch_config(slots=4, slot_capacity=1 << 17, max_panels=512,
                                scheme=dsim.EULER, prec=dvm.F32C))
    slots = []
    if with_small:
        slots += [b.admit(cpx, cpz, dsim.sheet_params(**kw), seed, 6)
                  for (cpx, cpz), kw, seed, _ in CASES[:2]]
    big = b.admit(px, pz, sp, 7, 6)
    b.set_state(big, snap["x"], snap["z"], snap["circ"], snap["sigma"])
    if fmm_min:
        b.set_fmm(fmm_min, order=16, theta=0.5)
    b.step(steps)
    out = {s: b.state(s) for s in slots + [big]}
    b.close()
    return out, slots, big, snap["x"].size


def test_fmm_slots(gpu_ctx):
    """dvm_batch_set_fmm: the large slot's velocities come from the fast
    multipole sum (within the FP32C form's 1e-5 normwise of the FP32C direct
    sum's at the same state), the small slots keep the FP32C sum bit for bit, and the large
    slot's result is the same alone and in the mix."""
    on, small, big, n0 = _fmm_batch(gpu_ctx, True, 20_000)
    off, _, _, _ = _fmm_batch(gpu_ctx, True, 0)
    for s in small:
        for f in ("x", "z", "circ", "sigma", "u", "w"):
            assert np.array_equal(on[s][f], off[s][f]), (s, f)
    vu = np.hypot(off[big]["u"][:n0], off[big]["w"][:n0])
    d = np.hypot(on[big]["u"][:n0] - off[big]["u"][:n0], on[big]["w"][:n0] - off[big]["w"][:n0])
    assert np.max(d) <= 1e-5 * np.max(vu), np.max(d) / np.max(vu)
    alone, _, big_alone, _ = _fmm_batch(gpu_ctx, False, 20_000)
    for f in ("x", "z", "circ", "sigma", "u", "w"):
        assert np.array_equal(alone[big_alone][f], on[big][f]), f
    # a second step (the slot has grown) runs as well
    two, _, big2, _ = _fmm_batch(gpu_ctx, False, 20_000, steps=2)
    assert two[big2]["x"].size > on[big]["x"].size
    assert np.all(np.isfinite(two[big2]["u"])) and np.all(np.isfinite(two[big2]["w"]))


def test_fmm_slots_need_euler_f32c(gpu_ctx):
    for scheme, prec in ((dsim.RK3, dvm.F32C), (dsim.EULER, dvm.F64)):
        b = Batch(gpu_ctx, batch_config(slots=2, slot_capacity=1 << 12, max_panels=64,
                                        scheme=scheme, prec=prec))
        with pytest.raises(dvm.DvmError):
            b.set_fmm(1000)
        b.set_fmm(0)  # off is always accepted
        b.close()
