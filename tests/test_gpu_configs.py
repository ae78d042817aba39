"""GPU parity on the benchmark configurations' own workloads.

C1: the plain cylinder (256 panels, Re 1000, Euler, dt 0.005, REFLECT, seed 0,
400 steps; src/DVMBase.cpp:86-133) in lock-step with re-sync against the
oracle: steps 1-5 from rest, then the GPU runs on alone and the oracle is
re-synchronised to the GPU state (blobs, step index and random-stream
position) before steps 100, 200 and 400, each compared after one step.

C3: the RK3 cylinder (512 panels, nu 5e-4, dt 0.0025, M4' remesh every 20
steps on a 0.01 grid; src/DVMBase.cpp:160-206) in lock-step across the remesh
at step 20, and the chosen FP32C sum against FP64 on the 1.5M-blob remeshed
wake of its plateau.

Bars: FP64 velocities, gamma and Cd/Cl within 1e-10 relative (Cd/Cl with an
absolute floor of 1e-10 x max(|Cd|, |Cl|)); circulations within 1e-10 (the
nascent ones are gamma ds / n); blob counts, IDs and core sizes bit-exact;
positions within 1e-12 of the coordinate scale (they carry the velocities'
last bits through one convection).
"""
import numpy as np
import pytest

import dvmpp_amd as dvm
import oracle
from conftest import rel_err
from dvmpp_amd import workloads
from dvmpp_amd import sim as dsim

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _params(**kw):
    """Matching oracle and library parameter sets (reference random stream)."""
    o = oracle.make_params(**kw)
    rm = dsim.remesh_params(frequency=o.remesh_frequency, kernel=o.remesh_kernel,
                            xmin=o.remesh_xmin, xmax=o.remesh_xmax, zmin=o.remesh_zmin,
                            zmax=o.remesh_zmax, dx=o.remesh_dx, dz=o.remesh_dz,
                            sigma=o.remesh_sigma, circ_threshold=o.remesh_circ_threshold)
    g = dsim.sim_params(scheme=o.scheme, crossing=o.crossing, seed=o.seed, prec=dvm.F64,
                        rng_mode=dsim.RNG_REFERENCE, remesh=rm, rho=o.rho, nu=o.nu, dt=o.dt,
                        ux=o.ux, uz=o.uz, cutoff_exp=o.cutoff_exp, max_npv=o.max_npv)
    return o, g


def _resync(osim, gsim):
    """The oracle takes the GPU simulation's state, step and stream position."""
    osim.set_state(gsim.get_state())
    osim.set_clock(gsim.step_index(), gsim.rng_draws())


def _step_and_compare(osim, gsim, tag):
    assert osim.step() == 0
    gsim.step()
    a, b = osim.get_state(), gsim.get_state()
    assert a["x"].size == b["x"].size, tag
    assert osim.rng_draws() == gsim.rng_draws(), tag
    assert osim.step_index() == gsim.step_index(), tag
    for f in ("id", "sigma"):
        assert np.array_equal(a[f], b[f]), (tag, f)
    # nascent circulations are gamma_i ds_i / n_i: gamma's last bits
    assert rel_err(b["circ"], a["circ"]) <= TOL, tag
    scale = max(1.0, float(np.max(np.abs(a["x"]))))
    assert np.max(np.abs(a["x"] - b["x"])) <= 1e-12 * scale, tag
    assert np.max(np.abs(a["z"] - b["z"])) <= 1e-12 * scale, tag
    for f in ("u", "w", "uvs", "wvs"):
        assert rel_err(b[f], a[f]) <= TOL, (tag, f)
    assert rel_err(gsim.gamma(), osim.gamma()) <= TOL, tag
    lo, lg = osim.loads(), gsim.loads()
    floor = TOL * max(abs(lo[2]), abs(lo[3]))
    assert abs(lg[2] - lo[2]) <= max(TOL * abs(lo[2]), floor), (tag, lo, lg)
    assert abs(lg[3] - lo[3]) <= max(TOL * abs(lo[3]), floor), (tag, lo, lg)
    return a["x"].size


def test_c1_plain_cylinder_lockstep(gpu_ctx):
    px, pz, sp = workloads.plain_cylinder()
    o, g = _params(rho=sp.rho, nu=sp.nu, dt=sp.dt, ux=sp.ux, uz=sp.uz, cutoff_exp=sp.cutoff_exp,
                   max_npv=sp.max_npv, scheme=0, crossing=2, seed=0)
    osim = oracle.Sim(px, pz, o)
    gsim = dsim.Sim(gpu_ctx, px, pz, g)
    try:
        sizes = {}
        for k in range(1, 6):
            _resync(osim, gsim)
            sizes[k] = _step_and_compare(osim, gsim, f"step {k}")
        for k in (100, 200, 400):
            gsim.run(k - 1 - gsim.step_index())
            assert gsim.step_index() == k - 1
            _resync(osim, gsim)
            sizes[k] = _step_and_compare(osim, gsim, f"step {k}")
        # every panel releases at least one blob per step (src/VortexSheet.cpp:287)
        assert sizes[400] >= 256 * 200, sizes
        print("C1 blob counts at the compared steps:", sizes)
    finally:
        gsim.close()


def test_c3_rk3_cylinder_lockstep_across_remesh(gpu_ctx):
    px, pz, sp, sig = workloads.rk3_cylinder()
    o, g = _params(rho=sp.rho, nu=sp.nu, dt=sp.dt, ux=sp.ux, uz=sp.uz, cutoff_exp=sp.cutoff_exp,
                   max_npv=sp.max_npv, scheme=1, crossing=2, seed=0, remesh_frequency=20,
                   remesh_kernel=0, remesh_xmin=-5.0, remesh_xmax=25.0, remesh_zmin=-10.0,
                   remesh_zmax=10.0, remesh_dx=0.01, remesh_dz=0.01, remesh_sigma=sig,
                   remesh_circ_threshold=1e-10)
    osim = oracle.Sim(px, pz, o)
    gsim = dsim.Sim(gpu_ctx, px, pz, g)
    try:
        _resync(osim, gsim)
        _step_and_compare(osim, gsim, "step 1")
        gsim.run(18 - gsim.step_index())
        sizes = []
        for k in (19, 20, 21):  # the remesh runs in step 20
            _resync(osim, gsim)
            sizes.append(_step_and_compare(osim, gsim, f"step {k}"))
        # the remesh put the ~15K-blob early wake on ~2.7K grid nodes
        assert sizes[1] < sizes[0] // 2, sizes
        print("C3 blob counts after steps 19, 20, 21:", sizes)
    finally:
        gsim.close()


def test_c3_plateau_f32c_against_f64(gpu_ctx):
    """The chosen FP32-compensated sum against FP64 on the remeshed wake of the
    RK3 cylinder's plateau (1.49M blobs on the 0.01 grid): per-target relative
    error p99 <= 1e-5 and normwise <= 1e-5."""
    import torch
    snap = workloads.wake_snapshot()
    x, z, c, s = snap["x"], snap["z"], snap["circ"], snap["sigma"]
    assert x.size > 1_400_000
    X, Z, C, S = (torch.from_numpy(np.ascontiguousarray(a)).to("cuda") for a in (x, z, c, s))
    u64, w64 = dvm.bs_self(gpu_ctx, X, Z, C, S, prec=dvm.F64)
    u32, w32 = dvm.bs_self(gpu_ctx, X, Z, C, S, prec=dvm.F32C)
    torch.cuda.synchronize()
    v64 = torch.sqrt(u64 * u64 + w64 * w64)
    dv = torch.sqrt((u32 - u64) ** 2 + (w32 - w64) ** 2)
    norm = float(torch.linalg.norm(torch.stack([u32 - u64, w32 - w64])) /
                 torch.linalg.norm(torch.stack([u64, w64])))
    rel = (dv / torch.clamp(v64, min=1e-300)).cpu().numpy()
    p99 = float(np.percentile(rel, 99))
    print(f"C3 plateau N={x.size}: FP32C normwise {norm:.2e}, p99 {p99:.2e}")
    assert norm <= 1e-5 and p99 <= 1e-5


def test_c4_ensemble_in_one_batch(gpu_ctx):
    """C4: the 128 uneven bodies (ellipses and rounded squares, 128-512
    panels, nu log-uniform in [2e-4, 2e-3]) admitted into one FP32C batch and
    stepped together.  Eight slots spread over the body sizes follow their
    single-simulation runs (same device random stream, FP32C sums: positions
    to 1e-6, Cd and Cl to 1e-6); two slots give bit for bit the same result
    alone as in the mix."""
    from dvmpp_amd.batch import Batch, batch_config
    ens = workloads.ensemble(128)
    steps = 3
    cfg = batch_config(slots=128, slot_capacity=1 << 13, max_panels=512, prec=dvm.F32C)
    b = Batch(gpu_ctx, cfg)
    # set up concurrently (dvm_batch_admit_many); the alone runs below admit
    # one body at a time on the context's stream
    slots = b.admit_many([(px, pz, sp, seed, steps) for px, pz, sp, seed, _ in ens])
    b.step(steps)
    done = dict(b.poll(256))
    assert len(done) == 128 and all(st == 0 for st in done.values())
    order = np.argsort([px.size for px, *_ in ens])
    picks = [int(order[q]) for q in np.linspace(0, 127, 8).round().astype(int)]
    mix = {k: (b.state(slots[k]), b.info(slots[k])) for k in picks}
    b.close()
    panels = sorted(ens[k][0].size - 1 for k in picks)
    assert panels[0] <= 160 and panels[-1] >= 480, panels
    for k in picks:
        px, pz, sp, seed, _ = ens[k]
        s = dsim.Sim(gpu_ctx, px, pz, dsim.SimParams(sp, dsim.EULER, dsim.REFLECT, seed, dvm.F32C,
                                                     dsim.RNG_DEVICE, dsim.remesh_params()))
        s.run(steps)
        ref, lo = s.get_state(), s.loads()
        s.close()
        got, info = mix[k]
        assert ref["x"].size == got["x"].size, k
        scale = max(1.0, float(np.max(np.abs(ref["x"]))))
        assert np.max(np.abs(ref["x"] - got["x"])) <= 1e-6 * scale, k
        assert np.max(np.abs(ref["z"] - got["z"])) <= 1e-6 * scale, k
        assert np.array_equal(ref["sigma"], got["sigma"]), k
        assert abs(lo[2] - info["loads"][2]) <= 1e-6 and abs(lo[3] - info["loads"][3]) <= 1e-6, k
    for k in (picks[0], picks[-1]):
        px, pz, sp, seed, _ = ens[k]
        b1 = Batch(gpu_ctx, cfg)
        s1 = b1.admit(px, pz, sp, seed, steps)
        b1.step(steps)
        alone = b1.state(s1)
        b1.close()
        for f in ("x", "z", "circ", "sigma", "u", "w"):
            assert np.array_equal(alone[f], mix[k][0][f]), (k, f)


def test_c5_served_requests_equal_alone(gpu_ctx):
    """C5: ten requests drawn from the ensemble distribution (128-512 panels),
    arriving faster than four slots can take them, are served by the
    continuous-batching loop; each result is bit for bit the one the same
    request gives when served alone."""
    from dvmpp_amd.batch import batch_config
    from dvmpp_amd.serve import Request, Server
    trace = workloads.request_trace(10, rate=500.0, seed=5, steps_range=(6, 14),
                                    panels_range=(128, 512))

    def reqs():
        return [Request(i, t, px, pz, sp, seed, st)
                for i, (t, (px, pz, sp, seed, st)) in enumerate(trace)]

    def serve(rs, slots):
        srv = Server(gpu_ctx, batch_config(slots=slots, slot_capacity=1 << 15, max_panels=512,
                                           prec=dvm.F32C), budget_bytes=200 << 30, keep_state=True)
        done, _ = srv.run(rs)
        srv.close()
        return {r.rid: r for r in done}

    mixed = serve(reqs(), slots=4)
    assert len(mixed) == 10 and all(r.status == 0 for r in mixed.values())
    panels = [r.px.size - 1 for r in mixed.values()]
    assert min(panels) >= 128 and max(panels) <= 512
    spans = sorted((r.admitted, r.finished) for r in mixed.values())
    overlaps = sum(spans[k + 1][0] < spans[k][1] for k in range(len(spans) - 1))
    assert overlaps >= 3, spans
    for req in reqs():
        alone = serve([Request(req.rid, 0.0, req.px, req.pz, req.params, req.seed, req.steps)],
                      slots=4)[req.rid]
        m = mixed[req.rid]
        for f in ("x", "z", "circ", "sigma", "u", "w"):
            assert np.array_equal(alone.state[f], m.state[f]), (req.rid, f)
        assert alone.loads == m.loads
