"""Continuous-batching server: requests arriving while others run are admitted
into free slots, every request completes, each result is bit-for-bit the one
the same request gives when served alone, and admission respects the HBM
budget."""
import numpy as np
import pytest

import dvmpp_amd as dvm
from dvmpp_amd.batch import batch_config
from dvmpp_amd.serve import Request, Server
from dvmpp_amd.workloads import request_trace

pytestmark = pytest.mark.gpu


def make_requests(count=6):
    trace = request_trace(count, rate=50.0, seed=3, steps_range=(6, 18), panels_range=(32, 64))
    return [Request(i, t, px, pz, sp, seed, st) for i, (t, (px, pz, sp, seed, st)) in enumerate(trace)]


def serve(ctx, reqs, slots, budget=200 << 30, fmm_min=0):
    srv = Server(ctx, batch_config(slots=slots, slot_capacity=1 << 14, max_panels=64,
                                   prec=dvm.F32C), budget_bytes=budget, keep_state=True,
                 fmm_min=fmm_min, fmm_order=16)
    done, wall = srv.run(reqs)
    srv.close()
    return {r.rid: r for r in done}, wall


@pytest.mark.parametrize("fmm_min", [0, 600])
def test_served_equals_alone(gpu_ctx, fmm_min):
    """fmm_min 600: once a request's slot passes 600 blobs (mid-run) its sum
    goes to the fast multipole form (or, where the tree declines such a small
    set, the direct sum): the result is still the request's alone."""
    mixed, _ = serve(gpu_ctx, make_requests(), slots=4, fmm_min=fmm_min)
    assert len(mixed) == 6 and all(r.status == 0 for r in mixed.values())
    # requests overlapped: some were admitted before others finished
    spans = sorted((r.admitted, r.finished) for r in mixed.values())
    assert any(spans[k + 1][0] < spans[k][1] for k in range(len(spans) - 1))
    for req in make_requests():
        alone, _ = serve(gpu_ctx, [Request(req.rid, 0.0, req.px, req.pz, req.params, req.seed,
                                           req.steps)], slots=1, fmm_min=fmm_min)
        a, m = alone[req.rid].state, mixed[req.rid].state
        for f in ("x", "z", "circ", "sigma", "u", "w"):
            assert np.array_equal(a[f], m[f]), (req.rid, f)
        assert alone[req.rid].loads == mixed[req.rid].loads


def test_budget_limits_admission(gpu_ctx):
    # a budget of one single-slot region: one request on the device at a time
    cfg = batch_config(slots=1, slot_capacity=1 << 14, max_panels=64, prec=dvm.F32C)
    srv = Server(gpu_ctx, cfg, budget_bytes=0)
    one = srv.class_bytes[1 << 14]
    srv.close()
    srv = Server(gpu_ctx, cfg, budget_bytes=one, keep_state=True)
    done, _ = srv.run(make_requests(4))
    assert srv.peak_bytes == one
    srv.close()
    spans = sorted((r.admitted, r.finished) for r in done)
    assert len(spans) == 4
    assert all(spans[k + 1][0] >= spans[k][1] for k in range(len(spans) - 1))
    assert all(r.status == 0 for r in done)


def test_capacity_classes(gpu_ctx):
    # small requests take the small class; results equal the single-class run
    reqs = make_requests(6)
    cfg = batch_config(slots=4, slot_capacity=1 << 14, max_panels=64, prec=dvm.F32C)
    srv = Server(gpu_ctx, cfg, budget_bytes=200 << 30, classes=[1 << 12, 1 << 14], keep_state=True)
    done, _ = srv.run(reqs)
    srv.close()
    single, _ = serve(gpu_ctx, make_requests(6), slots=4)
    assert all(r.status == 0 for r in done)
    for r in done:
        assert r.capacity == min(c for c in (1 << 12, 1 << 14) if srv.projected(r) <= c)
        for f in ("x", "z", "circ", "u", "w"):
            assert np.array_equal(r.state[f], single[r.rid].state[f]), (r.rid, f)
