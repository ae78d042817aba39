"""Simulation snapshot/restore, asynchronous copy-out through pinned host
buffers and events, and probes on a running simulation (include/dvm_hip.h).
The reference has no checkpoint (its output is per-step text,
src/DVMBase.cpp:151-158); the property checked is that a restored simulation
continues bit for bit as the original does, in both random-stream modes."""
import numpy as np
import pytest

import dvmpp_amd as dvm
import oracle
from conftest import to_dev
from dvmpp_amd import sim as dsim

pytestmark = pytest.mark.gpu


def _params(mode):
    rm = dsim.remesh_params(frequency=5, kernel=dsim.M4PRIME, xmin=-2.0, xmax=4.0, zmin=-2.0,
                            zmax=2.0, dx=0.05, dz=0.05, sigma=0.06, circ_threshold=1e-9)
    if mode == "reference":
        return dsim.sim_params(scheme=dsim.RK3, crossing=dsim.REFLECT, seed=4, prec=dvm.F64,
                               rng_mode=dsim.RNG_REFERENCE, remesh=rm)
    return dsim.sim_params(scheme=dsim.EULER, crossing=dsim.ABSORB, seed=9, prec=dvm.F32C,
                           rng_mode=dsim.RNG_DEVICE, remesh=rm)


def _same(a, b):
    sa, sb = a.get_state(), b.get_state()
    assert sa.keys() == sb.keys()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    assert np.array_equal(a.gamma(), b.gamma())
    assert a.loads() == b.loads()
    assert a.step_index() == b.step_index()
    assert a.rng_draws() == b.rng_draws()


@pytest.mark.parametrize("mode", ["reference", "device"])
def test_restore_continues_bit_exact(gpu_ctx, mode):
    px, pz = oracle.cylinder_points(96)
    a = dsim.Sim(gpu_ctx, px, pz, _params(mode))
    a.run(7)
    snap = a.snapshot()
    b = dsim.Sim(gpu_ctx, px, pz, _params(mode))
    b.restore(snap)
    _same(a, b)
    a.run(9)  # crosses two remesh steps
    b.run(9)
    _same(a, b)
    # restoring over a simulation that has run further rewinds it
    b.restore(snap)
    b.run(9)
    _same(a, b)
    assert a.size() > 0


def test_restore_rejects_foreign_snapshots(gpu_ctx):
    px, pz = oracle.cylinder_points(96)
    a = dsim.Sim(gpu_ctx, px, pz, _params("reference"))
    a.run(2)
    snap = a.snapshot()
    other_body = dsim.Sim(gpu_ctx, *oracle.cylinder_points(64), _params("reference"))
    with pytest.raises(dvm.DvmError):
        other_body.restore(snap)
    moved = dsim.Sim(gpu_ctx, px + 1e-3, pz, _params("reference"))
    with pytest.raises(dvm.DvmError):
        moved.restore(snap)
    other_params = dsim.Sim(gpu_ctx, px, pz, _params("device"))
    with pytest.raises(dvm.DvmError):
        other_params.restore(snap)
    fresh = dsim.Sim(gpu_ctx, px, pz, _params("reference"))
    with pytest.raises(dvm.DvmError) as e:
        fresh.restore(snap[:-8])
    assert e.value.status == 2
    bad = bytearray(snap)
    bad[0] ^= 0xFF
    with pytest.raises(dvm.DvmError):
        fresh.restore(bytes(bad))
    # a failed restore leaves the simulation untouched
    assert fresh.step_index() == 0 and fresh.size() == 0


def test_copy_out_through_pinned_buffers(gpu_ctx):
    px, pz = oracle.cylinder_points(96)
    s = dsim.Sim(gpu_ctx, px, pz, _params("reference"))
    s.run(6)
    n = s.size()
    cap = n - 17
    bufs = [dsim.HostBuffer(cap) for _ in range(3)]
    g, lo = dsim.HostBuffer(s.nl), dsim.HostBuffer(4)
    ev = dsim.Event(gpu_ctx)
    assert s.copy_out(bufs[0].array, bufs[1].array, bufs[2].array, g.array, lo.array) == n
    ev.record()
    ref = s.get_state()  # synchronous read of the same state
    ev.wait()
    assert ev.done()
    assert np.array_equal(bufs[0].array, ref["x"][:cap])
    assert np.array_equal(bufs[1].array, ref["z"][:cap])
    assert np.array_equal(bufs[2].array, ref["circ"][:cap])
    assert np.array_equal(g.array, s.gamma())
    assert tuple(lo.array) == s.loads()
    # the copies are ordered before later steps: stepping after the record
    # does not change what the buffers hold once the event is waited on
    snap_x = bufs[0].array.copy()
    assert s.copy_out(bufs[0].array, None, None, None, None) == n
    ev.record()
    s.step()
    ev.wait()
    assert np.array_equal(bufs[0].array, snap_x)
    for b in bufs + [g, lo]:
        b.close()
    ev.close()


def test_sim_probes_match_sheet_probes(gpu_ctx):
    px, pz = oracle.cylinder_points(96)
    p = _params("reference")
    s = dsim.Sim(gpu_ctx, px, pz, p)
    s.run(5)
    st = s.get_state()
    sh = dsim.Sheet(gpu_ctx, px, pz, p.sheet)
    sh.gamma = s.gamma()
    prx = np.linspace(-1.0, 3.0, 33)
    prz = np.linspace(-1.0, 1.0, 33)
    PX, PZ = to_dev(prx, prz)
    got = [t.cpu().numpy() for t in s.probe_quirk(PX, PZ)]
    ref = [t.cpu().numpy() for t in sh.probe_quirk(*to_dev(st["x"], st["z"], st["circ"],
                                                            st["sigma"]), PX, PZ)]
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)
