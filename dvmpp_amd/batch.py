"""Python handle for the batched ensemble (dvm_batch_* in include/dvm_hip.h)."""
import ctypes

import numpy as np

from ._lib import BatchConfig, check, load
from .sim import DELETE, EULER, REFLECT, remesh_params


def batch_config(slots=8, slot_capacity=1 << 16, max_panels=512, scheme=EULER, crossing=REFLECT,
                 prec=1, use_graph=True, remesh=None):
    return BatchConfig(slots, slot_capacity, max_panels, scheme, crossing, prec,
                       1 if use_graph else 0, remesh if remesh is not None else remesh_params())


def footprint(cfg):
    n = ctypes.c_size_t()
    st = load().dvm_batch_footprint(ctypes.byref(cfg), ctypes.byref(n))
    if st != 0:
        raise RuntimeError("dvm_batch_footprint failed")
    return n.value


class Batch:
    """dvm_batch_t: simulations stepped together on one device."""

    def __init__(self, ctx, cfg):
        self.ctx = ctx
        self.cfg = cfg
        h = ctypes.c_void_p()
        check(ctx.h, load().dvm_batch_create(ctx.h, ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        ctx.adopt(self)

    def close(self):
        if getattr(self, "h", None):
            if getattr(self.ctx, "h", None):
                load().dvm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def admit(self, px, pz, sheet_params, seed, steps):
        px = np.ascontiguousarray(px, dtype=np.float64)
        pz = np.ascontiguousarray(pz, dtype=np.float64)
        slot = ctypes.c_int()
        check(self.ctx.h, load().dvm_batch_admit(self.h, px.size, ctypes.c_void_p(px.ctypes.data),
                                                 ctypes.c_void_p(pz.ctypes.data),
                                                 ctypes.byref(sheet_params), seed, steps,
                                                 ctypes.byref(slot)))
        return slot.value

    def admit_many(self, jobs):
        """Admit (px, pz, sheet_params, seed, steps) bodies with their set-up
        run concurrently; returns their slots."""
        from ._lib import SheetParams
        k = len(jobs)
        pts = [(np.ascontiguousarray(px, dtype=np.float64), np.ascontiguousarray(pz, dtype=np.float64))
               for px, pz, *_ in jobs]
        npts = np.array([a.size for a, _ in pts], dtype=np.int32)
        pxp = (ctypes.c_void_p * k)(*[a.ctypes.data for a, _ in pts])
        pzp = (ctypes.c_void_p * k)(*[b.ctypes.data for _, b in pts])
        prm = (SheetParams * k)(*[j[2] for j in jobs])
        seeds = np.array([j[3] for j in jobs], dtype=np.int32)
        steps = np.array([j[4] for j in jobs], dtype=np.int32)
        slots = np.empty(k, dtype=np.int32)
        check(self.ctx.h, load().dvm_batch_admit_many(
            self.h, k, ctypes.c_void_p(npts.ctypes.data), pxp, pzp, prm,
            ctypes.c_void_p(seeds.ctypes.data), ctypes.c_void_p(steps.ctypes.data),
            ctypes.c_void_p(slots.ctypes.data)))
        return slots.tolist()

    def step(self, steps=1):
        check(self.ctx.h, load().dvm_batch_step(self.h, steps))

    def poll(self, max_n=1024):
        slots = np.empty(max_n, dtype=np.int32)
        status = np.empty(max_n, dtype=np.int32)
        cnt = ctypes.c_int()
        check(self.ctx.h, load().dvm_batch_poll(self.h, max_n, ctypes.c_void_p(slots.ctypes.data),
                                                ctypes.c_void_p(status.ctypes.data),
                                                ctypes.byref(cnt)))
        return list(zip(slots[:cnt.value].tolist(), status[:cnt.value].tolist()))

    def info(self, slot):
        n = ctypes.c_int64()
        step = ctypes.c_int()
        active = ctypes.c_int()
        status = ctypes.c_int()
        loads = np.empty(4)
        check(self.ctx.h, load().dvm_batch_slot_info(self.h, slot, ctypes.byref(n),
                                                     ctypes.byref(step), ctypes.byref(active),
                                                     ctypes.byref(status),
                                                     ctypes.c_void_p(loads.ctypes.data)))
        return dict(n=n.value, step=step.value, active=active.value, status=status.value,
                    loads=tuple(loads))

    def set_state(self, slot, x, z, circ, sigma):
        """Replace an admitted slot's blobs (velocities zero, IDs 0..n-1)."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, z, circ, sigma)]
        check(self.ctx.h, load().dvm_batch_set_state(
            self.h, slot, arrs[0].size, *[ctypes.c_void_p(a.ctypes.data) for a in arrs]))

    def set_fmm(self, min_blobs, order=12, theta=0.5):
        """Slots of at least min_blobs blobs take each step's blob-on-blob sum
        by the fast multipole form (order, theta); 0 turns it off."""
        check(self.ctx.h, load().dvm_batch_set_fmm(self.h, int(min_blobs), int(order),
                                                    float(theta)))

    def sort_order(self, slot, n):
        """Storage indices of the slot's first n blobs in its last FP32C pair
        sum's Morton order (n = the slot's count when that sum ran)."""
        out = np.empty(n, dtype=np.uint32)
        check(self.ctx.h, load().dvm_batch_sort_order(self.h, slot, n,
                                                      ctypes.c_void_p(out.ctypes.data)))
        return out

    def state(self, slot):
        n = self.info(slot)["n"]
        arrs = {k: np.empty(n) for k in ("x", "z", "circ", "sigma", "u", "w")}
        check(self.ctx.h, load().dvm_batch_get_state(
            self.h, slot, *[ctypes.c_void_p(arrs[k].ctypes.data)
                            for k in ("x", "z", "circ", "sigma", "u", "w")]))
        return arrs


__all__ = ["Batch", "batch_config", "footprint", "DELETE", "REFLECT"]
