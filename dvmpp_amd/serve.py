"""Continuous-batching request loop over the batched ensemble.

Requests (a body, its constants, a seed and a step count) arrive over time.
Device memory is reserved in capacity classes: a class is a batch region
(dvm_batch_create) whose slots each hold up to that class's blob count.  A
request goes to the smallest class that holds its projected blob count; it is
admitted into a free slot of a live batch of that class, or into a new batch
of that class when the new region fits in the HBM budget beside the regions
already allocated.  Every live batch advances together each poll
(dvm_batch_step), finished simulations are retired, and a batch whose slots
are all free is released, so the allocated footprint follows the admitted
work.  A request whose projection exceeds the largest class, or whose class
region alone exceeds the budget, is rejected on arrival (status
DVM_E_SIZE) instead of waiting forever.  All blob and sheet state stays
device-resident; results are read back once, on completion.
"""
import time
from dataclasses import dataclass, field

from ._lib import F32C, BatchConfig
from .batch import Batch, footprint
from .sim import EULER
from .workloads import projected_blobs

STATUS_REJECTED = 2  # DVM_E_SIZE: the request cannot fit any class or the budget


@dataclass
class Request:
    rid: int
    arrival: float      # seconds after the loop starts
    px: object
    pz: object
    params: object      # SheetParams
    seed: int
    steps: int
    admitted: float = -1.0
    finished: float = -1.0
    status: int = -1
    slot: int = -1
    capacity: int = 0   # blob capacity of the slot it ran in
    loads: tuple = ()
    state: dict = field(default_factory=dict)

    @property
    def latency(self):
        return self.finished - self.arrival


def _class_config(cfg, capacity):
    return BatchConfig(cfg.slots, capacity, cfg.max_panels, cfg.scheme, cfg.crossing, cfg.prec,
                       cfg.use_graph, cfg.remesh)


class Server:
    """`cfg` is the template of every batch (slots per batch, largest slot
    capacity, body size limit, scheme, crossing, arithmetic, graph replay,
    remesh); `classes` lists the slot capacities (default: the template's
    alone).  `budget_bytes` bounds the device bytes of all live batches.
    With fmm_min > 0 (FP32C Euler batches) every batch gives slots of at least
    fmm_min blobs the fast multipole sum (dvm_batch_set_fmm, fmm_order,
    fmm_theta); a request's result still depends on itself alone."""

    def __init__(self, ctx, cfg, budget_bytes, classes=None, steps_per_poll=1, keep_state=False,
                 fmm_min=0, fmm_order=12, fmm_theta=0.5):
        self.ctx = ctx
        self.cfg = cfg
        if fmm_min > 0 and (cfg.prec != F32C or cfg.scheme != EULER):
            raise ValueError("fast multipole slots need an FP32C Euler batch template")
        self.fmm = (fmm_min, fmm_order, fmm_theta)
        self.classes = sorted(classes) if classes else [int(cfg.slot_capacity)]
        self.class_bytes = {c: footprint(_class_config(cfg, c)) for c in self.classes}
        self.budget = budget_bytes
        self.steps_per_poll = steps_per_poll
        self.keep_state = keep_state
        self.batches = []        # [capacity, Batch, {slot: Request}]
        self.steps_done = 0      # sum over simulations of completed steps
        self.peak_bytes = 0

    def allocated_bytes(self):
        return sum(self.class_bytes[c] for c, _, _ in self.batches)

    def projected(self, req):
        """Blob count the request is expected to reach: without remesh every
        panel releases about 2.3 blobs per step; with remesh the wake is held
        on the grid, so at most the nodes of the strip it can cover (length
        U t plus the body, width 4 body lengths), plus one step of nascent
        blobs."""
        nl = req.px.size - 1
        n = projected_blobs(nl, req.steps)
        rm = self.cfg.remesh
        if rm.frequency > 0:
            sp = req.params
            length = abs(sp.ux) * sp.dt * req.steps + 2.0
            strip = int(length * 4.0 / (rm.dx * rm.dz))
            n = min(n, strip + 4 * nl * rm.frequency)
        return n

    def projected_bytes(self, req):
        """Device bytes of the slot class the request would occupy (a batch's
        region divided over its slots)."""
        c = self._class_of(req)
        if c is None:
            return None
        return self.class_bytes[c] // self.cfg.slots

    def _class_of(self, req):
        n = self.projected(req)
        for c in self.classes:
            if n <= c:
                return c
        return None

    def _place(self, req, reserved):
        """A (batch entry) with a free slot for the request's class (slots
        `reserved` in this admission pass count as taken), creating a batch if the budget
        allows; None if it must wait."""
        c = self._class_of(req)
        for entry in self.batches:
            if entry[0] == c and len(entry[2]) + reserved.get(id(entry), 0) < self.cfg.slots:
                return entry
        if self.allocated_bytes() + self.class_bytes[c] <= self.budget:
            entry = [c, Batch(self.ctx, _class_config(self.cfg, c)), {}]
            if self.fmm[0] > 0:
                entry[1].set_fmm(*self.fmm)
            self.batches.append(entry)
            self.peak_bytes = max(self.peak_bytes, self.allocated_bytes())
            return entry
        return None

    def _rejectable(self, req):
        c = self._class_of(req)
        return c is None or self.class_bytes[c] > self.budget

    def run(self, requests, clock=time.perf_counter):
        pending = sorted(requests, key=lambda r: r.arrival)
        done = []
        t0 = clock()
        while pending or any(e[2] for e in self.batches):
            now = clock() - t0
            waiting = []
            placed = {}      # id(entry) -> (entry, [requests])
            reserved = {}
            for req in pending:
                if req.arrival > now:
                    waiting.append(req)
                    continue
                if self._rejectable(req):
                    req.status = STATUS_REJECTED
                    req.admitted = req.finished = clock() - t0
                    done.append(req)
                    continue
                entry = self._place(req, reserved)
                if entry is None:
                    waiting.append(req)
                    continue
                placed.setdefault(id(entry), (entry, []))[1].append(req)
                reserved[id(entry)] = reserved.get(id(entry), 0) + 1
            # the requests placed in one region are set up together
            for entry, reqs in placed.values():
                slots = entry[1].admit_many([(r.px, r.pz, r.params, r.seed, r.steps)
                                             for r in reqs])
                t_adm = clock() - t0
                for req, slot in zip(reqs, slots):
                    req.slot = slot
                    req.capacity = entry[0]
                    req.admitted = t_adm
                    entry[2][slot] = req
            pending = waiting
            if not any(e[2] for e in self.batches):
                if not pending:
                    break
                now = clock() - t0
                if pending[0].arrival <= now:
                    # an arrived request waits while nothing runs: idle
                    # regions of other classes hold the budget; free them
                    # (with none left it would have been placed: its class
                    # region fits the budget, or it was rejected)
                    for _, b, _ in self.batches:
                        b.close()
                    if not self.batches:
                        for req in [r for r in pending if r.arrival <= now]:
                            req.status = STATUS_REJECTED
                            req.admitted = req.finished = now
                            done.append(req)
                            pending.remove(req)
                    self.batches = []
                    continue
                time.sleep(min(pending[0].arrival - now, 0.05))
                continue
            for entry in self.batches:
                if entry[2]:
                    entry[1].step(self.steps_per_poll)
            for entry in self.batches:
                if not entry[2]:
                    continue
                for slot, status in entry[1].poll():
                    req = entry[2].pop(slot)
                    info = entry[1].info(slot)
                    req.finished = clock() - t0
                    req.status = status
                    req.loads = info["loads"]
                    self.steps_done += info["step"]
                    if self.keep_state:
                        req.state = entry[1].state(slot)
                    done.append(req)
            # release the regions that hold nothing, keeping one empty batch
            # per class a waiting request will need
            needed = {self._class_of(r) for r in pending}
            for entry in [e for e in self.batches if not e[2]]:
                if entry[0] in needed:
                    needed.discard(entry[0])
                    continue
                entry[1].close()
                self.batches.remove(entry)
        return done, clock() - t0

    def close(self):
        for _, b, _ in self.batches:
            b.close()
        self.batches = []
