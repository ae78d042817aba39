"""Continuous-batching request loop over the batched ensemble.

Requests (a body, its constants, a seed and a step count) arrive over time.
The loop admits a waiting request as soon as a slot is free and the projected
HBM footprint of everything admitted stays within the budget, advances every
admitted simulation together (dvm_batch_step), and retires simulations as they
finish — so requests that arrive while others run share the device with them.
All blob and sheet state stays device-resident; results are read back once, on
completion.
"""
import time
from dataclasses import dataclass, field

from .batch import Batch, footprint
from .workloads import projected_blobs

# Per blob of slot capacity in an FP32C batch: two SoA sets (136), RK stores
# (16), sorted FP32 tile copies and FP64 positions (48), sort keys/values and
# permutation (16) plus sort scratch (8), chunk partials (64), flags (4).
BYTES_PER_BLOB = 292


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
    loads: tuple = ()
    state: dict = field(default_factory=dict)

    @property
    def latency(self):
        return self.finished - self.arrival


class Server:
    def __init__(self, ctx, cfg, budget_bytes, steps_per_poll=1, keep_state=False):
        self.ctx = ctx
        self.cfg = cfg
        self.batch = Batch(ctx, cfg)
        self.fixed_bytes = footprint(cfg)
        self.budget = budget_bytes
        self.steps_per_poll = steps_per_poll
        self.keep_state = keep_state
        self.running = {}       # slot -> Request
        self.projected = {}     # slot -> bytes
        self.steps_done = 0     # sum over simulations of completed steps

    def projected_bytes(self, req):
        nl = req.px.size - 1
        n = min(projected_blobs(nl, req.steps), self.cfg.slot_capacity)
        return n * BYTES_PER_BLOB

    def _can_admit(self, req):
        if len(self.running) >= self.cfg.slots:
            return False
        return sum(self.projected.values()) + self.projected_bytes(req) <= self.budget

    def run(self, requests, clock=time.perf_counter):
        pending = sorted(requests, key=lambda r: r.arrival)
        done = []
        t0 = clock()
        while pending or self.running:
            now = clock() - t0
            while pending and pending[0].arrival <= now and self._can_admit(pending[0]):
                req = pending.pop(0)
                req.slot = self.batch.admit(req.px, req.pz, req.params, req.seed, req.steps)
                req.admitted = clock() - t0
                self.running[req.slot] = req
                self.projected[req.slot] = self.projected_bytes(req)
            if not self.running:
                # idle until the next arrival
                wait = pending[0].arrival - (clock() - t0)
                if wait > 0:
                    time.sleep(min(wait, 0.05))
                continue
            self.batch.step(self.steps_per_poll)
            for slot, status in self.batch.poll():
                req = self.running.pop(slot)
                self.projected.pop(slot)
                info = self.batch.info(slot)
                req.finished = clock() - t0
                req.status = status
                req.loads = info["loads"]
                self.steps_done += info["step"]
                if self.keep_state:
                    req.state = self.batch.state(slot)
                done.append(req)
        return done, clock() - t0

    def close(self):
        self.batch.close()
