"""Continuous-batching serving loop for FDTD solve requests.

No counterpart in the reference, which runs one simulation per OS process
(FDTD/GPU/FDTD_2D_TM/s-parameter_extraction/simulation_launcher.py:26-35).
Requests (one run description each) arrive over time; the server

* prepares each arrived request on the host (setup, roughness, material map);
* admits requests first-come first-served while a batch slot is free and the
  projected device footprint of all resident runs stays within the HBM
  budget (a request larger than the whole budget is rejected, the others wait);
* keeps one device batch per (scheme, precision), each on its own HIP stream,
  so 2D and 3D work overlaps on the GPU, and keeps a few chunks enqueued ahead
  on every batch so the GPU never waits for the host;
* detects finished runs from completed chunks (non-blocking), reads their port
  phasors without waiting for the rest of the batch, and frees their slot.

A run's result does not depend on what else is resident: slots share no state,
each slot's DFT chunks start at its own step 0, and the per-cell arithmetic is
the same in every launch, so a request gives bit-identical phasors alone and in
the mix (tested).
"""
import heapq
import time
from dataclasses import dataclass, field
from typing import Any, Optional

import numpy as np

from . import configs, materials
from . import solver as S


@dataclass(order=True)
class Request:
    arrival: float
    rid: int = field(compare=False)
    desc: dict = field(compare=False, repr=False)
    setup: Any = field(default=None, compare=False, repr=False)
    fg: Any = field(default=None, compare=False, repr=False)
    footprint: int = field(default=0, compare=False)
    t_ready: float = field(default=0.0, compare=False)
    t_admit: float = field(default=0.0, compare=False)
    t_done: float = field(default=0.0, compare=False)
    handle: Any = field(default=None, compare=False, repr=False)
    result: Optional[dict] = field(default=None, compare=False, repr=False)
    error: Optional[str] = field(default=None, compare=False)
    waited_for_memory: bool = field(default=False, compare=False)

    @property
    def key(self):
        return (self.setup.scheme, np.dtype(self.setup.precision).name)

    @property
    def cell_updates(self):
        return self.setup.cells * self.setup.nt

    @property
    def latency(self):
        return self.t_done - self.arrival


class DeviceBackend:
    """The GPU: one oidt batch per (scheme, precision)."""

    def __init__(self, max_slots=64, xchunk=0, variant=0, launch=None):
        """``launch`` maps a scheme ('tm', 'te', '3d') to its (variant, xchunk[,
        grid]), e.g. the autotuner's picks; other schemes use ``variant`` and
        ``xchunk``."""
        self.max_slots = max_slots
        self.xchunk, self.variant = xchunk, variant
        self.launch = dict(launch or {})
        self.batches = {}
        self.resident = {}

    def footprint(self, setup):
        return S.footprint(setup)

    def batch(self, key, setup):
        if key not in self.batches:
            b = S.Batch(setup.scheme, setup.precision, self.max_slots)
            variant, xchunk, grid = (tuple(self.launch[setup.scheme]) + (0,))[:3] if setup.scheme in self.launch \
                else (self.variant, self.xchunk, 0)
            b.set_launch(xchunk, grid, variant)
            self.batches[key] = b
            self.resident[key] = 0
        return self.batches[key]

    def has_slot(self, key):
        return self.resident.get(key, 0) < self.max_slots

    def add(self, key, setup, fg):
        slot = self.batch(key, setup).add(setup, fg)
        self.resident[key] += 1
        return (key, slot)

    def in_flight(self, key):
        e, d = self.batches[key].progress()
        return e - d

    def advance(self, key, nchunks):
        self.batches[key].advance(nchunks)

    def done(self, handle, setup):
        key, slot = handle
        return self.batches[key].poll(slot) >= setup.nt

    def result(self, handle):
        key, slot = handle
        return self.batches[key].dft(slot)

    def remove(self, handle):
        key, slot = handle
        self.batches[key].remove(slot)
        self.resident[key] -= 1

    def sync(self):
        for b in self.batches.values():
            b.sync()

    def close(self):
        for b in self.batches.values():
            b.close()
        self.batches.clear()


class Server:
    def __init__(self, backend, hbm_budget, chunk_group=4, prepare=None, clock=time.perf_counter,
                 idle=lambda: time.sleep(0.0002)):
        self.backend = backend
        self.budget = int(hbm_budget)
        self.group = chunk_group
        self.prepare = prepare or prepare_request
        self.clock = clock
        self.idle = idle
        self.resident_bytes = 0
        self.peak_bytes = 0
        self.peak_resident = 0
        self.deferred_for_memory = 0   # requests that waited at the head of the queue for HBM

    def run(self, requests, t0=None):
        """Serve ``requests`` (arrival times relative to the start); returns them
        with t_admit / t_done / result filled in (times relative to the start)."""
        start = self.clock() if t0 is None else t0
        arrivals = sorted(requests)
        waiting = []              # prepared, FIFO by arrival
        resident = []
        finished = []
        ai = 0
        while ai < len(arrivals) or waiting or resident:
            now = self.clock() - start
            progressed = False
            while ai < len(arrivals) and arrivals[ai].arrival <= now:
                r = arrivals[ai]
                ai += 1
                self.prepare(r, self.backend)
                r.t_ready = self.clock() - start
                if r.footprint > self.budget:
                    r.error = 'footprint %d B exceeds the HBM budget %d B' % (r.footprint, self.budget)
                    r.t_done = r.t_ready
                    finished.append(r)
                else:
                    waiting.append(r)
                progressed = True
            while waiting:
                r = waiting[0]
                if self.resident_bytes + r.footprint > self.budget:
                    if not r.waited_for_memory:
                        r.waited_for_memory = True
                        self.deferred_for_memory += 1
                    break
                if not self.backend.has_slot(r.key):
                    break
                waiting.pop(0)
                r.handle = self.backend.add(r.key, r.setup, r.fg)
                r.t_admit = self.clock() - start
                self.resident_bytes += r.footprint
                resident.append(r)
                self.peak_bytes = max(self.peak_bytes, self.resident_bytes)
                self.peak_resident = max(self.peak_resident, len(resident))
                progressed = True
            keys = {r.key for r in resident}
            for key in keys:
                if self.backend.in_flight(key) < 2 * self.group:
                    self.backend.advance(key, self.group)
                    progressed = True
            still = []
            for r in resident:
                if self.backend.done(r.handle, r.setup):
                    r.result = self.backend.result(r.handle)
                    self.backend.remove(r.handle)
                    self.resident_bytes -= r.footprint
                    r.t_done = self.clock() - start
                    finished.append(r)
                    progressed = True
                else:
                    still.append(r)
            resident = still
            if not progressed:
                self.idle()
        self.backend.sync()
        return sorted(finished, key=lambda r: r.rid)


def prepare_request(r, backend):
    """Host preparation: setup, seeded roughness profiles, core mask, footprint."""
    s, up, lo = configs.realise(r.desc)
    r.setup = s
    r.fg = materials.fg_mask(s, up, lo)
    r.footprint = backend.footprint(s)


def poisson_arrivals(n, rate, seed=0):
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


def summarize(done, wall):
    ok = [r for r in done if r.error is None]
    lat = np.array([r.latency for r in ok]) if ok else np.zeros(1)
    cu = sum(r.cell_updates for r in ok)
    return {'requests': len(done), 'completed': len(ok), 'rejected': len(done) - len(ok),
            'p50_latency_s': float(np.percentile(lat, 50)), 'p99_latency_s': float(np.percentile(lat, 99)),
            'mean_latency_s': float(lat.mean()), 'wall_s': wall,
            'mcell_updates_per_s': cu / wall / 1e6 if wall > 0 else 0.0, 'cell_updates': cu}


def serve_plain(requests, launch=None):
    """Plain path: one solve at a time, each alone on the device (queueing in
    arrival order), with the kernel form of ``launch`` ({scheme: (variant,
    xchunk)}) where given.  Returns the requests with latencies."""
    clock = time.perf_counter
    start = clock()
    for r in sorted(requests):
        now = clock() - start
        if r.arrival > now:
            time.sleep(r.arrival - now)
        if r.setup is None:
            prepare_request(r, DeviceBackend())
        sv = S.Solver(r.setup, r.fg)
        if launch and r.setup.scheme in launch:
            variant, xchunk, grid = (tuple(launch[r.setup.scheme]) + (0,))[:3]
            sv.set_launch(xchunk, grid, variant)
        sv.run()
        r.result = sv.dft()
        sv.close()
        r.t_done = clock() - start
    return sorted(requests, key=lambda r: r.rid)


def make_requests(descs, arrivals):
    return [Request(arrival=float(t), rid=i, desc=d) for i, (d, t) in enumerate(zip(descs, arrivals))]
