"""Serving-loop host logic with a fake device: admission by footprint, FIFO
order, deferral, rejection of oversized requests (no GPU needed)."""
from types import SimpleNamespace

import numpy as np

from oidt_amd import serve


class FakeBackend:
    """Completes every run after its chunk count of advance calls."""

    def __init__(self, max_slots):
        self.max_slots = max_slots
        self.runs = {}
        self.next = 0
        self.admitted = []
        self.count = {}

    def footprint(self, setup):
        return setup.fp

    def has_slot(self, key):
        return self.count.get(key, 0) < self.max_slots

    def add(self, key, setup, fg):
        h = (key, self.next)
        self.next += 1
        self.runs[h] = [0, -(-setup.nt // 32)]
        self.count[key] = self.count.get(key, 0) + 1
        self.admitted.append(setup.rid)
        return h

    def in_flight(self, key):
        return 0

    def advance(self, key, n):
        for h, st in self.runs.items():
            if h[0] == key:
                st[0] += n

    def done(self, h, setup):
        return self.runs[h][0] >= self.runs[h][1]

    def result(self, h):
        return {'chunks': self.runs[h][0]}

    def remove(self, h):
        self.count[h[0]] -= 1
        del self.runs[h]

    def sync(self):
        pass


def fake_prepare(r, backend):
    d = r.desc
    r.setup = SimpleNamespace(scheme=d['scheme'], precision=np.float32, cells=d['cells'], nt=d['nt'], fp=d['fp'],
                              rid=r.rid)
    r.fg = None
    r.footprint = d['fp']


def test_admission_by_footprint_fifo_and_rejection():
    descs = [dict(scheme='tm', cells=1000, nt=64 + 32 * i, fp=40) for i in range(6)]
    descs.insert(3, dict(scheme='tm', cells=1000, nt=64, fp=150))
    reqs = serve.make_requests(descs, [0.0] * len(descs))
    be = FakeBackend(max_slots=8)
    srv = serve.Server(be, hbm_budget=100, chunk_group=1, prepare=fake_prepare, idle=lambda: None)
    done = srv.run(reqs)
    assert srv.peak_bytes <= 100
    assert srv.peak_resident == 2
    assert [r.rid for r in done if r.error] == [3]
    assert 'exceeds the HBM budget' in done[3].error
    assert be.admitted == [0, 1, 2, 4, 5, 6]          # first come, first served
    assert all(r.result is not None for r in done if r.error is None)
    s = serve.summarize(done, 1.0)
    assert s['completed'] == 6 and s['rejected'] == 1


def test_slot_limit_per_batch_and_separate_batches():
    descs = [dict(scheme='tm' if i % 2 else '3d', cells=10, nt=32, fp=1) for i in range(8)]
    be = FakeBackend(max_slots=2)
    srv = serve.Server(be, hbm_budget=10 ** 9, chunk_group=1, prepare=fake_prepare, idle=lambda: None)
    done = srv.run(serve.make_requests(descs, [0.0] * 8))
    assert srv.peak_resident == 4          # two per batch, two batches
    assert all(r.error is None for r in done)


def test_poisson_arrivals_rate():
    t = serve.poisson_arrivals(20000, 50.0, seed=1)
    assert abs(len(t) / t[-1] - 50.0) < 1.5
    assert np.all(np.diff(t) > 0)
