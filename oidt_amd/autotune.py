"""Launch-shape autotuner for the stencil kernels.

For a set of runs (one batch: scheme, precision, grid sizes) it times every
candidate (kernel form, x-chunk depth, grid cap) on a copy of the real batch
by replaying its chunk graph, re-times the two fastest over more chunks, and
keeps the one with the smallest time per step.  The copies last long enough
that every timed launch does a whole step for every run.  Results are cached
per workload class in a JSON file so a serving process tunes each class once.

Kernel forms: 1 = per-plane march, 2 = batched loads (two launches per step),
3 = fused Yee step (one launch per step), 4 = temporally blocked step
(x-chunk = steps per launch).  Otherwise the x-chunk depth is the number of
planes each thread or wave owns along x.
"""
import dataclasses
import json
import os
import time

import numpy as np

from . import _lib
from .solver import Batch

# (kernel form, x-chunk depth, grid cap; 0 = library default)
CANDIDATES = {
    'tm': [(2, 4, 0), (2, 4, 65536), (2, 8, 0), (3, 1, 0), (3, 2, 0), (3, 8, 0), (3, 12, 0), (3, 16, 0),
           (3, 20, 0), (3, 24, 0), (3, 32, 0), (4, 2, 0), (4, 4, 0), (4, 8, 0)],
    'te': [(2, 2, 0), (2, 4, 0), (2, 4, 65536), (2, 8, 0), (1, 4, 0), (3, 1, 0), (3, 2, 0), (3, 8, 0), (3, 16, 0),
           (4, 2, 0), (4, 4, 0), (4, 8, 0)],
    '3d': [(1, 2, 0), (1, 4, 0), (1, 4, 65536), (1, 8, 0), (2, 1, 0), (2, 1, 65536), (3, 8, 0), (4, 2, 0)],
}

DEFAULT_CACHE = os.path.join(os.path.expanduser('~'), '.cache', 'oidt_amd', 'autotune.json')


def workload_class(setups):
    """Key of a batch for the cache: scheme, precision, run count and total
    cells rounded to a power of two (the tuning depends on occupancy, not on
    exact sizes)."""
    s0 = setups[0]
    cells = sum(s.cells for s in setups)
    return '%s-%s-n%d-c%d' % (s0.scheme, np.dtype(s0.precision).name, len(setups),
                              int(round(np.log2(max(cells, 1)))))


# chunks replayed per candidate: a warm-up chunk, then TIMED_CHUNKS (first
# pass) or RETIME_CHUNKS (the two fastest again)
TIMED_CHUNKS = 2
RETIME_CHUNKS = 6


def with_steps(s, nt):
    """Copy of setup ``s`` lasting at least ``nt`` steps (J_SRC zero-padded):
    every launch of a timed chunk then does a full step of work."""
    if s.nt >= nt:
        return s
    J = np.zeros(nt, dtype=np.asarray(s.J_SRC).dtype)
    J[:len(s.J_SRC)] = s.J_SRC
    return dataclasses.replace(s, nt=int(nt), J_SRC=J)


def step_time(runs, variant, xchunk, grid=0, chunks=TIMED_CHUNKS):
    """Time per time step (ms) of one candidate on ``runs`` [(setup, fg), ...].

    The batch's own captured chunk graph is replayed ``chunks`` times after
    one warm-up chunk and timed around a stream synchronisation, so launch
    gaps count as they do in a real run.  The runs are copies lasting
    (1 + chunks) chunks at least, so no timed launch skips a finished run."""
    s0 = runs[0][0]
    chunk = _lib.chunk_steps()
    need = (1 + int(chunks)) * chunk
    b = Batch(s0.scheme, s0.precision, max_slots=len(runs))
    try:
        b.set_launch(xchunk, grid, variant)
        for s, fg in runs:
            b.add(with_steps(s, need), fg)
        b.prepare()
        b.advance(1)
        b.sync()
        t0 = time.perf_counter()
        b.advance(int(chunks))
        b.sync()
        return (time.perf_counter() - t0) * 1e3 / (int(chunks) * chunk)
    finally:
        b.close()


def tune(runs, candidates=None, cache_path=DEFAULT_CACHE, use_cache=True, verbose=False):
    """Best (variant, xchunk, grid) for ``runs``; returns (choice, {candidate: ms/step})."""
    setups = [s for s, _ in runs]
    key = workload_class(setups)
    cache = {}
    if use_cache and cache_path and os.path.exists(cache_path):
        try:
            with open(cache_path) as fh:
                cache = json.load(fh)
        except (OSError, ValueError):
            cache = {}
        if key in cache:
            return tuple(cache[key]['choice']), cache[key]['times']
    scheme = setups[0].scheme
    cands = candidates or CANDIDATES[scheme]
    short = [(s, fg) for s, fg in runs]
    times = {}
    for variant, xc, grid in cands:
        t0 = time.perf_counter()
        ms = step_time(short, variant, xc, grid)
        times['%d,%d,%d' % (variant, xc, grid)] = ms
        if verbose:
            print('autotune %s variant %d xchunk %d grid %d: %.4f ms/step (%.1f s)'
                  % (key, variant, xc, grid, ms, time.perf_counter() - t0), flush=True)
    # the two fastest again, over more chunks: near ties are common
    top = sorted(times, key=times.get)[:2]
    for k in top:
        variant, xc, grid = (int(v) for v in k.split(','))
        times[k] = step_time(short, variant, xc, grid, chunks=RETIME_CHUNKS)
    best = min(top, key=times.get)
    choice = tuple(int(v) for v in best.split(','))
    if cache_path:
        cache[key] = {'choice': list(choice), 'times': times}
        os.makedirs(os.path.dirname(cache_path), exist_ok=True)
        with open(cache_path, 'w') as fh:
            json.dump(cache, fh, indent=1, sort_keys=True)
    return choice, times
