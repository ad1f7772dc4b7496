#!/usr/bin/env python3
"""Benchmark of the FDTD time-stepping hot path on one MI355X.

Default workload: the batched 3D ensemble of 16 roughness realisations, FP32
fields, (x, y, z) = (96, 64, nz) with nz ~ U{192..512}, 64 DFT frequencies in
FP64 accumulators, all resident in one batch; a "step" is one Yee time step
of every realisation (E and H half-steps with CPML, source and port DFT).

Prints one JSON line with Mcell-updates/s (all grid cells per step, the
reference's MCells/s definition) over exactly K steps after W warm-up steps,
the same realisations run one at a time (the plain path), the roofline of the
dominant kernel measured with HIP events on the batch stream, the S21 error of
the full-length FP32 ensemble against FP64, the oracle timed on the host's
cores as the CPU baseline, and summaries of the 64-run 2D ensemble and of the
serving mix (p50 / p99 latency).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3|c0|c1|c2|c2te|c4|<study>]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E


def log(msg):
    """Progress of the legs on stderr (the JSON line is the only stdout)."""
    print('bench: ' + msg, file=sys.stderr, flush=True)


def algorithmic_bytes(scheme, es):
    """Field bytes per cell of the (E, H) half-steps: each component read twice
    and written once per Yee step (no coefficient maps, psi or DFT traffic)."""
    if scheme == 'tm':
        return (3 + 1) * es, (3 + 2) * es
    if scheme == 'te':
        return (3 + 2) * es, (3 + 1) * es
    return (6 + 3) * es, (6 + 3) * es


def cpu_baseline(p, seconds):
    """Oracle (FP64 NumPy restatement, one thread) on the same grid: whole
    steps until ``seconds`` of CPU work have run (a bounded sample), at most
    the run's own step count (a small 2D grid finishes all of it sooner)."""
    from oracle import fdtd as ofdtd
    from oracle import geometry as ogeo
    from oracle import setup as osetup
    c = osetup.explicit_setup(p)
    o = ofdtd.make(c, ogeo.fg_mask(c))
    o.run(1)
    steps = 0
    t0 = time.perf_counter()
    dt = 0.0
    while o.n < c.nt:
        o.run(1)
        steps += 1
        dt = time.perf_counter() - t0
        if dt >= seconds:
            break
    cells = c.nx * c.ny * (c.nz if c.scheme == '3d' else 1)
    return cells * steps / dt / 1e6, dt, steps


def committed_traffic(kernel, workload_key, variant, xchunk):
    """HBM bytes per launch of ``kernel`` from the rocprofv3 counter summaries
    under profiles/ (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by the
    copy-kernel calibration); None when no summary matches this launch form.
    A summary states it on a line
    ``traffic: kernel=K workload=W variant=V xchunk=X bytes_per_launch=B``."""
    d = os.path.join(ROOT, 'profiles')
    if not os.path.isdir(d):
        return None, None
    want = {'kernel': kernel, 'workload': workload_key, 'variant': str(variant), 'xchunk': str(xchunk)}
    for name in sorted(os.listdir(d)):
        if not name.endswith(('.md', '.txt')):
            continue
        with open(os.path.join(d, name)) as fh:
            for line in fh:
                line = line.strip().strip('`')
                if not line.startswith('traffic:'):
                    continue
                kv = dict(t.split('=', 1) for t in line[len('traffic:'):].split() if '=' in t)
                if all(kv.get(k) == v for k, v in want.items()):
                    return int(kv['bytes_per_launch']), 'profiles/' + name
    return None, None


def choose_launch(args, runs):
    """(variant, xchunk, grid, autotune record) for the batch of ``runs``: the
    command-line form, or the autotuner's pick when --variant is 0."""
    if args.variant:
        return args.variant, args.xchunk, args.grid, None
    from oidt_amd import autotune
    t0 = time.perf_counter()
    (variant, xc, grid), times = autotune.tune(runs, use_cache=False, cache_path=None)
    return variant, xc, grid, {'ms_per_step': {k: round(v, 5) for k, v in times.items()},
                               'seconds': round(time.perf_counter() - t0, 2)}


def workload(name):
    from oidt_amd import configs
    if name == 'c0':
        return [configs.config0()], '2D TMz FP64 slab 600x240, 32 freqs (config 0)'
    if name == 'c1':
        return [configs.config1()], '3D FP32 strip (x,y,z)=(128,96,320), 64 freqs (config 1)'
    if name == 'c2':
        return configs.config2(), '64 x 2D TMz FP32 slabs nx~U{800..4000} x 240, one batch (config 2)'
    if name == 'c2te':
        return configs.config2(scheme='te'), '64 x 2D TEz FP32 slabs nx~U{800..4000} x 240, one batch (config 2 in TEz)'
    if name == 'c3':
        return configs.config3(), '16 x 3D FP32 strips (96,64,nz~U{192..512}), one batch (config 3)'
    raise SystemExit('unknown workload ' + name)


STUDIES = {'sl_tm': ('tm', 'scattering_loss'), 'sl_te': ('te', 'scattering_loss'), 'sl_3d': ('3d', 'scattering_loss'),
           'wi_tm': ('tm', 'wave_impedance'), 'wi_te': ('te', 'wave_impedance'), 'wi_3d': ('3d', 'wave_impedance')}


def monitor_bytes(mons, es, after_h):
    """Algorithmic bytes of one monitor launch: per cell the field read once,
    plus the running sums read and written (DFT) or the record written."""
    total = 0
    for m in mons.values():
        if m.after_h != after_h:
            continue
        cells = int(np.prod(m.extent))
        total += cells * es * (5 if m.cs is not None else 2)
    return total


def bench_study(args):
    """A loss or wave-impedance study at the reference's default size, cut to
    warmup + steps time steps (field monitors included in every step)."""
    import ctypes as C
    from oidt_amd import _lib, variants
    scheme, study = STUDIES[args.workload]
    chunk = _lib.chunk_steps()
    warm = -(-args.warmup // chunk) * chunk
    K = -(-args.steps // chunk) * chunk
    c = variants.truncate(variants.config(scheme, study), warm + K + (1 if scheme == 'te' else 0))
    if args.realisations > 1:
        return bench_study_batch(args, scheme, study, warm, K)
    # two-kernel form and x-chunk: the given ones, or the fastest step
    # (E + H + monitor launches, HIP events, two chunks each); the 3D march
    # form is a candidate too
    forms = [(2, xc) for xc in (1, 2, 4, 8)] + ([(1, xc) for xc in (2, 4, 8)] if scheme == '3d' else [])
    forms = [(v, xc) for v, xc in forms if (not args.variant or v == args.variant) and
             (not args.xchunk or xc == args.xchunk)] or [(args.variant or 2, args.xchunk or 4)]
    # tuning and kernel profiles run on their own copy of the study, long
    # enough that a warm-up chunk and every timed chunk do whole steps
    nprof = max(2, int(args.profile_chunks))
    cp = variants.truncate(variants.config(scheme, study), (1 + nprof) * chunk + (1 if scheme == 'te' else 0))

    def profiled(form, nchunks):
        sol_, mons_, _ = variants.prepare(cp, (form[0], form[1], 0))
        out_ = (C.c_double * 6)()
        try:
            _lib.check(_lib.lib().oidt_solver_profile_all(sol_.h, 1, out_))
            _lib.check(_lib.lib().oidt_solver_profile_all(sol_.h, int(nchunks), out_))
        finally:
            sol_.close()
        return list(out_), mons_

    tune = {}
    for v, xc in forms:
        out, _ = profiled((v, xc), 2)
        tune[(v, xc)] = out[0] + out[1] + out[4] + out[5]
    variant, xchunk = min(tune, key=tune.get)
    sol, mons, ids = variants.prepare(c, (variant, xchunk, 0))
    try:
        sol.advance(warm).sync()
        t0 = time.perf_counter()
        sol.advance(K).sync()
        wall = time.perf_counter() - t0
    finally:
        sol.close()
    cells = c.nx * c.ny * (c.nz if scheme == '3d' else 1)
    value = cells * K / wall / 1e6
    out, mons = profiled((variant, xchunk), nprof)
    ms_e, ms_h, ms_dft, ms_adv, ms_me, ms_mh = out
    es = 4
    be, bh = algorithmic_bytes('tm' if scheme == 'tm' else scheme, es)
    stem = {'tm': 'k_tmzl', 'te': 'k_tez', '3d': 'k_3d'}[scheme]
    suffix = '' if scheme == 'tm' or variant == 1 else '2'
    cands = [(stem + '_e' + suffix, ms_e, be * cells),
             (stem + '_h' + suffix, ms_h, bh * cells),
             ('k_monitor (after E)', ms_me, monitor_bytes(mons, es, False)),
             ('k_monitor (after H)', ms_mh, monitor_bytes(mons, es, True))]
    kname, ms_k, bytes_k = max(cands, key=lambda t: t[1])
    achieved = bytes_k / (ms_k * 1e-3) / 1e9
    if achieved > HBM_PEAK_GBS:
        raise SystemExit('bench.py: %s measured %.1f GB/s of algorithmic bytes, above the %.0f GB/s HBM peak: '
                         'the kernel timing is not valid' % (kname, achieved, HBM_PEAK_GBS))
    traffic, traffic_src = committed_traffic(kname.split()[0], args.workload, variant, xchunk)
    roof = {'bound': 'hbm', 'kernel': kname, 'achieved': round(achieved, 1), 'peak': HBM_PEAK_GBS,
            'unit': 'GB/s', 'frac': round(achieved / HBM_PEAK_GBS, 4), 'traffic': traffic,
            'traffic_source': traffic_src, 'bytes_per_launch': bytes_k, 'ms_per_launch': round(ms_k, 5),
            'ms_e': round(ms_e, 5), 'ms_h': round(ms_h, 5), 'ms_monitor_e': round(ms_me, 5),
            'ms_monitor_h': round(ms_mh, 5)}
    line = {
        'metric': 'Mcell-updates/s', 'value': round(value, 1), 'unit': 'Mcell-updates/s', 'n_gpus': 1,
        'steps': K, 'warmup': warm, 'ms_per_step': round(wall / K * 1e3, 5), 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32', 'data': 'synthetic',
        'config': {'workload': '%s %s study, reference default configuration (%s cells), first %d steps of %d'
                               % (scheme, study, 'x'.join(str(n) for n in ((c.nx, c.ny, c.nz) if scheme == '3d'
                                                                           else (c.nx, c.ny))),
                                  warm + K, variants.config(scheme, study).nt),
                   'cells_per_step': cells, 'monitors': sorted(mons), 'variant': variant, 'xchunk': xchunk,
                   'form_ms_per_step': {'%d,%d' % k: round(v, 5) for k, v in tune.items()},
                   'parallelism': 'single GPU'},
        'roofline': roof,
    }
    if not args.no_cpu and scheme != '3d':
        line['cpu_baseline'] = study_cpu_baseline(scheme, study, args.cpu_seconds)
    print(json.dumps(line), flush=True)


def bench_study_batch(args, scheme, study, warm, K):
    """``--realisations R``: R roughness realisations of a study (profile
    numbers 0..R-1) as one batch, against the same R run one at a time."""
    from oidt_amd import variants
    R = args.realisations
    steps = warm + K + (1 if scheme == 'te' else 0)
    cs = [variants.truncate(variants.config(scheme, study, profile_number=p, rough_toggle=True), steps)
          for p in range(R)]
    b, runs = variants.prepare_batch(cs, (args.variant, args.xchunk, 0))
    variant, xc, _ = b.launch()
    try:
        b.advance(warm // 32)
        b.sync()
        t0 = time.perf_counter()
        b.advance(K // 32)
        b.sync()
        wall = time.perf_counter() - t0
    finally:
        b.close()
    cells = sum(c.nx * c.ny * (c.nz if scheme == '3d' else 1) for c in cs)
    value = cells * K / wall / 1e6
    # plain: the same realisations one at a time, each its own run (create, upload and read back included)
    t0 = time.perf_counter()
    for c in cs:
        sol, mons, ids = variants.prepare(c, (variant, xc, 0))
        try:
            sol.advance(warm + K).sync()
            for k in mons:
                sol.read_monitor(ids[k])
        finally:
            sol.close()
    plain_wall = time.perf_counter() - t0
    plain = cells * (warm + K) / plain_wall / 1e6
    line = {
        'metric': 'Mcell-updates/s', 'value': round(value, 1), 'unit': 'Mcell-updates/s', 'n_gpus': 1,
        'steps': K, 'warmup': warm, 'ms_per_step': round(wall / K * 1e3, 5), 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32', 'data': 'synthetic',
        'config': {'workload': '%d roughness realisations of the %s %s study, one batch, reference default '
                               'configuration, first %d steps' % (R, scheme, study, warm + K),
                   'cells_per_step': cells, 'variant': variant, 'xchunk': xc, 'parallelism': 'single GPU'},
        'plain_one_at_a_time': round(plain, 1), 'batched_over_plain': round(value / plain, 3),
    }
    print(json.dumps(line), flush=True)


def study_cpu_baseline(scheme, study, seconds):
    """The FP64 study oracle (oracle/studies.py, one thread) on the study's own
    grid, whole steps (recording included) for about ``seconds``, at most the
    study's own step count."""
    from oidt_amd import variants
    from oracle import studies as ost
    c = variants.config(scheme, study)
    o = ost.make(c)
    o.step(0)
    o.record(0)
    n, t0 = 1, time.perf_counter()
    dt = 0.0
    while n < c.nt:
        o.step(n)
        o.record(n)
        n += 1
        dt = time.perf_counter() - t0
        if dt >= seconds:
            break
    cells = c.nx * c.ny
    return {'value': round(cells * (n - 1) / dt / 1e6, 3), 'unit': 'Mcell-updates/s', 'cores': 1, 'kind': 'port',
            'sample': '%d steps (after 1 warm-up) of the FP64 NumPy study oracle on the %dx%d grid, %.1f s'
                      % (n - 1, c.nx, c.ny, dt)}


def bench_serving(args, count=None):
    """Serving mix (configs.serving_requests: 70 % 2D TMz, 30 % 3D, 2000..8000
    steps): the plain path (each request alone, back to back), continuous
    batching with every request queued at t = 0 (saturated), and Poisson
    arrivals at 0.8x the plain path's request capacity (latency percentiles),
    under an HBM budget that makes admission wait for memory.  Every leg
    includes each request's upload; host preparation (roughness, masks) is
    done once before all legs."""
    from oidt_amd import autotune, configs, serve
    from oidt_amd.solver import Batch
    n = int(count or args.requests)
    descs = configs.serving_requests(count=n)
    be = serve.DeviceBackend(max_slots=64)
    reqs = serve.make_requests(descs, [0.0] * len(descs))
    t0 = time.perf_counter()
    for r in reqs:
        serve.prepare_request(r, be)
    prep = time.perf_counter() - t0
    tuned = {}
    log('c4: %d requests prepared in %.1f s; autotuning' % (len(reqs), prep))
    for scheme in sorted({r.setup.scheme for r in reqs}):
        sub = [(r.setup, r.fg) for r in reqs if r.setup.scheme == scheme][:16]
        if args.variant:
            be.launch[scheme] = (args.variant, args.xchunk, args.grid)
            continue
        t1 = time.perf_counter()
        choice, times = autotune.tune(sub, use_cache=False, cache_path=None)
        be.launch[scheme] = choice
        tuned[scheme] = {'choice': list(choice), 'ms_per_step': {k: round(v, 5) for k, v in times.items()},
                         'seconds': round(time.perf_counter() - t1, 2)}
    log('c4: plain leg, %d requests one at a time' % len(reqs))
    t0 = time.perf_counter()
    serve.serve_plain(reqs, be.launch)
    plain_wall = time.perf_counter() - t0
    total_cu = sum(r.cell_updates for r in reqs)
    plain_rate = total_cu / plain_wall / 1e6
    plain_rps = len(reqs) / plain_wall
    # the plain leg's time per scheme (requests run back to back from t = 0)
    by_scheme, prev = {}, 0.0
    for r in sorted(reqs, key=lambda q: q.t_done):
        by_scheme[r.setup.scheme] = by_scheme.get(r.setup.scheme, 0.0) + (r.t_done - prev)
        prev = r.t_done
    # HBM budget: a quarter of the summed footprints of all requests, so that
    # admission waits for memory while many runs stay resident
    budget = int(sum(r.footprint for r in reqs) / 4)
    for r in reqs:
        r.result = None
    log('c4: saturated continuous batching')
    srv = serve.Server(be, hbm_budget=budget, chunk_group=4, prepare=lambda r, b: None)
    t0 = time.perf_counter()
    done = srv.run(reqs)
    sat_wall = time.perf_counter() - t0
    sat = serve.summarize(done, sat_wall)
    sat_deferred = srv.deferred_for_memory
    # Poisson arrivals at 0.8x the plain path's capacity
    rate = 0.8 * plain_rps
    arr = serve.poisson_arrivals(len(reqs), rate, seed=4)
    preqs = serve.make_requests(descs, arr)
    for a, b_ in zip(preqs, reqs):
        a.setup, a.fg, a.footprint = b_.setup, b_.fg, b_.footprint
    log('c4: Poisson arrivals at %.3f requests/s' % rate)
    srv2 = serve.Server(be, hbm_budget=budget, chunk_group=4, prepare=lambda r, b: None)
    t0 = time.perf_counter()
    pdone = srv2.run(preqs)
    pw = time.perf_counter() - t0
    pois = serve.summarize(pdone, pw)
    # each scheme's requests alone, saturated under the same budget: how much
    # continuous batching gains per scheme, against that scheme's share of
    # the plain leg's time
    by_scheme_served = {}
    for scheme in sorted(by_scheme):
        idx = [k for k, r in enumerate(reqs) if r.setup.scheme == scheme]
        sreqs = serve.make_requests([descs[k] for k in idx], [0.0] * len(idx))
        for a, k in zip(sreqs, idx):
            a.setup, a.fg, a.footprint = reqs[k].setup, reqs[k].fg, reqs[k].footprint
        log('c4: saturated, %s requests only' % scheme)
        srv3 = serve.Server(be, hbm_budget=budget, chunk_group=4, prepare=lambda r, b: None)
        t0 = time.perf_counter()
        srv3.run(sreqs)
        by_scheme_served[scheme] = time.perf_counter() - t0
    be.close()
    # dominant kernel: per scheme, a batch of (up to) 16 of its requests
    roofs = {}
    for scheme, (v, xc, g) in ((k, (tuple(c) + (0,))[:3]) for k, c in be.launch.items()):
        sub = [(r.setup, r.fg) for r in reqs if r.setup.scheme == scheme][:16]
        if not sub:
            continue
        cells = sum(s_.cells for s_, _ in sub)
        roofs[scheme] = roofline_of(scheme, 4, v, xc, cells, profile_kernels(sub, (v, xc, g), args.profile_chunks),
                                    'c4')
    # the dominant kernel: the scheme whose stencil launches take the most
    # device time over the whole mix (its cell-updates at its profiled rate)
    share = {}
    for scheme, rf in roofs.items():
        sub = [r for r in reqs if r.setup.scheme == scheme]
        prof_cells = sum(r.setup.cells for r in sub[:16])
        share[scheme] = sum(r.cell_updates for r in sub) / prof_cells * (rf['ms_e'] + rf['ms_h'])
    dom = dict(roofs[max(share, key=share.get)])
    n2d = sum(1 for r in reqs if r.setup.scheme == 'tm')
    rnd = lambda d: {k: (round(v, 4) if isinstance(v, float) else v) for k, v in d.items()}
    return {
        'metric': 'Mcell-updates/s', 'value': round(sat['mcell_updates_per_s'], 1), 'unit': 'Mcell-updates/s',
        'workload': 'serving mix: %d requests, %d 2D TMz (nx 400..4000, ny 200..240) and %d 3D '
                    '((128..384)x(64..128)x(48..96)), 2000..8000 steps, FP32' % (len(reqs), n2d, len(reqs) - n2d),
        'cell_updates': int(total_cu), 'hbm_budget_bytes': budget,
        'sum_footprint_bytes': int(sum(r.footprint for r in reqs)),
        'launch': {k: list(v) for k, v in be.launch.items()}, 'autotune': tuned or None,
        'plain_mcell_updates_per_s': round(plain_rate, 1), 'plain_requests_per_s': round(plain_rps, 3),
        'plain_wall_s': round(plain_wall, 3),
        'served_over_plain': round(sat['mcell_updates_per_s'] / plain_rate, 3),
        # per scheme: the plain leg's time, the same requests served alone
        # (saturated), and the ratio; the mix's ratio is bounded by the
        # schemes' ratios weighted by their plain time
        'plain_seconds_by_scheme': {k: round(v, 3) for k, v in by_scheme.items()},
        'served_seconds_by_scheme': {k: round(v, 3) for k, v in by_scheme_served.items()},
        'served_over_plain_by_scheme': {k: round(by_scheme[k] / v, 3) for k, v in by_scheme_served.items()},
        'served_over_plain_if_schemes_served_apart': round(plain_wall / sum(by_scheme_served.values()), 3),
        'saturated': dict(rnd(sat), admissions_deferred_for_memory=sat_deferred,
                          peak_resident_runs=srv.peak_resident, peak_resident_bytes=srv.peak_bytes),
        'poisson': dict(rnd(pois), rate_per_s=round(rate, 3),
                        admissions_deferred_for_memory=srv2.deferred_for_memory,
                        peak_resident_runs=srv2.peak_resident),
        'roofline': dom, 'roofline_per_scheme': roofs,
        'host_prepare_s': round(prep, 3),
    }


def cpu_pool_worker(args):
    """One CPU-baseline worker: the oracle on description ``p`` for about
    ``seconds`` of whole steps (after one warm-up step)."""
    p, seconds = args
    rate, dt, steps = cpu_baseline(p, seconds)
    return rate, dt, steps


def cpu_baseline_pool(p, seconds, label):
    """The oracle (FP64 NumPy, one thread per process) on one worker per host
    core of this process's affinity set, capped at 16 (the host share of one
    GPU): aggregate Mcell-updates/s of the independent solves."""
    import multiprocessing as mp
    aff = len(os.sched_getaffinity(0))
    n = max(1, min(aff, 16))
    ctx = mp.get_context('spawn')      # fresh interpreters: nothing of the HIP runtime is inherited
    with ctx.Pool(n) as pool:
        res = pool.map(cpu_pool_worker, [(p, seconds)] * n)
    rate = sum(r for r, _, _ in res)
    steps = sum(s for _, _, s in res)
    return {'value': round(rate, 3), 'unit': 'Mcell-updates/s', 'cores': n, 'affinity_cpus': aff,
            'kind': 'port',
            'sample': '%d independent FP64 NumPy oracle solves (one per core) of %s, %d whole steps in '
                      'total after one warm-up step each, about %.0f s each (at most the run\'s own steps)'
                      % (n, label, steps, max(dt for _, dt, _ in res))}


def ensemble_runs(descs, nt):
    """[(setup, fg)] of the descriptions with ``nt`` steps each."""
    from oidt_amd import configs, materials
    out = []
    for p in descs:
        s, up, lo = configs.realise(dict(p, nt=nt))
        out.append((s, materials.fg_mask(s, up, lo)))
    return out


def timed_batch(runs, launch, K):
    """Wall time (s) of the K steps of ``runs`` as one batch, uploads and
    graph capture done before the timed region."""
    from oidt_amd.solver import Batch
    s0 = runs[0][0]
    b = Batch(s0.scheme, s0.precision, max_slots=len(runs))
    try:
        b.set_launch(launch[1], launch[2], launch[0])
        for s, fg in runs:
            b.add(s, fg)
        b.prepare()
        b.sync()
        t0 = time.perf_counter()
        b.advance(-(-K // 32))
        b.sync()
        return time.perf_counter() - t0
    finally:
        b.close()


def profile_kernels(runs, launch, chunks):
    """Mean HIP-event durations (ms) of the E, H, DFT and advance launches on
    a batch of ``runs`` lasting 1 + ``chunks`` chunks: one warm-up chunk, then
    ``chunks`` timed chunks in which every launch does a full step of work."""
    import ctypes as C
    from oidt_amd import _lib
    from oidt_amd.autotune import with_steps
    from oidt_amd.solver import Batch
    need = (1 + int(chunks)) * 32
    s0 = runs[0][0]
    pb = Batch(s0.scheme, s0.precision, max_slots=len(runs))
    try:
        pb.set_launch(launch[1], launch[2], launch[0])
        for s, fg in runs:
            pb.add(with_steps(s, need), fg)
        out = (C.c_double * 4)()
        _lib.check(_lib.lib().oidt_batch_profile(pb.h, 1, out))
        _lib.check(_lib.lib().oidt_batch_profile(pb.h, int(chunks), out))
        return list(out)
    finally:
        pb.close()


def roofline_of(scheme, es, variant, xchunk, cells, ms, key):
    """Roofline of the dominant stencil kernel from its HIP-event duration.

    ``achieved`` is the algorithmic bytes of the cell-updates one launch
    performs (each field component read twice and written once per Yee
    step) over the launch's duration.  The temporally blocked steps
    (variant 4) perform ``xchunk`` steps per launch while moving each field
    value through HBM once, so its figure can exceed the HBM peak; for it the
    line also gives the bytes it must move at least (every field value read
    once and written once per launch) and their rate, which cannot."""
    ms_e, ms_h, ms_dft, ms_adv = ms
    be, bh = algorithmic_bytes(scheme, es)
    stem = {'tm': 'k_tmz', 'te': 'k_tez', '3d': 'k_3d'}[scheme]
    steps = 1
    if variant == 4:
        steps = int(xchunk)
        kname, ms_k, bytes_k = (stem + '_tb', ms_e, (be + bh) * cells * steps)
    elif variant == 3:
        kname, ms_k, bytes_k = (stem + ('_halo' if scheme == '3d' else '_step'), ms_e, (be + bh) * cells)
    elif ms_e >= ms_h:
        kname, ms_k, bytes_k = (stem + ('_e2' if variant == 2 else '_e'), ms_e, be * cells)
    else:
        kname, ms_k, bytes_k = (stem + ('_h2' if variant == 2 else '_h'), ms_h, bh * cells)
    achieved = bytes_k / (ms_k * 1e-3) / 1e9
    frac = achieved / HBM_PEAK_GBS
    ncomp = 6 if scheme == '3d' else 3
    min_bytes = 2 * ncomp * es * cells            # each field value read once and written once per launch
    min_frac = min_bytes / (ms_k * 1e-3) / 1e9 / HBM_PEAK_GBS
    if (frac if steps == 1 else min_frac) > 1.0:
        raise SystemExit('bench.py: %s measured %.1f GB/s of algorithmic bytes, above the %.0f GB/s HBM peak: '
                         'the kernel timing is not valid' % (kname, achieved, HBM_PEAK_GBS))
    traffic, traffic_src = committed_traffic(kname, key, variant, xchunk)
    out = {'bound': 'hbm', 'kernel': kname, 'achieved': round(achieved, 1), 'peak': HBM_PEAK_GBS,
           'unit': 'GB/s', 'frac': round(frac, 4), 'traffic': traffic, 'traffic_source': traffic_src,
           'bytes_per_launch': bytes_k, 'ms_per_launch': round(ms_k, 5),
           'ms_e': round(ms_e, 5), 'ms_h': round(ms_h, 5), 'ms_dft': round(ms_dft, 5),
           'ms_advance': round(ms_adv, 5)}
    if steps > 1:
        out['steps_per_launch'] = steps
        out['min_hbm_bytes_per_launch'] = min_bytes
        out['min_hbm_frac'] = round(min_frac, 4)
    return out


def s21_accuracy(descs, launch=None):
    """Per-realisation S21 / S11 of the full-length ensemble in FP32 against
    the same ensemble in FP64 (one shared incident run each, device
    extraction), both in the launch form the timed leg uses: max |dB| and max
    |degrees| differences over the frequencies."""
    from oidt_amd import ensemble
    t0 = time.perf_counter()
    e32 = ensemble.Ensemble(descs, precision='float32', launch=launch).run()
    e64 = ensemble.Ensemble(descs, precision='float64', launch=launch).run()
    rows = []
    for r in range(len(descs)):
        a, b = e32.s_params(r), e64.s_params(r)
        d21 = ensemble.s_error(a['S21'], b['S21'])
        d11 = ensemble.s_error(a['S11'], b['S11'])
        rows.append({'r': r, 's21_db_mean': round(float(np.mean(ensemble.db(b['S21']))), 4),
                     's21_err_db': float('%.3g' % d21[0]), 's21_err_deg': float('%.3g' % d21[1]),
                     's11_err_db': float('%.3g' % d11[0])})
    saved = e32.cell_updates_unshared() - e32.cell_updates
    e32.close()
    e64.close()
    return {'reference': 'the same ensemble in FP64 on the device (FP64 kernels pinned to the FP64 oracle at 1e-10)',
            'launch': list(launch) if launch else None, 'steps': int(e32.nt), 'incident_runs': len(e32.incident),
            'cell_updates_saved_by_shared_incident': int(saved),
            'max_s21_err_db': max(r['s21_err_db'] for r in rows),
            'max_s21_err_deg': max(r['s21_err_deg'] for r in rows),
            'per_realisation': rows, 'seconds': round(time.perf_counter() - t0, 1)}


def bench_ensemble(args, descs, label):
    """An ensemble (or a single run) as one batch: K timed steps after W
    warm-up steps, the same runs one at a time (plain path), the dominant
    kernel's roofline, the S21 accuracy of the full-length ensemble and the
    CPU baseline."""
    from oidt_amd import _lib
    W, K = int(args.warmup), int(args.steps)
    runs = ensemble_runs(descs, K)
    scheme, prec = runs[0][0].scheme, runs[0][0].precision
    es = 4 if prec == np.float32 else 8
    log('%s: %d runs, autotuning' % (args.workload, len(runs)))
    variant, xchunk, grid, tuned = choose_launch(args, runs)
    launch = (variant, xchunk, grid)
    log('%s: launch form %r; warm-up and %d timed steps' % (args.workload, launch, K))
    # warm-up: the same runs for W steps in their own batch
    if W > 0:
        timed_batch(ensemble_runs(descs, W), launch, W)
    wall = timed_batch(runs, launch, K)
    cells = sum(s.cells for s, _ in runs)
    value = cells * K / wall / 1e6
    line = {
        'metric': 'Mcell-updates/s', 'value': round(value, 1), 'unit': 'Mcell-updates/s', 'n_gpus': 1,
        'steps': K, 'warmup': W, 'ms_per_step': round(wall / K * 1e3, 5), 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32' if es == 4 else 'f64', 'data': 'synthetic',
        'config': {'workload': label, 'runs': len(runs), 'cells_per_step': cells, 'steps_per_run': K,
                   'xchunk': xchunk, 'variant': variant, 'grid_cap': grid, 'parallelism': 'single GPU',
                   'timed_region': 'K steps of every run; uploads and graph capture before it',
                   'autotune': tuned},
    }
    be, bh = algorithmic_bytes(scheme, es)
    line['hbm_roofline_frac_step'] = round(value * 1e6 * (be + bh) / (HBM_PEAK_GBS * 1e9), 4)
    line['algorithmic_bytes_per_cell_update'] = be + bh
    if len(runs) > 1 and not args.no_plain:
        # plain path: each run alone, in its own batch, tuned for one run;
        # uploads and graph capture outside the timed region as above
        log('%s: plain leg (each run alone)' % args.workload)
        pv, pxc, pg, ptuned = choose_launch(args, runs[:1])
        pwall = sum(timed_batch([r], (pv, pxc, pg), K) for r in runs)
        plain = cells * K / pwall / 1e6
        line['plain_one_at_a_time'] = round(plain, 1)
        line['plain_launch'] = [pv, pxc, pg]
        line['batched_over_plain'] = round(value / plain, 3)
    log('%s: kernel profile' % args.workload)
    line['roofline'] = roofline_of(scheme, es, variant, xchunk, cells,
                                   profile_kernels(runs, launch, args.profile_chunks), args.workload)
    if args.workload in ('c3', 'c2', 'c2te') and not args.no_accuracy:
        log('%s: full-length S21 accuracy, FP32 against FP64' % args.workload)
        line['s21_vs_fp64'] = s21_accuracy(descs, launch)
    if not args.no_cpu:
        p0 = dict(descs[0])
        lbl = 'the %s grid of the first run' % 'x'.join(str(p0[k]) for k in ('nx', 'ny', 'nz') if k in p0)
        log('%s: CPU baseline' % args.workload)
        line['cpu_baseline'] = cpu_baseline_pool(p0, args.cpu_seconds, lbl)
    return line


def extra_ensemble(args, name, nt):
    """Summary of another ensemble at its full step count (batched against
    one at a time, dominant kernel roofline)."""
    descs, label = workload(name)
    sub = argparse.Namespace(**dict(vars(args), steps=nt, warmup=32, workload=name, no_cpu=True,
                                    no_accuracy=True, no_plain=False))
    t0 = time.perf_counter()
    line = bench_ensemble(sub, descs, label)
    keep = ('value', 'steps', 'ms_per_step', 'hbm_roofline_frac_step', 'plain_one_at_a_time',
            'batched_over_plain', 'roofline')
    out = {k: line[k] for k in keep if k in line}
    out['workload'] = label
    out['launch'] = [line['config'][k] for k in ('variant', 'xchunk', 'grid_cap')]
    out['seconds'] = round(time.perf_counter() - t0, 1)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=1024)
    ap.add_argument('--warmup', type=int, default=64)
    ap.add_argument('--workload', default='c3',
                    help='c3 (default), c0, c1, c2, c2te, c4 (serving), or a study: sl_tm, sl_te, sl_3d, wi_tm, '
                         'wi_te, wi_3d')
    ap.add_argument('--xchunk', type=int, default=0, help='0: library default for the scheme')
    ap.add_argument('--variant', type=int, default=0,
                    help='kernel form: 1 march, 2 batched loads, 3 fused step, 4 blocked step (xchunk steps per launch); '
                         '0: autotuner picks form and xchunk')
    ap.add_argument('--grid', type=int, default=0, help='persistent grid size cap of the stencil launches (0: 8192)')
    ap.add_argument('--profile-chunks', type=int, default=4)
    ap.add_argument('--cpu-seconds', type=float, default=15.0, help='length of the CPU baseline sample')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-plain', action='store_true', help='ensembles: skip the one-at-a-time leg')
    ap.add_argument('--no-accuracy', action='store_true', help='ensembles: skip the full-length S21 accuracy leg')
    ap.add_argument('--no-extras', action='store_true',
                    help='default workload: skip the config 2 ensemble and config 4 serving summaries')
    ap.add_argument('--requests', type=int, default=256, help='serving workload: number of requests')
    ap.add_argument('--realisations', type=int, default=1,
                    help='study workloads: roughness realisations in one batch (against the same run one at a time)')
    ap.add_argument('--copy-probe', type=float, default=0.0,
                    help='only run the copy kernel on this many MiB (counter calibration) and exit')
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) > 1:
        print('bench.py: measures one GPU only; --gpus %d is not supported '
              '(runs are independent and shard across GPUs without communication)' % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    from oidt_amd import _lib

    if args.copy_probe > 0:
        nbytes = int(args.copy_probe * 2 ** 20)
        ms = _lib.copy_probe(nbytes, 10)
        print(json.dumps({'copy_probe_bytes': nbytes, 'ms': ms, 'GBps_rw': 2 * nbytes / ms / 1e6}))
        return

    if args.workload == 'c4':
        c4 = bench_serving(args)
        line = {'metric': 'Mcell-updates/s', 'value': c4.pop('value'), 'unit': 'Mcell-updates/s', 'n_gpus': 1,
                'steps': None, 'warmup': 0, 'ms_per_step': None, 'higher_is_better': True, 'scaling': 'weak',
                'vs_baseline': None, 'dtype': 'f32', 'data': 'synthetic',
                'config': {'workload': c4.pop('workload'), 'parallelism': 'single GPU'}}
        line.update(c4)
        print(json.dumps(line), flush=True)
        return
    if args.workload in STUDIES:
        bench_study(args)
        return
    descs, label = workload(args.workload)
    line = bench_ensemble(args, descs, label)
    if args.workload == 'c3' and not args.no_extras:
        # the other batched configurations, each at its own full length
        line['c2'] = extra_ensemble(args, 'c2', 5000)
        line['c4'] = bench_serving(args)
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
