#!/usr/bin/env python3
"""Benchmark of the FDTD time-stepping hot path on one MI355X.

Default workload: the plain 3D FP32 run (strip waveguide, (x, y, z) =
(128, 96, 320) cells, 64 DFT frequencies in FP64 accumulators), a "step" being
one Yee time step (E and H half-steps with CPML, source and port DFT).

Prints one JSON line with Mcell-updates/s (all grid cells per step, the
reference's MCells/s definition), the roofline of the dominant kernel measured
with HIP events on the library stream, and the oracle timed on the host as
the CPU baseline.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c1|c0|c2|c3]
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
    steps until ``seconds`` of CPU work have run (a bounded sample)."""
    from oracle import fdtd as ofdtd
    from oracle import geometry as ogeo
    from oracle import setup as osetup
    c = osetup.explicit_setup(p)
    o = ofdtd.make(c, ogeo.fg_mask(c))
    o.run(1)
    steps = 0
    t0 = time.perf_counter()
    while True:
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
    tune = {}
    for v, xc in forms:
        sol, mons, ids = variants.prepare(c, (v, xc, 0))
        out = (C.c_double * 6)()
        try:
            _lib.check(_lib.lib().oidt_solver_profile_all(sol.h, 2, out))
        finally:
            sol.close()
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
    sol, mons, ids = variants.prepare(c, (variant, xchunk, 0))
    out = (C.c_double * 6)()
    try:
        _lib.check(_lib.lib().oidt_solver_profile_all(sol.h, int(args.profile_chunks), out))
    finally:
        sol.close()
    ms_e, ms_h, ms_dft, ms_adv, ms_me, ms_mh = list(out)
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
    grid, whole steps (recording included) for about ``seconds``."""
    from oidt_amd import variants
    from oracle import studies as ost
    c = variants.config(scheme, study)
    o = ost.make(c)
    o.step(0)
    o.record(0)
    n, t0 = 1, time.perf_counter()
    while True:
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


def bench_serving(args):
    """Serving mix: plain one-at-a-time path, saturated continuous batching, and
    Poisson arrivals at 0.8x the saturated capacity (latency percentiles)."""
    from oidt_amd import configs, serve
    descs = configs.serving_requests(count=args.requests)
    be = serve.DeviceBackend(max_slots=64, xchunk=args.xchunk, variant=args.variant or 2)
    reqs = serve.make_requests(descs, [0.0] * len(descs))
    t0 = time.perf_counter()
    for r in reqs:
        serve.prepare_request(r, be)
    prep = time.perf_counter() - t0
    tuned = {}
    if not args.variant:
        # autotune each scheme once on (up to) 16 of its requests
        for scheme in sorted({r.setup.scheme for r in reqs}):
            sub = [(r.setup, r.fg) for r in reqs if r.setup.scheme == scheme][:16]
            variant, xchunk, grid, rec = choose_launch(args, sub)
            be.launch[scheme] = (variant, xchunk, grid)
            tuned[scheme] = dict(rec, choice=[variant, xchunk, grid])
    # plain: each request alone, back to back, same kernel forms
    t0 = time.perf_counter()
    plain = serve.serve_plain(reqs, be.launch)
    plain_wall = time.perf_counter() - t0
    total_cu = sum(r.cell_updates for r in reqs)
    plain_rate = total_cu / plain_wall / 1e6
    # saturated continuous batching (everything queued at t = 0)
    hbm = int(0.8 * __import__('oidt_amd._lib', fromlist=['x']).device_info()['hbm_free_bytes'])
    for r in reqs:
        r.result = None
    srv = serve.Server(be, hbm_budget=hbm, chunk_group=4, prepare=lambda r, b: None)
    t0 = time.perf_counter()
    done = srv.run(reqs)
    sat_wall = time.perf_counter() - t0
    sat = serve.summarize(done, sat_wall)
    cap_rps = len(reqs) / sat_wall
    # Poisson arrivals at 0.8x the saturated request capacity
    rate = 0.8 * cap_rps
    arr = serve.poisson_arrivals(len(reqs), rate, seed=4)
    preqs = serve.make_requests(descs, arr)
    for a, b_ in zip(preqs, reqs):
        a.setup, a.fg, a.footprint = b_.setup, b_.fg, b_.footprint
    srv2 = serve.Server(be, hbm_budget=hbm, chunk_group=4, prepare=lambda r, b: None)
    t0 = time.perf_counter()
    pdone = srv2.run(preqs)
    pw = time.perf_counter() - t0
    pois = serve.summarize(pdone, pw)
    be.close()
    n2d = sum(1 for r in reqs if r.setup.scheme == 'tm')
    line = {
        'metric': 'Mcell-updates/s', 'value': round(sat['mcell_updates_per_s'], 1), 'unit': 'Mcell-updates/s',
        'n_gpus': 1, 'steps': int(sum(r.setup.nt for r in reqs)), 'warmup': 0,
        'ms_per_step': None, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32',
        'data': 'synthetic',
        'config': {'workload': 'serving mix: %d requests, %d 2D TMz (nx 400..4000, ny 200..240) and %d 3D '
                               '((128..384)x(64..128)x(48..96)), 2000..8000 steps, FP32' % (len(reqs), n2d, len(reqs) - n2d),
                   'max_slots_per_batch': 64, 'hbm_budget_bytes': hbm, 'launch': be.launch or None,
                   'autotune': tuned or None},
        'plain_mcell_updates_per_s': round(plain_rate, 1),
        'served_over_plain': round(sat['mcell_updates_per_s'] / plain_rate, 3),
        'saturated': {k: (round(v, 4) if isinstance(v, float) else v) for k, v in sat.items()},
        'poisson': dict({k: (round(v, 4) if isinstance(v, float) else v) for k, v in pois.items()},
                        rate_per_s=round(rate, 3)),
        'peak_resident_runs': srv.peak_resident, 'host_prepare_s': round(prep, 3),
    }
    print(json.dumps(line), flush=True)


def bench_plain_ensemble(args, descs, label):
    """The ensemble one run at a time (each alone on the device), same steps,
    same kernel form as the batched path (autotuned on the first run when
    --variant is 0)."""
    from oidt_amd import configs, materials
    from oidt_amd.solver import Solver
    K = args.steps
    runs = []
    for p in descs:
        p = dict(p, nt=K)
        s, up, lo = configs.realise(p)
        runs.append((s, materials.fg_mask(s, up, lo)))
    variant, xchunk, grid, tuned = choose_launch(args, runs[:1])
    # warm the device on the first run
    Solver(*runs[0]).set_launch(xchunk, grid, variant).advance(64).sync().close()
    t0 = time.perf_counter()
    for s, fg in runs:
        sv = Solver(s, fg)
        sv.set_launch(xchunk, grid, variant)
        sv.run()
        sv.close()
    wall = time.perf_counter() - t0
    cells = sum(s.cells for s, _ in runs)
    print(json.dumps({'metric': 'Mcell-updates/s', 'value': round(cells * K / wall / 1e6, 1),
                      'unit': 'Mcell-updates/s', 'n_gpus': 1, 'steps': K, 'warmup': 0,
                      'ms_per_step': round(wall / K * 1e3, 5), 'higher_is_better': True, 'scaling': 'weak',
                      'vs_baseline': None, 'dtype': 'f32', 'data': 'synthetic',
                      'config': {'workload': label + ', plain: one run at a time incl. create/upload',
                                 'runs': len(runs), 'variant': variant, 'xchunk': xchunk, 'grid_cap': grid,
                                 'autotune': tuned}}),
          flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=6016)
    ap.add_argument('--warmup', type=int, default=256)
    ap.add_argument('--workload', default='c1',
                    help='c0, c1, c2, c2te, c3, c4 (serving), or a study: sl_tm, sl_te, sl_3d, wi_tm, wi_te, wi_3d')
    ap.add_argument('--xchunk', type=int, default=0, help='0: library default for the scheme')
    ap.add_argument('--variant', type=int, default=0,
                    help='kernel form: 1 march, 2 batched loads, 3 fused step; 0: autotuner picks form and xchunk')
    ap.add_argument('--grid', type=int, default=0, help='persistent grid size cap of the stencil launches (0: 8192)')
    ap.add_argument('--profile-chunks', type=int, default=4)
    ap.add_argument('--cpu-seconds', type=float, default=15.0, help='length of the CPU baseline sample')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--plain', action='store_true', help='ensembles: one run at a time instead of one batch')
    ap.add_argument('--requests', type=int, default=48, help='serving workload: number of requests')
    ap.add_argument('--realisations', type=int, default=1,
                    help='study workloads: roughness realisations in one batch (against the same run one at a time)')
    ap.add_argument('--copy-probe', type=float, default=0.0,
                    help='only run the copy kernel on this many MiB (counter calibration) and exit')
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) > 1:
        print('bench.py: this build measures one GPU only; --gpus %d is not supported '
              '(runs are independent and shard across GPUs without communication)' % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    from oidt_amd import _lib, configs, materials
    from oidt_amd.solver import Batch

    if args.copy_probe > 0:
        nbytes = int(args.copy_probe * 2 ** 20)
        ms = _lib.copy_probe(nbytes, 10)
        print(json.dumps({'copy_probe_bytes': nbytes, 'ms': ms, 'GBps_rw': 2 * nbytes / ms / 1e6}))
        return

    if args.workload == 'c4':
        bench_serving(args)
        return
    if args.workload in STUDIES:
        bench_study(args)
        return
    descs, label = workload(args.workload)
    if args.plain:
        bench_plain_ensemble(args, descs, label)
        return
    chunk = _lib.chunk_steps()
    warm = -(-args.warmup // chunk) * chunk
    K = args.steps
    runs = []
    for p in descs:
        p = dict(p, nt=warm + K)
        s, up, lo = configs.realise(p)
        runs.append((p, s, materials.fg_mask(s, up, lo)))
    scheme, prec = runs[0][1].scheme, runs[0][1].precision
    variant, xchunk, grid, tuned = choose_launch(args, [(s, fg) for _, s, fg in runs])
    b = Batch(scheme, prec, max_slots=len(runs))
    b.set_launch(xchunk, grid, variant)
    variant, xchunk, grid = b.launch()
    for p, s, fg in runs:
        b.add(s, fg)
    b.advance(warm // chunk)
    b.sync()
    nch = -(-K // chunk)
    t0 = time.perf_counter()
    b.advance(nch)
    b.sync()
    wall = time.perf_counter() - t0
    cells = sum(s.cells for _, s, _ in runs)
    value = cells * K / wall / 1e6

    # Per-kernel durations with HIP events on the batch stream (a second batch of
    # the same runs, so the timed run above is untouched).
    pb = Batch(scheme, prec, max_slots=len(runs))
    pb.set_launch(xchunk, grid, variant)
    for p, s, fg in runs:
        pb.add(s, fg)
    import ctypes as C
    out = (C.c_double * 4)()
    _lib.check(_lib.lib().oidt_batch_profile(pb.h, int(args.profile_chunks), out))
    ms_e, ms_h, ms_dft, ms_adv = list(out)
    pb.close()
    es = 4 if prec == np.float32 else 8
    be, bh = algorithmic_bytes(scheme, es)
    stem = {'tm': 'k_tmz', 'te': 'k_tez', '3d': 'k_3d'}[scheme]
    if variant == 3:
        kname, ms_k, bytes_k = (stem + '_step', ms_e, (be + bh) * cells)
    elif ms_e >= ms_h:
        kname, ms_k, bytes_k = (stem + ('_e2' if variant == 2 else '_e'), ms_e, be * cells)
    else:
        kname, ms_k, bytes_k = (stem + ('_h2' if variant == 2 else '_h'), ms_h, bh * cells)
    achieved = bytes_k / (ms_k * 1e-3) / 1e9
    traffic, traffic_src = committed_traffic(kname, args.workload, variant, xchunk)
    roof = {'bound': 'hbm', 'kernel': kname, 'achieved': round(achieved, 1), 'peak': HBM_PEAK_GBS,
            'unit': 'GB/s', 'frac': round(achieved / HBM_PEAK_GBS, 4), 'traffic': traffic,
            'traffic_source': traffic_src,
            'bytes_per_launch': bytes_k, 'ms_per_launch': round(ms_k, 5),
            'ms_e': round(ms_e, 5), 'ms_h': round(ms_h, 5), 'ms_dft': round(ms_dft, 5), 'ms_advance': round(ms_adv, 5)}
    line = {
        'metric': 'Mcell-updates/s', 'value': round(value, 1), 'unit': 'Mcell-updates/s', 'n_gpus': 1,
        'steps': K, 'warmup': warm, 'ms_per_step': round(wall / K * 1e3, 5), 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32' if es == 4 else 'f64', 'data': 'synthetic',
        'config': {'workload': label, 'runs': len(runs), 'cells_per_step': cells,
                   'steps_per_run': K, 'xchunk': xchunk, 'variant': variant, 'grid_cap': grid,
                   'parallelism': 'single GPU',
                   'autotune': tuned},
        'hbm_roofline_frac_step': round(value * 1e6 * (be + bh) / (HBM_PEAK_GBS * 1e9), 4),
        'algorithmic_bytes_per_cell_update': be + bh,
        'roofline': roof,
    }
    b.close()
    if not args.no_cpu:
        # CPU leg: the oracle on the first run's grid, bounded sample.
        p0 = dict(descs[0])
        rate, secs, nsteps = cpu_baseline(p0, args.cpu_seconds)
        line['cpu_baseline'] = {'value': round(rate, 3), 'unit': 'Mcell-updates/s', 'cores': 1, 'kind': 'port',
                                'sample': '%d steps (after 1 warm-up) of the FP64 NumPy oracle on the %dx%dx%d grid '
                                          'of the first run, %.1f s' % (nsteps, p0['nx'], p0['ny'],
                                                                         p0.get('nz', 1), secs)}
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
