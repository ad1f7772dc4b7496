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


def cpu_baseline(p, steps):
    """Oracle (FP64 NumPy restatement, one thread) on the same grid, bounded sample."""
    from oracle import fdtd as ofdtd
    from oracle import geometry as ogeo
    from oracle import setup as osetup
    c = osetup.explicit_setup(p)
    o = ofdtd.make(c, ogeo.fg_mask(c))
    o.run(1)
    t0 = time.perf_counter()
    o.run(steps)
    dt = time.perf_counter() - t0
    cells = c.nx * c.ny * (c.nz if c.scheme == '3d' else 1)
    return cells * steps / dt / 1e6, dt


def workload(name):
    from oidt_amd import configs
    if name == 'c0':
        return [configs.config0()], '2D TMz FP64 slab 600x240, 32 freqs (config 0)'
    if name == 'c1':
        return [configs.config1()], '3D FP32 strip (x,y,z)=(128,96,320), 64 freqs (config 1)'
    if name == 'c2':
        return configs.config2(), '64 x 2D TMz FP32 slabs nx~U{800..4000} x 240, one batch (config 2)'
    if name == 'c3':
        return configs.config3(), '16 x 3D FP32 strips (96,64,nz~U{192..512}), one batch (config 3)'
    raise SystemExit('unknown workload ' + name)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=6016)
    ap.add_argument('--warmup', type=int, default=256)
    ap.add_argument('--workload', default='c1')
    ap.add_argument('--xchunk', type=int, default=8)
    ap.add_argument('--variant', type=int, default=2)
    ap.add_argument('--profile-chunks', type=int, default=4)
    ap.add_argument('--cpu-steps', type=int, default=20)
    ap.add_argument('--no-cpu', action='store_true')
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

    descs, label = workload(args.workload)
    chunk = _lib.chunk_steps()
    warm = -(-args.warmup // chunk) * chunk
    K = args.steps
    runs = []
    for p in descs:
        p = dict(p, nt=warm + K)
        s, up, lo = configs.realise(p)
        runs.append((p, s, materials.fg_mask(s, up, lo)))
    scheme, prec = runs[0][1].scheme, runs[0][1].precision
    b = Batch(scheme, prec, max_slots=len(runs))
    b.set_launch(args.xchunk, 0, args.variant)
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
    pb.set_launch(args.xchunk, 0, args.variant)
    for p, s, fg in runs:
        pb.add(s, fg)
    import ctypes as C
    out = (C.c_double * 4)()
    _lib.check(_lib.lib().oidt_batch_profile(pb.h, int(args.profile_chunks), out))
    ms_e, ms_h, ms_dft, ms_adv = list(out)
    pb.close()
    es = 4 if prec == np.float32 else 8
    be, bh = algorithmic_bytes(scheme, es)
    if ms_e >= ms_h:
        kname, ms_k, bytes_k = ('E half-step', ms_e, be * cells)
    else:
        kname, ms_k, bytes_k = ('H half-step', ms_h, bh * cells)
    achieved = bytes_k / (ms_k * 1e-3) / 1e9
    roof = {'bound': 'hbm', 'kernel': kname, 'achieved': round(achieved, 1), 'peak': HBM_PEAK_GBS,
            'unit': 'GB/s', 'frac': round(achieved / HBM_PEAK_GBS, 4), 'traffic': None,
            'bytes_per_launch': bytes_k, 'ms_per_launch': round(ms_k, 5),
            'ms_e': round(ms_e, 5), 'ms_h': round(ms_h, 5), 'ms_dft': round(ms_dft, 5), 'ms_advance': round(ms_adv, 5)}
    line = {
        'metric': 'Mcell-updates/s', 'value': round(value, 1), 'unit': 'Mcell-updates/s', 'n_gpus': 1,
        'steps': K, 'warmup': warm, 'ms_per_step': round(wall / K * 1e3, 5), 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32' if es == 4 else 'f64', 'data': 'synthetic',
        'config': {'workload': label, 'runs': len(runs), 'cells_per_step': cells,
                   'steps_per_run': K, 'xchunk': args.xchunk, 'variant': args.variant, 'parallelism': 'single GPU'},
        'hbm_roofline_frac_step': round(value * 1e6 * (be + bh) / (HBM_PEAK_GBS * 1e9), 4),
        'algorithmic_bytes_per_cell_update': be + bh,
        'roofline': roof,
    }
    b.close()
    if not args.no_cpu:
        # CPU leg: the oracle on the first run's grid, bounded sample.
        p0 = dict(descs[0])
        rate, secs = cpu_baseline(p0, args.cpu_steps)
        line['cpu_baseline'] = {'value': round(rate, 3), 'unit': 'Mcell-updates/s', 'cores': 1, 'kind': 'port',
                                'sample': '%d steps (after 1 warm-up) of the FP64 NumPy oracle on the %dx%dx%d grid '
                                          'of the first run, %.1f s' % (args.cpu_steps, p0['nx'], p0['ny'],
                                                                         p0.get('nz', 1), secs)}
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
