"""The HIP path on the five benchmark configurations at their real sizes.

* config 0 (2D TMz, 600 x 240, FP64, 4000 steps, 32 frequencies): the whole
  four-run S-parameter protocol against four oracle runs at full length --
  fields and port DFT within 1e-10 (FP64), S11/S21/S12/S22 within 0.01 dB and
  0.1 degree -- and the same grid in FP32 against the same oracle runs (the
  FP32 error after 4000 steps, S within 0.01 dB / 0.1 degree).
* config 1 (3D, 128 x 96 x 320, FP32): the full grid against the FP64 C
  oracle over the first 1200 steps (the wave packet peaks at step ~1120):
  fields and port DFT within relative L2 1e-4; the full 6000 steps against
  the FP64 device run, whose kernels the small-grid tests pin to the oracle at
  1e-10 (S21/S11 within 0.01 dB / 0.1 degree, power balance stated).
* config 2 (64 x 2D FP32) and config 3 (16 x 3D FP32): the whole batch on a
  truncated step count, every run bit-identical to itself alone; config 2's
  two shortest realisations against the oracle; config 3's S21 / S11 per
  realisation against the FP64 C oracle at full length (5000 steps) for its
  two shortest realisations, with one shared incident run.
* config 4 (serving mix): requests drawn from the mix at their real sizes and
  step counts, bit-identical alone and in the mix.

Reference loops followed by the oracle: FDTD/GPU/FDTD_2D_TM/s-parameter_
extraction/fdtd_main.py:275-304 and FDTD/GPU/FDTD_3D/s-parameter_extraction/
fdtd_main.py:357-406; extraction extract_sparams.py:52-104 (3D :52-144)."""
import numpy as np
import pytest

from oidt_amd import configs, ensemble, extract, materials, serve
from oidt_amd.solver import Batch, Solver
from oracle import extract as oextract
from oracle import fdtd_c
from oracle import geometry as ogeo
from oracle import setup as osetup

from helpers import oracle_run, rel_l2, window

COMP_ORDER = {'tm': ('e', 'h'), 'te': ('e', 'h'), '3d': ('ex', 'ey', 'hx', 'hy')}
FILE_COMP = {'tm': ('ez', 'hy'), 'te': ('ey', 'hz'), '3d': ('ex', 'ey', 'hx', 'hy')}
RUNS = (('f', 'i'), ('f', 'r'), ('b', 'i'), ('b', 'r'))


def files_from(scheme, dfts):
    files = {}
    for (d, k), dft in dfts.items():
        for port, tag in extract.run_tags(d, k).items():
            files[(tag, 1)] = tuple(dft[(key, port)] for key in COMP_ORDER[scheme])
    return files


def db_phase_err(a, b):
    return ensemble.s_error(a, b)


def protocol(p):
    return {('f', 'i'): configs.incident(p), ('b', 'i'): configs.incident(p), ('f', 'r'): p, ('b', 'r'): p}


@pytest.fixture(scope='module')
def config0_oracle():
    """The four FP64 oracle runs of config 0's protocol (about a minute)."""
    p = configs.config0()
    out = {}
    for key, q in protocol(p).items():
        print('oracle config 0 run', key, flush=True)
        s, up, lo = configs.realise(q)
        o, FG = oracle_run(q, up, lo, key[0])
        out[key] = (o.fields() if key == ('f', 'r') else None, o.dft(), FG)
    S = oextract.s_matrix('tm', osetup.explicit_setup(p), files_from('tm', {k: v[1] for k, v in out.items()}), 1,
                          p['nf'])
    return p, out, S


def device_protocol(p, precision, launch=None):
    b = Batch(p['scheme'], np.float32 if precision == 'float32' else np.float64, max_slots=4)
    if launch:
        b.set_launch(launch[1], launch[2], launch[0])
    slots, setups, fgs = {}, {}, {}
    for (d, k), q in protocol(dict(p, precision=precision)).items():
        s, up, lo = configs.realise(q)
        fg = materials.fg_mask(s, up, lo)
        setups[(d, k)], fgs[(d, k)] = s, fg
        slots[(d, k)] = b.add(s, fg, d, 0)
    b.advance(-(-p['nt'] // 32))
    b.sync()
    return b, slots, setups, fgs


@pytest.mark.gpu
@pytest.mark.parametrize('precision,tol,launch', [('float64', 1e-10, None), ('float32', None, None),
                                                 ('float64', 1e-10, (4, 8, 0)), ('float32', None, (4, 4, 0))])
def test_config0_protocol_full_length(config0_oracle, precision, tol, launch):
    p, ora, S_ora = config0_oracle
    print('case config 0', precision, 'launch', launch, flush=True)
    b, slots, setups, fgs = device_protocol(p, precision, launch)
    try:
        for key in RUNS:
            assert np.array_equal(fgs[key], ora[key][2]), key
        sr = slots[('f', 'r')]
        c = osetup.explicit_setup(p)
        w = window('tm', c)
        errs = {}
        for name, ref in ora[('f', 'r')][0].items():
            errs[name] = rel_l2(b.field(sr, name), ref)
        for key in RUNS:
            od = ora[key][1]
            for k2, z in b.dft(slots[key]).items():
                errs['dft %s%s %s%d' % (key + k2)] = rel_l2(z[w], od[k2][w])
        for k, e in sorted(errs.items()):
            print(k, 'rel L2 %.3e' % e, flush=True)
        if tol is not None:
            assert max(errs.values()) < tol, errs
        dev = {key: b.dft(slot) for key, slot in slots.items()}
        files = files_from('tm', dev)
        S = extract.s_matrix('tm', setups[('f', 'r')],
                             lambda comp, t, l: files[(t, l)][FILE_COMP['tm'].index(comp)], 1, p['nf'])
        for i, j in ((0, 0), (1, 0), (0, 1), (1, 1)):
            e_db, e_ph = db_phase_err(S[:, i, j], S_ora[:, i, j])
            print('S%d%d mean |S| %.4f: %.2e dB, %.2e deg' % (i + 1, j + 1, np.abs(S_ora[:, i, j]).mean(), e_db, e_ph),
                  flush=True)
            assert e_db <= 0.01 and e_ph <= 0.1, (i, j, e_db, e_ph)
    finally:
        b.close()


def c_oracle(p, direction='f', nsteps=None):
    s, up, lo = configs.realise(p)
    c = osetup.explicit_setup(p)
    FG = ogeo.fg_mask(c, up, lo)
    o = fdtd_c.make(c, FG, direction)
    o.run(nsteps)
    return o, FG


@pytest.mark.gpu
def test_config1_full_grid_against_c_oracle():
    """128 x 96 x 320 cells, FP32 device against the FP64 C oracle over the
    first 1200 steps, in the default form and the blocked step."""
    p = dict(configs.config1(), nt=1200)
    print('case config 1, 1200 steps', flush=True)
    o, FG = c_oracle(p)
    s, up, lo = configs.realise(p)
    fg = materials.fg_mask(s, up, lo)
    assert np.array_equal(fg, FG)
    w = window('3d', o.c)
    od = o.dft()
    for launch in (None, (4, 2, 0)):
        print('launch', launch, flush=True)
        dev = Solver(s, fg)
        if launch:
            dev.set_launch(launch[1], launch[2], launch[0])
        dev.run()
        try:
            for name, ref in o.fields().items():
                assert np.abs(ref).max() > 0, name
                err = rel_l2(dev.field(name), ref)
                print(name, 'rel L2 %.3e' % err, flush=True)
                assert err < 1e-4, (launch, name, err)
            for key, z in dev.dft().items():
                err = rel_l2(z[w], od[key][w])
                print('dft', key, 'rel L2 %.3e' % err, flush=True)
                assert err < 1e-4, (launch, key, err)
        finally:
            dev.close()


@pytest.mark.gpu
def test_config1_full_length_fp32_against_fp64():
    """All 6000 steps: S21 / S11 of the FP32 run (with its incident run)
    within 0.01 dB / 0.1 degree of the FP64 device run; the power balance
    |S11|^2 + |S21|^2 stated.  The fields after 6000 steps are the small
    remainder of a pulse that has left the grid, so their FP32 error is
    stated against each field's norm at step 1200 (near the pulse peak):
    within 1e-4."""
    p = configs.config1()
    res = {}
    names = ('ex', 'ey', 'ez', 'hx', 'hy', 'hz')
    for prec in ('float32', 'float64'):
        e = ensemble.Ensemble([p], precision=prec).run()
        res[prec] = (e.s_params(0), {k: e.batch.field(e.rough_slot[(0, 'f')], k) for k in names})
        e.close()
    s32, f32 = res['float32']
    s64, f64 = res['float64']
    for k in ('S11', 'S21'):
        e_db, e_ph = db_phase_err(s32[k], s64[k])
        print(k, 'mean %.3f dB: %.2e dB, %.2e deg' % (np.mean(ensemble.db(s64[k])), e_db, e_ph), flush=True)
        assert e_db <= 0.01 and e_ph <= 0.1
    s, up, lo = configs.realise(dict(p, precision='float64', nt=1200))
    mid = Solver(s, materials.fg_mask(s, up, lo)).run()
    for name in names:
        scale = np.linalg.norm(mid.field(name))
        err = np.linalg.norm((f32[name] - f64[name]).ravel()) / scale
        print(name, 'after 6000 steps: |f32 - f64| / |f64 at step 1200| = %.3e (final |f64| / that = %.3e)'
              % (err, np.linalg.norm(f64[name]) / scale), flush=True)
        assert err < 1e-4, (name, err)
    mid.close()
    # S = sqrt(Pref / Pinc) of complex port powers, with Pinc from the separate
    # shortened incident run (extract_sparams.py:88-104): not passive by
    # construction, so the power balance is stated, and bounded loosely
    bal = np.abs(s64['S11']) ** 2 + np.abs(s64['S21']) ** 2
    print('|S11|^2 + |S21|^2: min %.4f, max %.4f' % (bal.min(), bal.max()), flush=True)
    assert np.all(bal < 1.1) and np.all(bal > 0.5)


def batch_vs_alone(descs, nt, launch=None):
    runs = []
    for p in descs:
        s, up, lo = configs.realise(dict(p, nt=nt))
        runs.append((s, materials.fg_mask(s, up, lo)))
    s0 = runs[0][0]
    b = Batch(s0.scheme, s0.precision, max_slots=len(runs))
    if launch:
        b.set_launch(launch[1], launch[2], launch[0])
    slots = [b.add(s, fg) for s, fg in runs]
    b.advance(-(-nt // 32))
    b.sync()
    names = {'3d': ('ex', 'ey', 'ez', 'hx', 'hy', 'hz'), 'tm': ('ez', 'hx', 'hy'), 'te': ('ex', 'ey', 'hz')}[s0.scheme]
    try:
        for r, ((s, fg), slot) in enumerate(zip(runs, slots)):
            a = Batch(s.scheme, s.precision, max_slots=1)
            if launch:
                a.set_launch(launch[1], launch[2], launch[0])
            sa = a.add(s, fg)
            a.advance(-(-nt // 32))
            a.sync()
            for name in names:
                assert np.array_equal(a.field(sa, name), b.field(slot, name)), (r, name)
            da, db_ = a.dft(sa), b.dft(slot)
            for k in da:
                assert np.array_equal(da[k], db_[k]), (r, k)
            a.close()
    finally:
        b.close()


@pytest.mark.gpu
@pytest.mark.parametrize('launch', [None, (3, 16, 0), (2, 4, 0), (4, 4, 0)])
def test_config2_batch_members_identical_alone(launch):
    print('case config 2, 64 runs, launch', launch, flush=True)
    batch_vs_alone(configs.config2(), 160, launch)


@pytest.mark.gpu
@pytest.mark.parametrize('launch', [(4, 8, 0), (3, 16, 0)])
def test_config2_tez_batch_members_identical_alone(launch):
    print('case config 2 in TEz, 64 runs, launch', launch, flush=True)
    batch_vs_alone(configs.config2(scheme='te'), 160, launch)


@pytest.mark.gpu
@pytest.mark.parametrize('launch', [None, (2, 1, 0), (3, 16, 0), (4, 2, 0)])
def test_config3_batch_members_identical_alone(launch):
    print('case config 3, 16 runs, launch', launch, flush=True)
    batch_vs_alone(configs.config3(), 96, launch)


@pytest.mark.gpu
def test_config2_shortest_realisations_against_oracle():
    """The two shortest of the 64 realisations, 1500 steps, in one batch,
    against the oracle: fields and port DFT within relative L2 1e-4."""
    descs = sorted(configs.config2(), key=lambda p: p['nx'])[:2]
    descs = [dict(p, nt=1500) for p in descs]
    b = Batch('tm', np.float32, max_slots=2)
    slots = []
    for p in descs:
        s, up, lo = configs.realise(p)
        slots.append(b.add(s, materials.fg_mask(s, up, lo)))
    b.advance(-(-1500 // 32))
    b.sync()
    try:
        for p, slot in zip(descs, slots):
            print('case config 2 realisation nx', p['nx'], flush=True)
            s, up, lo = configs.realise(p)
            o, _ = oracle_run(p, up, lo)
            for name, ref in o.fields().items():
                err = rel_l2(b.field(slot, name), ref)
                print(name, 'rel L2 %.3e' % err, flush=True)
                assert err < 1e-4, (name, err)
            w = window('tm', o.c)
            od = o.dft()
            for key, z in b.dft(slot).items():
                err = rel_l2(z[w], od[key][w])
                print('dft', key, 'rel L2 %.3e' % err, flush=True)
                assert err < 1e-4, (key, err)
    finally:
        b.close()


@pytest.mark.gpu
def test_config3_s21_per_realisation_against_c_oracle():
    """Config 3's two shortest realisations at full length (5000 steps, 64
    frequencies), FP32 on the device with one shared incident run, against
    the FP64 C oracle's incident and rough runs: S21 and S11 within 0.01 dB
    and 0.1 degree, per realisation."""
    all_descs = configs.config3()
    order = sorted(range(len(all_descs)), key=lambda r: all_descs[r]['nz'])[:2]
    descs = [all_descs[r] for r in order]
    e = ensemble.Ensemble(descs, precision='float32').run()
    try:
        assert len(e.incident) == 1
        dev = [e.s_params(r) for r in range(len(descs))]
    finally:
        e.close()
    inc_desc = configs.incident(descs[0])
    print('oracle incident run', flush=True)
    oi, _ = c_oracle(inc_desc)
    nf = descs[0]['nf']
    for r, p in zip(order, descs):
        print('oracle realisation', r, 'nz', p['nz'], flush=True)
        o, _ = c_oracle(p)
        c = osetup.explicit_setup(p)
        inc = tuple(oi.dft()[(k, 1)] for k in COMP_ORDER['3d'])
        ft = tuple(o.dft()[(k, 1)] for k in COMP_ORDER['3d'])
        br = tuple(o.dft()[(k, 2)] for k in COMP_ORDER['3d'])
        # the backward runs are not simulated: their slots repeat the forward
        # ones, which leaves S[0, 0] = S11 (ft - fi over fi) and S[1, 0] = S21
        # (br over fi) as extract_sparams.py forms them
        files = {('fi', 1): inc, ('bi', 1): inc, ('ft', 1): ft, ('bt', 1): br, ('br', 1): br, ('fr', 1): ft}
        S = oextract.s_matrix('3d', c, files, 1, nf)
        S11, S21 = S[:, 0, 0], S[:, 1, 0]
        d = dev[descs.index(p)]
        for name, ref, got in (('S11', S11, d['S11']), ('S21', S21, d['S21'])):
            e_db, e_ph = db_phase_err(got, ref)
            print('realisation %d %s mean %.3f dB: %.2e dB, %.2e deg' % (r, name, np.mean(ensemble.db(ref)), e_db,
                                                                         e_ph), flush=True)
            assert e_db <= 0.01 and e_ph <= 0.1, (r, name, e_db, e_ph)
    assert nf == 64


@pytest.mark.gpu
def test_config4_requests_identical_alone_and_in_mix():
    """Eight requests of the serving mix (real sizes and step counts), served
    with continuous batching under a budget that makes admission wait, each
    bit-identical to the same request solved alone."""
    descs = configs.serving_requests(count=256)
    pick = [i for i, p in enumerate(descs) if p['scheme'] == 'tm'][:5] + \
           [i for i, p in enumerate(descs) if p['scheme'] == '3d'][:3]
    sub = [descs[i] for i in pick]
    reqs = serve.make_requests(sub, [0.0, 0.001, 0.002, 0.01, 0.02, 0.03, 0.05, 0.08])
    be = serve.DeviceBackend(max_slots=8)
    fps = [be.footprint(configs.realise(d)[0]) for d in sub]
    srv = serve.Server(be, hbm_budget=max(fps) + sorted(fps)[-2], chunk_group=4)
    done = srv.run(reqs)
    be.close()
    assert all(r.error is None for r in done)
    assert srv.deferred_for_memory > 0
    for r in done:
        print('request', pick[r.rid], r.setup.scheme, r.setup.shape, r.setup.nt, 'latency %.3f' % r.latency,
              flush=True)
        alone = Solver(r.setup, r.fg).run()
        ref = alone.dft()
        alone.close()
        for k, z in ref.items():
            assert np.array_equal(z, r.result[k]), (r.rid, k)
