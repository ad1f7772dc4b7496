"""Every stencil kernel form and x-chunk depth gives the oracle's answer
(FP32, relative L2 1e-4 on fields and port DFT)."""
import numpy as np
import pytest

from oidt_amd import configs
from oidt_amd.solver import Batch

from helpers import device_inputs, oracle_run, rel_l2, window

FORMS = [(1, 1), (1, 2), (1, 4), (1, 8), (1, 16), (1, 32), (2, 1), (2, 2), (2, 4), (2, 8), (2, 16)]
FUSED = [(3, 1), (3, 2), (3, 4), (3, 8)]          # fused Yee step
FUSED_LONG = {'tm': [(3, 12), (3, 16), (3, 20), (3, 24), (3, 32), (4, 2), (4, 4), (4, 8)],
              '3d': [(3, 16), (4, 2)],
              'te': [(3, 16), (3, 32), (4, 2), (4, 4), (4, 8)]}   # also the temporally blocked step (variant 4)


@pytest.mark.gpu
@pytest.mark.parametrize('scheme', ['tm', 'te', '3d'])
def test_all_launch_forms_match_oracle(scheme):
    p = configs.small(scheme, 'float32')
    p.update(rough_std=15e-9, rough_acl=40e-9, seed=11)
    s, up, lo, fg = device_inputs(p)
    o, _ = oracle_run(p, up, lo)
    ref_f = o.fields()
    od = o.dft()
    w = window(scheme, o.c)
    for variant, xc in FORMS + FUSED + FUSED_LONG[scheme]:
        print('case', scheme, 'variant', variant, 'xchunk', xc, flush=True)
        b = Batch(scheme, np.float32, max_slots=1)
        b.set_launch(xc, 0, variant)
        slot = b.add(s, fg)
        b.advance(-(-s.nt // 32))
        b.sync()
        assert b.poll(slot) == s.nt
        for name, ref in ref_f.items():
            err = rel_l2(b.field(slot, name), ref)
            assert err < 1e-4, (variant, xc, name, err)
        for key, z in b.dft(slot).items():
            err = rel_l2(z[w], od[key][w])
            assert err < 1e-4, (variant, xc, key, err)
        b.close()


@pytest.mark.gpu
@pytest.mark.parametrize('scheme,xc,direction,extra', [
    ('tm', 1, 'b', {}), ('tm', 4, 'f', {'nt': 101}), ('te', 2, 'f', {}), ('te', 16, 'b', {'nt': 101}),
    ('3d', 1, 'f', {}), ('3d', 2, 'b', {}), ('3d', 4, 'f', {'nt': 101}), ('3d', 8, 'f', {'source_condition': 'TM'}),
    ('3d', 16, 'b', {'source_condition': 'TM'}),
])
def test_fused_step_fp64(scheme, xc, direction, extra):
    """The fused step in FP64 equals the oracle to 1e-10 (same precision, only
    the order of operations differs): both field sets alternate correctly,
    also after an odd step count and with the 3D TM source on Hy."""
    p = configs.small(scheme, 'float64', **extra)
    p.update(rough_std=15e-9, rough_acl=40e-9, seed=5)
    s, up, lo, fg = device_inputs(p)
    o, _ = oracle_run(p, up, lo, direction)
    print('case', scheme, 'xchunk', xc, direction, extra, flush=True)
    b = Batch(scheme, np.float64, max_slots=1)
    b.set_launch(xc, 0, 3)
    slot = b.add(s, fg, direction)
    b.advance(-(-s.nt // 32))
    b.sync()
    assert b.poll(slot) == s.nt
    for name, ref in o.fields().items():
        err = rel_l2(b.field(slot, name), ref)
        assert err < 1e-10, (name, err)
    w = window(scheme, o.c)
    od = o.dft()
    for key, z in b.dft(slot).items():
        err = rel_l2(z[w], od[key][w])
        assert err < 1e-10, (key, err)
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize('scheme,tb,direction,extra', [('tm', 2, 'b', {'nt': 101}), ('tm', 4, 'f', {'nt': 101}),
                                                       ('tm', 8, 'f', {}), ('tm', 8, 'b', {'nt': 99}),
                                                       ('te', 2, 'f', {'nt': 101}), ('te', 4, 'b', {}),
                                                       ('te', 8, 'b', {'nt': 99}),
                                                       ('3d', 2, 'f', {'nt': 101}), ('3d', 2, 'b', {}),
                                                       ('3d', 2, 'f', {'source_condition': 'TM', 'nt': 99})])
def test_blocked_step_fp64(scheme, tb, direction, extra):
    """The temporally blocked step (tb steps per launch) in FP64 equals
    the oracle to 1e-10, also when the step count is not a multiple of tb (the
    last launch takes fewer steps and the state ends in the other field set)."""
    p = configs.small(scheme, 'float64', **extra)
    p.update(rough_std=15e-9, rough_acl=40e-9, seed=5)
    s, up, lo, fg = device_inputs(p)
    o, _ = oracle_run(p, up, lo, direction)
    print('case blocked', scheme, 'fp64 tb', tb, direction, extra, flush=True)
    b = Batch(scheme, np.float64, max_slots=1)
    b.set_launch(tb, 0, 4)
    slot = b.add(s, fg, direction)
    b.advance(-(-s.nt // 32))
    b.sync()
    assert b.poll(slot) == s.nt
    for name, ref in o.fields().items():
        err = rel_l2(b.field(slot, name), ref)
        assert err < 1e-10, (name, err)
    w = window(scheme, o.c)
    od = o.dft()
    for key, z in b.dft(slot).items():
        err = rel_l2(z[w], od[key][w])
        assert err < 1e-10, (key, err)
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize('scheme,tb', [('3d', 2), ('tm', 8), ('te', 4)])
def test_blocked_step_incident_run_fp64(scheme, tb):
    """The protocol's incident run (smooth guide cut to the two port lengths,
    both port planes on the same cell, fdtd_auto_setup.py:82-91): the blocked
    step writes every sample plane of both ports from the one plane, and its
    fields and port DFT equal the oracle's to 1e-10."""
    inc = configs.incident(configs.small(scheme, 'float64', nt=101))
    assert inc['p1'] == inc['p2']
    s, up, lo, fg = device_inputs(inc)
    o, _ = oracle_run(inc, up, lo)
    print('case blocked incident', scheme, 'tb', tb, flush=True)
    b = Batch(scheme, np.float64, max_slots=1)
    b.set_launch(tb, 0, 4)
    slot = b.add(s, fg)
    b.advance(-(-s.nt // 32))
    b.sync()
    for name, ref in o.fields().items():
        err = rel_l2(b.field(slot, name), ref)
        assert err < 1e-10, (name, err)
    w = window(scheme, o.c)
    od = o.dft()
    for key, z in b.dft(slot).items():
        assert np.abs(od[key][w]).max() > 0, key
        err = rel_l2(z[w], od[key][w])
        assert err < 1e-10, (key, err)
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize('scheme,tb,direction', [('tm', 8, 'f'), ('tm', 4, 'b'), ('te', 8, 'b'), ('te', 2, 'f')])
def test_blocked_step_interior_tiles_fp64(scheme, tb, direction):
    """A grid wide enough (200 x 160) for tiles whose extended region lies
    wholly inside the CPML-free bulk: those take the blocked step's interior
    path; fields and port DFT still equal the oracle to 1e-10."""
    p = configs.slab2d(200, ny=160, nt=301, nf=4, precision='float64', scheme=scheme, core=12, port_len=30,
                       source_type=1, gauss_pulse_deg=-1.0)
    p.update(rough_std=15e-9, rough_acl=40e-9, seed=9)
    s, up, lo, fg = device_inputs(p)
    o, _ = oracle_run(p, up, lo, direction)
    print('case blocked interior', scheme, 'tb', tb, direction, flush=True)
    b = Batch(scheme, np.float64, max_slots=1)
    b.set_launch(tb, 0, 4)
    slot = b.add(s, fg, direction)
    b.advance(-(-s.nt // 32))
    b.sync()
    for name, ref in o.fields().items():
        err = rel_l2(b.field(slot, name), ref)
        assert err < 1e-10, (name, err)
    w = window(scheme, o.c)
    od = o.dft()
    for key, z in b.dft(slot).items():
        err = rel_l2(z[w], od[key][w])
        assert err < 1e-10, (key, err)
    b.close()


@pytest.mark.gpu
@pytest.mark.parametrize('scheme', ['tm', 'te'])
@pytest.mark.parametrize('precision', ['float32', 'float64'])
@pytest.mark.parametrize('tb', [2, 4, 8])
def test_blocked_step_matches_fused(scheme, precision, tb):
    """The blocked step applies the fused step's per-cell arithmetic: on an
    uneven pair of runs in one batch (tile edges fall at different cells) its
    fields and port DFT equal the one-step fused form's to rounding (the two
    kernels may contract multiply-adds differently): relative L2 1e-12 in
    FP64, 2e-6 in FP32."""
    tol = 1e-12 if precision == 'float64' else 2e-6
    runs = []
    for nt, nx, seed in ((101, 140, 5), (96, 181, 6)):
        p = configs.small(scheme, precision, nt=nt, nx=nx)
        p.update(rough_std=15e-9, rough_acl=40e-9, seed=seed)
        s, up, lo, fg = device_inputs(p)
        runs.append((s, fg))
    out = {}
    for form in ((3, 2), (4, tb)):
        print('case blocked vs fused', scheme, precision, 'form', form, flush=True)
        b = Batch(scheme, np.float32 if precision == 'float32' else np.float64, max_slots=2)
        b.set_launch(form[1], 0, form[0])
        slots = [b.add(s, fg) for s, fg in runs]
        b.advance(-(-max(s.nt for s, _ in runs) // 32))
        b.sync()
        names = ('ez', 'hx', 'hy') if scheme == 'tm' else ('ex', 'ey', 'hz')
        out[form] = [({n: b.field(sl, n) for n in names}, b.dft(sl)) for sl in slots]
        b.close()
    for (fa, da), (fb, dbb) in zip(out[(3, 2)], out[(4, tb)]):
        errs = {n: rel_l2(fb[n], fa[n]) for n in fa}
        errs.update({k: rel_l2(dbb[k], da[k]) for k in da})
        print({k: '%.2e' % e for k, e in errs.items()}, flush=True)
        assert max(errs.values()) < tol, errs


@pytest.mark.gpu
@pytest.mark.parametrize('precision,tb', [('float32', 2), ('float64', 2)])
def test_blocked_3d_matches_march(precision, tb):
    """The blocked 3D step applies the march's per-cell arithmetic: on an
    uneven pair of runs in one batch (plane-segment, row and z-tile edges
    fall at different cells; 70 planes make three segments) its fields and
    port DFT equal the two-kernel march's to rounding: relative L2 1e-12 in
    FP64, 2e-6 in FP32."""
    tol = 1e-12 if precision == 'float64' else 2e-6
    runs = []
    for nz, nx, ny, nt, seed in ((56, 36, 34, 101, 5), (90, 70, 40, 96, 6)):
        p = configs.strip3d(nz, nx=nx, ny=ny, nt=nt, nf=6, precision=precision, width=12, height=8, port_len=10,
                            buffer_z=6, source_type=1, gauss_pulse_deg=-1.0)
        p.update(rough_std=15e-9, rough_acl=40e-9, seed=seed)
        s, up, lo, fg = device_inputs(p)
        runs.append((s, fg))
    out = {}
    for form in ((1, 4), (4, tb)):
        print('case blocked 3d vs march', precision, 'form', form, flush=True)
        b = Batch('3d', np.float32 if precision == 'float32' else np.float64, max_slots=2)
        b.set_launch(form[1], 0, form[0])
        slots = [b.add(s, fg) for s, fg in runs]
        b.advance(-(-max(s.nt for s, _ in runs) // 32))
        b.sync()
        names = ('ex', 'ey', 'ez', 'hx', 'hy', 'hz')
        out[form] = [({n: b.field(sl, n) for n in names}, b.dft(sl)) for sl in slots]
        b.close()
    for (fa, da), (fb, dbb) in zip(out[(1, 4)], out[(4, tb)]):
        errs = {n: rel_l2(fb[n], fa[n]) for n in fa}
        errs.update({k: rel_l2(dbb[k], da[k]) for k in da})
        print({k: '%.2e' % e for k, e in errs.items()}, flush=True)
        assert max(errs.values()) < tol, errs


@pytest.mark.gpu
@pytest.mark.parametrize('precision,variant,xc,direction,extra', [
    ('float64', 2, 1, 'f', {}),                                # runs [64,128) and [128,192) outside the slabs
    ('float64', 2, 2, 'b', {'source_condition': 'TM'}),
    ('float64', 2, 1, 'f', {'sz': 70}),                        # source slab inside the first run
    ('float32', 2, 4, 'f', {'p1': 63, 'p2': 127}),             # H samples planes 64 and 128
    ('float32', 2, 1, 'b', {}),
    ('float64', 1, 4, 'f', {'sz': 70}),                        # the march form
    ('float64', 1, 8, 'b', {'source_condition': 'TM'}),
    ('float32', 1, 16, 'f', {'p1': 63, 'p2': 127}),
    ('float64', 3, 4, 'f', {'sz': 70}),                        # the fused step: 63-cell z runs
    ('float64', 3, 8, 'b', {'source_condition': 'TM'}),
    ('float32', 3, 2, 'f', {'p1': 62, 'p2': 125}),             # H samples at the run boundaries 63 and 126
    ('float32', 3, 16, 'b', {'nz': 300, 'p2': 270}),
    ('float64', 3, 1, 'f', {'nz': 300, 'p2': 270}),
    ('float64', 4, 2, 'f', {'sz': 70}),                        # the blocked step: 60-cell inner z runs
    ('float64', 4, 2, 'b', {'source_condition': 'TM', 'p1': 59, 'p2': 119}),   # H samples at 60 and 120
    ('float32', 4, 2, 'f', {'nz': 300, 'p2': 270}),
    ('float32', 4, 2, 'f', {'p1': 59, 'p2': 119}),
])
def test_3d_long_z_runs(precision, variant, xc, direction, extra):
    """A guide long enough in z (nz = 250) for four 64-cell z runs per row:
    runs wholly outside the z-CPML slabs, runs that hold a source slab or a
    port plane (E samples plane p, H samples p and p+1, here on both sides of
    a run boundary), and the partial last run all give the oracle's fields and
    port DFT in the two-kernel 3D forms, the fused step (whose waves own
    63 z-cells each and compute the cell below their run) and the blocked
    step (60 inner z-cells per tile)."""
    p = configs.strip3d(250, nx=36, ny=34, nt=200, nf=6, precision=precision, width=12, height=8, port_len=10,
                        buffer_z=6, source_type=1, gauss_pulse_deg=-1.0)
    p.update(extra)
    p.update(rough_std=15e-9, rough_acl=40e-9, seed=3)
    s, up, lo, fg = device_inputs(p)
    o, _ = oracle_run(p, up, lo, direction)
    print('case', precision, 'variant', variant, 'xchunk', xc, direction, extra, flush=True)
    tol = 1e-10 if precision == 'float64' else 1e-4
    b = Batch('3d', np.dtype(precision).type, max_slots=1)
    b.set_launch(xc, 0, variant)
    slot = b.add(s, fg, direction)
    b.advance(-(-s.nt // 32))
    b.sync()
    assert b.poll(slot) == s.nt
    for name, ref in o.fields().items():
        err = rel_l2(b.field(slot, name), ref)
        assert err < tol, (name, err)
    w = window('3d', o.c)
    od = o.dft()
    for key, z in b.dft(slot).items():
        err = rel_l2(z[w], od[key][w])
        assert err < tol, (key, err)
    b.close()


@pytest.mark.gpu
def test_launch_form_validation():
    """Tile depths outside a form's set are rejected, also when kept from an
    earlier call."""
    from oidt_amd._lib import OidtError
    b = Batch('tm', np.float32, max_slots=1)
    try:
        b.set_launch(12, 0, 3)
        assert b.launch()[:2] == (3, 12)
        with pytest.raises(OidtError):
            b.set_launch(0, 0, 2)       # keeps xchunk 12, which the two-kernel forms do not take
        with pytest.raises(OidtError):
            b.set_launch(12, 0, 1)
        b.set_launch(4, 0, 2)
        assert b.launch()[:2] == (2, 4)
    finally:
        b.close()
    t = Batch('te', np.float32, max_slots=1)
    try:
        t.set_launch(20, 0, 3)
        with pytest.raises(OidtError):
            t.set_launch(0, 0, 1)
    finally:
        t.close()
    d = Batch('3d', np.float32, max_slots=1)
    try:
        d.set_launch(16, 0, 3)
        for xc in (12, 20, 32):         # the fused 3D step takes 1, 2, 4, 8 or 16 planes
            with pytest.raises(OidtError):
                d.set_launch(xc, 0, 3)
        assert d.launch()[:2] == (3, 16)
        with pytest.raises(OidtError):
            d.set_launch(0, 0, 4)       # the blocked 3D step takes 2 steps per launch
        d.set_launch(2, 0, 4)
        assert d.launch()[:2] == (4, 2)
        for xc in (1, 4, 8):
            with pytest.raises(OidtError):
                d.set_launch(xc, 0, 4)
    finally:
        d.close()
    # the blocked TMz step takes 2, 4 or 8 steps per launch
    q = Batch('tm', np.float64, max_slots=1)
    try:
        q.set_launch(8, 0, 4)
        for xc in (1, 16):
            with pytest.raises(OidtError):
                q.set_launch(xc, 0, 4)
        q.set_launch(4, 0, 2)
    finally:
        q.close()


@pytest.mark.gpu
def test_default_launch_forms():
    """A new batch starts in its scheme's default form (3D: the march over
    4-plane tiles; 2D: batched loads, 4 cells per thread), and 0 keeps any
    knob as it is."""
    d = Batch('3d', np.float32, max_slots=1)
    t = Batch('tm', np.float32, max_slots=1)
    try:
        assert d.launch() == (1, 4, 0)
        assert t.launch() == (2, 4, 0)
        d.set_launch(0, 0, 0)
        assert d.launch() == (1, 4, 0)
        d.set_launch(2, 0, 0)
        assert d.launch() == (1, 2, 0)
        d.set_launch(0, 4096, 2)
        assert d.launch() == (2, 2, 4096)
    finally:
        d.close()
        t.close()

