"""Roughness ensembles with shared incident runs (oidt_amd.ensemble).

The incident run of the reference protocol (a smooth, shortened guide,
FDTD/GPU/FDTD_2D_TM/s-parameter_extraction/fdtd_auto_setup.py:82-91) does not
depend on the roughness, so an ensemble simulates it once per distinct
incident description.  S from the shared incident run equals S with one
incident run per realisation bit for bit, and equals the FP64 oracle's
(S21, S11 within 0.01 dB and 0.1 degree)."""
import numpy as np
import pytest

from oidt_amd import configs, ensemble
from oracle import extract as oextract
from oracle import setup as osetup

from helpers import oracle_run


def _descs(scheme, count=3):
    out = []
    for r in range(count):
        p = configs.small(scheme, 'float32')
        if scheme == '3d':
            p['nz'] += 8 * r
            p['p2'] += 8 * r
        else:
            p['nx'] += 20 * r
            p['p2'] += 20 * r
        p.update(rough_std=(10 + 3 * r) * 1e-9, rough_acl=40e-9, seed=50 + r)
        out.append(p)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize('scheme', ['tm', 'te', '3d'])
def test_shared_incident_equals_one_per_realisation(scheme):
    descs = _descs(scheme)
    shared = ensemble.Ensemble(descs, directions=('f', 'b')).run()
    alone = ensemble.Ensemble(descs, directions=('f', 'b'), share_incident=False).run()
    try:
        assert len(shared.incident) == 1 and len(alone.incident) == len(descs)
        assert shared.cell_updates < alone.cell_updates == shared.cell_updates_unshared()
        for r in range(len(descs)):
            a, b = shared.s_params(r), alone.s_params(r)
            for k in ('S11', 'S21', 'S22', 'S12'):
                assert np.array_equal(a[k], b[k]), (r, k)
        print(scheme, 'cell-updates saved by sharing the incident run:',
              shared.cell_updates_unshared() - shared.cell_updates, flush=True)
    finally:
        shared.close()
        alone.close()


@pytest.mark.gpu
@pytest.mark.parametrize('scheme', ['tm', '3d'])
def test_ensemble_s_params_match_oracle(scheme):
    """Forward S21 and S11 of each realisation against the oracle's incident
    and rough runs through the oracle's extraction (extract_sparams.py:52-104,
    3D :52-144)."""
    descs = _descs(scheme)
    e = ensemble.Ensemble(descs).run()
    try:
        dev = [e.s_params(r) for r in range(len(descs))]
    finally:
        e.close()
    inc = configs.incident(descs[0])
    s, up, lo = configs.realise(inc)
    oi, _ = oracle_run(inc, up, lo)
    keys = ('ex', 'ey', 'hx', 'hy') if scheme == '3d' else ('e', 'h')
    for r, p in enumerate(descs):
        s, up, lo = configs.realise(p)
        o, _ = oracle_run(p, up, lo)
        fi = tuple(oi.dft()[(k, 1)] for k in keys)
        ft = tuple(o.dft()[(k, 1)] for k in keys)
        br = tuple(o.dft()[(k, 2)] for k in keys)
        # backward runs not simulated: their slots repeat the forward ones,
        # leaving S[0, 0] = S11 and S[1, 0] = S21 (3D; 2D mirrors S[0, 1])
        files = {('fi', 1): fi, ('bi', 1): fi, ('ft', 1): ft, ('bt', 1): br, ('br', 1): br, ('fr', 1): br}
        S = oextract.s_matrix(scheme, osetup.explicit_setup(p), files, 1, p['nf'])
        for name, got, ref in (('S11', dev[r]['S11'], S[:, 0, 0]), ('S21', dev[r]['S21'], S[:, 1, 0])):
            e_db, e_ph = ensemble.s_error(got, ref)
            print(scheme, r, name, '%.2e dB %.2e deg' % (e_db, e_ph), flush=True)
            assert e_db <= 0.01 and e_ph <= 0.1, (r, name, e_db, e_ph)
