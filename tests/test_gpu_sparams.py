"""S-parameters from the device path match the FP64 oracle.

The reference protocol for one line: forward/backward incident runs on the
shortened guide and forward/backward runs on the full rough guide
(simulation_launcher.py:26-35), phasors extracted to S = sqrt(Pref/Pinc)
(extract_sparams.py:52-104).  The device runs all four as one uneven batch.
Bar: |S| within 0.01 dB and angle within 0.1 degree of the
oracle for S11, S21, S12, S22 (FP32 fields, FP64 DFT and normalisation)."""
import numpy as np
import pytest

from oidt_amd import configs, extract, materials, params
from oidt_amd.solver import Batch
from oracle import extract as oextract
from oracle import setup as osetup

from helpers import oracle_run

COMP_ORDER = {'tm': ('e', 'h'), 'te': ('e', 'h'), '3d': ('ex', 'ey', 'hx', 'hy')}
FILE_COMP = {'tm': ('ez', 'hy'), 'te': ('ey', 'hz'), '3d': ('ex', 'ey', 'hx', 'hy')}


def protocol(p):
    return {('f', 'i'): configs.incident(p), ('b', 'i'): configs.incident(p), ('f', 'r'): p, ('b', 'r'): p}


def files_from(scheme, dfts):
    """{(tag, line): component tuple} from {(direction, kind): dft dict}."""
    files = {}
    for (d, k), dft in dfts.items():
        for port, tag in extract.run_tags(d, k).items():
            files[(tag, 1)] = tuple(dft[(key, port)] for key in COMP_ORDER[scheme])
    return files


def db_phase_err(a, b):
    dbe = np.abs(20 * np.log10(np.abs(a)) - 20 * np.log10(np.abs(b)))
    ph = np.abs(np.angle(a / b, deg=True))
    return dbe.max(), ph.max()


@pytest.mark.gpu
@pytest.mark.parametrize('scheme,precision', [('tm', 'float32'), ('te', 'float32'), ('3d', 'float32'),
                                              ('tm', 'float64')])
def test_sparams_match_oracle(scheme, precision):
    p = configs.small(scheme, precision)
    p.update(rough_std=15e-9, rough_acl=40e-9, seed=5)
    runs = protocol(p)
    print('case', scheme, precision, flush=True)
    # device: all four runs in one batch
    setups = {}
    b = Batch(scheme, np.float32 if precision == 'float32' else np.float64, max_slots=4)
    slots = {}
    for (d, k), q in runs.items():
        s, up, lo = configs.realise(q)
        setups[(d, k)] = s
        slots[(d, k)] = b.add(s, materials.fg_mask(s, up, lo), d, 0)
    b.advance(-(-p['nt'] // 32))
    b.sync()
    dev = {key: b.dft(slot) for key, slot in slots.items()}
    b.close()
    nf = p['nf']
    files = files_from(scheme, dev)
    S_dev = extract.s_matrix(scheme, setups[('f', 'r')], lambda c, t, l: files[(t, l)][FILE_COMP[scheme].index(c)],
                             1, nf)
    # oracle: four sequential FP64 runs
    odft = {}
    for (d, k), q in runs.items():
        s, up, lo = configs.realise(q)
        o, _ = oracle_run(q, up, lo, d)
        odft[(d, k)] = o.dft()
    c = osetup.explicit_setup(p)
    S_ora = oextract.s_matrix(scheme, c, files_from(scheme, odft), 1, nf)
    for (i, j) in ((0, 0), (1, 0), (0, 1), (1, 1)):
        e_db, e_ph = db_phase_err(S_dev[:, i, j], S_ora[:, i, j])
        print('S%d%d |S|=%.4f err %.2e dB %.2e deg' % (i + 1, j + 1, np.abs(S_ora[:, i, j]).mean(), e_db, e_ph),
              flush=True)
        assert e_db <= 0.01 and e_ph <= 0.1, (i, j, e_db, e_ph)
    # the same extraction restated twice agrees to rounding
    S_dev_o = oextract.s_matrix(scheme, c, files, 1, nf)
    assert np.allclose(S_dev_o, S_dev, rtol=1e-12, atol=0)
