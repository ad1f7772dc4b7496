"""Setup arithmetic: the oracle against independently derived integers, and the
product host code against the oracle (bit for bit)."""
import json
import os

import numpy as np
import pytest

from oidt_amd import configs, materials, params
from oidt_amd.solver import RunArrays
from oracle import geometry as ogeo
from oracle import setup as osetup

HERE = os.path.dirname(os.path.abspath(__file__))
USR = json.load(open(os.path.join(HERE, 'golden', 'reference_user_configs.json')))

# Grid integers of the reference defaults (user_config.py of each body), derived
# by hand from fdtd_auto_setup.py with numpy-1.20 promotion: there a Python float
# divided by an np.float32 scalar is a float64 division, so int(length / dx) uses
# the float32-rounded dx in float64 (numpy 2 would divide in float32).
GOLDEN = {
    ('tm', 'r'): dict(nx=1480, ny=430, nt=9239, p1=506, p2=974),
    ('tm', 'i'): dict(nx=1013, ny=430, nt=9239, p1=506, p2=507),
    ('te', 'r'): dict(nx=1480, ny=244, nt=9239),
    ('3d', 'r'): dict(nx=291, ny=226, nz=1480, nt=11316, p1=226, p2=1254),
    ('3d', 'i'): dict(nz=453, nt=11316),
}


def usr(scheme, nf=7):
    u = dict(USR[scheme])
    u['precision'] = np.float32
    u['sparam_num_freqs'] = nf
    return u


@pytest.mark.parametrize('key', sorted(GOLDEN))
def test_reference_default_grid_integers(key):
    scheme, kind = key
    for c in (osetup.auto_setup(usr(scheme), scheme, kind), params.auto_setup(usr(scheme), scheme, kind)):
        g = GOLDEN[key]
        got = dict(nx=c.nx, ny=c.ny, nz=getattr(c, 'nz', None), nt=c.nt,
                   p1=c.p1z if scheme == '3d' else c.p1x, p2=c.p2z if scheme == '3d' else c.p2x)
        for k, v in g.items():
            assert got[k] == v, (key, k, got[k], v)


def test_numpy1_promotion_matters_for_grid_integers():
    """float32 dx with numpy-2 float32 division gives a different cell count for
    some lengths; the setup must use float64 division (numpy 1.20)."""
    c = osetup.auto_setup(usr('tm'), 'tm', 'r')
    dx32 = np.float32(c.delta_x)
    L = USR['tm']['sgl_wg_length']
    assert int(L / float(dx32)) == c.nx_swg
    # the float32 quotient differs from the float64 one (so the choice is observable)
    assert float(np.float32(L) / dx32) != L / float(dx32)


def test_dt_on_cladding_cfl_and_rounding():
    c = osetup.auto_setup(usr('tm'), 'tm', 'r')
    # dt = dx / (c_bg sqrt 2) before rounding (fdtd_auto_setup.py:74), both rounded to float32
    assert c.delta_t.dtype == np.float32 and c.delta_x.dtype == np.float32
    assert abs(float(c.delta_t) * float(c.c_bg) * np.sqrt(2) / float(c.delta_x) - 1) < 1e-6


def test_cpml_coefficient_properties():
    c = osetup.auto_setup(usr('tm'), 'tm', 'r')
    g = osetup.cpml_grading(20, 4, 1, 5.0, 0.05, 1.2, c.eta, c.delta_x, c.delta_t, c.eps, np.float64)
    BE, BH, KE, KH, CE, CH = (g[k] for k in ('BE', 'BH', 'KE', 'KH', 'CE', 'CH'))
    assert np.all((BE > 0) & (BE < 1)) and np.all((BH > 0) & (BH < 1))
    assert KH[0] == 1.0 and np.all(np.diff(KE) > 0) and KE[-1] < 5.0
    # sigma_H[0] = 0 -> c_H[0] = 0; b decreases into the layer (stronger absorption)
    assert CH[0] == 0.0 and np.all(np.diff(BE) < 0)
    # closed form for alpha = 0 limit is not applicable (alpha>0); check c < 0 and the
    # analytic relation c = sigma (b - 1) / (sigma kappa + kappa^2 alpha)
    assert np.all(CE < 0)


@pytest.mark.parametrize('scheme', ['tm', 'te', '3d'])
@pytest.mark.parametrize('precision', ['float32', 'float64'])
def test_product_host_arrays_equal_oracle(scheme, precision):
    p = configs.small(scheme, precision)
    p.update(rough_std=15e-9, rough_acl=40e-9, seed=3)
    s, up, lo = configs.realise(p)
    fg = materials.fg_mask(s, up, lo)
    ra = RunArrays(s, fg)
    c = osetup.explicit_setup(p)
    FG = ogeo.fg_mask(c, up, lo)
    osetup.coefficients(c)
    assert np.array_equal(fg, FG)
    # vectorised vs scalar exp: last-bit differences only
    np.testing.assert_allclose(s.J_SRC, c.J_SRC, rtol=1e-14, atol=1e-300)
    for k in ('BE', 'CE', 'BH', 'CH', 'KE', 'KH'):
        np.testing.assert_array_max_ulp(ra.grading[k], getattr(c, k), maxulp=1)
    axes = ('x', 'y', 'z') if scheme == '3d' else ('x', 'y')
    for a, (ke, kh) in zip(axes, ra.kappa):
        np.testing.assert_array_max_ulp(ke, getattr(c, 'k%se' % a), maxulp=1)
        np.testing.assert_array_max_ulp(kh, getattr(c, 'k%sh' % a), maxulp=1)
    for got, ref in zip(ra.cb, ogeo.cb_maps(c, FG)):
        assert np.array_equal(got, ref)
    assert float(ra.db) == float(c.DB)
    assert ra.db.dtype == c.DB.dtype


def test_source_waveform_types():
    for st in (1, 2, 3):
        J, nt = params.source_waveform(np.float64, 400, 3e-17, st, 5.0, 2e14, 0.5, -6)
        Jo, nto = osetup.waveforms(np.float64, 400, 3e-17, st, 5.0, 2e14, 0.5, -6)
        assert nt == nto
        np.testing.assert_allclose(J, Jo, rtol=1e-12, atol=1e-15)
    # wave packet: |J| peaks at tpk = 9 tsp, amplitude source_amp
    J, nt = params.source_waveform(np.float64, 4000, 3e-17, 2, 5.0, 2e14, 0.5)
    tsp = 2 * np.sqrt(np.log(2)) / (0.5 * 2 * np.pi * 2e14)
    assert abs(np.argmax(np.abs(J)) * 3e-17 - 9 * tsp) < 3 / 2e14
    assert abs(np.abs(J).max() - 5.0) < 0.05
    # the ramp extension nt += int(tpk/nt) is a no-op for types 1, 2 (fdtd_auto_setup.py:130-133)
    assert nt == 4000
    with pytest.raises(Exception):
        params.source_waveform(np.float64, 10, 3e-17, 4, 5.0, 2e14, 0.5)


def test_setup_guards_raise():
    u = usr('tm')
    u['port1_length'] = u['port2_length'] = 9e-6
    with pytest.raises(Exception, match='Port settling ranges overlap'):
        params.auto_setup(u, 'tm', 'r')
    with pytest.raises(Exception, match='Port settling ranges overlap'):
        osetup.auto_setup(u, 'tm', 'r')
    u = usr('tm')
    u['pitch'] = 20e-6
    with pytest.raises(Exception, match='Boundary conditions too small'):
        params.auto_setup(u, 'tm', 'r')
