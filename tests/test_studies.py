"""Loss and wave-impedance study drivers: host logic and the oracle, on the CPU.

* The study configuration shares the S-parameter driver's derivation of dx,
  dt and the grid (FDTD_2D_TE/scattering_loss/fdtd_config.py:95-133 against
  FDTD_2D_TE/s-parameter_extraction/fdtd_auto_setup.py:59-119), so it is pinned
  against ``params.auto_setup`` on the same physical parameters.
* The TEz study loop (H first, +J) is the S-parameter TEz loop (E first, -J)
  shifted by half a step: the FP64 study oracle and the FP64 S-parameter
  oracle, stepped one step longer with the source negated, agree to rounding.
  This is the identity the device driver relies on.
"""
import numpy as np
import pytest

from oidt_amd import params, variants
from oracle import fdtd as ofdtd
from oracle import studies as ost

SMALL_TE = dict(sgl_wg_length=2.0e-6, port1_length=0.5e-6, points_per_wl=16, ny_clad_wl=0.5, buffer_xhat=10,
                num_cpml=8)
SMALL_3D = dict(sgl_wg_length=1.5e-6, port1_length=0.3e-6, points_per_wl=10, sgl_wg_width=0.5e-6, sgl_wg_height=0.3e-6,
                nx_clad_wl=0.3, ny_clad_wl=0.3, buffer_zhat=10, num_cpml=8)


def _usr_like(c, scheme):
    """A reference user dictionary with the study's physical parameters."""
    usr = dict(precision=np.float32, eps_rel_bg=c.eps_rel_bg, eps_rel_fg=c.eps_rel_fg * c.eps_rel_bg, f0=c.f0,
               harmonics=c.harmonics, sgl_wg_length=c.sgl_wg_length, points_per_wl=c.points_per_wl,
               num_flights=c.num_flights, pitch=0.0, num_cpml=c.num_cpml, num_lines=1, sparam_fmin=c.f0,
               sparam_fmax=c.f0, sparam_num_freqs=1, rough_std=c.rough_std, rough_acl=c.rough_acl,
               source_type=c.source_type, source_amp=c.source_amp, wave_packet_bw=c.wave_packet_bw,
               gauss_pulse_deg=c.gauss_pulse_deg, port1_length=c.port1_length, port2_length=c.port2_length,
               auto_condition=False)
    if scheme == 'te':
        usr.update(ny_clad_wl=c.ny_clad_wl, sgl_wg_height=c.sgl_wg_width, buffer_xhat=c.buffer_xhat)
    else:
        usr.update(nx_clad_wl=c.nx_clad_wl, ny_clad_wl=c.ny_clad_wl, sgl_wg_height=c.sgl_wg_height,
                   sgl_wg_width=c.sgl_wg_width, buffer_zhat=c.buffer_zhat, source_condition=c.source_condition)
    return usr


@pytest.mark.parametrize('scheme,study', sorted(variants.DEFAULTS))
def test_study_config_matches_sparameter_derivation(scheme, study):
    c = variants.config(scheme, study)
    if scheme == 'tm':
        # same grid arithmetic as the TM S-parameter setup; its own waveform
        # (peak at 5 spreads) and port rule p2x = nx - p1x
        usr = _usr_like(variants.config('te', 'scattering_loss'), 'te')
        usr.update(sgl_wg_length=c.sgl_wg_length, sgl_wg_height=c.sgl_wg_width, port1_length=c.port_length,
                   port2_length=c.port_length, buffer_xhat=c.buffer_xhat, ny_clad_wl=c.ny_clad_wl,
                   source_amp=c.source_amp, wave_packet_bw=c.wave_packet_bw)
        s = params.auto_setup(usr, 'tm', 'r')
        assert (c.delta_x, c.delta_t, c.nx, c.ny, c.bx, c.by, c.ny_swg, c.nt) == \
            (s.delta_x, s.delta_t, s.nx, s.ny, s.bx, s.by, s.ny_swg, s.nt)
        assert (c.sx, c.p1x, c.p2x) == (s.sx, s.p1x, s.p2x)
        t = np.arange(c.nt) * float(c.delta_t)
        tsp = np.sqrt(-2 * np.log(10 ** (c.gauss_pulse_deg / 20))) / (2 * np.pi * c.f0)
        assert np.argmax(c.J_SRC) == np.argmin(np.abs(t - 5 * tsp))
        return
    s = params.auto_setup(_usr_like(c, scheme), scheme, 'r')
    assert (c.delta_x, c.delta_t) == (s.delta_x, s.delta_t)
    assert (c.nx, c.ny, c.bx, c.by, c.ny_swg) == (s.nx, s.ny, s.bx, s.by, s.ny_swg)
    # source type 1: the reference's extension int(gp_tpk / nt) is 0 at these nt
    assert c.nt == s.nt
    np.testing.assert_array_equal(c.J_SRC, s.J_SRC)
    if scheme == 'te':
        assert (c.sx, c.p1x, c.p2x) == (s.sx + 10, s.p1x + 10, s.p2x + 10)   # source 10 cells in (:132)
    else:
        assert (c.nz, c.bz, c.sz, c.p1z, c.p2z) == (s.nz, s.bz, s.sz, s.p1z, s.p2z)


def test_default_study_sizes():
    """Grid sizes of the shipped study configurations (cells, steps)."""
    got = {k: (variants.config(*k).nx, variants.config(*k).ny, getattr(variants.config(*k), 'nz', 1),
               variants.config(*k).nt) for k in variants.DEFAULTS}
    assert got[('te', 'scattering_loss')] == (4170, 430, 1, 26997)
    assert got[('3d', 'scattering_loss')] == (157, 280, 1716, 13226)


def test_te_material_maps_match_literal_loop():
    c = variants.config('te', 'scattering_loss', **SMALL_TE)
    s, fg, cb = variants.build(c)
    ref = ost.te_cb_maps(ost.namespace(c), fg)
    for a, b in zip(cb, ref):
        np.testing.assert_array_equal(a, b)
    assert len(np.unique(cb[0])) == 3       # cladding, core and their average


def test_phase_tables_follow_float32_evaluation():
    c = variants.config('3d', 'scattering_loss', **SMALL_3D)
    for n in (0, 1, 77, c.nt - 1):
        ph = np.float32(2 * np.pi * c.f0 * (n + 0.5) * float(c.delta_t))
        assert c.cosH[n] == np.cos(ph, dtype=np.float32) and c.sinH[n] == np.sin(ph, dtype=np.float32)
    # the float32 phase loses about 1e-4 rad here: the tables differ from float64 cos
    n = np.arange(c.nt)
    exact = np.cos(2 * np.pi * c.f0 * n * float(c.delta_t))
    assert 0 < np.abs(c.cosE - exact).max() < 1e-3


def test_te_study_oracle_is_the_shifted_sparameter_loop():
    c = variants.config('te', 'scattering_loss', **SMALL_TE)
    c.nt = 150
    up, lo = variants._profiles(c, c.p2x - c.p1x)
    st = ost.TEzStudy(c, up, lo)
    for n in range(c.nt):
        st.step(n)
    # S-parameter oracle on the same arrays, source negated and padded
    o_ns = ost.namespace(c)
    o_ns.nt = c.nt + 1
    o_ns.J_SRC = np.append(-np.asarray(c.J_SRC[:c.nt], np.float32), np.float32(0))
    o = ofdtd.TEz(o_ns, st.FG)
    o.cb = [st.cbx, st.cby]
    for n in range(c.nt):
        o.step(n)
    hz = o.Hz.copy()
    o.step(c.nt)
    assert np.abs(st.Hz).max() > 0 and np.abs(st.Ey).max() > 0
    for a, b in ((st.Hz, hz), (st.Ex, o.Ex), (st.Ey, o.Ey)):
        err = np.linalg.norm(a - b) / np.linalg.norm(b)
        assert err < 1e-12, err


def test_study_monitors_shapes():
    for k in variants.DEFAULTS:
        c = variants.config(*k, **{"te": SMALL_TE, "3d": SMALL_3D, "tm": SMALL_TM}[k[0]])
        mons = variants.monitors(c)
        names = variants.output_names(c)
        assert set(mons) == set(names)
        for m in mons.values():
            assert m.nrec == c.nt
            assert (m.cs is None) == (k[1] == 'wave_impedance')
            assert m.after_h == (k[0] != 'te')


# ---- lossy TMz studies -------------------------------------------------------

SMALL_TM = dict(sgl_wg_length=1.5e-6, port_length=0.3e-6, points_per_wl=16, ny_clad_wl=0.4, buffer_xhat=5,
                num_cpml=6)


def _loops_tm_step(o, n):
    """The TM study kernels transcribed cell by cell (tiny grids only):
    FDTD_2D_TM/scattering_loss/fdtd_funcs.py:77-152 in the launch order of
    fdtd_main.py:197-213."""
    c, k, R, dx = o.cfg, o.k, o.R, o.dx
    Ez, Hx, Hy = o.Ez, o.Hx, o.Hy
    ie, je = c.nx, c.ny
    for i in range(ie):
        for j in range(je):
            if 0 < i < ie - 1 and 0 < j < je - 1:
                Ez[i, j] = k['MOD_E'][i, j] * Ez[i, j] + k['CURL_H'][i, j] * (
                    (Hy[i, j] - Hy[i - 1, j]) / k['DEN_EX'][i] + (Hx[i, j - 1] - Hx[i, j]) / k['DEN_EY'][j])
    for i in range(c.sx, c.sx + 1):
        for j in range(c.by, c.ny - c.by):
            Ez[i, j] = Ez[i, j] - o.J[n]
    be, ce, bh, ch = k['BE'], k['CE'], k['BH'], k['CH']
    for i in range(1, R):
        for j in range(1, je - 1):
            o.p_ez_xlo[i, j] = be[i] * o.p_ez_xlo[i, j] + ce[i] / dx * (Hy[i, j] - Hy[i - 1, j])
            Ez[i, j] = Ez[i, j] + k['CURL_H'][i, j] * dx * o.p_ez_xlo[i, j]
            o.p_ez_xhi[i, j] = be[i] * o.p_ez_xhi[i, j] + ce[i] / dx * (Hy[ie - 1 - i, j] - Hy[ie - 2 - i, j])
            Ez[ie - 1 - i, j] = Ez[ie - 1 - i, j] + k['CURL_H'][ie - 1 - i, j] * dx * o.p_ez_xhi[i, j]
    for i in range(1, ie - 1):
        for j in range(1, R):
            o.p_ez_ylo[i, j] = be[j] * o.p_ez_ylo[i, j] + ce[j] / dx * (Hx[i, j - 1] - Hx[i, j])
            Ez[i, j] = Ez[i, j] + k['CURL_H'][i, j] * dx * o.p_ez_ylo[i, j]
            o.p_ez_yhi[i, j] = be[j] * o.p_ez_yhi[i, j] + ce[j] / dx * (Hx[i, je - 2 - j] - Hx[i, je - 1 - j])
            Ez[i, je - 1 - j] = Ez[i, je - 1 - j] + k['CURL_H'][i, je - 1 - j] * dx * o.p_ez_yhi[i, j]
    for i in range(ie - 1):
        for j in range(je - 1):
            Hx[i, j] = Hx[i, j] + k['CURL_E'] * (Ez[i, j] - Ez[i, j + 1]) / k['DEN_HY'][j]
            Hy[i, j] = Hy[i, j] + k['CURL_E'] * (Ez[i + 1, j] - Ez[i, j]) / k['DEN_HX'][i]
    dtmu = float(np.float32(o.dt) / np.float32(o.mu))
    for i in range(1, R):
        for j in range(je - 1):
            o.p_hy_xlo[i, j] = bh[i] * o.p_hy_xlo[i, j] + ch[i] / dx * (Ez[i + 1, j] - Ez[i, j])
            Hy[i, j] = Hy[i, j] + dtmu * o.p_hy_xlo[i, j]
            o.p_hy_xhi[i, j] = bh[i] * o.p_hy_xhi[i, j] + ch[i] / dx * (Ez[ie - 1 - i, j] - Ez[ie - 2 - i, j])
            Hy[ie - 2 - i, j] = Hy[ie - 2 - i, j] + dtmu * o.p_hy_xhi[i, j]
    for i in range(ie - 1):
        for j in range(1, R):
            o.p_hx_ylo[i, j] = bh[j] * o.p_hx_ylo[i, j] + ch[j] / dx * (Ez[i, j] - Ez[i, j + 1])
            Hx[i, j] = Hx[i, j] + dtmu * o.p_hx_ylo[i, j]
            o.p_hx_yhi[i, j] = bh[j] * o.p_hx_yhi[i, j] + ch[j] / dx * (Ez[i, je - 2 - j] - Ez[i, je - 1 - j])
            Hx[i, je - 2 - j] = Hx[i, je - 2 - j] + dtmu * o.p_hx_yhi[i, j]


def test_lossy_tm_oracle_matches_cell_loops():
    c = variants.config('tm', 'scattering_loss', **SMALL_TM)
    up, lo = variants._profiles(c, c.p2x - c.p1x)
    a, b = ost.TMzLossyStudy(c, up, lo), ost.TMzLossyStudy(c, up, lo)
    for n in range(40):
        a.step(n)
        _loops_tm_step(b, n)
    assert np.abs(a.Ez).max() > 0 and np.abs(a.Hy[1:c.cpml_range]).max() > 0
    for f in ('Ez', 'Hx', 'Hy'):
        x, y = getattr(a, f), getattr(b, f)
        assert np.linalg.norm(x - y) <= 1e-13 * np.linalg.norm(y), f


def test_lossy_tm_mirror_symmetry():
    """Smooth guide centred in y, source row spanning [by, ny-by): Ez is even
    in y about the centre and, with the source on the centre column of an odd
    grid, even in x.  A shifted stencil or CPML index breaks either."""
    c = variants.config('tm', 'scattering_loss', rough_toggle=False, **SMALL_TM)
    if c.nx % 2 == 0:
        c.nx += 1
    c.sx = c.nx // 2
    c.p2x = c.nx - c.p1x
    o = ost.TMzLossyStudy(c)
    for n in range(160):
        o.step(n)
    Ez = o.Ez
    assert np.abs(Ez).max() > 0
    np.testing.assert_allclose(Ez, Ez[:, ::-1], rtol=0, atol=1e-12 * np.abs(Ez).max())
    np.testing.assert_allclose(Ez, Ez[::-1, :], rtol=0, atol=1e-12 * np.abs(Ez).max())


def test_lossy_tm_host_arrays_match_oracle():
    c = variants.config('tm', 'scattering_loss', **SMALL_TM)
    up, lo = variants._profiles(c, c.p2x - c.p1x)
    fg = variants.tm_mask(c, up, lo)
    np.testing.assert_array_equal(fg, ost.tm_mask(c, up, lo))
    a = variants.lossy_tm_coefficients(c, fg)
    b = ost.lossy_tm_coefficients(c, fg)
    for k in b:
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)
    assert np.all(a['MOD_E'] == 1) and a['BE'][-1] == 0 and a['CE'][-1] == 0
