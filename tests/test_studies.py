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
        c = variants.config(*k, **(SMALL_TE if k[0] == 'te' else SMALL_3D))
        mons = variants.monitors(c)
        names = variants.output_names(c)
        assert set(mons) == set(names)
        for m in mons.values():
            assert m.nrec == c.nt
            assert (m.cs is None) == (k[1] == 'wave_impedance')
            assert m.after_h == (k[0] == '3d')
