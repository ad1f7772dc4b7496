"""Loss and wave-impedance study drivers on the device library.

The reference ships, beside the S-parameter driver, two study drivers per
scheme: ``scattering_loss`` (a single-frequency DFT of whole field planes) and
``wave_impedance`` (a time record of fields along a line).  Each is a
``fdtd_config.py`` of module-level constants and a ``fdtd_main.py`` that steps
the same Yee/CPML kernels as the S-parameter driver and records with its own
kernels.  Here

* ``config(scheme, study, **over)`` restates fdtd_config.py and returns its
  names on a namespace:
    FDTD/GPU/FDTD_2D_TE/scattering_loss/fdtd_config.py:95-180
    FDTD/GPU/FDTD_2D_TE/wave_impedance/fdtd_config.py (same derivation)
    FDTD/GPU/FDTD_3D/scattering_loss/fdtd_config.py:113-223
    FDTD/GPU/FDTD_3D/wave_impedance/fdtd_config.py:105-205
* ``run(c)`` restates fdtd_main.py: geometry and coefficients as the reference
  builds them, then the library's two-kernel Yee step with field monitors
  (include/oidt_fdtd.h, ``oidt_monitor_desc``) in place of simul_fft /
  map_*_zwave, and the reference's output arrays (``save`` writes the .npy
  files under their reference names).

TEz order.  The 2D TE study drivers update H first (Hz, its CPML, +J_SRC on
Hz, then Ex, Ey and their CPML; FDTD_2D_TE/scattering_loss/fdtd_main.py:202-223)
while the library steps E first.  From zero fields the E update of step 0 is
an identity, so the study's state after its step n is the library's state
right after the E half-step of step n + 1: the run is stepped nt + 1 times
with J_SRC negated (the study adds the source, the library's update_source
subtracts it) and padded with one zero, and the monitors are taken after the
E half-step with record index step - 1.  The study's Ey x-CPML kernel
(fdtd_funcs.py:135-144) negates both the difference and the correction of
the S-parameter kernel; psi changes sign and E is unchanged.

TMz.  The 2D TM study drivers use their own discretisation (lossy update
form, CPML graded over the whole border with a kappa divisor,
FDTD/GPU/FDTD_2D_TM/scattering_loss/fdtd_funcs.py:77-152): the library's
lossy TMz scheme (``oidt_scheme`` OIDT_TMZ_LOSSY) with the arrays of
fdtd_main.py:38-119, built by ``LossyTmArrays``.  Its config:
FDTD/GPU/FDTD_2D_TM/scattering_loss/fdtd_config.py:95-177.
"""
import os
from types import SimpleNamespace

import numpy as np

import ctypes as C

from . import _lib
from . import materials as M
from . import params
from . import roughness
from ._lib import ptr
from .solver import Monitor, Solver

STUDIES = ('scattering_loss', 'wave_impedance')

_TE = dict(eps_rel_bg=1.5 ** 2, eps_rel_fg=3.5 ** 2, sgl_wg_length=45.0e-6, sgl_wg_width=200e-9,
           port1_length=5e-6, port2_length=0e-6, rough_std=15.0e-9, rough_acl=700.0e-9, tol_std=10.0,
           tol_acl=10.0, rough_toggle=True, ctype=3, output_dir='Results', hz_fft_name='hz_fft',
           ex_fft_name='ex_fft', ey_fft_name='ey_fft', profile_number=0, source_type=1, source_amp=20.0,
           wave_packet_bw=0.10, gauss_pulse_deg=-6, f0=194.8e12, harmonics=1.0, num_flights=2.0,
           points_per_wl=40, num_cpml=20, buffer_xhat=20, ny_clad_wl=2.0)
_3D = dict(eps_rel_bg=1.5 ** 2, eps_rel_fg=3.5 ** 2, sgl_wg_length=24.0e-6, sgl_wg_width=2.0e-6,
           sgl_wg_height=0.2e-6, port1_length=4.0e-6, port2_length=0.0e-6, rough_std=15.0e-9,
           rough_acl=700.0e-9, tol_std=10.0, tol_acl=10.0, rough_toggle=True, ctype=3, output_dir='Results',
           ex1_fft_name='ex1_fft', ey1_fft_name='ey1_fft', ez1_fft_name='ez1_fft', hx1_fft_name='hx1_fft',
           hy1_fft_name='hy1_fft', hz1_fft_name='hz1_fft', ex2_fft_name='ex2_fft', ey2_fft_name='ey2_fft',
           ez2_fft_name='ez2_fft', hx2_fft_name='hx2_fft', hy2_fft_name='hy2_fft', hz2_fft_name='hz2_fft',
           profile_number=0, source_type=1, source_amp=5.0, wave_packet_bw=0.5, gauss_pulse_deg=-6,
           f0=194.8e12, source_condition='TE', harmonics=1.0, num_flights=2.0, points_per_wl=30,
           num_cpml=20, buffer_zhat=20, nx_clad_wl=0.75, ny_clad_wl=0.75)

_TM = dict(eps_rel_bg=1.5 ** 2, eps_rel_fg=3.5 ** 2, sgl_wg_length=30e-6, sgl_wg_width=200e-9, port_length=5e-6,
           rough_std=15.0e-9, rough_acl=700.0e-9, tol_std=10.0, tol_acl=10.0, rough_toggle=True, ctype=3,
           output_dir='Results', ez_fft_name='ez_fft', hx_fft_name='hx_fft', hy_fft_name='hy_fft',
           profile_number=0, source_type=1, source_amp=5.0, wave_packet_bw=0.8, gauss_pulse_deg=-6,
           f0=194.8e12, harmonics=1.0, num_flights=2.0, points_per_wl=40, num_cpml=20, buffer_xhat=10,
           ny_clad_wl=2.0)

DEFAULTS = {
    ('tm', 'scattering_loss'): _TM,
    ('tm', 'wave_impedance'): {**{k: v for k, v in _TM.items() if not k.endswith('_fft_name')},
                               'rough_toggle': False},
    ('te', 'scattering_loss'): _TE,
    ('te', 'wave_impedance'): {**{k: v for k, v in _TE.items() if not k.endswith('_fft_name')},
                               'sgl_wg_length': 25.0e-6, 'rough_toggle': False},
    ('3d', 'scattering_loss'): _3D,
    ('3d', 'wave_impedance'): {**{k: v for k, v in _3D.items() if not k.endswith('_fft_name')},
                               'sgl_wg_length': 14.0e-6, 'rough_toggle': False, 'source_condition': 'TM',
                               'points_per_wl': 40, 'ey_zwave_name': 'ey_te_zwave',
                               'ez_zwave_name': 'ez_tm_zwave', 'hy_zwave_name': 'hy_tm_zwave',
                               'hz_zwave_name': 'hz_te_zwave'},
}


def config(scheme, study, **over):
    """Namespace with the names of the study's fdtd_config.py (float32 dx, dt)."""
    if (scheme, study) not in DEFAULTS:
        raise ValueError('no study driver for scheme %r, study %r' % (scheme, study))
    c = SimpleNamespace(**DEFAULTS[(scheme, study)])
    for k, v in over.items():
        if not hasattr(c, k):
            raise KeyError('unknown study parameter %r' % k)
        setattr(c, k, v)
    c.scheme, c.study = scheme, study
    c.eps = c.eps_rel_bg / (36.0e9 * np.pi)
    c.mu = 4.0e-7 * np.pi
    c.c_bg = 1 / np.sqrt(c.eps * c.mu)
    c.eta = np.sqrt(c.mu / c.eps)
    c.eps_rel_fg = c.eps_rel_fg / c.eps_rel_bg
    c.f_max = c.harmonics * c.f0
    c.vp = c.c_bg / np.sqrt(c.eps_rel_fg)
    c.wl = c.vp / c.f_max
    c.wl_clad = c.c_bg / c.f_max
    c.flight_time = c.sgl_wg_length / c.vp
    c.sim_time = c.num_flights * c.flight_time
    dx64 = c.wl / c.points_per_wl
    D = 3 if scheme == '3d' else 2
    c.delta_x = np.float32(dx64)
    c.delta_t = np.float32(dx64 / (c.c_bg * np.sqrt(D)))
    # numpy 1.20: a float32 scalar met by a Python or float64 scalar gives float64
    dx, dt = float(c.delta_x), float(c.delta_t)
    nt = int(c.sim_time / dt)
    N = c.num_cpml
    if scheme == 'tm':
        c.buffer_yhat = int(c.ny_clad_wl * c.wl_clad / dx)
        c.cpml_range = N + 1
        c.nx_swg = int(c.sgl_wg_length / dx)
        c.ny_swg = int(c.sgl_wg_width / dx)
        c.bx = N + c.buffer_xhat
        c.by = N + c.buffer_yhat
        c.nx = 2 * c.bx + c.nx_swg
        c.ny = 2 * c.by + c.ny_swg
        c.nt = nt
        c.cx, c.cy = c.nx // 2, c.ny // 2
        c.cx_swg, c.cy_swg = c.nx_swg // 2, c.ny_swg // 2
        c.sx = c.bx + 0
        c.sy = c.by + c.cy_swg
        c.J_SRC = tm_waveform(c)
        c.p1x = c.sx + int(c.port_length / dx)
        c.p2x = c.nx - c.p1x
        if c.p2x <= c.p1x:
            raise Exception('Port settling ranges overlap. Make the WG longer or decrease port settling range!')
    elif scheme == 'te':
        c.buffer_yhat = int(c.ny_clad_wl * c.wl_clad / dx)
        c.nx_swg = int(c.sgl_wg_length / dx)
        c.ny_swg = int(c.sgl_wg_width / dx)
        c.bx = N + c.buffer_xhat
        c.by = N + c.buffer_yhat
        c.nx = 2 * c.bx + c.nx_swg
        c.ny = 2 * c.by + c.ny_swg
        c.cx, c.cy = c.nx // 2, c.ny // 2
        c.cx_swg, c.cy_swg = c.nx_swg // 2, c.ny_swg // 2
        c.sx = c.bx + 10
        c.sy = c.by + c.cy_swg
        c.J_SRC, c.nt = params.source_waveform(np.float32, nt, c.delta_t, c.source_type, c.source_amp, c.f0,
                                               c.wave_packet_bw, c.gauss_pulse_deg, extend=False)
        c.p1x = c.sx + int(c.port1_length / dx)
        c.p2x = c.sx + c.nx_swg - int(c.port2_length / dx)
        if c.p2x <= c.p1x:
            raise Exception('Port settling ranges overlap. Make the WG longer or decrease port settling range!')
    else:
        c.buffer_xhat = int(c.nx_clad_wl * c.wl_clad / dx)
        c.buffer_yhat = int(c.ny_clad_wl * c.wl_clad / dx)
        c.nx_swg = int(c.sgl_wg_height / dx)
        c.ny_swg = int(c.sgl_wg_width / dx)
        c.nz_swg = int(c.sgl_wg_length / dx)
        c.bx, c.by, c.bz = N + c.buffer_xhat, N + c.buffer_yhat, N + c.buffer_zhat
        c.nx, c.ny, c.nz = 2 * c.bx + c.nx_swg, 2 * c.by + c.ny_swg, 2 * c.bz + c.nz_swg
        c.cx, c.cy, c.cz = c.nx // 2, c.ny // 2, c.nz // 2
        c.cx_swg, c.cy_swg, c.cz_swg = c.nx_swg // 2, c.ny_swg // 2, c.nz_swg // 2
        c.sx, c.sy, c.sz = c.bx + c.cx_swg, c.by + c.cy_swg, c.bz
        c.J_SRC, c.nt = params.source_waveform(np.float32, nt, c.delta_t, c.source_type, c.source_amp, c.f0,
                                               c.wave_packet_bw, c.gauss_pulse_deg, extend=True)
        c.p1z = c.sz + int(c.port1_length / dx)
        c.p2z = c.sz + c.nz_swg - int(c.port2_length / dx)
        if c.p2z <= c.p1z:
            raise Exception('Port settling ranges overlap. Make the WG longer or decrease port settling range!')
        if study == 'scattering_loss':
            c.cosE, c.sinE, c.cosH, c.sinH = phase_tables_f32(c.f0, c.delta_t, c.nt)
    c.rstd = c.rough_std / dx
    c.racl = c.rough_acl / dx
    c.tol_std = c.tol_std / 100
    c.tol_acl = c.tol_acl / 100
    c.roughness_profile = '_s{}nm_lc{}nm_r{}'.format(str(round(c.rough_std * 1e9)), str(round(c.rough_acl * 1e9)),
                                                     c.profile_number)
    c.up = './rough_profiles/profile_upper' + c.roughness_profile
    c.lp = './rough_profiles/profile_lower' + c.roughness_profile
    return c


def tm_waveform(c):
    """J_SRC of FDTD_2D_TM/scattering_loss/fdtd_config.py:145-174: pulse peaks
    at 5 spreads, no step extension, the ramped harmonic scaled by dx**2."""
    dt = float(c.delta_t)
    gp_tsp = np.sqrt(-2 * np.log(10 ** (c.gauss_pulse_deg / 20))) / (2 * np.pi * c.f0)
    gp_tpk = 5 * gp_tsp
    wp_tsp = 2 * np.sqrt(np.log(2)) / (c.wave_packet_bw * 2 * np.pi * c.f0)
    wp_tpk = 5 * wp_tsp
    t = np.arange(c.nt, dtype=np.float64) * dt
    f32 = np.float32
    if c.source_type == 1:
        return (c.source_amp * np.exp(-0.5 * ((t - gp_tpk) / gp_tsp) ** 2).astype(f32)).astype(f32)
    if c.source_type == 2:
        wp = (np.exp(-0.5 * ((t - wp_tpk) / wp_tsp) ** 2) * np.exp(2j * np.pi * c.f0 * (t - wp_tpk)).real).astype(f32)
        return (c.source_amp * wp).astype(f32)
    if c.source_type != 3:
        raise Exception('Pick a valid source type')
    gp = np.exp(-0.5 * ((t - gp_tpk) / gp_tsp) ** 2).astype(f32)
    ramp = np.zeros(c.nt, f32)
    acc = f32(0)
    for n in range(c.nt):
        acc = f32(acc + c.delta_t * gp[n])
        ramp[n] = acc
    ramp = ramp / np.max(ramp)
    th = np.exp(2j * np.pi * c.f0 * (t - gp_tpk)).real.astype(f32)
    amp = f32(c.source_amp * float(c.delta_x * c.delta_x))   # float64 scalar product, applied in float32
    return ((amp * ramp) * th).astype(f32)


def lossy_tm_coefficients(c, fg):
    """CURL_E, CURL_H, MOD_E, BE, CE, BH, CH, DEN_EX/EY/HX/HY of
    FDTD_2D_TM/scattering_loss/fdtd_main.py:38-119 with numpy 1.20 element
    types (float32 arrays met by scalars stay float32; scalar-scalar
    arithmetic is float64)."""
    f32 = np.float32
    dt, dx = f32(c.delta_t), f32(c.delta_x)
    sigma = np.zeros((c.nx, c.ny), f32)
    eps_mask = (c.eps * (c.eps_rel_fg * fg + 1 * ~fg)).astype(f32)
    out = {'CURL_E': f32(float(dt) / (c.mu * float(dx)))}
    den = (dx * (f32(2) * eps_mask + sigma * dt)).astype(f32)
    out['CURL_H'] = (f32(2 * float(dt)) / den).astype(f32)
    out['MOD_E'] = ((f32(2) * eps_mask - sigma * dt) / (f32(2) * eps_mask + sigma * dt)).astype(f32)
    N, R = c.num_cpml, c.cpml_range
    m_grade, m_alpha = 4, 1
    sigma_max = 1.2 * (0.8 * (m_grade + 1) / (float(c.eta) * float(dx)))
    alpha_max, kappa_max = 0.05, 5
    A = {k: np.zeros(R, f32) for k in ('AE', 'KE', 'SE', 'AH', 'KH', 'SH', 'BE', 'CE', 'BH', 'CH')}
    den_ex, den_ey = np.ones(c.nx, f32), np.ones(c.ny, f32)
    den_hx, den_hy = np.ones(c.nx, f32), np.ones(c.ny, f32)
    for q in range(N):
        for t, pos in (('E', q), ('H', q + 0.5)):
            off = 1 if t == 'E' else 1.5
            A['A' + t][q] = alpha_max * (pos / N) ** m_alpha
            A['K' + t][q] = 1 + (kappa_max - 1) * ((R - q - off) / N) ** m_grade
            A['S' + t][q] = sigma_max * ((R - q - off) / N) ** m_grade
            S_, K_, A_ = A['S' + t][q], A['K' + t][q], A['A' + t][q]
            A['B' + t][q] = np.exp(-1.0 * float(S_ / K_ + A_) * float(dt) / c.eps)
            A['C' + t][q] = float(S_ / (S_ + K_ * A_) / K_) * (float(A['B' + t][q]) - 1.0)
        den_ex[q] = den_ex[c.nx - 1 - q] = A['KE'][q]
        den_ey[q] = den_ey[c.ny - 1 - q] = A['KE'][q]
        den_hx[q] = den_hx[c.nx - 2 - q] = A['KH'][q]
        den_hy[q] = den_hy[c.ny - 2 - q] = A['KH'][q]
    out.update({k: A[k] for k in ('BE', 'CE', 'BH', 'CH')})
    out.update(DEN_EX=den_ex, DEN_EY=den_ey, DEN_HX=den_hx, DEN_HY=den_hy)
    return out


class LossyTmArrays:
    """Run description and arrays of a lossy TMz study run (C ABI scheme
    OIDT_TMZ_LOSSY); kept alive while the library reads them."""

    def __init__(self, c, fg):
        self.c = c
        self.k = lossy_tm_coefficients(c, fg)
        self.cb = [np.ascontiguousarray(self.k['CURL_H']), np.ascontiguousarray(self.k['MOD_E'])]
        self.jsrc = np.ascontiguousarray(c.J_SRC[:c.nt], np.float32)

    def desc(self):
        c = self.c
        d = _lib.RunDesc()
        d.abi_version = _lib.ABI_VERSION
        d.scheme = _lib.TMZ_LOSSY
        d.precision = _lib.F32
        d.nx, d.ny, d.nz = c.nx, c.ny, 1
        d.num_cpml = c.cpml_range
        d.offset = 1
        d.nt = c.nt
        d.nf = 0
        d.src_field = _lib.EZ
        box = (c.sx, c.sx + 1, c.by, c.ny - c.by, 0, 1)   # update_source of fdtd_main.py:202
        for a in range(6):
            d.src_box[a] = box[a]
        d.port[0], d.port[1] = c.p1x, c.p2x
        d.dx = float(c.delta_x)
        d.dt = float(c.delta_t)
        d.db = float(self.k['CURL_E'])
        return d

    def inputs(self):
        k = self.k
        r = _lib.RunInputs()
        r.cb[0], r.cb[1] = ptr(self.cb[0]), ptr(self.cb[1])
        r.kappa_e[0], r.kappa_e[1] = ptr(k['DEN_EX']), ptr(k['DEN_EY'])
        r.kappa_h[0], r.kappa_h[1] = ptr(k['DEN_HX']), ptr(k['DEN_HY'])
        r.be, r.ce, r.bh, r.ch = ptr(k['BE']), ptr(k['CE']), ptr(k['BH']), ptr(k['CH'])
        r.jsrc = ptr(self.jsrc)
        r.freqs = None
        return r


def phase_tables_f32(f0, delta_t, nt):
    """cosE, sinE, cosH, sinH of FDTD_3D/scattering_loss/fdtd_config.py:215-223:
    the float64 phase 2 pi f0 (n + 0 | 0.5) dt rounded to float32, then cos / sin
    evaluated in float32."""
    n = np.arange(nt, dtype=np.float64)
    pe = (2 * np.pi * f0 * (n + 0.0) * float(delta_t)).astype(np.float32)
    ph = (2 * np.pi * f0 * (n + 0.5) * float(delta_t)).astype(np.float32)
    return np.cos(pe), np.sin(pe), np.cos(ph), np.sin(ph)


def phases_f64(f0, delta_t, steps):
    """(cos, sin) rows of simul_fft's weight cos(2 pi freq step dt) with freq,
    step and dt passed as float32 and the product and cos / sin in float64
    (FDTD_2D_TE/scattering_loss/fdtd_funcs.py:113-120)."""
    ph = 2 * np.pi * float(np.float32(f0)) * np.asarray(steps, dtype=np.float32).astype(np.float64) * float(delta_t)
    return np.stack([np.cos(ph), np.sin(ph)], axis=1)


def _profiles(c, length):
    if not c.rough_toggle:
        return np.zeros(length), np.zeros(length)
    return roughness.profiles(length, c.rstd, c.racl, seed=c.profile_number, correlation=c.ctype)


def te_cb_maps(c, fg):
    """CBX, CBY of FDTD_2D_TE/scattering_loss/fdtd_main.py:37-47: the float32
    EPS_MASK averaged with float32 sums, then delta_t / EPS."""
    mask = (c.eps * (c.eps_rel_fg * fg + 1 * ~fg)).astype(np.float32)
    ex, ey = mask.copy(), mask.copy()
    ex[1:, 1:] = (mask[1:, 1:] + mask[:-1, 1:]) * np.float32(0.5)
    ey[1:, 1:] = (mask[1:, 1:] + mask[1:, :-1]) * np.float32(0.5)
    return c.delta_t / ex, c.delta_t / ey


def setup_of(c):
    """The library Setup for a study config (no port DFT frequencies)."""
    common = dict(precision=np.float32, delta_x=c.delta_x, delta_t=c.delta_t, eps=c.eps, mu=c.mu, eta=c.eta,
                  eps_rel_fg=c.eps_rel_fg, num_cpml=c.num_cpml, sparam_freqs=np.zeros(0), pp=0, ph=0,
                  bx=c.bx, by=c.by, rstd=c.rstd, racl=c.racl)
    if c.scheme == 'te':
        # stepped nt + 1 times, source negated and padded (module docstring)
        J = np.append(-c.J_SRC[:c.nt], np.float32(0)).astype(np.float32)
        return params.Setup(scheme='te', nx=c.nx, ny=c.ny, nt=c.nt + 1, J_SRC=J, ny_swg=c.ny_swg,
                            cy_swg=c.cy_swg, sx=c.sx, p1x=c.p1x, p2x=c.p2x, **common)
    return params.Setup(scheme='3d', nx=c.nx, ny=c.ny, nz=c.nz, nt=c.nt, J_SRC=c.J_SRC[:c.nt], bz=c.bz,
                        nx_swg=c.nx_swg, ny_swg=c.ny_swg, cx_swg=c.cx_swg, cy_swg=c.cy_swg, sz=c.sz,
                        p1z=c.p1z, p2z=c.p2z, source_condition=c.source_condition, **common)


# Stencils of the reference's recording kernels: (points, scale).
E_AVG = {'ex': ([(0, 0, 0), (1, 0, 0)], 0.5), 'ey': ([(0, 0, 0), (0, 1, 0)], 0.5),
         'ez': ([(0, 0, 0), (0, 0, 1)], 0.5)}
H_AVG = {'hx': ([(0, 0, 0), (0, 1, 0), (0, 0, 1), (0, 1, 1)], 0.25),
         'hy': ([(0, 0, 0), (1, 0, 0), (0, 0, 1), (1, 0, 1)], 0.25),
         'hz': ([(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0)], 0.25)}


def monitors(c):
    """{output name: Monitor} of the study, with the reference's array names."""
    nt = c.nt
    if c.scheme == 'tm':
        if c.study == 'scattering_loss':
            # simul_fft of Ez at n + 0.5, Hx and Hy at n + 1.0 after the whole step
            # (FDTD_2D_TM/scattering_loss/fdtd_main.py:216-218)
            box = [(0, c.nx), (0, c.ny)]
            return {k: Monitor(k, [(0, 0)], 1.0, box, nt, after_h=True,
                               cs=phases_f64(c.f0, c.delta_t, np.arange(nt) + (0.5 if k == 'ez' else 1.0)))
                    for k in ('ez', 'hx', 'hy')}
        # map_efield_zwave / map_hfield_zwave at x = cx (FDTD_2D_TM/wave_impedance/
        # fdtd_funcs.py:155-168); Hx[cx, -1] wraps to Hx[cx, ny-1], a cell no update writes
        box = [(c.cx, c.cx + 1), (0, c.ny)]
        return {'ez': Monitor('ez', [(0, 0)], 1.0, box, nt, after_h=True),
                'hx': Monitor('hx', [(0, 0), (0, -1)], 0.5, box, nt, after_h=True)}
    if c.scheme == 'te':
        if c.study == 'scattering_loss':
            # simul_fft of Hz at step n + 0.0 and Ey at n + 0.5 over the whole grid
            # (FDTD_2D_TE/scattering_loss/fdtd_main.py:222-223)
            box = [(0, c.nx), (0, c.ny)]
            return {'hz': Monitor('hz', [(0, 0)], 1.0, box, nt, after_h=False, shift=-1,
                                  cs=phases_f64(c.f0, c.delta_t, np.arange(nt) + 0.0)),
                    'ey': Monitor('ey', [(0, 0)], 1.0, box, nt, after_h=False, shift=-1,
                                  cs=phases_f64(c.f0, c.delta_t, np.arange(nt) + 0.5))}
        # map_efield_zwave / map_hfield_zwave at x = cx (fdtd_funcs.py:169-181);
        # the reference's Hz[cx, ny] neighbour is a cell no update writes (zero)
        box = [(c.cx, c.cx + 1), (0, c.ny)]
        return {'ex': Monitor('ex', [(0, 0), (1, 0)], 0.5, box, nt, after_h=False, shift=-1),
                'hz': Monitor('hz', [(0, 0), (1, 0), (0, 1), (1, 1)], 0.25, box, nt, after_h=False, shift=-1)}
    if c.study == 'scattering_loss':
        # simul_fft_efield / _hfield at both port planes over [N, nx-N) x [N, ny-N)
        # (FDTD_3D/scattering_loss/fdtd_main.py:341-344, fdtd_funcs.py:217-242)
        N = c.num_cpml
        csE = np.stack([c.cosE, c.sinE], axis=1).astype(np.float64)
        csH = np.stack([c.cosH, c.sinH], axis=1).astype(np.float64)
        out = {}
        for p, kc in ((1, c.p1z), (2, c.p2z)):
            box = [(N, c.nx - N), (N, c.ny - N), (kc, kc + 1)]
            for k, (pts, sc) in E_AVG.items():
                out[k + str(p)] = Monitor(k, pts, sc, box, nt, after_h=True, cs=csE)
            for k, (pts, sc) in H_AVG.items():
                out[k + str(p)] = Monitor(k, pts, sc, box, nt, after_h=True, cs=csH)
        return out
    # map_efield_zwave / map_hfield_zwave along x at (cy, cz), i < ie = nx - 2
    # (FDTD_3D/wave_impedance/fdtd_main.py:298-299, fdtd_funcs.py:321-335);
    # only the pair the source condition saves (:341-346)
    box = [(0, c.nx - 2), (c.cy, c.cy + 1), (c.cz, c.cz + 1)]
    keys = ('ey', 'hz') if c.source_condition == 'TE' else ('ez', 'hy')
    out = {}
    for k in keys:
        pts, sc = E_AVG[k] if k[0] == 'e' else H_AVG[k]
        out[k] = Monitor(k, pts, sc, box, nt, after_h=True)
    return out


def build(c):
    """(setup, fg mask, cb maps or None) of a study config."""
    if c.scheme == 'tm':
        up, lo = _profiles(c, c.p2x - c.p1x)
        return None, tm_mask(c, up, lo), None
    s = setup_of(c)
    if c.scheme == 'te':
        up, lo = _profiles(c, c.p2x - c.p1x)
        fg = M.fg_mask(s, up, lo)
        return s, fg, te_cb_maps(c, fg)
    up, lo = _profiles(c, c.p2z - c.p1z)
    return s, M.fg_mask(s, up, lo), None


def tm_mask(c, upper, lower):
    """FG_REG of gen_fg_mask_sgl (FDTD_2D_TM/scattering_loss/fdtd_funcs.py:
    32-74): core iff by + lower[i] <= j < by + ny_swg + upper[i], profiles
    padded by p1x zeros at both ends."""
    L = c.p2x - c.p1x
    up = M.padded(np.zeros(L) if upper is None else upper, c.p1x, c.nx - c.p2x)
    lo = M.padded(np.zeros(L) if lower is None else lower, c.p1x, c.nx - c.p2x)
    j = np.arange(c.ny, dtype=np.float64)[None, :]
    return (j >= c.by + lo[:, None]) & (j < c.by + c.ny_swg + up[:, None])


def truncate(c, nt):
    """The study cut to its first nt steps (source and phase tables too)."""
    c.nt = int(nt)
    c.J_SRC = c.J_SRC[:c.nt]
    for k in ('cosE', 'sinE', 'cosH', 'sinH'):
        if hasattr(c, k):
            setattr(c, k, getattr(c, k)[:c.nt])
    return c


def prepare(c, launch=(0, 0, 0)):
    """(solver, monitors, monitor ids): the study's run on the device with its
    monitors attached, not yet advanced."""
    variant, xchunk, grid = launch
    if variant == 3:
        raise ValueError('the study drivers need a two-kernel form (monitors sit between the half-steps)')
    s, fg, cb = build(c)
    sol = Solver.from_arrays(LossyTmArrays(c, fg), c.nt) if c.scheme == 'tm' else Solver(s, fg, cb=cb)
    try:
        sol.set_launch(xchunk, grid, variant)
        mons = monitors(c)
        ids = {k: sol.add_monitor(m) for k, m in mons.items()}
    except Exception:
        sol.close()
        raise
    return sol, mons, ids


def run(c, launch=(0, 0, 0)):
    """Step the study on the device; returns {name: array} in the reference's
    output shapes and types (complex64 phasors, float32 time records)."""
    sol, mons, ids = prepare(c, launch)
    try:
        sol.run()
        return {k: _output(c, k, mons[k], sol.read_monitor(ids[k])) for k in mons}
    finally:
        sol.close()


def prepare_batch(cs, launch=(0, 0, 0)):
    """(batch, runs): several studies of one scheme in one device batch with
    their monitors attached, not yet advanced; runs = [(config, slot,
    monitors, monitor ids, steps)]."""
    from .solver import Batch
    schemes = {c.scheme for c in cs}
    if len(schemes) != 1:
        raise ValueError('a batch holds studies of one scheme')
    scheme = schemes.pop()
    variant, xchunk, grid = launch
    if variant == 3:
        raise ValueError('the study drivers need a two-kernel form (monitors sit between the half-steps)')
    b = Batch('tm_lossy' if scheme == 'tm' else scheme, np.float32, max_slots=len(cs))
    try:
        b.set_launch(xchunk, grid, variant)
        runs = []
        for c in cs:
            s, fg, cb = build(c)
            if scheme == 'tm':
                meta = SimpleNamespace(nt=c.nt, precision=np.float32, shape=(c.nx, c.ny), sparam_freqs=np.zeros(0))
                slot = b.add_arrays(LossyTmArrays(c, fg), meta)
            else:
                slot = b.add(s, fg, cb=cb)
            mons = monitors(c)
            ids = {k: b.add_monitor(slot, m) for k, m in mons.items()}
            runs.append((c, slot, mons, ids, s.nt if s is not None else c.nt))
    except Exception:
        b.close()
        raise
    return b, runs


def run_batch(cs, launch=(0, 0, 0)):
    """Several studies of one scheme (e.g. the roughness realisations that
    automated_results_collection.py runs one process each) as one device
    batch; returns one output dict per configuration, as ``run``."""
    b, runs = prepare_batch(cs, launch)
    try:
        chunk = _lib.chunk_steps()
        b.advance(-(-max(r[4] for r in runs) // chunk))
        b.sync()
        return [{k: _output(c, k, mons[k], b.read_monitor(slot, ids[k])) for k in mons}
                for c, slot, mons, ids, _ in runs]
    finally:
        b.close()


def _output(c, key, mon, val):
    if c.scheme == 'tm':
        if c.study == 'scattering_loss':
            re, im = val
            # dEZfr.copy_to_host() - 1j*dEZfi.copy_to_host()  (fdtd_main.py:260-262)
            return (re - 1j * im.astype(np.complex64)).astype(np.complex64)[:, :, 0]
        return val.reshape(c.nt, c.ny)
    if c.scheme == 'te':
        if c.study == 'scattering_loss':
            re, im = val
            # dHZfr.copy_to_host() - 1j*dHZfi.copy_to_host()  (fdtd_main.py:265-266)
            return (re - 1j * im.astype(np.complex64)).astype(np.complex64)[:, :, 0]
        return val.reshape(c.nt, c.ny)
    if c.study == 'scattering_loss':
        re, im = val
        full = np.zeros((c.nx, c.ny), dtype=np.complex64)
        (x0, x1), (y0, y1) = mon.box[0], mon.box[1]
        # EX1 = EX1r + 1j*EX1i over the zero-initialised [nx, ny] arrays (fdtd_main.py:424-435)
        full[x0:x1, y0:y1] = (re + 1j * im.astype(np.complex64))[:, :, 0]
        return full
    out = np.zeros((c.nt, c.nx), dtype=np.float32)
    out[:, :c.nx - 2] = val.reshape(c.nt, c.nx - 2)
    return out


def output_names(c):
    """{output key: .npy file stem} as the reference's np.save calls."""
    if c.scheme == 'tm':
        if c.study == 'scattering_loss':
            return {k: getattr(c, k + '_fft_name') + c.roughness_profile for k in ('ez', 'hx', 'hy')}
        return {'ez': 'ez_zwave', 'hx': 'hx_zwave'}
    if c.scheme == 'te':
        if c.study == 'scattering_loss':
            return {'hz': c.hz_fft_name + c.roughness_profile, 'ey': c.ey_fft_name + c.roughness_profile}
        return {'ex': 'ex_zwave', 'hz': 'hz_zwave'}
    if c.study == 'scattering_loss':
        return {k: getattr(c, k + '_fft_name') + c.roughness_profile
                for k in ('ex1', 'ey1', 'ez1', 'hx1', 'hy1', 'hz1', 'ex2', 'ey2', 'ez2', 'hx2', 'hy2', 'hz2')}
    return {k: getattr(c, k + '_zwave_name') for k in (('ey', 'hz') if c.source_condition == 'TE' else ('ez', 'hy'))}


def save(c, outputs, workdir='.'):
    """np.save each output under <output_dir>/<reference name>.npy."""
    d = os.path.join(workdir, c.output_dir)
    os.makedirs(d, exist_ok=True)
    names = output_names(c)
    paths = []
    for k, v in outputs.items():
        p = os.path.join(d, names[k])
        np.save(p, v)
        paths.append(p + '.npy')
    return paths


def main(argv=None):
    """``python -m oidt_amd.variants <tm|te|3d> <scattering_loss|wave_impedance> [key=value ...]``."""
    import sys
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) < 2:
        raise SystemExit(main.__doc__)
    import ast
    over = {}
    for kv in argv[2:]:
        k, v = kv.split('=', 1)
        try:
            over[k] = ast.literal_eval(v)
        except (ValueError, SyntaxError):
            over[k] = v
    c = config(argv[0], argv[1], **over)
    print('The simulation size will be: {} cells over {} time steps'.format(
        'x'.join(str(n) for n in ((c.nx, c.ny, c.nz) if c.scheme == '3d' else (c.nx, c.ny))), c.nt))
    for p in save(c, run(c)):
        print('saved', p)


if __name__ == '__main__':
    main()
