"""Run parameters: the reference's parameter interface and its derived constants.

* ``dict_from_module`` / ``load_user_dict`` -- the user_config -> user_dict.pkl
  protocol of FDTD/GPU/FDTD_2D_TM/s-parameter_extraction/aux_funcs.py:64-71 and
  fdtd_auto_setup.py:26-53.
* ``auto_setup`` -- fdtd_auto_setup.py (TM/TE :59-170, 3D :59-177): dx, dt, grid,
  ports, waveform, from the user dictionary, for one run kind ('i' or 'r').
* ``explicit_setup`` -- the same ``Setup`` built from explicit cell counts, dx and
  Courant number (the benchmark workloads, whose grids are given in cells).

numpy 1.20 (the reference's pinned version) promoted numpy-float32-scalar /
Python-float operations to float64; numpy 2 keeps float32.  The integer grid
sizes depend on it, so every such expression takes the float64 value of the
rounded dx / dt explicitly.
"""
import os
import pickle
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

EPS_SCALE = 36.0e9 * np.pi   # eps = eps_rel / (36e9 pi)  (fdtd_auto_setup.py:62)
MU0 = 4.0e-7 * np.pi         # fdtd_auto_setup.py:63


def dict_from_module(module):
    """aux_funcs.py:64-71: public float/str/int/type/bool attributes of a module."""
    out = {}
    for item in dir(module):
        it = type(getattr(module, item))
        if it is float or it is str or it is int or it is type or it is bool:
            if item[0] != '_':
                out[item] = getattr(module, item)
    return out


def load_user_dict(line_arg=None, module_name='user_config', workdir='.'):
    """fdtd_auto_setup.py:26-40: ``user_dict_<argv1>.pkl`` if present, else the
    ``user_config`` module re-pickled to ``user_dict.pkl``."""
    if line_arg is not None:
        p = os.path.join(workdir, 'user_dict_{}.pkl'.format(line_arg))
        if os.path.exists(p):
            with open(p, 'rb') as f:
                return pickle.load(f)
    import importlib
    mod = importlib.import_module(module_name)
    usr = dict_from_module(mod)
    with open(os.path.join(workdir, 'user_dict.pkl'), 'wb') as f:
        pickle.dump(usr, f)
    return usr


@dataclass
class Setup:
    """Derived constants of one run; attribute names follow fdtd_auto_setup.py."""
    scheme: str                      # 'tm' | 'te' | '3d'
    precision: type                  # np.float32 | np.float64
    delta_x: np.floating
    delta_t: np.floating
    eps: float
    mu: float
    eta: float
    eps_rel_fg: float                # core permittivity relative to cladding
    num_cpml: int
    nx: int
    ny: int
    nt: int
    J_SRC: np.ndarray
    sparam_freqs: np.ndarray
    pp: int
    ph: int
    bx: int
    by: int
    num_lines: int = 1
    offset: int = 2
    m_grade: int = 4
    m_alpha: int = 1
    kappa_max: float = 5.0
    alpha_max: float = 0.05
    sigma_factor: float = 1.2
    # 2D
    ny_swg: int = 0
    cy_swg: int = 0
    sx: int = 0
    p1x: int = 0
    p2x: int = 0
    # 3D
    nz: int = 1
    bz: int = 0
    nx_swg: int = 0
    cx_swg: int = 0
    sz: int = 0
    p1z: int = 0
    p2z: int = 0
    source_condition: str = 'TE'
    rstd: float = 0.0
    racl: float = 0.0
    extra: dict = field(default_factory=dict)

    @property
    def ports(self):
        return (self.p1z, self.p2z) if self.scheme == '3d' else (self.p1x, self.p2x)

    @property
    def shape(self):
        return (self.nx, self.ny, self.nz) if self.scheme == '3d' else (self.nx, self.ny)

    @property
    def cells(self):
        return self.nx * self.ny * (self.nz if self.scheme == '3d' else 1)


def source_waveform(precision, nt, delta_t, source_type, source_amp, f0, wave_packet_bw, gauss_pulse_deg=-6,
                    extend=True):
    """J_SRC and nt after the ramp extension (fdtd_auto_setup.py:121-159).
    ``extend=False``: no extension (the 2D loss and impedance study configs,
    FDTD_2D_TE/scattering_loss/fdtd_config.py:115, 148-167)."""
    dt = float(delta_t)
    gp_tsp = np.sqrt(-2 * np.log(10 ** (gauss_pulse_deg / 20))) / (2 * np.pi * f0)
    gp_tpk = 9 * gp_tsp
    wp_tsp = 2 * np.sqrt(np.log(2)) / (wave_packet_bw * 2 * np.pi * f0)
    wp_tpk = 9 * wp_tsp
    ext = {1: lambda: int(gp_tpk / nt), 2: lambda: int(wp_tpk / nt), 3: lambda: int((gp_tpk + gp_tsp) / dt)}
    if source_type not in ext:
        raise Exception('Pick a valid source type')
    if extend:
        nt = nt + ext[source_type]()
    t = np.arange(nt, dtype=np.float64) * dt
    if source_type == 1:
        shape = np.exp(-0.5 * ((t - gp_tpk) / gp_tsp) ** 2).astype(precision)
        return (source_amp * shape).astype(precision), nt
    if source_type == 2:
        shape = (np.exp(-0.5 * ((t - wp_tpk) / wp_tsp) ** 2) *
                 np.exp(2j * np.pi * f0 * (t - wp_tpk)).real).astype(precision)
        return (source_amp * shape).astype(precision), nt
    gp = np.exp(-0.5 * ((t - gp_tpk) / gp_tsp) ** 2).astype(precision)
    ramp = np.zeros(nt, precision)
    acc = precision(0)
    dts = precision(delta_t)
    for n in range(nt):       # running sum in the run precision (RAMP[n-1] + dt*GP[n])
        acc = precision(acc + dts * gp[n])
        ramp[n] = acc
    ramp = ramp / np.max(ramp)
    th = np.exp(2j * np.pi * f0 * (t - gp_tpk)).real.astype(precision)
    return (source_amp * ramp * th).astype(precision), nt


def _precision(p):
    if p in (np.float32, 'float32', 'f4'):
        return np.float32
    if p in (np.float64, 'float64', 'f8', float):
        return np.float64
    raise ValueError('precision must be float32 or float64')


def auto_setup(usr_in, scheme, kind='r'):
    """Setup from a reference user dictionary for run kind 'i' or 'r'."""
    usr = dict(usr_in)
    if usr.get('auto_condition'):
        usr.update(f0=0.5 * (usr['sparam_fmin'] + usr['sparam_fmax']), harmonics=2.0, points_per_wl=20,
                   num_flights=2.0)
    prec = _precision(usr['precision'])
    eps = usr['eps_rel_bg'] / EPS_SCALE
    mu = MU0
    c_bg = 1 / np.sqrt(eps * mu)
    eta = np.sqrt(mu / eps)
    eps_rel_fg = usr['eps_rel_fg'] / usr['eps_rel_bg']
    f_max = usr['harmonics'] * usr['f0']
    vp = c_bg / np.sqrt(eps_rel_fg)
    wl = vp / f_max
    wl_clad = c_bg / f_max
    sim_time = usr['num_flights'] * (usr['sgl_wg_length'] / vp)
    dx64 = wl / usr['points_per_wl']
    D = 3 if scheme == '3d' else 2
    delta_x = prec(dx64)
    delta_t = prec(dx64 / (c_bg * np.sqrt(D)))
    dx, dt = float(delta_x), float(delta_t)
    length = usr['port1_length'] + usr['port2_length'] if kind == 'i' else usr['sgl_wg_length']
    nt = int(sim_time / dt)
    pp = int(usr['pitch'] / dx)
    ph = int(np.ceil(pp / 2))
    N = usr['num_cpml']
    J, nt = None, nt
    common = dict(scheme=scheme, precision=prec, delta_x=delta_x, delta_t=delta_t, eps=eps, mu=mu, eta=eta,
                  eps_rel_fg=eps_rel_fg, num_cpml=N, pp=pp, ph=ph, num_lines=usr['num_lines'],
                  sparam_freqs=np.linspace(usr['sparam_fmin'], usr['sparam_fmax'], num=usr['sparam_num_freqs']),
                  rstd=usr['rough_std'] / dx, racl=usr['rough_acl'] / dx)
    J, nt = source_waveform(prec, nt, delta_t, usr['source_type'], usr['source_amp'], usr['f0'],
                            usr['wave_packet_bw'], usr['gauss_pulse_deg'])
    if scheme in ('tm', 'te'):
        buffer_yhat = int(usr['ny_clad_wl'] * wl_clad / dx)
        nx_swg = int(length / dx)
        ny_swg = int(usr['sgl_wg_height'] / dx)
        bx = N + usr['buffer_xhat']
        by = N + buffer_yhat
        if buffer_yhat < pp // 2:
            raise Exception('Boundary conditions too small for WG pitch setting. Reduce pitch or increase ny_clad_wl')
        sx = bx
        p1x = sx + int(usr['port1_length'] / dx)
        p2x = sx + nx_swg - int(usr['port2_length'] / dx)
        if p2x < p1x:
            raise Exception('Port settling ranges overlap. Make the WG longer or decrease port settling range!')
        return Setup(nx=2 * bx + nx_swg, ny=2 * by + (usr['num_lines'] - 1) * pp + ny_swg, nt=nt, J_SRC=J,
                     bx=bx, by=by, ny_swg=ny_swg, cy_swg=ny_swg // 2, sx=sx, p1x=p1x, p2x=p2x,
                     extra={'nx_swg': nx_swg, 'buffer_yhat': buffer_yhat}, **common)
    buffer_xhat = int(usr['nx_clad_wl'] * wl_clad / dx)
    buffer_yhat = int(usr['ny_clad_wl'] * wl_clad / dx)
    nx_swg = int(usr['sgl_wg_height'] / dx)
    ny_swg = int(usr['sgl_wg_width'] / dx)
    nz_swg = int(length / dx)
    bx, by, bz = N + buffer_xhat, N + buffer_yhat, N + usr['buffer_zhat']
    if buffer_xhat < pp // 2:
        raise Exception('Boundary conditions too small for WG pitch setting. Reduce pitch or increase nx_clad_wl')
    sz = bz
    p1z = sz + int(usr['port1_length'] / dx)
    p2z = sz + nz_swg - int(usr['port2_length'] / dx)
    if p2z <= p1z:
        raise Exception('Port settling ranges overlap. Make the WG longer or decrease port settling range!')
    return Setup(nx=2 * bx + (usr['num_lines'] - 1) * pp + nx_swg, ny=2 * by + ny_swg, nz=2 * bz + nz_swg,
                 nt=nt, J_SRC=J, bx=bx, by=by, bz=bz, nx_swg=nx_swg, ny_swg=ny_swg, cx_swg=nx_swg // 2,
                 cy_swg=ny_swg // 2, sz=sz, p1z=p1z, p2z=p2z,
                 source_condition=usr.get('source_condition', 'TE'),
                 extra={'nz_swg': nz_swg, 'buffer_xhat': buffer_xhat, 'buffer_yhat': buffer_yhat}, **common)


def explicit_setup(p):
    """Setup from an explicit description (keys as in ``oidt_amd.configs``)."""
    prec = _precision(p['precision'])
    eps = p['eps_rel_bg'] / EPS_SCALE
    mu = MU0
    c_bg = 1 / np.sqrt(eps * mu)
    delta_x = prec(p['delta_x'])
    delta_t = prec(p['courant'] * p['delta_x'] / c_bg)
    J, nt = source_waveform(prec, p['nt'], delta_t, p['source_type'], p['source_amp'], p['f0'],
                            p['wave_packet_bw'], p.get('gauss_pulse_deg', -6))
    common = dict(scheme=p['scheme'], precision=prec, delta_x=delta_x, delta_t=delta_t, eps=eps, mu=mu,
                  eta=np.sqrt(mu / eps), eps_rel_fg=p['eps_rel_fg'] / p['eps_rel_bg'], num_cpml=p['num_cpml'],
                  nx=p['nx'], ny=p['ny'], nt=nt, J_SRC=J,
                  sparam_freqs=np.linspace(p['fmin'], p['fmax'], num=p['nf']), pp=p['pp'],
                  ph=int(np.ceil(p['pp'] / 2)), bx=p['bx'], by=p['by'], num_lines=p.get('num_lines', 1),
                  m_grade=p['m_grade'], m_alpha=p['m_alpha'], kappa_max=p['kappa_max'],
                  alpha_max=p['alpha_max'], sigma_factor=p.get('sigma_factor', 1.2),
                  rstd=p.get('rough_std', 0.0) / float(delta_x), racl=p.get('rough_acl', 0.0) / float(delta_x))
    if p['scheme'] in ('tm', 'te'):
        return Setup(ny_swg=p['ny_swg'], cy_swg=p['ny_swg'] // 2, sx=p['sx'], p1x=p['p1'], p2x=p['p2'], **common)
    return Setup(nz=p['nz'], bz=p['bz'], nx_swg=p['nx_swg'], ny_swg=p['ny_swg'], cx_swg=p['nx_swg'] // 2,
                 cy_swg=p['ny_swg'] // 2, sz=p['sz'], p1z=p['p1'], p2z=p['p2'],
                 source_condition=p.get('source_condition', 'TE'), **common)


def kind_of(direction: str, kind: str) -> Optional[str]:
    if direction not in ('f', 'b') or kind not in ('i', 'r'):
        raise ValueError('direction must be f|b and kind i|r')
    return kind
