"""Workload definitions: the benchmark configurations and small parity cases.

All are explicit-grid descriptions consumed by ``params.explicit_setup``.
Shared conventions: eps0 = 1/(36 pi 1e9), mu0 = 4 pi 1e-7 (c0 = 3e8 exactly,
fdtd_auto_setup.py:62-63); Si core n = 3.47 in SiO2 n = 1.44; dx = 10 nm;
dt = S dx / c_clad with S = 0.99/sqrt(D); CPML 10 cells, polynomial grading
m = 3, m_alpha = 1, kappa_max = 5, alpha_max = 0.05; wave-packet source at
1550 nm with 40 % fractional bandwidth, amplitude 5; DFT frequencies
linspace(0.8 f0, 1.2 f0, nf).  A single line is simulated; the extraction
window (pp rows) spans the whole region between the CPML slabs.
"""
import numpy as np

from . import params
from . import roughness

C0 = 3.0e8
F0 = C0 / 1550e-9
N_CORE, N_CLAD = 3.47, 1.44
DX = 10e-9

_COMMON = dict(delta_x=DX, eps_rel_bg=N_CLAD ** 2, eps_rel_fg=N_CORE ** 2, num_cpml=10, m_grade=3, m_alpha=1,
               kappa_max=5.0, alpha_max=0.05, source_type=2, source_amp=5.0, f0=F0, wave_packet_bw=0.40,
               fmin=0.8 * F0, fmax=1.2 * F0, num_lines=1)


def slab2d(nx, ny=240, nt=4000, nf=32, precision='float64', scheme='tm', core=22, port_len=100, buffer_x=10,
           **over):
    """2D slab waveguide: core rows centred in y, source at x = bx, ports port_len
    cells inside the guide ends."""
    N = _COMMON['num_cpml']
    bx = N + buffer_x
    by = (ny - core) // 2
    p = dict(_COMMON, scheme=scheme, precision=precision, nx=nx, ny=ny, courant=0.99 / np.sqrt(2), bx=bx,
             by=by, ny_swg=core, pp=ny - 2 * N, sx=bx, p1=bx + port_len, p2=nx - bx - port_len, nt=nt, nf=nf)
    p.update(over)
    return p


def strip3d(nz, nx=128, ny=96, nt=6000, nf=64, precision='float32', width=50, height=22, port_len=80,
            buffer_z=10, source_condition='TE', **over):
    """3D strip waveguide: x across the 500 nm width (the roughened axis of the
    reference's 3D mask), y across the 220 nm height, z along propagation."""
    N = _COMMON['num_cpml']
    bz = N + buffer_z
    p = dict(_COMMON, scheme='3d', precision=precision, nx=nx, ny=ny, nz=nz, courant=0.99 / np.sqrt(3),
             bx=(nx - width) // 2, by=(ny - height) // 2, bz=bz, nx_swg=width, ny_swg=height, pp=nx - 2 * N,
             sz=bz, p1=bz + port_len, p2=nz - bz - port_len, nt=nt, nf=nf, source_condition=source_condition)
    p.update(over)
    return p


# --- benchmark configurations ----------------------------------------------

def config0():
    """Plain 2D TMz, FP64, 600 x 240, 4000 steps, 32 frequencies; sigma 5 nm, Lc 50 nm, seed 0."""
    return dict(slab2d(600, precision='float64'), rough_std=5e-9, rough_acl=50e-9, seed=0)


def config1():
    """Plain 3D FP32 strip, (x, y, z) = (128, 96, 320), 6000 steps, 64 frequencies; sigma 5 nm, Lc 50 nm, seed 1."""
    return dict(strip3d(320), rough_std=5e-9, rough_acl=50e-9, seed=1)


def config2(count=64, nt=5000, scheme='tm'):
    """Ensemble of 2D FP32 realisations: nx ~ U{800..4000}, ny = 240, sigma ~ U(2,10) nm, Lc ~ U(20,100) nm
    (TMz; the same draws in TEz with ``scheme='te'``)."""
    rng = np.random.default_rng(2)
    out = []
    for r in range(count):
        nx = int(rng.integers(800, 4001))
        out.append(dict(slab2d(nx, precision='float32', nt=nt, scheme=scheme), rough_std=float(rng.uniform(2e-9, 10e-9)),
                        rough_acl=float(rng.uniform(20e-9, 100e-9)), seed=1000 + r))
    return out


def config3(count=16, nt=5000):
    """Ensemble of 3D FP32 realisations: (96, 64, nz ~ U{192..512}), sigma ~ U(3,8) nm, Lc ~ U(30,80) nm."""
    rng = np.random.default_rng(3)
    out = []
    for r in range(count):
        nz = int(rng.integers(192, 513))
        out.append(dict(strip3d(nz, nx=96, ny=64, nt=nt, port_len=40), rough_std=float(rng.uniform(3e-9, 8e-9)),
                        rough_acl=float(rng.uniform(30e-9, 80e-9)), seed=2000 + r))
    return out


def serving_requests(count=256, seed=4):
    """Request mix: 70 % 2D (nx 400..4000, ny 200..240), 30 % 3D ((128,64,48)..(384,128,96)),
    2000..8000 steps, random roughness.  3D cross-sections keep x as the width axis."""
    rng = np.random.default_rng(seed)
    out = []
    for r in range(count):
        steps = int(rng.integers(2000, 8001))
        if rng.random() < 0.7:
            ny = int(rng.integers(200, 241))
            nx = int(rng.integers(400, 4001))
            p = dict(slab2d(nx, ny=ny, precision='float32', nt=steps, port_len=min(100, nx // 8)),
                     rough_std=float(rng.uniform(2e-9, 10e-9)), rough_acl=float(rng.uniform(20e-9, 100e-9)))
        else:
            nx = int(rng.integers(128, 385))
            ny = int(rng.integers(64, 129))
            nz = int(rng.integers(48, 97))
            p = dict(strip3d(nz, nx=nx, ny=ny, nt=steps, port_len=max(4, nz // 6), buffer_z=4, nf=64),
                     rough_std=float(rng.uniform(3e-9, 8e-9)), rough_acl=float(rng.uniform(30e-9, 80e-9)))
        p['seed'] = 4000 + r
        out.append(p)
    return out


# --- small parity cases ----------------------------------------------------

def small(scheme, precision='float64', nt=None, **over):
    """Tiny grids that the FP64 NumPy oracle finishes in seconds.  The source is a
    broadband Gaussian pulse (type 1, -1 dB at f0) so fields are large early."""
    if scheme == '3d':
        p = strip3d(56, nx=36, ny=34, nt=nt or 260, nf=6, precision=precision, width=12, height=8, port_len=10,
                    buffer_z=6, source_type=1, gauss_pulse_deg=-1.0)
    else:
        p = slab2d(140, ny=72, nt=nt or 700, nf=8, precision=precision, scheme=scheme, core=12, port_len=25,
                   source_type=1, gauss_pulse_deg=-1.0)
    p.update(over)
    return p


def incident(p):
    """Description of the incident (kind 'i') run paired with ``p``: the guide is
    shortened to the two port settling lengths and both port planes fall on
    the same cell, while nt is kept (fdtd_auto_setup.py:82-91, :161-162)."""
    q = dict(p)
    q.pop('rough_std', None)
    q.pop('rough_acl', None)
    port_len = p['p1'] - (p['sz'] if p['scheme'] == '3d' else p['sx'])
    b = p['bz'] if p['scheme'] == '3d' else p['bx']
    n = 2 * b + 2 * port_len
    q['nz' if p['scheme'] == '3d' else 'nx'] = n
    q['p1'] = q['p2'] = b + port_len
    return q


# --- building a realisation ------------------------------------------------

def realise(p):
    """(Setup, upper profile, lower profile) of a description; profiles are None
    for a smooth guide (no rough_std or rough_std == 0)."""
    s = params.explicit_setup(p)
    if not p.get('rough_std'):
        return s, None, None
    length = (s.p2z - s.p1z) if s.scheme == '3d' else (s.p2x - s.p1x)
    up, lo = roughness.profiles(length, p['rough_std'] / float(s.delta_x), p['rough_acl'] / float(s.delta_x),
                                p.get('seed', 0))
    return s, up, lo


def cell_updates(p):
    n = p['nx'] * p['ny'] * (p['nz'] if p['scheme'] == '3d' else 1)
    return n * p['nt']
