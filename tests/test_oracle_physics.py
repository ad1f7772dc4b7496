"""Pins of the oracle time stepping by what the mathematics and physics fix.

The reference ships no golden vectors (there are no tests in it), so the
FP64 oracle is checked against: zero-source and linearity invariants, the
discrete divergence identities of the Yee curls (a sign or index error in any
curl term breaks them), exact forward/backward mirror symmetry of the TMz
scheme (pins the CPML and kappa index conventions), CPML reflection against
an enlarged domain, the analytic slab-waveguide effective index (the
reference's own validation quantity, aux_funcs.find_neff), and a per-cell loop
transcription of the reference kernels on tiny grids."""
import numpy as np
import pytest
from scipy.optimize import brentq

from oidt_amd import configs
from oracle import fdtd as ofdtd
from oracle import geometry as ogeo
from oracle import setup as osetup


def run(p, direction='f', up=None, lo=None, nsteps=None):
    c = osetup.explicit_setup(p)
    o = ofdtd.make(c, ogeo.fg_mask(c, up, lo), direction)
    o.run(nsteps)
    return o


@pytest.mark.parametrize('scheme', ['tm', 'te', '3d'])
def test_zero_source_zero_fields(scheme):
    o = run(configs.small(scheme, source_amp=0.0, nt=80))
    for a in o.fields().values():
        assert not np.any(a)


@pytest.mark.parametrize('scheme', ['tm', 'te', '3d'])
def test_linear_in_source_amplitude(scheme):
    p = configs.small(scheme, nt=120)
    a = run(dict(p, source_amp=1.0))
    b = run(dict(p, source_amp=3.0))
    for k, v in a.fields().items():
        assert np.abs(v).max() > 0
        np.testing.assert_allclose(b.fields()[k], 3.0 * v, rtol=1e-9, atol=1e-12 * np.abs(v).max())


def _interior(c, margin):
    N = c.num_cpml
    lo = 2 + N + margin
    return slice(lo, c.nx - lo), slice(lo, c.ny - lo)


def test_3d_discrete_div_h_vanishes():
    """div H (backward differences) is identically zero where kappa = 1 and the
    source drives E only (TE source on Ey)."""
    p = configs.small('3d', nt=150)
    o = run(p)
    Hx, Hy, Hz = o.Hx, o.Hy, o.Hz
    div = (Hx[1:, 1:, 1:] - Hx[:-1, 1:, 1:]) + (Hy[1:, 1:, 1:] - Hy[1:, :-1, 1:]) + (Hz[1:, 1:, 1:] - Hz[1:, 1:, :-1])
    c = o.c
    m = 2 + c.num_cpml + 2
    inner = div[m:c.nx - m - 1, m:c.ny - m - 1, m:c.nz - m - 1]
    scale = max(np.abs(Hx).max(), np.abs(Hy).max(), np.abs(Hz).max())
    assert scale > 0
    assert np.abs(inner).max() < 1e-11 * scale


def test_3d_div_h_detects_a_sign_error():
    """The same identity fails for a plausible mistake planted in the oracle:
    the x-derivative term of the Hz update with the wrong sign."""
    p = configs.small('3d', nt=150)
    c = osetup.explicit_setup(p)
    o = ofdtd.make(c, ogeo.fg_mask(c))
    real = o.step

    def bad_step(n):
        real(n)
        ib = jb = kb = 2
        ie, je, ke = c.nx - 2, c.ny - 2, c.nz - 2
        s = (slice(ib + 1, ie), slice(jb + 1, je), slice(kb, ke))
        # Hz += db (dEx/dy + dEy/dx) instead of db (dEx/dy - dEy/dx)
        o.Hz[s] += 2 * o.db * (o.Ey[s] - o.Ey[ib:ie - 1, jb + 1:je, kb:ke]) / o.kxh[ib + 1:ie, None, None]
    o.step = bad_step
    o.run(150)
    div = (o.Hx[1:, 1:, 1:] - o.Hx[:-1, 1:, 1:]) + (o.Hy[1:, 1:, 1:] - o.Hy[1:, :-1, 1:]) + \
          (o.Hz[1:, 1:, 1:] - o.Hz[1:, 1:, :-1])
    m = 2 + c.num_cpml + 2
    scale = max(np.abs(o.Hx).max(), np.abs(o.Hz).max())
    # the correct scheme keeps this below 1e-11 * scale
    assert np.abs(div[m:c.nx - m - 1, m:c.ny - m - 1, m:c.nz - m - 1]).max() > 1e-8 * scale


def test_tmz_discrete_div_h_vanishes():
    o = run(configs.small('tm', nt=300))
    d = (o.Hx[1:, 1:] - o.Hx[:-1, 1:]) + (o.Hy[1:, 1:] - o.Hy[1:, :-1])
    c = o.c
    m = 2 + c.num_cpml + 2
    assert np.abs(o.Hx).max() > 0
    assert np.abs(d[m:c.nx - m - 1, m:c.ny - m - 1]).max() < 1e-11 * np.abs(o.Hy).max()


def test_tez_discrete_div_d_vanishes():
    """div(E / cb) = 0 away from the CPML: the TE source drives Hz only, and the
    per-component averaged permittivities enter through cb."""
    p = configs.small('te', nt=300)
    p.update(rough_std=15e-9, rough_acl=40e-9, seed=2)
    s, up, lo = configs.realise(p)
    o = run(p, up=up, lo=lo)
    cbx, cby = o.cb
    Dx, Dy = o.Ex / cbx, o.Ey / cby
    # E is updated from forward differences of Hz, so div D uses forward differences
    d = (Dx[1:, :-1] - Dx[:-1, :-1]) + (Dy[:-1, 1:] - Dy[:-1, :-1])
    c = o.c
    m = 2 + c.num_cpml + 2
    scale = max(np.abs(Dx).max(), np.abs(Dy).max())
    assert np.abs(d[m:c.nx - m - 1, m:c.ny - m - 1]).max() < 1e-10 * scale


def test_tmz_forward_backward_mirror_symmetry():
    """Smooth guide with ports mirrored about the grid centre: the backward run
    is the exact mirror image of the forward run (Ez even, Hy odd)."""
    p = configs.small('tm', nt=500)
    p['p2'] = p['nx'] - 1 - p['p1']
    f = run(p, 'f')
    b = run(p, 'b')
    nx = p['nx']
    np.testing.assert_allclose(b.Ez, f.Ez[::-1, :], rtol=0, atol=1e-12 * np.abs(f.Ez).max())
    # Hy[i] sits at x = i - 1/2; its mirror is -Hy[nx - i]
    np.testing.assert_allclose(b.Hy[1:, :], -f.Hy[1:, :][::-1, :], rtol=0, atol=1e-12 * np.abs(f.Hy).max())
    fd, bd = f.dft(), b.dft()
    np.testing.assert_allclose(bd[('e', 2)], fd[('e', 1)], rtol=1e-9, atol=1e-12 * np.abs(fd[('e', 1)]).max())
    np.testing.assert_allclose(bd[('h', 2)], -fd[('h', 1)], rtol=1e-9, atol=1e-12 * np.abs(fd[('h', 1)]).max())


def test_cpml_reflection_below_minus_40db():
    """Homogeneous 2D TMz: probe near the CPML in a small grid against the same
    probe in a grid large enough that nothing returns within the window."""
    base = configs.slab2d(100, ny=100, nt=320, nf=1, precision='float64', core=4, source_type=1,
                          gauss_pulse_deg=-1.0, eps_rel_fg=1.44 ** 2)
    base.update(sx=50, by=48, p1=20, p2=80, bx=20)
    big = dict(base, nx=400, ny=400, sx=200, by=198, p1=170, p2=230, bx=170)
    cs, cb_ = osetup.explicit_setup(base), osetup.explicit_setup(big)
    small = ofdtd.make(cs, ogeo.fg_mask(cs))
    large = ofdtd.make(cb_, ogeo.fg_mask(cb_))
    probe_s, probe_l = (86, 50), (86 + 150, 50 + 150)
    ts, tl = [], []
    for n in range(320):
        small.step(n)
        large.step(n)
        ts.append(small.Ez[probe_s])
        tl.append(large.Ez[probe_l])
    ts, tl = np.array(ts), np.array(tl)
    assert np.abs(tl).max() > 0
    refl = np.abs(ts - tl).max() / np.abs(tl).max()
    assert refl < 1e-2, refl          # below -40 dB


def test_cpml_reflection_test_detects_a_bad_layer():
    """With the psi correction sign flipped the layer reflects strongly."""
    base = configs.slab2d(100, ny=100, nt=320, nf=1, precision='float64', core=4, source_type=1,
                          gauss_pulse_deg=-1.0, eps_rel_fg=1.44 ** 2)
    base.update(sx=50, by=48, p1=20, p2=80, bx=20)
    big = dict(base, nx=400, ny=400, sx=200, by=198, p1=170, p2=230, bx=170)
    cs, cb_ = osetup.explicit_setup(base), osetup.explicit_setup(big)
    small = ofdtd.make(cs, ogeo.fg_mask(cs))
    small.ce = -small.ce
    small.ch = -small.ch
    large = ofdtd.make(cb_, ogeo.fg_mask(cb_))
    ts, tl = [], []
    for n in range(320):
        small.step(n)
        large.step(n)
        ts.append(small.Ez[86, 50])
        tl.append(large.Ez[236, 200])
    refl = np.abs(np.array(ts) - np.array(tl)).max() / np.abs(tl).max()
    assert refl > 1e-2


def slab_te0_neff(n1, n2, t, lam):
    """Fundamental even mode of a symmetric slab, E parallel to the interfaces:
    kappa t = 2 atan(gamma / kappa)."""
    k0 = 2 * np.pi / lam

    def f(neff):
        kap = k0 * np.sqrt(n1 ** 2 - neff ** 2)
        gam = k0 * np.sqrt(neff ** 2 - n2 ** 2)
        return kap * t - 2 * np.arctan(gam / kap)
    return brentq(f, n2 + 1e-9, n1 - 1e-9)


@pytest.mark.slow
def test_slab_mode_phase_between_ports_matches_analytic_neff():
    """E phasor phase between the two ports on the core axis equals
    -beta L of the analytic TE0 slab mode (TMz: Ez parallel to the slab)."""
    p = configs.slab2d(500, ny=160, nt=3600, nf=6, precision='float64', port_len=130)
    o = run(p)
    c = o.c
    jc = c.by + c.ny_swg // 2
    d = o.dft()
    L = (c.p2x - c.p1x) * float(c.delta_x)
    for fi, f in enumerate(c.sparam_freqs):
        lam = 3e8 / f
        neff = slab_te0_neff(3.47, 1.44, c.ny_swg * float(c.delta_x), lam)
        expect = -2 * np.pi * neff / lam * L
        got = np.angle(d[('e', 2)][fi, jc] / d[('e', 1)][fi, jc])
        err = np.angle(np.exp(1j * (got - expect)))
        assert abs(err) < 0.15, (f, neff, err)


def test_tmz_matches_per_cell_loop_transcription():
    """The vectorised oracle equals a per-thread loop transcription of the
    reference kernels (fdtd_funcs.py:44-144) in launch order."""
    p = configs.slab2d(40, ny=30, nt=25, nf=2, precision='float64', core=6, port_len=5, buffer_x=2,
                       source_type=1, gauss_pulse_deg=-1.0, num_cpml=4)
    p.update(by=12, pp=22, bx=6, sx=6, p1=11, p2=28)
    c = osetup.explicit_setup(p)
    FG = ogeo.fg_mask(c)
    o = ofdtd.make(c, FG)
    o.run(25)
    nx, ny, N = c.nx, c.ny, c.num_cpml
    ib = jb = 2
    ie, je = nx - 2, ny - 2
    Ez, Hx, Hy = np.zeros((nx, ny)), np.zeros((nx, ny)), np.zeros((nx, ny))
    PEzxl, PEzxh, PHyxl, PHyxh = (np.zeros((N, ny)) for _ in range(4))
    PEzyl, PEzyh, PHxyl, PHxyh = (np.zeros((nx, N)) for _ in range(4))
    be, ce, bh, ch = o.be, o.ce, o.bh, o.ch
    kxe, kye, kxh, kyh = o.kxe, o.kye, o.kxh, o.kyh
    (cb,) = o.cb
    db, dx = o.db, o.dx
    J = o.J
    for n in range(25):
        for i in range(nx):
            for j in range(ny):
                if ib <= i < ie and jb <= j < je:
                    Ez[i, j] = Ez[i, j] + cb[i, j] * ((Hy[i + 1, j] - Hy[i, j]) / kxe[i] - (Hx[i, j + 1] - Hx[i, j]) / kye[j])
        for i in range(c.sx, c.sx + 1):
            for j in range(c.by + 1, c.by + c.ny_swg - 1):
                Ez[i, j] = Ez[i, j] - J[n]
        for i in range(N):
            for j in range(jb, je):
                PEzxl[i, j] = be[-(i + 1)] * PEzxl[i, j] + ce[-(i + 1)] * (Hy[ib + i + 1, j] - Hy[ib + i, j]) / dx
                Ez[ib + i, j] = Ez[ib + i, j] + cb[ib + i, j] * PEzxl[i, j]
                PEzxh[i, j] = be[i] * PEzxh[i, j] + ce[i] * (Hy[ie - N + i + 1, j] - Hy[ie - N + i, j]) / dx
                Ez[ie - N + i, j] = Ez[ie - N + i, j] + cb[ie - N + i, j] * PEzxh[i, j]
        for i in range(ib, ie):
            for j in range(N):
                PEzyl[i, j] = be[-(j + 1)] * PEzyl[i, j] + ce[-(j + 1)] * (Hx[i, jb + j + 1] - Hx[i, jb + j]) / dx
                Ez[i, jb + j] = Ez[i, jb + j] - cb[i, jb + j] * PEzyl[i, j]
                PEzyh[i, j] = be[j] * PEzyh[i, j] + ce[j] * (Hx[i, je - N + j + 1] - Hx[i, je - N + j]) / dx
                Ez[i, je - N + j] = Ez[i, je - N + j] - cb[i, je - N + j] * PEzyh[i, j]
        for i in range(ib, ie):
            for j in range(jb + 1, je):
                Hx[i, j] = Hx[i, j] + db * (Ez[i, j - 1] - Ez[i, j]) / kyh[j]
        for i in range(ib + 1, ie):
            for j in range(jb, je):
                Hy[i, j] = Hy[i, j] + db * (Ez[i, j] - Ez[i - 1, j]) / kxh[i]
        for i in range(ib, ie):
            for j in range(1, N):
                PHxyl[i, j] = bh[-j] * PHxyl[i, j] + ch[-j] * (Ez[i, jb + j] - Ez[i, jb + j - 1]) / dx
                Hx[i, jb + j] = Hx[i, jb + j] - db * PHxyl[i, j]
                PHxyh[i, j] = bh[j] * PHxyh[i, j] + ch[j] * (Ez[i, je - N + j] - Ez[i, je - N + j - 1]) / dx
                Hx[i, je - N + j] = Hx[i, je - N + j] - db * PHxyh[i, j]
        for i in range(1, N):
            for j in range(jb, je):
                PHyxl[i, j] = bh[-i] * PHyxl[i, j] + ch[-i] * (Ez[ib + i, j] - Ez[ib + i - 1, j]) / dx
                Hy[ib + i, j] = Hy[ib + i, j] + db * PHyxl[i, j]
                PHyxh[i, j] = bh[i] * PHyxh[i, j] + ch[i] * (Ez[ie - N + i, j] - Ez[ie - N + i - 1, j]) / dx
                Hy[ie - N + i, j] = Hy[ie - N + i, j] + db * PHyxh[i, j]
    assert np.abs(Ez).max() > 0
    np.testing.assert_allclose(o.Ez, Ez, rtol=0, atol=1e-13 * np.abs(Ez).max())
    np.testing.assert_allclose(o.Hx, Hx, rtol=0, atol=1e-13 * np.abs(Hx).max())
    np.testing.assert_allclose(o.Hy, Hy, rtol=0, atol=1e-13 * np.abs(Hy).max())
