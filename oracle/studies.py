"""Oracle of the loss and wave-impedance study drivers (test infrastructure only).

FP64 NumPy restatement of the study time loops, each kernel one statement
group in the reference launch order:

* TEz studies  FDTD/GPU/FDTD_2D_TE/scattering_loss/fdtd_funcs.py:79-167, loop
               fdtd_main.py:202-223 (H first: update_hz, update_pml_hz_yinc,
               update_pml_hz_xinc, update_source (+J), update_ex, update_ey,
               update_pml_ex_yinc, update_pml_ey_xinc, then recording);
               recording simul_fft (:113-120) or map_*_zwave
               (FDTD_2D_TE/wave_impedance/fdtd_funcs.py:169-181, loop :220-221).
* 3D studies   the S-parameter 3D step (``fdtd.Yee3D``; the study kernels are
               the same statements, FDTD_3D/scattering_loss/fdtd_funcs.py:
               15-215) followed by simul_fft_efield / simul_fft_hfield at both
               port planes (:217-242, loop fdtd_main.py:341-344) or
               map_efield_zwave / map_hfield_zwave along x
               (FDTD_3D/wave_impedance/fdtd_funcs.py:321-335, loop :298-299).

The coefficient arrays are those of the study drivers: CPML grading and kappa
axes as the S-parameter driver (FDTD_2D_TE/scattering_loss/fdtd_main.py:
75-124 holds the same statements), TEz material maps from the float32
EPS_MASK (:37-47).  Inputs are a study configuration namespace (the names of
the study's fdtd_config.py) and the two roughness profiles.  Recorded values
and DFT sums are float64 here; the reference keeps float32 accumulators.
"""
from types import SimpleNamespace

import numpy as np

from . import fdtd as F
from . import geometry as G
from . import setup as S


def namespace(cfg):
    """Oracle setup namespace (the attribute set ``setup.coefficients`` and
    ``geometry.fg_mask`` read) from a study configuration."""
    c = SimpleNamespace(scheme=cfg.scheme, precision=np.float32, eps=cfg.eps, mu=cfg.mu, eta=cfg.eta,
                        eps_rel_fg=cfg.eps_rel_fg, delta_x=np.float32(cfg.delta_x),
                        delta_t=np.float32(cfg.delta_t), num_cpml=cfg.num_cpml, num_lines=1, offset=2,
                        m_grade=4, m_alpha=1, kappa_max=5, alpha_max=0.05, sigma_factor=1.2,
                        sparam_freqs=np.zeros(0), pp=0, ph=0, bx=cfg.bx, by=cfg.by, nx=cfg.nx, ny=cfg.ny,
                        nt=cfg.nt, J_SRC=np.asarray(cfg.J_SRC, np.float32), ny_swg=cfg.ny_swg)
    if cfg.scheme == 'te':
        c.__dict__.update(sx=cfg.sx, p1x=cfg.p1x, p2x=cfg.p2x, cy_swg=cfg.cy_swg)
    else:
        c.__dict__.update(nz=cfg.nz, bz=cfg.bz, nx_swg=cfg.nx_swg, cx_swg=cfg.cx_swg, cy_swg=cfg.cy_swg,
                          sz=cfg.sz, p1z=cfg.p1z, p2z=cfg.p2z, source_condition=cfg.source_condition)
    return c


def te_cb_maps(c, FG):
    """CBX, CBY of FDTD_2D_TE/scattering_loss/fdtd_main.py:37-47, literal loop."""
    EPS_MASK = (c.eps * (c.eps_rel_fg * FG + 1 * ~FG)).astype(np.float32)
    EPS_EX = EPS_MASK.copy()
    EPS_EY = EPS_MASK.copy()
    for j in range(1, c.ny):
        for i in range(1, c.nx):
            EPS_EX[i, j] = 0.5 * float(EPS_MASK[i, j] + EPS_MASK[i - 1, j])
            EPS_EY[i, j] = 0.5 * float(EPS_MASK[i, j] + EPS_MASK[i, j - 1])
    return c.delta_t / EPS_EX, c.delta_t / EPS_EY


def simul_fft_weight(f0, step, dt):
    """cos / sin argument of simul_fft: freq, step, dt arrive as float32 and
    2*np.pi*freq*step*dt is evaluated in float64."""
    return 2 * np.pi * float(np.float32(f0)) * float(np.float32(step)) * float(np.float32(dt))


class TEzStudy:
    """2D TEz study loop (H first), FP64."""

    def __init__(self, cfg, upper=None, lower=None):
        c = namespace(cfg)
        S.coefficients(c)
        self.cfg, self.c = cfg, c
        f = np.float64
        self.FG = G.fg_mask(c, upper, lower)
        self.cbx, self.cby = [a.astype(f) for a in te_cb_maps(c, self.FG)]
        self.N, self.dx, self.db = c.num_cpml, float(c.delta_x), float(c.DB)
        self.J = c.J_SRC.astype(f)
        self.be, self.ce, self.bh, self.ch = (c.BE.astype(f), c.CE.astype(f), c.BH.astype(f), c.CH.astype(f))
        self.kxe, self.kxh, self.kye, self.kyh = (c.kxe.astype(f), c.kxh.astype(f), c.kye.astype(f),
                                                  c.kyh.astype(f))
        nx, ny, N = c.nx, c.ny, self.N
        z = lambda *s: np.zeros(s)
        self.Ex, self.Ey, self.Hz = z(nx, ny), z(nx, ny), z(nx, ny)
        self.p_ex_ylo, self.p_ex_yhi = z(nx, N), z(nx, N)
        self.p_ey_xlo, self.p_ey_xhi = z(N, ny), z(N, ny)
        self.p_hz_xlo, self.p_hz_xhi = z(N, ny), z(N, ny)
        self.p_hz_ylo, self.p_hz_yhi = z(nx, N), z(nx, N)
        self.rec = {}

    def step(self, n):
        c, N, dx, db = self.c, self.N, self.dx, self.db
        Ex, Ey, Hz = self.Ex, self.Ey, self.Hz
        ib, jb, ie, je = 2, 2, c.nx - 2, c.ny - 2
        I, J, I1, J1 = slice(ib, ie), slice(jb, je), slice(ib + 1, ie), slice(jb + 1, je)
        # update_hz: i in (ib,ie), j in (jb,je)
        Hz[I1, J1] += db * ((Ex[I1, J1] - Ex[I1, jb:je - 1]) / self.kyh[None, J1]
                            - (Ey[I1, J1] - Ey[ib:ie - 1, J1]) / self.kxh[I1, None])
        jj = np.arange(1, N)
        # update_pml_hz_yinc: i in (ib,ie), j in (0,pmle)
        lo, hi = jb + jj, je - N + jj
        self.p_hz_ylo[I1, 1:] = (self.bh[N - jj][None, :] * self.p_hz_ylo[I1, 1:]
                                 + self.ch[N - jj][None, :] * (Ex[I1, lo] - Ex[I1, lo - 1]) / dx)
        Hz[I1, lo] += db * self.p_hz_ylo[I1, 1:]
        self.p_hz_yhi[I1, 1:] = (self.bh[jj][None, :] * self.p_hz_yhi[I1, 1:]
                                 + self.ch[jj][None, :] * (Ex[I1, hi] - Ex[I1, hi - 1]) / dx)
        Hz[I1, hi] += db * self.p_hz_yhi[I1, 1:]
        # update_pml_hz_xinc: i in (0,pmle), j in (jb,je)
        lo, hi = ib + jj, ie - N + jj
        self.p_hz_xlo[1:, J1] = (self.bh[N - jj][:, None] * self.p_hz_xlo[1:, J1]
                                 + self.ch[N - jj][:, None] * (Ey[lo, J1] - Ey[lo - 1, J1]) / dx)
        Hz[lo, J1] -= db * self.p_hz_xlo[1:, J1]
        self.p_hz_xhi[1:, J1] = (self.bh[jj][:, None] * self.p_hz_xhi[1:, J1]
                                 + self.ch[jj][:, None] * (Ey[hi, J1] - Ey[hi - 1, J1]) / dx)
        Hz[hi, J1] -= db * self.p_hz_xhi[1:, J1]
        # update_source: Hz[sx, by+2 : by+ny_swg-2] += J_SRC[n]
        Hz[c.sx, c.by + 2:c.by + c.ny_swg - 2] += self.J[n]
        # update_ex: i in (ib,ie), j in [jb,je)
        Ex[I1, J] += self.cbx[I1, J] * (Hz[I1, jb + 1:je + 1] - Hz[I1, J]) / self.kye[None, J]
        # update_ey: i in [ib,ie), j in (jb,je)
        Ey[I, J1] += self.cby[I, J1] * (Hz[I, J1] - Hz[ib + 1:ie + 1, J1]) / self.kxe[I, None]
        d = np.arange(N)
        # update_pml_ex_yinc: i in (ib,ie), j in [0,pmle)
        lo, hi = jb + d, je - N + d
        self.p_ex_ylo[I1, :] = (self.be[N - 1 - d][None, :] * self.p_ex_ylo[I1, :]
                                + self.ce[N - 1 - d][None, :] * (Hz[I1, lo + 1] - Hz[I1, lo]) / dx)
        Ex[I1, lo] += self.cbx[I1, lo] * self.p_ex_ylo[I1, :]
        self.p_ex_yhi[I1, :] = (self.be[d][None, :] * self.p_ex_yhi[I1, :]
                                + self.ce[d][None, :] * (Hz[I1, hi + 1] - Hz[I1, hi]) / dx)
        Ex[I1, hi] += self.cbx[I1, hi] * self.p_ex_yhi[I1, :]
        # update_pml_ey_xinc (study form: reversed difference, added correction)
        lo, hi = ib + d, ie - N + d
        self.p_ey_xlo[:, J1] = (self.be[N - 1 - d][:, None] * self.p_ey_xlo[:, J1]
                                + self.ce[N - 1 - d][:, None] * (Hz[lo, J1] - Hz[lo + 1, J1]) / dx)
        Ey[lo, J1] += self.cby[lo, J1] * self.p_ey_xlo[:, J1]
        self.p_ey_xhi[:, J1] = (self.be[d][:, None] * self.p_ey_xhi[:, J1]
                                + self.ce[d][:, None] * (Hz[hi, J1] - Hz[hi + 1, J1]) / dx)
        Ey[hi, J1] += self.cby[hi, J1] * self.p_ey_xhi[:, J1]

    def record(self, n):
        cfg, c = self.cfg, self.c
        if cfg.study == 'scattering_loss':
            if not self.rec:
                self.rec = {k: np.zeros((c.nx, c.ny), complex) for k in ('hz', 'ey')}
            for k, F_, st in (('hz', self.Hz, n + 0.0), ('ey', self.Ey, n + 0.5)):
                w = simul_fft_weight(cfg.f0, st, c.delta_t)
                # Re += F cos, Im += F sin; the file holds Re - 1j*Im (fdtd_main.py:265-266)
                self.rec[k] += F_ * np.cos(w) - 1j * (F_ * np.sin(w))
            return
        if not self.rec:
            self.rec = {k: np.zeros((c.nt, c.ny)) for k in ('ex', 'hz')}
        cx = cfg.cx
        Hz = self.Hz
        self.rec['ex'][n] = 0.5 * (self.Ex[cx, :] + self.Ex[cx + 1, :])
        # Hz[cx, ny] reads past the row: a cell no update writes (zero)
        hz1 = np.append(Hz[cx, 1:], 0.0)
        hz2 = np.append(Hz[cx + 1, 1:], 0.0)
        self.rec['hz'][n] = 0.25 * (Hz[cx, :] + Hz[cx + 1, :] + hz1 + hz2)

    def run(self):
        for n in range(self.c.nt):
            self.step(n)
            self.record(n)
        return self.rec


class Yee3DStudy(F.Yee3D):
    """3D study loop: the S-parameter step followed by the study's recording."""

    def __init__(self, cfg, upper=None, lower=None):
        c = namespace(cfg)
        FG = G.fg_mask(c, upper, lower)
        super().__init__(c, FG)
        self.cfg = cfg
        self.FG = FG
        self.rec = {}

    def record(self, n):
        cfg, c, N = self.cfg, self.c, self.N
        Ex, Ey, Ez, Hx, Hy, Hz = self.Ex, self.Ey, self.Ez, self.Hx, self.Hy, self.Hz
        if cfg.study == 'scattering_loss':
            if not self.rec:
                self.rec = {k + p: np.zeros((c.nx, c.ny), complex)
                            for k in ('ex', 'ey', 'ez', 'hx', 'hy', 'hz') for p in '12'}
            wi, wj = slice(N, c.nx - N), slice(N, c.ny - N)
            wi1, wj1 = slice(N + 1, c.nx - N + 1), slice(N + 1, c.ny - N + 1)
            eE = float(cfg.cosE[n]) + 1j * float(cfg.sinE[n])
            eH = float(cfg.cosH[n]) + 1j * float(cfg.sinH[n])
            for p, kc in (('1', c.p1z), ('2', c.p2z)):
                k1 = kc + 1
                vals = {
                    'ex': 0.5 * (Ex[wi, wj, kc] + Ex[wi1, wj, kc]),
                    'ey': 0.5 * (Ey[wi, wj, kc] + Ey[wi, wj1, kc]),
                    'ez': 0.5 * (Ez[wi, wj, kc] + Ez[wi, wj, k1]),
                    'hx': 0.25 * (Hx[wi, wj, kc] + Hx[wi, wj1, kc] + Hx[wi, wj, k1] + Hx[wi, wj1, k1]),
                    'hy': 0.25 * (Hy[wi, wj, kc] + Hy[wi1, wj, kc] + Hy[wi, wj, k1] + Hy[wi1, wj, k1]),
                    'hz': 0.25 * (Hz[wi, wj, kc] + Hz[wi1, wj, kc] + Hz[wi, wj1, kc] + Hz[wi1, wj1, kc]),
                }
                for k, v in vals.items():
                    # Rx += v cos, Ix += v sin; the file holds R + 1j*I (fdtd_main.py:424-435)
                    self.rec[k + p][wi, wj] += v * (eE if k[0] == 'e' else eH)
            return
        if not self.rec:
            self.rec = {k: np.zeros((c.nt, c.nx)) for k in ('ey', 'ez', 'hy', 'hz')}
        jc, kc, ie = cfg.cy, cfg.cz, c.nx - 2
        i, i1 = slice(0, ie), slice(1, ie + 1)
        self.rec['ey'][n, :ie] = 0.5 * (Ey[i, jc, kc] + Ey[i, jc + 1, kc])
        self.rec['ez'][n, :ie] = 0.5 * (Ez[i, jc, kc] + Ez[i, jc, kc + 1])
        self.rec['hy'][n, :ie] = 0.25 * (Hy[i, jc, kc] + Hy[i1, jc, kc] + Hy[i, jc, kc + 1] + Hy[i1, jc, kc + 1])
        self.rec['hz'][n, :ie] = 0.25 * (Hz[i, jc, kc] + Hz[i1, jc, kc] + Hz[i, jc + 1, kc] + Hz[i1, jc + 1, kc])

    def run(self, nsteps=None):
        for n in range(self.c.nt if nsteps is None else nsteps):
            self.step(n)
            self.record(n)
        return self.rec


def make(cfg, upper=None, lower=None):
    return (TEzStudy if cfg.scheme == 'te' else Yee3DStudy)(cfg, upper, lower)
