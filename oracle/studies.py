"""Oracle of the loss and wave-impedance study drivers (test infrastructure only).

FP64 NumPy restatement of the study time loops, each kernel one statement
group in the reference launch order:

* TEz studies  FDTD/GPU/FDTD_2D_TE/scattering_loss/fdtd_funcs.py:79-167, loop
               fdtd_main.py:202-223 (H first: update_hz, update_pml_hz_yinc,
               update_pml_hz_xinc, update_source (+J), update_ex, update_ey,
               update_pml_ex_yinc, update_pml_ey_xinc, then recording);
               recording simul_fft (:113-120) or map_*_zwave
               (FDTD_2D_TE/wave_impedance/fdtd_funcs.py:169-181, loop :220-221).
* TMz studies  FDTD/GPU/FDTD_2D_TM/scattering_loss/fdtd_funcs.py:77-152 (lossy
               update form), loop fdtd_main.py:197-218, recording simul_fft
               (:138-144) or map_*_zwave (FDTD_2D_TM/wave_impedance/
               fdtd_funcs.py:155-168, loop :210-211); arrays fdtd_main.py:38-119.
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




def lossy_tm_coefficients(cfg, FG):
    """Arrays of FDTD_2D_TM/scattering_loss/fdtd_main.py:38-119, literal loop
    (numpy 1.20 element types written out with float())."""
    f32 = np.float32
    dt, dx = f32(cfg.delta_t), f32(cfg.delta_x)
    SIGMA_MASK = np.zeros([cfg.nx, cfg.ny], dtype=f32)
    EPS_MASK = (cfg.eps * (cfg.eps_rel_fg * FG + 1 * ~FG)).astype(f32)
    CURL_E = f32(float(dt) / (cfg.mu * float(dx)))
    CURL_H = f32(2 * float(dt)) / (dx * (f32(2) * EPS_MASK + SIGMA_MASK * dt))
    MOD_E = (f32(2) * EPS_MASK - SIGMA_MASK * dt) / (f32(2) * EPS_MASK + SIGMA_MASK * dt)
    N, R = cfg.num_cpml, cfg.num_cpml + 1
    sigma_max = 1.2 * (0.8 * (4 + 1) / (float(cfg.eta) * float(dx)))
    AE, KE, SE, AH, KH, SH, BE, CE, BH, CH = (np.zeros(R, f32) for _ in range(10))
    DEN_EX, DEN_EY, DEN_HX, DEN_HY = (np.ones(cfg.nx, f32), np.ones(cfg.ny, f32), np.ones(cfg.nx, f32),
                                      np.ones(cfg.ny, f32))
    for q in range(N):
        AE[q] = 0.05 * (q / N) ** 1
        KE[q] = 1 + (5 - 1) * ((R - q - 1) / N) ** 4
        SE[q] = sigma_max * ((R - q - 1) / N) ** 4
        BE[q] = np.exp(-1 * float(SE[q] / KE[q] + AE[q]) * float(dt) / cfg.eps)
        CE[q] = float(SE[q] / (SE[q] + KE[q] * AE[q]) / KE[q]) * (float(BE[q]) - 1)
        AH[q] = 0.05 * ((q + 0.5) / N) ** 1
        KH[q] = 1 + (5 - 1) * ((R - q - 1.5) / N) ** 4
        SH[q] = sigma_max * ((R - q - 1.5) / N) ** 4
        BH[q] = np.exp(-1 * float(SH[q] / KH[q] + AH[q]) * float(dt) / cfg.eps)
        CH[q] = float(SH[q] / (SH[q] + KH[q] * AH[q]) / KH[q]) * (float(BH[q]) - 1)
        DEN_EX[q] = KE[q]
        DEN_EX[cfg.nx - 1 - q] = KE[q]
        DEN_EY[q] = KE[q]
        DEN_EY[cfg.ny - 1 - q] = KE[q]
        DEN_HX[q] = KH[q]
        DEN_HX[cfg.nx - 2 - q] = KH[q]
        DEN_HY[q] = KH[q]
        DEN_HY[cfg.ny - 2 - q] = KH[q]
    return dict(CURL_E=CURL_E, CURL_H=CURL_H, MOD_E=MOD_E, BE=BE, CE=CE, BH=BH, CH=CH, DEN_EX=DEN_EX,
                DEN_EY=DEN_EY, DEN_HX=DEN_HX, DEN_HY=DEN_HY)


def tm_mask(cfg, upper=None, lower=None):
    """gen_fg_mask_sgl (FDTD_2D_TM/scattering_loss/fdtd_funcs.py:32-74), literal loop."""
    neff = cfg.nx - 2 * cfg.p1x
    upper = np.zeros(neff) if upper is None else np.asarray(upper, np.float64)
    lower = np.zeros(neff) if lower is None else np.asarray(lower, np.float64)
    up = np.append(np.append(np.zeros(cfg.p1x), upper), np.zeros(cfg.p1x))
    lo = np.append(np.append(np.zeros(cfg.p1x), lower), np.zeros(cfg.p1x))
    fg = np.zeros([cfg.nx, cfg.ny], dtype=bool)
    for j in range(cfg.ny):
        for i in range(cfg.nx):
            if j >= (cfg.by + lo[i]) and j < (cfg.by + cfg.ny_swg + up[i]):
                fg[i, j] = True
    return fg


class TMzLossyStudy:
    """2D TM study loop (FDTD_2D_TM/scattering_loss/fdtd_main.py:197-218,
    kernels fdtd_funcs.py:77-152), FP64, vectorised."""

    def __init__(self, cfg, upper=None, lower=None, FG=None):
        self.cfg = cfg
        self.FG = tm_mask(cfg, upper, lower) if FG is None else FG
        k = lossy_tm_coefficients(cfg, self.FG)
        f = np.float64
        self.k = {n: (np.asarray(v).astype(f) if np.ndim(v) else float(v)) for n, v in k.items()}
        self.R = cfg.num_cpml + 1
        nx, ny, R = cfg.nx, cfg.ny, self.R
        z = lambda *s: np.zeros(s)
        self.Ez, self.Hx, self.Hy = z(nx, ny), z(nx, ny), z(nx, ny)
        self.p_ez_xlo, self.p_ez_xhi, self.p_hy_xlo, self.p_hy_xhi = z(R, ny), z(R, ny), z(R, ny), z(R, ny)
        self.p_ez_ylo, self.p_ez_yhi, self.p_hx_ylo, self.p_hx_yhi = z(nx, R), z(nx, R), z(nx, R), z(nx, R)
        self.J = np.asarray(cfg.J_SRC, np.float32).astype(f)
        self.dx, self.dt, self.mu = float(np.float32(cfg.delta_x)), float(np.float32(cfg.delta_t)), float(
            np.float32(cfg.mu))
        self.rec = {}

    def step(self, n):
        cfg, k, R, dx = self.cfg, self.k, self.R, self.dx
        Ez, Hx, Hy = self.Ez, self.Hx, self.Hy
        nx, ny = cfg.nx, cfg.ny
        ch, me = k['CURL_H'], k['MOD_E']
        I, J = slice(1, nx - 1), slice(1, ny - 1)
        # update_ez
        Ez[I, J] = me[I, J] * Ez[I, J] + ch[I, J] * ((Hy[I, J] - Hy[0:nx - 2, J]) / k['DEN_EX'][I, None]
                                                     + (Hx[I, 0:ny - 2] - Hx[I, J]) / k['DEN_EY'][None, J])
        # update_source: Ez[sx, by:ny-by] -= J_SRC[n]
        Ez[cfg.sx, cfg.by:ny - cfg.by] -= self.J[n]
        q = np.arange(1, R)
        be, ce = k['BE'][q], k['CE'][q]
        # update_ez_cpml_x: cells q and nx-1-q, j in (0, ny-1)
        lo, hi = q, nx - 1 - q
        self.p_ez_xlo[q, J] = be[:, None] * self.p_ez_xlo[q, J] + ce[:, None] / dx * (Hy[lo, J] - Hy[lo - 1, J])
        Ez[lo, J] += ch[lo, J] * dx * self.p_ez_xlo[q, J]
        self.p_ez_xhi[q, J] = be[:, None] * self.p_ez_xhi[q, J] + ce[:, None] / dx * (Hy[hi, J] - Hy[hi - 1, J])
        Ez[hi, J] += ch[hi, J] * dx * self.p_ez_xhi[q, J]
        # update_ez_cpml_y: cells q and ny-1-q, i in (0, nx-1)
        lo, hi = q, ny - 1 - q
        self.p_ez_ylo[I, q] = be[None, :] * self.p_ez_ylo[I, q] + ce[None, :] / dx * (Hx[I, lo - 1] - Hx[I, lo])
        Ez[I, lo] += ch[I, lo] * dx * self.p_ez_ylo[I, q]
        self.p_ez_yhi[I, q] = be[None, :] * self.p_ez_yhi[I, q] + ce[None, :] / dx * (Hx[I, hi - 1] - Hx[I, hi])
        Ez[I, hi] += ch[I, hi] * dx * self.p_ez_yhi[I, q]
        # update_hx_hy on i < nx-1, j < ny-1
        I0, J0 = slice(0, nx - 1), slice(0, ny - 1)
        ce_ = k['CURL_E']
        Hx[I0, J0] += ce_ * (Ez[I0, J0] - Ez[I0, 1:ny]) / k['DEN_HY'][None, J0]
        Hy[I0, J0] += ce_ * (Ez[1:nx, J0] - Ez[I0, J0]) / k['DEN_HX'][I0, None]
        bh, chh = k['BH'][q], k['CH'][q]
        dtmu = float(np.float32(self.dt) / np.float32(self.mu))
        # update_hy_cpml_x: cells q and nx-2-q, j < ny-1
        lo, hi = q, nx - 2 - q
        self.p_hy_xlo[q, J0] = bh[:, None] * self.p_hy_xlo[q, J0] + chh[:, None] / dx * (Ez[lo + 1, J0] - Ez[lo, J0])
        Hy[lo, J0] += dtmu * self.p_hy_xlo[q, J0]
        self.p_hy_xhi[q, J0] = bh[:, None] * self.p_hy_xhi[q, J0] + chh[:, None] / dx * (Ez[hi + 1, J0] - Ez[hi, J0])
        Hy[hi, J0] += dtmu * self.p_hy_xhi[q, J0]
        # update_hx_cpml_y: cells q and ny-2-q, i < nx-1
        lo, hi = q, ny - 2 - q
        self.p_hx_ylo[I0, q] = bh[None, :] * self.p_hx_ylo[I0, q] + chh[None, :] / dx * (Ez[I0, lo] - Ez[I0, lo + 1])
        Hx[I0, lo] += dtmu * self.p_hx_ylo[I0, q]
        self.p_hx_yhi[I0, q] = bh[None, :] * self.p_hx_yhi[I0, q] + chh[None, :] / dx * (Ez[I0, hi] - Ez[I0, hi + 1])
        Hx[I0, hi] += dtmu * self.p_hx_yhi[I0, q]

    def record(self, n):
        cfg = self.cfg
        if cfg.study == 'scattering_loss':
            if not self.rec:
                self.rec = {k: np.zeros((cfg.nx, cfg.ny), complex) for k in ('ez', 'hx', 'hy')}
            for key, F_, st in (('ez', self.Ez, n + 0.5), ('hx', self.Hx, n + 1.0), ('hy', self.Hy, n + 1.0)):
                w = simul_fft_weight(cfg.f0, st, cfg.delta_t)
                self.rec[key] += F_ * np.cos(w) - 1j * (F_ * np.sin(w))
            return
        if not self.rec:
            self.rec = {k: np.zeros((cfg.nt, cfg.ny)) for k in ('ez', 'hx')}
        cx = cfg.cx
        self.rec['ez'][n] = self.Ez[cx, :]
        # Hx[cx, j-1] at j = 0 wraps to Hx[cx, ny-1], which no update writes
        self.rec['hx'][n] = 0.5 * (self.Hx[cx, :] + np.roll(self.Hx[cx, :], 1))

    def run(self, nsteps=None):
        for n in range(self.cfg.nt if nsteps is None else nsteps):
            self.step(n)
            self.record(n)
        return self.rec


def make(cfg, upper=None, lower=None):
    return {'tm': TMzLossyStudy, 'te': TEzStudy, '3d': Yee3DStudy}[cfg.scheme](cfg, upper, lower)
