"""Oracle time stepping: FP64 NumPy restatement of the reference CUDA kernels
(test infrastructure only).

Every reference kernel is one NumPy statement group below, applied in the
reference launch order:

* TMz  FDTD/GPU/FDTD_2D_TM/s-parameter_extraction/fdtd_funcs.py:44-144, order fdtd_main.py:275-304
* TEz  FDTD/GPU/FDTD_2D_TE/s-parameter_extraction/fdtd_funcs.py:44-144, order fdtd_main.py:278-307
* 3D   FDTD/GPU/FDTD_3D/s-parameter_extraction/fdtd_funcs.py:44-292,   order fdtd_main.py:357-406

Arithmetic is float64 throughout; the coefficient arrays are the run's
precision-rounded values (cb maps, CPML b/c, kappa*dx, J_SRC, dx, db) widened
to float64, so a float32 device run differs from this only by the rounding
of its own stepping arithmetic.
"""
import numpy as np

from . import geometry as G
from . import setup as S


class _Base:
    def __init__(self, c, FG, direction='f', line=0, ports=None):
        S.coefficients(c)
        self.c = c
        self.direction = direction
        self.shift = line * c.pp
        f = np.float64
        self.N = c.num_cpml
        self.dx = float(c.delta_x)
        self.db = float(c.DB)
        self.dt = float(c.delta_t)
        self.J = c.J_SRC.astype(f)
        self.be, self.ce = c.BE.astype(f), c.CE.astype(f)
        self.bh, self.ch = c.BH.astype(f), c.CH.astype(f)
        self.kxe, self.kxh = c.kxe.astype(f), c.kxh.astype(f)
        self.kye, self.kyh = c.kye.astype(f), c.kyh.astype(f)
        self.cb = [a.astype(f) for a in G.cb_maps(c, FG)]
        self.w = 2 * np.pi * np.asarray(c.sparam_freqs, f)
        self.nf = len(self.w)
        self.n = 0
        # The sign with which each psi array corrects its field, and the
        # grading order of each slab side, follow the reference kernels.  Both
        # are instance data so that the pins in tests/ can plant a mistake
        # (a flipped sign, one side graded in the other side's order) and
        # show that they catch it.
        self.psi_sign = dict(self.PSI_SIGN)
        self.psi_flip = set()

    PSI_SIGN = {}

    def grading(self, name, side, kind):
        """(b, c) per slab depth for one side of psi array ``name``.  E psi,
        depth d in [0, N): grading index N-1-d on the low side, d on the high
        side (update_pml_e*, fdtd_funcs.py:80-100).  H psi, depth d in [1, N):
        N-d low, d high (update_pml_h*, :104-124)."""
        N = self.N
        low = (side == 0) != ((name, side) in self.psi_flip)
        if kind == 'e':
            d = np.arange(N)
            g = N - 1 - d if low else d
            return self.be[g], self.ce[g]
        d = np.arange(1, N)
        g = N - d if low else d
        return self.bh[g], self.ch[g]

    def correct(self, F, idx, t, name):
        """F[idx] += t or -= t, by the sign of psi array ``name``."""
        if self.psi_sign[name] > 0:
            F[idx] += t
        else:
            F[idx] -= t

    def phasors(self, n):
        # map_sfft_coeffs (fdtd_funcs.py:230-237): E at n*dt, H at (n+0.5)*dt
        wn = self.w * float(n)
        wh = self.w * (n + 0.5)
        return np.cos(wn * self.dt), np.sin(wn * self.dt), np.cos(wh * self.dt), np.sin(wh * self.dt)

    def run(self, nsteps=None):
        nsteps = self.c.nt - self.n if nsteps is None else nsteps
        for _ in range(nsteps):
            self.step(self.n)
            self.n += 1
        return self

    def step(self, n):
        """One time step: the field updates in launch order, then the port DFT."""
        self.update(n)
        self.accumulate(n)


class TMz(_Base):
    """2D TMz (Ez, Hx, Hy)."""

    def __init__(self, c, FG, direction='f', line=0):
        super().__init__(c, FG, direction, line)
        nx, ny, N = c.nx, c.ny, self.N
        z = lambda *s: np.zeros(s)
        self.Ez, self.Hx, self.Hy = z(nx, ny), z(nx, ny), z(nx, ny)
        self.p_ez_xlo, self.p_ez_xhi = z(N, ny), z(N, ny)
        self.p_ez_ylo, self.p_ez_yhi = z(nx, N), z(nx, N)
        self.p_hx_ylo, self.p_hx_yhi = z(nx, N), z(nx, N)
        self.p_hy_xlo, self.p_hy_xhi = z(N, ny), z(N, ny)
        self.ports = (c.p1x, c.p2x)
        self.acc = {k: z(self.nf, ny) for k in ('e1r', 'e1i', 'h1r', 'h1i', 'e2r', 'e2i', 'h2r', 'h2i')}

    def source_box(self):
        c = self.c
        syb, sye = c.by + self.shift + 1, c.by + self.shift + c.ny_swg - 1
        if self.direction == 'f':
            return c.sx, c.sx + 1, syb, sye
        return c.nx - c.sx - 1, c.nx - c.sx, syb, sye

    PSI_SIGN = {'ezx': +1, 'ezy': -1, 'hxy': -1, 'hyx': +1}

    def update(self, n):
        c, N, dx, db = self.c, self.N, self.dx, self.db
        Ez, Hx, Hy = self.Ez, self.Hx, self.Hy
        (cbz,) = self.cb
        ib, jb, ie, je = 2, 2, c.nx - 2, c.ny - 2
        I, J = slice(ib, ie), slice(jb, je)
        # update_ez
        Ez[I, J] += cbz[I, J] * ((Hy[ib + 1:ie + 1, J] - Hy[I, J]) / self.kxe[I, None]
                                 - (Hx[I, jb + 1:je + 1] - Hx[I, J]) / self.kye[None, J])
        # update_source (Ez)
        x0, x1, y0, y1 = self.source_box()
        Ez[max(x0, 0):x1, max(y0, 0):y1] -= self.J[n]
        d = np.arange(N)
        # update_pml_ez_xinc
        for side, cells, p in ((0, ib + d, self.p_ez_xlo), (1, ie - N + d, self.p_ez_xhi)):
            b_, c_ = self.grading('ezx', side, 'e')
            p[:, J] = b_[:, None] * p[:, J] + c_[:, None] * (Hy[cells + 1, J] - Hy[cells, J]) / dx
            self.correct(Ez, (cells, J), cbz[cells, J] * p[:, J], 'ezx')
        # update_pml_ez_yinc
        for side, cells, p in ((0, jb + d, self.p_ez_ylo), (1, je - N + d, self.p_ez_yhi)):
            b_, c_ = self.grading('ezy', side, 'e')
            p[I, :] = b_[None, :] * p[I, :] + c_[None, :] * (Hx[I, cells + 1] - Hx[I, cells]) / dx
            self.correct(Ez, (I, cells), cbz[I, cells] * p[I, :], 'ezy')
        # update_hx, update_hy
        Hx[I, jb + 1:je] += db * (Ez[I, jb:je - 1] - Ez[I, jb + 1:je]) / self.kyh[None, jb + 1:je]
        Hy[ib + 1:ie, J] += db * (Ez[ib + 1:ie, J] - Ez[ib:ie - 1, J]) / self.kxh[ib + 1:ie, None]
        dd = np.arange(1, N)
        # update_pml_hx_yinc
        for side, cells, p in ((0, jb + dd, self.p_hx_ylo), (1, je - N + dd, self.p_hx_yhi)):
            b_, c_ = self.grading('hxy', side, 'h')
            p[I, 1:] = b_[None, :] * p[I, 1:] + c_[None, :] * (Ez[I, cells] - Ez[I, cells - 1]) / dx
            self.correct(Hx, (I, cells), db * p[I, 1:], 'hxy')
        # update_pml_hy_xinc
        for side, cells, p in ((0, ib + dd, self.p_hy_xlo), (1, ie - N + dd, self.p_hy_xhi)):
            b_, c_ = self.grading('hyx', side, 'h')
            p[1:, J] = b_[:, None] * p[1:, J] + c_[:, None] * (Ez[cells, J] - Ez[cells - 1, J]) / dx
            self.correct(Hy, (cells, J), db * p[1:, J], 'hyx')

    def port_samples(self, ic):
        """simul_fft_efield / simul_fft_hfield samples at port column ic over
        the window [N, ny-N): Ez, and Hy averaged over the two columns ic,
        ic+1 (fdtd_funcs.py:128-144)."""
        W = slice(self.N, self.c.ny - self.N)
        return {'e': self.Ez[ic, W], 'h': 0.5 * (self.Hy[ic, W] + self.Hy[ic + 1, W])}

    def accumulate(self, n):
        cE, sE, cH, sH = self.phasors(n)
        W = slice(self.N, self.c.ny - self.N)
        for p, ic in zip('12', self.ports):
            smp = self.port_samples(ic)
            e, h = smp['e'], smp['h']
            a = self.acc
            a['e' + p + 'r'][:, W] += e[None, :] * cE[:, None]
            a['e' + p + 'i'][:, W] -= e[None, :] * sE[:, None]
            a['h' + p + 'r'][:, W] += h[None, :] * cH[:, None]
            a['h' + p + 'i'][:, W] -= h[None, :] * sH[:, None]

    def fields(self):
        return {'ez': self.Ez, 'hx': self.Hx, 'hy': self.Hy}

    def dft(self):
        """Complex port phasors {('e'|'h', port 1|2): [nf, ny]} as fdtd_main.py:359-363."""
        a = self.acc
        return {(k, p): a[k + str(p) + 'r'] + 1j * a[k + str(p) + 'i'] for k in 'eh' for p in (1, 2)}


class TEz(_Base):
    """2D TEz (Ex, Ey, Hz)."""

    def __init__(self, c, FG, direction='f', line=0):
        super().__init__(c, FG, direction, line)
        nx, ny, N = c.nx, c.ny, self.N
        z = lambda *s: np.zeros(s)
        self.Ex, self.Ey, self.Hz = z(nx, ny), z(nx, ny), z(nx, ny)
        self.p_ex_ylo, self.p_ex_yhi = z(nx, N), z(nx, N)
        self.p_ey_xlo, self.p_ey_xhi = z(N, ny), z(N, ny)
        self.p_hz_xlo, self.p_hz_xhi = z(N, ny), z(N, ny)
        self.p_hz_ylo, self.p_hz_yhi = z(nx, N), z(nx, N)
        self.ports = (c.p1x, c.p2x)
        self.acc = {k: z(self.nf, ny) for k in ('e1r', 'e1i', 'h1r', 'h1i', 'e2r', 'e2i', 'h2r', 'h2i')}

    def source_box(self):
        c = self.c
        syb, sye = c.by + self.shift + 2, c.by + self.shift + c.ny_swg - 2
        if self.direction == 'f':
            return c.sx, c.sx + 1, syb, sye
        return c.nx - c.sx - 1, c.nx - c.sx, syb, sye

    PSI_SIGN = {'exy': +1, 'eyx': -1, 'hzx': -1, 'hzy': +1}

    def update(self, n):
        c, N, dx, db = self.c, self.N, self.dx, self.db
        Ex, Ey, Hz = self.Ex, self.Ey, self.Hz
        cbx, cby = self.cb
        ib, jb, ie, je = 2, 2, c.nx - 2, c.ny - 2
        I, J, I1, J1 = slice(ib, ie), slice(jb, je), slice(ib + 1, ie), slice(jb + 1, je)
        # update_ex: i in (ib,ie), j in [jb,je)
        Ex[I1, J] += cbx[I1, J] * (Hz[I1, jb + 1:je + 1] - Hz[I1, J]) / self.kye[None, J]
        # update_ey: i in [ib,ie), j in (jb,je)
        Ey[I, J1] += cby[I, J1] * (Hz[I, J1] - Hz[ib + 1:ie + 1, J1]) / self.kxe[I, None]
        d = np.arange(N)
        # update_pml_ex_yinc: i in (ib,ie)
        for side, cells, p in ((0, jb + d, self.p_ex_ylo), (1, je - N + d, self.p_ex_yhi)):
            b_, c_ = self.grading('exy', side, 'e')
            p[I1, :] = b_[None, :] * p[I1, :] + c_[None, :] * (Hz[I1, cells + 1] - Hz[I1, cells]) / dx
            self.correct(Ex, (I1, cells), cbx[I1, cells] * p[I1, :], 'exy')
        # update_pml_ey_xinc: j in (jb,je)
        for side, cells, p in ((0, ib + d, self.p_ey_xlo), (1, ie - N + d, self.p_ey_xhi)):
            b_, c_ = self.grading('eyx', side, 'e')
            p[:, J1] = b_[:, None] * p[:, J1] + c_[:, None] * (Hz[cells + 1, J1] - Hz[cells, J1]) / dx
            self.correct(Ey, (cells, J1), cby[cells, J1] * p[:, J1], 'eyx')
        # update_hz: i in (ib,ie), j in (jb,je)
        Hz[I1, J1] += db * ((Ex[I1, J1] - Ex[I1, jb:je - 1]) / self.kyh[None, J1]
                            - (Ey[I1, J1] - Ey[ib:ie - 1, J1]) / self.kxh[I1, None])
        # update_source (Hz), after update_hz
        x0, x1, y0, y1 = self.source_box()
        Hz[max(x0, 0):x1, max(y0, 0):y1] -= self.J[n]
        dd = np.arange(1, N)
        # update_pml_hz_xinc: j in (jb,je)
        for side, cells, p in ((0, ib + dd, self.p_hz_xlo), (1, ie - N + dd, self.p_hz_xhi)):
            b_, c_ = self.grading('hzx', side, 'h')
            p[1:, J1] = b_[:, None] * p[1:, J1] + c_[:, None] * (Ey[cells, J1] - Ey[cells - 1, J1]) / dx
            self.correct(Hz, (cells, J1), db * p[1:, J1], 'hzx')
        # update_pml_hz_yinc: i in (ib,ie)
        for side, cells, p in ((0, jb + dd, self.p_hz_ylo), (1, je - N + dd, self.p_hz_yhi)):
            b_, c_ = self.grading('hzy', side, 'h')
            p[I1, 1:] = b_[None, :] * p[I1, 1:] + c_[None, :] * (Ex[I1, cells] - Ex[I1, cells - 1]) / dx
            self.correct(Hz, (I1, cells), db * p[I1, 1:], 'hzy')

    def port_samples(self, ic):
        """Port samples at column ic over the window [N, ny-N): Ey averaged
        over rows j, j+1; Hz over the 2x2 corners of columns ic, ic+1 and rows
        j, j+1 (FDTD_2D_TE/.../fdtd_funcs.py:128-144)."""
        N, ny = self.N, self.c.ny
        W, W1 = slice(N, ny - N), slice(N + 1, ny - N + 1)
        Ey, Hz = self.Ey, self.Hz
        return {'e': 0.5 * (Ey[ic, W] + Ey[ic, W1]),
                'h': 0.25 * (Hz[ic, W] + Hz[ic + 1, W] + Hz[ic, W1] + Hz[ic + 1, W1])}

    def accumulate(self, n):
        cE, sE, cH, sH = self.phasors(n)
        W = slice(self.N, self.c.ny - self.N)
        for p, ic in zip('12', self.ports):
            smp = self.port_samples(ic)
            e, h = smp['e'], smp['h']
            a = self.acc
            a['e' + p + 'r'][:, W] += e[None, :] * cE[:, None]
            a['e' + p + 'i'][:, W] -= e[None, :] * sE[:, None]
            a['h' + p + 'r'][:, W] += h[None, :] * cH[:, None]
            a['h' + p + 'i'][:, W] -= h[None, :] * sH[:, None]

    def fields(self):
        return {'ex': self.Ex, 'ey': self.Ey, 'hz': self.Hz}

    def dft(self):
        a = self.acc
        return {(k, p): a[k + str(p) + 'r'] + 1j * a[k + str(p) + 'i'] for k in 'eh' for p in (1, 2)}


class Yee3D(_Base):
    """3D (Ex, Ey, Ez, Hx, Hy, Hz) with 24 CPML psi arrays."""

    def __init__(self, c, FG, direction='f', line=0):
        super().__init__(c, FG, direction, line)
        f = np.float64
        self.kze, self.kzh = c.kze.astype(f), c.kzh.astype(f)
        nx, ny, nz, N = c.nx, c.ny, c.nz, self.N
        z = lambda *s: np.zeros(s)
        for k in ('Ex', 'Ey', 'Ez', 'Hx', 'Hy', 'Hz'):
            setattr(self, k, z(nx, ny, nz))
        shp = {'x': (N, ny, nz), 'y': (nx, N, nz), 'z': (nx, ny, N)}
        self.psi = {}
        for comp, axes in (('ex', 'yz'), ('ey', 'xz'), ('ez', 'xy'), ('hx', 'yz'), ('hy', 'xz'), ('hz', 'xy')):
            for a in axes:
                self.psi[comp + a] = (z(*shp[a]), z(*shp[a]))
        self.ports = (c.p1z, c.p2z)
        self.acc = {(k, p): np.zeros((self.nf, nx, ny), dtype=complex)
                    for k in ('ex', 'ey', 'hx', 'hy') for p in (1, 2)}

    def source_box(self):
        c = self.c
        m = 1 if c.source_condition == 'TE' else 2
        x0, x1 = c.bx + self.shift + m, c.bx + self.shift + c.nx_swg - m
        y0, y1 = c.by + m, c.by + c.ny_swg - m
        if self.direction == 'f':
            z0, z1 = c.sz, c.sz + 2
        else:
            z0, z1 = c.nz - c.sz - 1, c.nz - c.sz
        return x0, x1, y0, y1, z0, z1

    PSI_SIGN = {'exy': +1, 'exz': -1, 'eyx': -1, 'eyz': +1, 'ezx': +1, 'ezy': -1,
                'hxy': -1, 'hxz': +1, 'hyx': +1, 'hyz': -1, 'hzx': -1, 'hzy': +1}

    def update(self, n):
        c, N, dx, db = self.c, self.N, self.dx, self.db
        Ex, Ey, Ez, Hx, Hy, Hz = self.Ex, self.Ey, self.Ez, self.Hx, self.Hy, self.Hz
        cbx, cby, cbz = self.cb
        kxe, kye, kze = self.kxe, self.kye, self.kze
        kxh, kyh, kzh = self.kxh, self.kyh, self.kzh
        ib = jb = kb = 2
        ie, je, ke = c.nx - 2, c.ny - 2, c.nz - 2
        I, J, K = slice(ib, ie), slice(jb, je), slice(kb, ke)
        I1, J1, K1 = slice(ib + 1, ie), slice(jb + 1, je), slice(kb + 1, ke)
        # update_ex: i (ib,ie), j [jb,je), k [kb,ke)
        s = (I1, J, K)
        Ex[s] += cbx[s] * ((Hz[I1, jb + 1:je + 1, K] - Hz[s]) / kye[None, J, None]
                           - (Hy[I1, J, kb + 1:ke + 1] - Hy[s]) / kze[None, None, K])
        # update_ey: i [ib,ie), j (jb,je), k [kb,ke)
        s = (I, J1, K)
        Ey[s] += cby[s] * ((Hx[I, J1, kb + 1:ke + 1] - Hx[s]) / kze[None, None, K]
                           - (Hz[ib + 1:ie + 1, J1, K] - Hz[s]) / kxe[I, None, None])
        # update_ez: i [ib,ie), j [jb,je), k (kb,ke)
        s = (I, J, K1)
        Ez[s] += cbz[s] * ((Hy[ib + 1:ie + 1, J, K1] - Hy[s]) / kxe[I, None, None]
                           - (Hx[I, jb + 1:je + 1, K1] - Hx[s]) / kye[None, J, None])
        # update_source: Ey (TE) or Hy (TM), right after the E updates
        x0, x1, y0, y1, z0, z1 = self.source_box()
        F = Ey if c.source_condition == 'TE' else Hy
        F[max(x0, 0):x1, max(y0, 0):y1, max(z0, 0):z1] -= self.J[n]
        d = np.arange(N)
        ps = self.psi
        # E CPML: each kernel's low side (cells b + d), then its high side (cells e - N + d)
        for side, cells in ((0, jb + d), (1, je - N + d)):
            # ex_y (+), i (ib,ie), k [kb,ke)
            b_, c_ = self.grading('exy', side, 'e')
            p = ps['exy'][side]
            p[I1, :, K] = b_[None, :, None] * p[I1, :, K] + c_[None, :, None] * (Hz[I1, cells + 1, K] - Hz[I1, cells, K]) / dx
            self.correct(Ex, (I1, cells, K), cbx[I1, cells, K] * p[I1, :, K], 'exy')
        for side, cells in ((0, kb + d), (1, ke - N + d)):
            # ex_z (-), i (ib,ie), j [jb,je)
            b_, c_ = self.grading('exz', side, 'e')
            p = ps['exz'][side]
            p[I1, J, :] = b_[None, None, :] * p[I1, J, :] + c_[None, None, :] * (Hy[I1, J, cells + 1] - Hy[I1, J, cells]) / dx
            self.correct(Ex, (I1, J, cells), cbx[I1, J, cells] * p[I1, J, :], 'exz')
        for side, cells in ((0, ib + d), (1, ie - N + d)):
            # ey_x (-), j (jb,je), k [kb,ke)
            b_, c_ = self.grading('eyx', side, 'e')
            p = ps['eyx'][side]
            p[:, J1, K] = b_[:, None, None] * p[:, J1, K] + c_[:, None, None] * (Hz[cells + 1, J1, K] - Hz[cells, J1, K]) / dx
            self.correct(Ey, (cells, J1, K), cby[cells, J1, K] * p[:, J1, K], 'eyx')
        for side, cells in ((0, kb + d), (1, ke - N + d)):
            # ey_z (+), i [ib,ie), j (jb,je)
            b_, c_ = self.grading('eyz', side, 'e')
            p = ps['eyz'][side]
            p[I, J1, :] = b_[None, None, :] * p[I, J1, :] + c_[None, None, :] * (Hx[I, J1, cells + 1] - Hx[I, J1, cells]) / dx
            self.correct(Ey, (I, J1, cells), cby[I, J1, cells] * p[I, J1, :], 'eyz')
        for side, cells in ((0, ib + d), (1, ie - N + d)):
            # ez_x (+), j [jb,je), k (kb,ke)
            b_, c_ = self.grading('ezx', side, 'e')
            p = ps['ezx'][side]
            p[:, J, K1] = b_[:, None, None] * p[:, J, K1] + c_[:, None, None] * (Hy[cells + 1, J, K1] - Hy[cells, J, K1]) / dx
            self.correct(Ez, (cells, J, K1), cbz[cells, J, K1] * p[:, J, K1], 'ezx')
        for side, cells in ((0, jb + d), (1, je - N + d)):
            # ez_y (-), i [ib,ie), k (kb,ke)
            b_, c_ = self.grading('ezy', side, 'e')
            p = ps['ezy'][side]
            p[I, :, K1] = b_[None, :, None] * p[I, :, K1] + c_[None, :, None] * (Hx[I, cells + 1, K1] - Hx[I, cells, K1]) / dx
            self.correct(Ez, (I, cells, K1), cbz[I, cells, K1] * p[I, :, K1], 'ezy')
        # update_hx: i [ib,ie), j (jb,je), k (kb,ke)
        s = (I, J1, K1)
        Hx[s] += db * ((Ey[s] - Ey[I, J1, kb:ke - 1]) / kzh[None, None, K1]
                       - (Ez[s] - Ez[I, jb:je - 1, K1]) / kyh[None, J1, None])
        # update_hy: i (ib,ie), j [jb,je), k (kb,ke)
        s = (I1, J, K1)
        Hy[s] += db * ((Ez[s] - Ez[ib:ie - 1, J, K1]) / kxh[I1, None, None]
                       - (Ex[s] - Ex[I1, J, kb:ke - 1]) / kzh[None, None, K1])
        # update_hz: i (ib,ie), j (jb,je), k [kb,ke)
        s = (I1, J1, K)
        Hz[s] += db * ((Ex[s] - Ex[I1, jb:je - 1, K]) / kyh[None, J1, None]
                       - (Ey[s] - Ey[ib:ie - 1, J1, K]) / kxh[I1, None, None])
        dd = np.arange(1, N)
        for side, cells in ((0, jb + dd), (1, je - N + dd)):
            # hx_y (-), i [ib,ie), k (kb,ke)
            b_, c_ = self.grading('hxy', side, 'h')
            p = ps['hxy'][side]
            p[I, 1:, K1] = b_[None, :, None] * p[I, 1:, K1] + c_[None, :, None] * (Ez[I, cells, K1] - Ez[I, cells - 1, K1]) / dx
            self.correct(Hx, (I, cells, K1), db * p[I, 1:, K1], 'hxy')
        for side, cells in ((0, kb + dd), (1, ke - N + dd)):
            # hx_z (+), i [ib,ie), j (jb,je)
            b_, c_ = self.grading('hxz', side, 'h')
            p = ps['hxz'][side]
            p[I, J1, 1:] = b_[None, None, :] * p[I, J1, 1:] + c_[None, None, :] * (Ey[I, J1, cells] - Ey[I, J1, cells - 1]) / dx
            self.correct(Hx, (I, J1, cells), db * p[I, J1, 1:], 'hxz')
        for side, cells in ((0, ib + dd), (1, ie - N + dd)):
            # hy_x (+), j [jb,je), k (kb,ke)
            b_, c_ = self.grading('hyx', side, 'h')
            p = ps['hyx'][side]
            p[1:, J, K1] = b_[:, None, None] * p[1:, J, K1] + c_[:, None, None] * (Ez[cells, J, K1] - Ez[cells - 1, J, K1]) / dx
            self.correct(Hy, (cells, J, K1), db * p[1:, J, K1], 'hyx')
        for side, cells in ((0, kb + dd), (1, ke - N + dd)):
            # hy_z (-), i (ib,ie), j [jb,je)
            b_, c_ = self.grading('hyz', side, 'h')
            p = ps['hyz'][side]
            p[I1, J, 1:] = b_[None, None, :] * p[I1, J, 1:] + c_[None, None, :] * (Ex[I1, J, cells] - Ex[I1, J, cells - 1]) / dx
            self.correct(Hy, (I1, J, cells), db * p[I1, J, 1:], 'hyz')
        for side, cells in ((0, ib + dd), (1, ie - N + dd)):
            # hz_x (-), j (jb,je), k [kb,ke)
            b_, c_ = self.grading('hzx', side, 'h')
            p = ps['hzx'][side]
            p[1:, J1, K] = b_[:, None, None] * p[1:, J1, K] + c_[:, None, None] * (Ey[cells, J1, K] - Ey[cells - 1, J1, K]) / dx
            self.correct(Hz, (cells, J1, K), db * p[1:, J1, K], 'hzx')
        for side, cells in ((0, jb + dd), (1, je - N + dd)):
            # hz_y (+), i (ib,ie), k [kb,ke)
            b_, c_ = self.grading('hzy', side, 'h')
            p = ps['hzy'][side]
            p[I1, 1:, K] = b_[None, :, None] * p[I1, 1:, K] + c_[None, :, None] * (Ex[I1, cells, K] - Ex[I1, cells - 1, K]) / dx
            self.correct(Hz, (I1, cells, K), db * p[I1, 1:, K], 'hzy')

    def port_samples(self, kc):
        """Port samples on plane k = kc over the window i in [N, nx-N), j in
        [N, ny-N), each averaged to the cell corner (i, j, kc): Ex over i,
        i+1; Ey over j, j+1; Hx over (j, k) pairs; Hy over (i, k) pairs
        (simul_fft_efield / simul_fft_hfield, fdtd_funcs.py:270-292)."""
        c, N = self.c, self.N
        wi, wj = slice(N, c.nx - N), slice(N, c.ny - N)
        wi1, wj1 = slice(N + 1, c.nx - N + 1), slice(N + 1, c.ny - N + 1)
        Ex, Ey, Hx, Hy = self.Ex, self.Ey, self.Hx, self.Hy
        return {'ex': 0.5 * (Ex[wi, wj, kc] + Ex[wi1, wj, kc]),
                'ey': 0.5 * (Ey[wi, wj, kc] + Ey[wi, wj1, kc]),
                'hx': 0.25 * (Hx[wi, wj, kc] + Hx[wi, wj1, kc] + Hx[wi, wj, kc + 1] + Hx[wi, wj1, kc + 1]),
                'hy': 0.25 * (Hy[wi, wj, kc] + Hy[wi1, wj, kc] + Hy[wi, wj, kc + 1] + Hy[wi1, wj, kc + 1])}

    def accumulate(self, n):
        c, N = self.c, self.N
        cE, sE, cH, sH = self.phasors(n)
        wi, wj = slice(N, c.nx - N), slice(N, c.ny - N)
        eE = (cE - 1j * sE)[:, None, None]
        eH = (cH - 1j * sH)[:, None, None]
        for p, kc in zip((1, 2), self.ports):
            smp = self.port_samples(kc)
            for k in ('ex', 'ey'):
                self.acc[(k, p)][:, wi, wj] += smp[k][None] * eE
            for k in ('hx', 'hy'):
                self.acc[(k, p)][:, wi, wj] += smp[k][None] * eH

    def fields(self):
        return {k.lower(): getattr(self, k) for k in ('Ex', 'Ey', 'Ez', 'Hx', 'Hy', 'Hz')}

    def dft(self):
        return dict(self.acc)


def make(c, FG, direction='f', line=0):
    return {'tm': TMz, 'te': TEz, '3d': Yee3D}[c.scheme](c, FG, direction, line)
