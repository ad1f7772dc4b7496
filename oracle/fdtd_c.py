"""ctypes front of the FP64 C restatement of the 3D time step (oracle/fdtd3d.c).

TEST INFRASTRUCTURE ONLY, like the rest of ``oracle/``.  ``Yee3DC`` is the
NumPy oracle's ``Yee3D`` (same arrays, coefficients and state layout) whose
time stepping runs in C with OpenMP, for config-scale 3D checks.  Built by
``oracle.build`` (called from ``__graft_entry__.build``, or on first use) into
``oracle/liboracle_fdtd3d.so``.
"""
import ctypes as C
import os

import numpy as np

from .fdtd import Yee3D

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'liboracle_fdtd3d.so')

_P = C.POINTER(C.c_double)
PSI_ORDER = ('exy', 'exz', 'eyx', 'eyz', 'ezx', 'ezy', 'hxy', 'hxz', 'hyx', 'hyz', 'hzx', 'hzy')


class _State(C.Structure):
    _fields_ = [('nx', C.c_int32), ('ny', C.c_int32), ('nz', C.c_int32), ('N', C.c_int32), ('ib', C.c_int32),
                ('nf', C.c_int32), ('src_is_hy', C.c_int32), ('src', C.c_int32 * 6), ('port', C.c_int32 * 2),
                ('dx', C.c_double), ('db', C.c_double), ('dt', C.c_double),
                ('cbx', _P), ('cby', _P), ('cbz', _P),
                ('kxe', _P), ('kye', _P), ('kze', _P), ('kxh', _P), ('kyh', _P), ('kzh', _P),
                ('be', _P), ('ce', _P), ('bh', _P), ('ch', _P), ('J', _P), ('w', _P),
                ('f', _P * 6), ('psi', (_P * 2) * 12), ('acc', (_P * 2) * 4)]


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            from . import build
            build.build()
        L = C.CDLL(LIB_PATH)
        L.o3d_run.restype = C.c_int
        L.o3d_run.argtypes = [C.POINTER(_State), C.c_int32, C.c_int32]
        L.o3d_state_size.restype = C.c_int
        if L.o3d_state_size() != C.sizeof(_State):
            raise RuntimeError('oracle/liboracle_fdtd3d.so layout mismatch')
        _LIB = L
    return _LIB


def _p(a):
    assert a.dtype == np.float64 and a.flags['C_CONTIGUOUS']
    return a.ctypes.data_as(_P)


class Yee3DC(Yee3D):
    """``oracle.fdtd.Yee3D`` stepped by the C restatement."""

    def __init__(self, c, FG, direction='f', line=0):
        super().__init__(c, FG, direction, line)
        self.keep = []
        st = _State()
        st.nx, st.ny, st.nz, st.N, st.ib = c.nx, c.ny, c.nz, self.N, 2
        st.nf = self.nf
        st.src_is_hy = int(c.source_condition != 'TE')
        x0, x1, y0, y1, z0, z1 = self.source_box()
        lim = (c.nx, c.ny, c.nz)
        box = []
        for a, (lo, hi) in enumerate(((x0, x1), (y0, y1), (z0, z1))):   # numpy slice clipping
            lo, hi = min(max(lo, 0), lim[a]), min(max(hi, 0), lim[a])
            box += [lo, max(hi, lo)]
        for a in range(6):
            st.src[a] = box[a]
        st.port[0], st.port[1] = self.ports
        st.dx, st.db, st.dt = self.dx, self.db, self.dt
        arr = lambda a: self._keep(np.ascontiguousarray(a, dtype=np.float64))
        st.cbx, st.cby, st.cbz = (_p(arr(a)) for a in self.cb)
        st.kxe, st.kye, st.kze = _p(arr(self.kxe)), _p(arr(self.kye)), _p(arr(self.kze))
        st.kxh, st.kyh, st.kzh = _p(arr(self.kxh)), _p(arr(self.kyh)), _p(arr(self.kzh))
        st.be, st.ce, st.bh, st.ch = (_p(arr(a)) for a in (self.be, self.ce, self.bh, self.ch))
        st.J = _p(arr(self.J))
        st.w = _p(arr(self.w))
        for k, name in enumerate(('Ex', 'Ey', 'Ez', 'Hx', 'Hy', 'Hz')):
            st.f[k] = _p(getattr(self, name))
        for k, name in enumerate(PSI_ORDER):
            lo, hi = self.psi[name]
            st.psi[k][0], st.psi[k][1] = _p(lo), _p(hi)
        for k, comp in enumerate(('ex', 'ey', 'hx', 'hy')):
            for p in (1, 2):
                a = self.acc[(comp, p)]
                assert a.dtype == np.complex128 and a.flags['C_CONTIGUOUS']
                st.acc[k][p - 1] = a.ctypes.data_as(_P)
        self.st = st

    def _keep(self, a):
        self.keep.append(a)
        return a

    def run(self, nsteps=None):
        nsteps = self.c.nt - self.n if nsteps is None else nsteps
        if nsteps > 0:
            lib().o3d_run(C.byref(self.st), int(self.n), int(nsteps))
            self.n += nsteps
        return self

    def step(self, n):
        if n != self.n:
            raise ValueError('Yee3DC steps in order')
        return self.run(1)


def make(c, FG, direction='f', line=0):
    """The C-stepped oracle for 3D, the NumPy one otherwise."""
    if c.scheme == '3d':
        return Yee3DC(c, FG, direction, line)
    from .fdtd import make as make_np
    return make_np(c, FG, direction, line)
