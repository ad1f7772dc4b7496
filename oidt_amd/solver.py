"""Device runs through the C ABI: one run (``Solver``) or many (``Batch``).

``run_arrays`` assembles exactly the arrays the reference driver hands its
kernels (FDTD/GPU/FDTD_2D_TM/s-parameter_extraction/fdtd_main.py:83-178,
FDTD_3D fdtd_main.py:86-221) plus the source box of update_source
(fdtd_main.py:280-286; TE :291-297; 3D :364-374).  ``Solver.dft`` returns the
port phasors in the reference's layout (``EZF1 = EZ1r + 1j*EZ1i``,
fdtd_main.py:359-363; 3D [nf, nx, ny], FDTD_3D fdtd_main.py:462-470).
"""
import ctypes as C
from types import SimpleNamespace

import numpy as np

from . import _lib
from . import materials as M
from ._lib import check, lib, ptr


class RunArrays:
    """Host arrays of one run, kept alive while the library reads them."""

    def __init__(self, s, fg, direction='f', line=0, cb=None):
        if direction not in ('f', 'b'):
            raise ValueError("direction must be 'f' or 'b'")
        P = s.precision
        self.setup = s
        self.direction = direction
        self.line = line
        g = M.cpml_grading(s.num_cpml, s.m_grade, s.m_alpha, s.kappa_max, s.alpha_max, s.sigma_factor,
                           s.eta, s.delta_x, s.delta_t, s.eps, P)
        self.grading = g
        o = s.offset
        dims = (s.nx, s.ny, s.nz) if s.scheme == '3d' else (s.nx, s.ny)
        self.kappa = [M.kappa_axes(n, s.num_cpml, o, g['KE'], g['KH'], s.delta_x, P) for n in dims]
        self.cb = [np.ascontiguousarray(a, dtype=P) for a in (M.cb_maps(s, fg) if cb is None else cb)]
        self.jsrc = np.ascontiguousarray(s.J_SRC, dtype=P)
        self.freqs = np.ascontiguousarray(s.sparam_freqs, dtype=np.float64)
        self.db = P(float(s.delta_t) / s.mu)   # DB = delta_t / mu (fdtd_main.py:86), narrowed at launch
        self.src_field, self.src_box = source_box(s, direction, line)

    def desc(self):
        s = self.setup
        d = _lib.RunDesc()
        d.abi_version = _lib.ABI_VERSION
        d.scheme = _lib.SCHEME_ID[s.scheme]
        d.precision = _lib.F32 if s.precision == np.float32 else _lib.F64
        d.nx, d.ny = s.nx, s.ny
        d.nz = s.nz if s.scheme == '3d' else 1
        d.num_cpml = s.num_cpml
        d.offset = s.offset
        d.nt = s.nt
        d.nf = len(self.freqs)
        d.src_field = self.src_field
        for a in range(6):
            d.src_box[a] = self.src_box[a]
        d.port[0], d.port[1] = s.ports
        d.dx = float(s.delta_x)
        d.dt = float(s.delta_t)
        d.db = float(self.db)
        return d

    def inputs(self):
        r = _lib.RunInputs()
        for c, a in enumerate(self.cb):
            r.cb[c] = ptr(a)
        for a, (ke, kh) in enumerate(self.kappa):
            r.kappa_e[a] = ptr(ke)
            r.kappa_h[a] = ptr(kh)
        g = self.grading
        r.be, r.ce, r.bh, r.ch = ptr(g['BE']), ptr(g['CE']), ptr(g['BH']), ptr(g['CH'])
        r.jsrc = ptr(self.jsrc)
        r.freqs = ptr(self.freqs)
        return r


def source_box(s, direction, line):
    """(field id, [x0,x1,y0,y1,z0,z1]) of update_source for this run."""
    shift = line * s.pp
    if s.scheme in ('tm', 'te'):
        m = 1 if s.scheme == 'tm' else 2
        y0, y1 = s.by + shift + m, s.by + shift + s.ny_swg - m
        x0, x1 = (s.sx, s.sx + 1) if direction == 'f' else (s.nx - s.sx - 1, s.nx - s.sx)
        box = [x0, x1, y0, y1, 0, 1]
        fid = _lib.EZ if s.scheme == 'tm' else _lib.HZ
        lim = (s.nx, s.ny, 1)
    else:
        m = 1 if s.source_condition == 'TE' else 2
        x0, x1 = s.bx + shift + m, s.bx + shift + s.nx_swg - m
        y0, y1 = s.by + m, s.by + s.ny_swg - m
        z0, z1 = (s.sz, s.sz + 2) if direction == 'f' else (s.nz - s.sz - 1, s.nz - s.sz)
        box = [x0, x1, y0, y1, z0, z1]
        fid = _lib.EY if s.source_condition == 'TE' else _lib.HY
        lim = (s.nx, s.ny, s.nz)
    for a in range(3):   # the reference kernel guard only reaches cells inside the grid
        box[2 * a] = min(max(box[2 * a], 0), lim[a])
        box[2 * a + 1] = min(max(box[2 * a + 1], box[2 * a]), lim[a])
    return fid, box


def _window(s):
    N = s.num_cpml
    if s.scheme == '3d':
        return (N, s.nx - N), (N, s.ny - N)
    return None, (N, s.ny - N)


def expand_dft(s, comp_arrays):
    """Compact window phasors -> the reference's full-plane arrays (zeros outside)."""
    nf = len(s.sparam_freqs)
    wx, wy = _window(s)
    out = {}
    for key, z in comp_arrays.items():
        if s.scheme == '3d':
            full = np.zeros((nf, s.nx, s.ny), dtype=np.complex128)
            full[:, wx[0]:wx[1], wy[0]:wy[1]] = z.reshape(nf, wx[1] - wx[0], wy[1] - wy[0])
        else:
            full = np.zeros((nf, s.ny), dtype=np.complex128)
            full[:, wy[0]:wy[1]] = z.reshape(nf, wy[1] - wy[0])
        out[key] = full
    return out


COMP_KEYS = {'tm': ('e', 'h'), 'te': ('e', 'h'), '3d': ('ex', 'ey', 'hx', 'hy')}
FIELDS = {'tm': ('ez', 'hx', 'hy'), 'te': ('ex', 'ey', 'hz'), '3d': ('ex', 'ey', 'ez', 'hx', 'hy', 'hz')}


def _window_cells(s):
    wx, wy = _window(s)
    return (wy[1] - wy[0]) * ((wx[1] - wx[0]) if wx else 1)


class Solver:
    """One run on the device (a batch of one inside the library)."""

    def __init__(self, setup, fg, direction='f', line=0, cb=None):
        self.arrays = RunArrays(setup, fg, direction, line, cb)
        self.setup = setup
        self._desc = self.arrays.desc()
        self._in = self.arrays.inputs()
        h = C.c_void_p()
        check(lib().oidt_solver_create(C.byref(self._desc), C.byref(self._in), C.byref(h)))
        self.h = h

    @classmethod
    def from_arrays(cls, arrays, nt, precision=np.float32, shape=None):
        """A run from any object with ``desc()`` / ``inputs()`` (the study
        drivers' own array sets); ``setup`` then only carries nt, precision
        and the grid shape."""
        self = cls.__new__(cls)
        self.arrays = arrays
        self._desc = arrays.desc()
        self._in = arrays.inputs()
        d = self._desc
        self.setup = SimpleNamespace(nt=nt, precision=precision,
                                     shape=shape or ((d.nx, d.ny, d.nz) if d.nz > 1 else (d.nx, d.ny)),
                                     scheme=None, sparam_freqs=np.zeros(0))
        h = C.c_void_p()
        check(lib().oidt_solver_create(C.byref(self._desc), C.byref(self._in), C.byref(h)))
        self.h = h
        return self

    def set_launch(self, xchunk=0, grid_blocks=0, variant=0):
        """Kernel form and x-chunk depth (before the first advance); 0 keeps
        the current value (the library default for the scheme)."""
        check(lib().oidt_solver_set_launch(self.h, int(xchunk), int(grid_blocks), int(variant)))
        return self

    def advance(self, nsteps):
        check(lib().oidt_solver_advance(self.h, int(nsteps)))
        return self

    def sync(self):
        check(lib().oidt_solver_sync(self.h))
        return self

    def run(self):
        return self.advance(self.setup.nt).sync()

    def steps(self):
        n = C.c_int32()
        check(lib().oidt_solver_steps(self.h, C.byref(n)))
        return n.value

    def read_dft(self, port, comp):
        n = len(self.setup.sparam_freqs) * _window_cells(self.setup)
        re = np.zeros(n)
        im = np.zeros(n)
        check(lib().oidt_solver_read_dft(self.h, port, comp, ptr(re), ptr(im)))
        return re + 1j * im

    def dft(self):
        """{(component key, port 1|2): complex [nf, ny] (2D) or [nf, nx, ny] (3D)}."""
        keys = COMP_KEYS[self.setup.scheme]
        raw = {(k, p + 1): self.read_dft(p, c) for p in (0, 1) for c, k in enumerate(keys)}
        return expand_dft(self.setup, raw)

    def field(self, name):
        s = self.setup
        a = np.zeros(s.shape, dtype=s.precision)
        check(lib().oidt_solver_read_field(self.h, _lib.FIELD_ID[name], ptr(a), a.nbytes))
        return a

    def add_monitor(self, mon):
        """Attach a ``Monitor`` before the first advance; returns its index."""
        d = mon.desc()
        idx = C.c_int32()
        check(lib().oidt_solver_add_monitor(self.h, C.byref(d), C.byref(idx)))
        self.monitors = getattr(self, 'monitors', []) + [mon]
        return idx.value

    def read_monitor(self, idx):
        """DFT monitor: complex re + 1j*im over the box (run precision parts);
        capture monitor: [nrec, *box shape] records."""
        mon = self.monitors[idx]
        re, im = mon.buffers(self.setup.precision)
        check(lib().oidt_solver_read_monitor(self.h, idx, ptr(re), ptr(im), re.nbytes))
        return mon.result(re, im)

    def fields(self):
        return {k: self.field(k) for k in FIELDS[self.setup.scheme]}

    def close(self):
        if getattr(self, 'h', None):
            check(lib().oidt_solver_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Many runs of one scheme/precision advanced by the same launches."""

    def __init__(self, scheme, precision, max_slots=64):
        self.scheme = scheme
        self.precision = precision
        h = C.c_void_p()
        check(lib().oidt_batch_create(_lib.SCHEME_ID[scheme], _lib.F32 if precision == np.float32 else _lib.F64,
                                      int(max_slots), C.byref(h)))
        self.h = h
        self.runs = {}

    def add(self, setup, fg, direction='f', line=0, cb=None):
        return self.add_arrays(RunArrays(setup, fg, direction, line, cb), setup)

    def add_arrays(self, arrays, setup):
        """Add a run from any object with ``desc()`` / ``inputs()``; ``setup``
        describes it to the Python side (nt, precision, shape, frequencies)."""
        d, r = arrays.desc(), arrays.inputs()
        slot = C.c_int32()
        check(lib().oidt_batch_add(self.h, C.byref(d), C.byref(r), C.byref(slot)))
        self.runs[slot.value] = setup
        self.monitors = getattr(self, 'monitors', {})
        self.monitors[slot.value] = []
        return slot.value

    def add_monitor(self, slot, mon):
        """Attach a ``Monitor`` to run ``slot`` before it starts; returns its index."""
        d = mon.desc()
        idx = C.c_int32()
        check(lib().oidt_batch_add_monitor(self.h, int(slot), C.byref(d), C.byref(idx)))
        self.monitors[slot].append(mon)
        return idx.value

    def read_monitor(self, slot, idx):
        mon = self.monitors[slot][idx]
        re, im = mon.buffers(self.runs[slot].precision)
        check(lib().oidt_batch_read_monitor(self.h, int(slot), int(idx), ptr(re), ptr(im), re.nbytes))
        return mon.result(re, im)

    def set_launch(self, xchunk=0, grid_blocks=0, variant=0):
        """Kernel form, x-chunk depth and grid cap; 0 keeps the current value."""
        check(lib().oidt_batch_set_launch(self.h, int(xchunk), int(grid_blocks), int(variant)))

    def launch(self):
        """Current (variant, xchunk, grid_blocks) of the batch."""
        xc, g, v = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().oidt_batch_get_launch(self.h, C.byref(xc), C.byref(g), C.byref(v)))
        return v.value, xc.value, g.value

    def advance(self, nchunks):
        check(lib().oidt_batch_advance(self.h, int(nchunks)))

    def prepare(self):
        """Build the launch lists and capture the chunk graph now (the next
        advance then only replays it)."""
        check(lib().oidt_batch_prepare(self.h))

    def sync(self):
        check(lib().oidt_batch_sync(self.h))

    def progress(self):
        """(chunks enqueued, chunks completed by the GPU), non-blocking."""
        e, d = C.c_int64(), C.c_int64()
        check(lib().oidt_batch_progress(self.h, C.byref(e), C.byref(d)))
        return e.value, d.value

    def poll(self, slot):
        n = C.c_int32()
        check(lib().oidt_batch_poll(self.h, int(slot), C.byref(n)))
        return n.value

    def read_dft(self, slot, port, comp):
        s = self.runs[slot]
        n = len(s.sparam_freqs) * _window_cells(s)
        re = np.zeros(n)
        im = np.zeros(n)
        check(lib().oidt_batch_read_dft(self.h, int(slot), port, comp, ptr(re), ptr(im)))
        return re + 1j * im

    def port_power(self, slot, port, row0, nrows, sub=None):
        """Complex port power [nf] summed on the device over DFT-window rows
        [row0, row0 + nrows) of ``port`` (0 or 1) of run ``slot``, after
        subtracting the phasors of ``sub`` = (slot, port) if given."""
        nf = len(self.runs[slot].sparam_freqs)
        re = np.zeros(nf)
        im = np.zeros(nf)
        ss, sp = sub if sub is not None else (-1, 0)
        check(lib().oidt_batch_port_power(self.h, int(slot), int(port), int(ss), int(sp), int(row0), int(nrows),
                                          ptr(re), ptr(im)))
        return re + 1j * im

    def dft(self, slot):
        s = self.runs[slot]
        keys = COMP_KEYS[s.scheme]
        raw = {(k, p + 1): self.read_dft(slot, p, c) for p in (0, 1) for c, k in enumerate(keys)}
        return expand_dft(s, raw)

    def field(self, slot, name):
        s = self.runs[slot]
        a = np.zeros(s.shape, dtype=s.precision)
        check(lib().oidt_batch_read_field(self.h, int(slot), _lib.FIELD_ID[name], ptr(a), a.nbytes))
        return a

    def remove(self, slot):
        check(lib().oidt_batch_remove(self.h, int(slot)))
        self.runs.pop(slot, None)

    def close(self):
        if getattr(self, 'h', None):
            check(lib().oidt_batch_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Monitor:
    """A field monitor (include/oidt_fdtd.h, oidt_monitor_desc): the stencil
    sum of ``field`` over ``points`` (cell offsets), times ``scale``, at every
    cell of ``box`` ((x0, x1), (y0, y1)[, (z0, z1)]), taken after the E
    (``after_h`` False) or H half-step of every step n, as record r = n + shift.
    ``cs`` ([nrec, 2] cos/sin weights) makes it a single-frequency DFT;
    without it every record is kept."""

    def __init__(self, field, points, scale, box, nrec, after_h=True, shift=0, cs=None):
        self.field = field
        self.points = [tuple(p) + (0,) * (3 - len(p)) for p in points]
        self.scale = float(scale)
        self.box = [tuple(b) for b in box] + [(0, 1)] * (3 - len(box))
        self.nrec = int(nrec)
        self.after_h = bool(after_h)
        self.shift = int(shift)
        self.cs = None if cs is None else np.ascontiguousarray(cs, dtype=np.float64).reshape(-1, 2)
        if self.cs is not None and len(self.cs) != self.nrec:
            raise ValueError('cs needs one (cos, sin) row per record')

    @property
    def extent(self):
        return tuple(b1 - b0 for b0, b1 in self.box)

    def desc(self):
        d = _lib.MonitorDesc()
        d.field = _lib.FIELD_ID[self.field]
        d.after_h = int(self.after_h)
        d.mode = _lib.MON_CAPTURE if self.cs is None else _lib.MON_DFT
        d.npts = len(self.points)
        for q, p in enumerate(self.points):
            for a in range(3):
                d.off[q][a] = p[a]
        d.scale = self.scale
        for a, (b0, b1) in enumerate(self.box):
            d.box[2 * a], d.box[2 * a + 1] = b0, b1
        d.shift = self.shift
        d.nrec = self.nrec
        d.cs = None if self.cs is None else self.cs.ctypes.data_as(C.c_void_p)
        return d

    def buffers(self, precision):
        cells = int(np.prod(self.extent))
        if self.cs is None:
            return np.zeros(self.nrec * cells, dtype=precision), None
        return np.zeros(cells, dtype=precision), np.zeros(cells, dtype=precision)

    def result(self, re, im):
        shape = tuple(e for e in self.extent)
        if im is None:
            return re.reshape((self.nrec,) + shape)
        return re.reshape(shape), im.reshape(shape)


def footprint(setup, direction='f', line=0):
    """Device bytes a run of ``setup`` holds (admission by footprint)."""
    fg = np.zeros(setup.shape, dtype=bool)
    arr = RunArrays(setup, fg, direction, line)
    d = arr.desc()
    b = C.c_uint64()
    check(lib().oidt_run_footprint(C.byref(d), C.byref(b)))
    return b.value
