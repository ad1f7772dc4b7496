"""oidt_batch_port_power_noalloc (the allocation-free variant) through the C
ABI gives exactly what oidt_batch_port_power gives: two calls in a row, the
second on a smaller input (fewer frequencies and window rows), so the fixed
device array keeps the first call's values beyond the second call's range;
with and without a subtracted run; in 2D and 3D.  Runs with more
frequencies than the array holds are refused."""
import ctypes as C

import numpy as np
import pytest

from oidt_amd import _lib, configs, materials
from oidt_amd.solver import Batch

MAX_NF = 16384


def _power(fn, b, slot, port, sub, row0, nrows, nf):
    re = np.full(nf, np.nan)
    im = np.full(nf, np.nan)
    ss, sp = sub if sub is not None else (-1, 0)
    _lib.check(getattr(_lib.lib(), fn)(b.h, int(slot), int(port), int(ss), int(sp), int(row0), int(nrows),
                                       _lib.ptr(re), _lib.ptr(im)))
    return re, im


def _add(b, p):
    s, up, lo = configs.realise(p)
    return b.add(s, materials.fg_mask(s, up, lo)), s


@pytest.mark.gpu
@pytest.mark.parametrize('scheme', ['tm', '3d'])
def test_noalloc_port_power_equals_port_power(scheme):
    big = configs.small(scheme, 'float32', nf=8)
    big.update(rough_std=15e-9, rough_acl=40e-9, seed=3)
    inc = dict(configs.incident(big), nf=8)
    small = configs.small(scheme, 'float32', nf=3, nt=200)
    if scheme == '3d':
        small.update(nx=30, bx=9, pp=10)
    else:
        small.update(ny=60, by=24, pp=40)
    b = Batch(scheme, np.float32, max_slots=3)
    try:
        sa, s_big = _add(b, big)
        si, _ = _add(b, inc)
        sb, s_small = _add(b, small)
        b.advance(-(-max(big['nt'], small['nt']) // 32))
        b.sync()
        rows_big = s_big.pp
        rows_small = s_small.pp
        calls = [(sa, 0, (si, 0), 0, rows_big, 8),         # first: all rows, incident subtracted
                 (sb, 1, None, 2, rows_small - 4, 3)]     # second: fewer frequencies and rows
        for slot, port, sub, row0, nrows, nf in calls:
            print('case', scheme, 'slot', slot, 'port', port, 'sub', sub, 'rows', row0, nrows, 'nf', nf, flush=True)
            want = _power('oidt_batch_port_power', b, slot, port, sub, row0, nrows, nf)
            got = _power('oidt_batch_port_power_noalloc', b, slot, port, sub, row0, nrows, nf)
            assert np.all(np.isfinite(want[0])) and np.any(want[0] != 0)
            assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    finally:
        b.close()


@pytest.mark.gpu
def test_noalloc_port_power_refuses_too_many_frequencies():
    p = configs.small('tm', 'float32', nf=MAX_NF + 1, nt=32)
    b = Batch('tm', np.float32, max_slots=1)
    try:
        slot, s = _add(b, p)
        re = np.zeros(MAX_NF + 1)
        im = np.zeros(MAX_NF + 1)
        rc = _lib.lib().oidt_batch_port_power_noalloc(b.h, slot, 0, -1, 0, 0, s.pp, _lib.ptr(re), _lib.ptr(im))
        assert rc == _lib.EINVAL
        assert b'at most 16384 frequencies' in _lib.lib().oidt_last_error()
    finally:
        b.close()
