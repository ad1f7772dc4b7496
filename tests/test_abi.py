"""The C ABI library loads and exports every symbol include/oidt_fdtd.h
declares; host-side validation errors need no GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from oidt_amd import _lib, configs, params
from oidt_amd.solver import RunArrays, footprint

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, 'include', 'oidt_fdtd.h')).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return set(re.findall(r'\b(oidt_[a-z_0-9]+)\s*\(', text))


def test_header_and_binding_agree():
    assert declared() == set(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    for name in declared():
        assert hasattr(L, name), name
    assert L.oidt_abi_version() == _lib.ABI_VERSION
    assert _lib.chunk_steps() == 32


def _desc(kind='tm', **over):
    p = configs.small(kind, 'float32')
    s = params.explicit_setup(p)
    ra = RunArrays(s, np.zeros(s.shape, bool))
    d = ra.desc()
    for k, v in over.items():
        setattr(d, k, v)
    return d


@pytest.mark.parametrize('field,value,msg', [
    ('abi_version', 99, 'abi_version'), ('scheme', 7, 'scheme'), ('precision', 5, 'precision'),
    ('nz', 4, 'nz == 1'), ('num_cpml', 60, 'CPML'), ('nt', -1, 'negative')])
def test_validation_errors(field, value, msg):
    d = _desc(**{field: value})
    b = C.c_uint64()
    rc = _lib.lib().oidt_run_footprint(C.byref(d), C.byref(b))
    assert rc == _lib.EINVAL
    assert msg in _lib.lib().oidt_last_error().decode()


def test_bad_port_and_source_rejected():
    d = _desc()
    d.port[1] = d.nx - 1
    assert _lib.lib().oidt_run_footprint(C.byref(d), C.byref(C.c_uint64())) == _lib.EINVAL
    d = _desc()
    d.src_field = _lib.HZ          # TMz drives Ez only
    assert _lib.lib().oidt_run_footprint(C.byref(d), C.byref(C.c_uint64())) == _lib.EINVAL


def test_3d_tm_source_box_in_cpml_stencil_rejected():
    p = configs.small('3d', 'float32', source_condition='TM')
    s = params.explicit_setup(p)
    ra = RunArrays(s, np.zeros(s.shape, bool))
    d = ra.desc()
    assert _lib.lib().oidt_run_footprint(C.byref(d), C.byref(C.c_uint64())) == _lib.OK
    d.src_box[4], d.src_box[5] = 3, 5      # inside the low z CPML slab
    assert _lib.lib().oidt_run_footprint(C.byref(d), C.byref(C.c_uint64())) == _lib.EINVAL
    assert 'CPML' in _lib.lib().oidt_last_error().decode()


def test_footprint_scales_with_grid():
    a = footprint(params.explicit_setup(configs.slab2d(800, precision='float32', nt=100)))
    b = footprint(params.explicit_setup(configs.slab2d(1600, precision='float32', nt=100)))
    assert 1.6 < b / a < 2.1      # fixed per-run state (DFT ring and accumulators) on top
    c = footprint(params.explicit_setup(configs.config1()))
    # two sets of 6 FP32 fields and psi (the fused step alternates them), the
    # 1-byte material map and the DFT state of the 128x96x320 run
    assert 250e6 < c < 400e6


def test_padded_field_size_limit():
    """The kernels address a field with a 32-bit byte offset into the padded
    layout (unit-stride axis rounded up to 128 bytes): a grid whose FP32 field
    fits 4 GiB only before padding is refused before anything is allocated."""
    d = _desc(ny=33, nx=17_000_000)          # 2.2 GB of cells, 4.35 GB once y pads to 64
    d.port[0], d.port[1] = 20, 40
    rc = _lib.lib().oidt_run_footprint(C.byref(d), C.byref(C.c_uint64()))
    assert rc == _lib.EINVAL
    assert 'padding' in _lib.lib().oidt_last_error().decode()
    d = _desc(ny=64, nx=15_000_000)           # 3.84 GB per field, no padding needed
    d.port[0], d.port[1] = 20, 40
    assert _lib.lib().oidt_run_footprint(C.byref(d), C.byref(C.c_uint64())) == _lib.OK


def test_3d_psi_slab_pair_limit():
    """The 3D march addresses a high-side psi slab from the low-side array of
    the same component by a 32-bit offset: a grid whose fields fit 4 GiB but
    whose x slabs (with the other side and the second set between them) do
    not is refused before anything is allocated."""
    d = _desc('3d', nx=25, ny=2000, nz=20000)       # 4.0 GB per field, 3 x 1.6 GB of x slab
    rc = _lib.lib().oidt_run_footprint(C.byref(d), C.byref(C.c_uint64()))
    assert rc == _lib.EINVAL
    assert 'psi slab pair' in _lib.lib().oidt_last_error().decode()
    d = _desc('3d', nx=31, ny=1700, nz=20000)       # 4.2 GB per field, 3 x 1.36 GB of x slab
    assert _lib.lib().oidt_run_footprint(C.byref(d), C.byref(C.c_uint64())) == _lib.OK


def test_struct_layouts_match_the_header(tmp_path):
    """ctypes mirrors of the ABI structs have the C compiler's sizes and field
    offsets (compiled against include/oidt_fdtd.h with the host gcc)."""
    import shutil
    import subprocess
    if shutil.which('gcc') is None:
        pytest.skip('no host C compiler')
    structs = {'oidt_run_desc': _lib.RunDesc, 'oidt_run_inputs': _lib.RunInputs,
               'oidt_monitor_desc': _lib.MonitorDesc, 'oidt_device_info': _lib.DeviceInfo}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "oidt_fdtd.h"', 'int main(void) {']
    for cname, py in structs.items():
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for f in py._fields_:
            lines.append('printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, f[0], cname, f[0]))
    lines.append('return 0; }')
    src = tmp_path / 'layout.c'
    src.write_text('\n'.join(lines))
    exe = tmp_path / 'layout'
    subprocess.run(['gcc', '-I', os.path.join(ROOT, 'include'), str(src), '-o', str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split('\n')
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in structs.items():
        assert got[(cname, 'size')] == C.sizeof(py), cname
        for f in py._fields_:
            assert got[(cname, f[0])] == getattr(py, f[0]).offset, (cname, f[0])


def test_lossy_tm_description_checks():
    """The lossy TMz scheme: offset 1, no port DFT, source inside the update range."""
    from oidt_amd import variants
    c = variants.config('tm', 'wave_impedance', sgl_wg_length=1.5e-6, port_length=0.3e-6, points_per_wl=16,
                        ny_clad_wl=0.4, buffer_xhat=5, num_cpml=6)
    a = variants.LossyTmArrays(c, variants.tm_mask(c, None, None))
    fp = lambda d: _lib.lib().oidt_run_footprint(C.byref(d), C.byref(C.c_uint64()))
    assert fp(a.desc()) == _lib.OK
    for field, value in (('offset', 2), ('nf', 1)):
        d = a.desc()
        setattr(d, field, value)
        assert fp(d) == _lib.EINVAL
    d = a.desc()
    d.src_box[3] = c.ny
    assert fp(d) == _lib.EINVAL
