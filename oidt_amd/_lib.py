"""ctypes binding of liboidt_fdtd.so (the C ABI of include/oidt_fdtd.h).

The library is the only compute path: if it is missing or fails to load,
``lib()`` raises; there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'liboidt_fdtd.so')

ABI_VERSION = 1
OK, EINVAL, EHIP, ENOMEM, ESTATE, EFULL = 0, -1, -2, -3, -4, -5
TMZ, TEZ, YEE3D, TMZ_LOSSY = 0, 1, 2, 3
F32, F64 = 0, 1
EX, EY, EZ, HX, HY, HZ = range(6)
SCHEME_ID = {'tm': TMZ, 'te': TEZ, '3d': YEE3D, 'tm_lossy': TMZ_LOSSY}
FIELD_ID = {'ex': EX, 'ey': EY, 'ez': EZ, 'hx': HX, 'hy': HY, 'hz': HZ}


class RunDesc(C.Structure):
    _fields_ = [('abi_version', C.c_int32), ('scheme', C.c_int32), ('precision', C.c_int32),
                ('nx', C.c_int32), ('ny', C.c_int32), ('nz', C.c_int32), ('num_cpml', C.c_int32),
                ('offset', C.c_int32), ('nt', C.c_int32), ('nf', C.c_int32), ('src_field', C.c_int32),
                ('src_box', C.c_int32 * 6), ('port', C.c_int32 * 2),
                ('dx', C.c_double), ('dt', C.c_double), ('db', C.c_double)]


class RunInputs(C.Structure):
    _fields_ = [('cb', C.c_void_p * 3), ('kappa_e', C.c_void_p * 3), ('kappa_h', C.c_void_p * 3),
                ('be', C.c_void_p), ('ce', C.c_void_p), ('bh', C.c_void_p), ('ch', C.c_void_p),
                ('jsrc', C.c_void_p), ('freqs', C.c_void_p)]


MON_DFT, MON_CAPTURE = 0, 1


class MonitorDesc(C.Structure):
    _fields_ = [('field', C.c_int32), ('after_h', C.c_int32), ('mode', C.c_int32), ('npts', C.c_int32),
                ('off', (C.c_int32 * 3) * 4), ('scale', C.c_double), ('box', C.c_int32 * 6),
                ('shift', C.c_int32), ('nrec', C.c_int32), ('cs', C.c_void_p)]


class DeviceInfo(C.Structure):
    _fields_ = [('name', C.c_char * 64), ('arch', C.c_char * 32), ('compute_units', C.c_int32),
                ('device_id', C.c_int32), ('hbm_total_bytes', C.c_uint64), ('hbm_free_bytes', C.c_uint64)]


# name -> (restype, argtypes); every symbol declared in include/oidt_fdtd.h
SIGNATURES = {
    'oidt_abi_version': (C.c_int32, []),
    'oidt_last_error': (C.c_char_p, []),
    'oidt_device_info_get': (C.c_int, [C.POINTER(DeviceInfo)]),
    'oidt_run_footprint': (C.c_int, [C.POINTER(RunDesc), C.POINTER(C.c_uint64)]),
    'oidt_chunk_steps': (C.c_int32, []),
    'oidt_solver_create': (C.c_int, [C.POINTER(RunDesc), C.POINTER(RunInputs), C.POINTER(C.c_void_p)]),
    'oidt_solver_advance': (C.c_int, [C.c_void_p, C.c_int32]),
    'oidt_solver_set_launch': (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]),
    'oidt_solver_sync': (C.c_int, [C.c_void_p]),
    'oidt_solver_steps': (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    'oidt_solver_read_dft': (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p]),
    'oidt_solver_read_field': (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_uint64]),
    'oidt_solver_add_monitor': (C.c_int, [C.c_void_p, C.POINTER(MonitorDesc), C.POINTER(C.c_int32)]),
    'oidt_solver_read_monitor': (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_uint64]),
    'oidt_solver_destroy': (C.c_int, [C.c_void_p]),
    'oidt_batch_create': (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    'oidt_batch_add': (C.c_int, [C.c_void_p, C.POINTER(RunDesc), C.POINTER(RunInputs), C.POINTER(C.c_int32)]),
    'oidt_batch_advance': (C.c_int, [C.c_void_p, C.c_int32]),
    'oidt_batch_prepare': (C.c_int, [C.c_void_p]),
    'oidt_batch_sync': (C.c_int, [C.c_void_p]),
    'oidt_batch_poll': (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]),
    'oidt_batch_progress': (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    'oidt_batch_read_dft': (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p]),
    'oidt_batch_read_field': (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_uint64]),
    'oidt_batch_add_monitor': (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(MonitorDesc), C.POINTER(C.c_int32)]),
    'oidt_batch_read_monitor': (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                                          C.c_uint64]),
    'oidt_batch_remove': (C.c_int, [C.c_void_p, C.c_int32]),
    'oidt_batch_set_launch': (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]),
    'oidt_batch_port_power': (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                        C.c_int32, C.c_void_p, C.c_void_p]),
    'oidt_batch_port_power_noalloc': (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                                C.c_int32, C.c_void_p, C.c_void_p]),
    'oidt_batch_get_launch': (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    'oidt_batch_destroy': (C.c_int, [C.c_void_p]),
    'oidt_batch_profile': (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    'oidt_solver_profile': (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    'oidt_batch_profile_all': (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    'oidt_solver_profile_all': (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    'oidt_copy_probe': (C.c_int, [C.c_uint64, C.c_int32, C.POINTER(C.c_double)]),
}

_LIB = None


class OidtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__('oidt error %d: %s' % (code, msg))
        self.code = code


def lib():
    """Load liboidt_fdtd.so (built in-tree by oidt_amd.build); raise if absent."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise OidtError(ESTATE, 'liboidt_fdtd.so not built: run `python -m oidt_amd.build`')
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        if L.oidt_abi_version() != ABI_VERSION:
            raise OidtError(ESTATE, 'liboidt_fdtd.so ABI version mismatch')
        _LIB = L
    return _LIB


def check(rc):
    if rc != OK:
        raise OidtError(rc, lib().oidt_last_error().decode(errors='replace'))
    return rc


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def device_info():
    d = DeviceInfo()
    check(lib().oidt_device_info_get(C.byref(d)))
    return {'name': d.name.decode(), 'arch': d.arch.decode(), 'compute_units': d.compute_units,
            'hbm_total_bytes': d.hbm_total_bytes, 'hbm_free_bytes': d.hbm_free_bytes}


def copy_probe(nbytes, reps=20):
    ms = C.c_double()
    check(lib().oidt_copy_probe(int(nbytes), int(reps), C.byref(ms)))
    return ms.value


def chunk_steps():
    return int(lib().oidt_chunk_steps())
