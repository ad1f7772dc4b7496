"""Shared test helpers: build the same run for the device path and the oracle."""
import numpy as np

from oidt_amd import configs, materials
from oracle import fdtd as ofdtd
from oracle import geometry as ogeo
from oracle import setup as osetup


def oracle_run(p, up=None, lo=None, direction='f', line=0, nsteps=None):
    c = osetup.explicit_setup(p)
    FG = ogeo.fg_mask(c, up, lo)
    o = ofdtd.make(c, FG, direction, line)
    o.run(nsteps)
    return o, FG


def device_inputs(p):
    s, up, lo = configs.realise(p)
    fg = materials.fg_mask(s, up, lo)
    return s, up, lo, fg


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.complex128 if np.iscomplexobj(a) or np.iscomplexobj(b) else np.float64)
    b = np.asarray(b, dtype=a.dtype)
    nb = np.linalg.norm(b.ravel())
    return np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0)


def window(scheme, c):
    N = c.num_cpml
    if scheme == '3d':
        return (slice(None), slice(N, c.nx - N), slice(N, c.ny - N))
    return (slice(None), slice(N, c.ny - N))


