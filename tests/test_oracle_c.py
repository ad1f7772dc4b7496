"""The oracle's C restatement of the 3D step (oracle/fdtd3d.c) against the
NumPy oracle (oracle/fdtd.py, Yee3D), which the physics pins check: the two
must agree to the last bit on every field, psi array and DFT accumulator
(same expressions, same operand order, no fused multiply-add)."""
import numpy as np
import pytest

from oidt_amd import configs
from oracle import fdtd as ofdtd
from oracle import fdtd_c
from oracle import geometry as ogeo
from oracle import setup as osetup


def _pair(p, direction):
    s, up, lo = configs.realise(p)
    out = []
    for make in (ofdtd.make, fdtd_c.make):
        c = osetup.explicit_setup(p)
        out.append(make(c, ogeo.fg_mask(c, up, lo), direction).run())
    return out


@pytest.mark.parametrize('source,direction', [('TE', 'f'), ('TE', 'b'), ('TM', 'f'), ('TM', 'b')])
def test_c_oracle_equals_numpy_oracle(source, direction):
    p = configs.small('3d', nt=90, source_condition=source)
    p.update(rough_std=12e-9, rough_acl=40e-9, seed=4)
    a, b = _pair(p, direction)
    assert isinstance(b, fdtd_c.Yee3DC)
    for k, v in a.fields().items():
        assert np.abs(v).max() > 0, k
        assert np.array_equal(v, b.fields()[k]), k
    for k, (lo, hi) in a.psi.items():
        assert np.array_equal(lo, b.psi[k][0]) and np.array_equal(hi, b.psi[k][1]), k
    for k, v in a.dft().items():
        np.testing.assert_allclose(b.dft()[k], v, rtol=0, atol=1e-14 * np.abs(v).max())


def test_c_oracle_steps_in_pieces():
    p = configs.small('3d', nt=40)
    s, up, lo = configs.realise(p)
    c = osetup.explicit_setup(p)
    a = fdtd_c.make(c, ogeo.fg_mask(c, up, lo)).run()
    c = osetup.explicit_setup(p)
    b = fdtd_c.make(c, ogeo.fg_mask(c, up, lo))
    b.run(13)
    b.run(27)
    for k, v in a.fields().items():
        assert np.array_equal(v, b.fields()[k]), k
