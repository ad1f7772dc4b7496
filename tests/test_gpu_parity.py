"""Parity of the HIP path (through the C ABI) against the FP64 oracle.

Tolerances: FP32 fields and DFT phasors within relative L2 1e-4
of the oracle; FP64 runs agree to 1e-10 (same algorithm, same precision; only
the order of floating-point operations differs)."""
import numpy as np
import pytest

from oidt_amd import configs
from oidt_amd.solver import Solver

from helpers import device_inputs, oracle_run, rel_l2, window

TOL = {'float64': 1e-10, 'float32': 1e-4}

CASES = [(s, p, d) for s in ('tm', 'te', '3d') for p in ('float64', 'float32') for d in ('f', 'b')]


def _case(scheme, precision, rough=True, **over):
    p = configs.small(scheme, precision, **over)
    if rough:
        p.update(rough_std=15e-9, rough_acl=40e-9, seed=7)
    return p


@pytest.mark.gpu
@pytest.mark.parametrize('scheme,precision,direction', CASES)
def test_fields_and_dft_match_oracle(scheme, precision, direction):
    p = _case(scheme, precision)
    print('case', scheme, precision, direction, p['nx'], p['ny'], p.get('nz'), p['nt'], flush=True)
    s, up, lo, fg = device_inputs(p)
    o, FG = oracle_run(p, up, lo, direction)
    assert np.array_equal(fg, FG)
    dev = Solver(s, fg, direction).run()
    assert dev.steps() == s.nt
    tol = TOL[precision]
    for name, ref in o.fields().items():
        got = dev.field(name)
        assert np.abs(ref).max() > 0, name
        err = rel_l2(got, ref)
        print(name, 'rel L2', err, flush=True)
        assert err < tol, (name, err)
    w = window(scheme, o.c)
    od = o.dft()
    for key, z in dev.dft().items():
        err = rel_l2(z[w], od[key][w])
        print(key, 'dft rel L2', err, flush=True)
        assert err < tol, (key, err)
        outside = np.ones(z.shape, bool)
        outside[w] = False
        assert not np.any(z[outside])
    dev.close()


@pytest.mark.gpu
@pytest.mark.parametrize('scheme', ['tm', 'te', '3d'])
def test_partial_chunk_and_advance_in_steps(scheme):
    """nt not a multiple of the chunk; advancing in pieces equals one run."""
    p = _case(scheme, 'float64', nt=101)
    s, up, lo, fg = device_inputs(p)
    o, _ = oracle_run(p, up, lo)
    dev = Solver(s, fg)
    dev.advance(40).advance(40).advance(40).sync()
    assert dev.steps() == 101
    for name, ref in o.fields().items():
        assert rel_l2(dev.field(name), ref) < 1e-10, name


@pytest.mark.gpu
def test_3d_tm_source_on_hy():
    p = _case('3d', 'float64', source_condition='TM')
    s, up, lo, fg = device_inputs(p)
    o, _ = oracle_run(p, up, lo)
    dev = Solver(s, fg).run()
    for name, ref in o.fields().items():
        assert rel_l2(dev.field(name), ref) < 1e-10, name


@pytest.mark.gpu
def test_zero_source_gives_zero_fields():
    p = _case('tm', 'float32', source_amp=0.0)
    s, up, lo, fg = device_inputs(p)
    dev = Solver(s, fg).run()
    for name in ('ez', 'hx', 'hy'):
        assert not np.any(dev.field(name))
