"""Loss and wave-impedance study drivers on the device against the FP64 study
oracle (oracle/studies.py), through the C ABI's field monitors (and, for the
TM studies, the lossy TMz scheme).

Tolerance: the study drivers run in float32 (their kernels are float32 only);
device outputs are within relative L2 1e-4 of the FP64 oracle, the FP32 field
bar.  The single-frequency DFTs also keep float32 running sums, as the
reference's accumulators do."""
import ctypes as C

import numpy as np
import pytest

from oidt_amd import _lib, variants
from oidt_amd.solver import Monitor, Solver
from oracle import studies as ost

from helpers import rel_l2
from test_studies import SMALL_3D, SMALL_TE, SMALL_TM

SMALL = {'tm': SMALL_TM, 'te': SMALL_TE, '3d': SMALL_3D}

CASES = sorted(variants.DEFAULTS)


@pytest.mark.gpu
@pytest.mark.parametrize('scheme,study', CASES)
def test_study_outputs_match_oracle(scheme, study):
    c = variants.config(scheme, study, **SMALL[scheme])
    print('case', scheme, study, c.nx, c.ny, getattr(c, 'nz', 1), c.nt, flush=True)
    dev = variants.run(c)
    length = (c.p2z - c.p1z) if scheme == '3d' else (c.p2x - c.p1x)
    up, lo = variants._profiles(c, length)
    ref = ost.make(c, up, lo).run()
    for k, v in dev.items():
        r = ref[k]
        assert v.shape == r.shape, (k, v.shape, r.shape)
        assert v.dtype == (np.complex64 if study == 'scattering_loss' else np.float32), (k, v.dtype)
        assert np.abs(r).max() > 0, k
        err = rel_l2(v, r)
        print(k, 'rel L2', err, flush=True)
        assert err < 1e-4, (k, err)


@pytest.mark.gpu
def test_study_record_edges():
    """Records outside the reference's loops stay zero: 3D lines end at
    i = nx - 3 (i < ie), 3D plane DFTs cover [N, nx-N) x [N, ny-N) only."""
    c = variants.config('3d', 'scattering_loss', **SMALL_3D)
    c.nt = 96
    c.cosE, c.sinE, c.cosH, c.sinH = (a[:96] for a in (c.cosE, c.sinE, c.cosH, c.sinH))
    out = variants.run(c)
    N = c.num_cpml
    for v in out.values():
        inner = np.zeros(v.shape, bool)
        inner[N:c.nx - N, N:c.ny - N] = True
        assert not np.any(v[~inner])
    c = variants.config('3d', 'wave_impedance', **SMALL_3D)
    c.nt = 96
    for v in variants.run(c).values():
        assert v.shape == (96, c.nx) and not np.any(v[:, c.nx - 2:])


@pytest.mark.gpu
def test_monitor_argument_checks():
    c = variants.config('te', 'wave_impedance', **SMALL_TE)
    s, fg, cb = variants.build(c)
    sol = Solver(s, fg, cb=cb)
    try:
        bad_box = Monitor('ex', [(0, 0)], 1.0, [(0, c.nx + 1), (0, c.ny)], 4)
        with pytest.raises(_lib.OidtError) as e:
            sol.add_monitor(bad_box)
        assert e.value.code == _lib.EINVAL
        with pytest.raises(_lib.OidtError) as e:
            sol.add_monitor(Monitor('ez', [(0, 0)], 1.0, [(0, 4), (0, 4)], 4))   # no Ez in TEz
        assert e.value.code == _lib.EINVAL
        with pytest.raises(_lib.OidtError) as e:
            sol.add_monitor(Monitor('ex', [(0, 2)], 1.0, [(0, 4), (0, 4)], 4))   # offset beyond +-1
        assert e.value.code == _lib.EINVAL
        sol.add_monitor(Monitor('ex', [(0, 0)], 1.0, [(0, 4), (0, 4)], 4))
        with pytest.raises(_lib.OidtError) as e:
            sol.set_launch(8, 0, 3)                 # fused step has no point between the half-steps
        assert e.value.code == _lib.ESTATE
        sol.advance(32).sync()
        with pytest.raises(_lib.OidtError) as e:
            sol.add_monitor(Monitor('ex', [(0, 0)], 1.0, [(0, 4), (0, 4)], 4))   # run already started
        assert e.value.code == _lib.ESTATE
        re = np.zeros(4 * 16 + 1, np.float32)
        rc = _lib.lib().oidt_solver_read_monitor(sol.h, 0, re.ctypes.data_as(C.c_void_p), None, re.nbytes)
        assert rc == _lib.EINVAL                    # wrong buffer size
    finally:
        sol.close()


@pytest.mark.gpu
def test_lossy_tm_rejects_fused_and_port_dft():
    c = variants.config('tm', 'wave_impedance', **SMALL_TM)
    arrays = variants.LossyTmArrays(c, variants.tm_mask(c, None, None))
    sol = Solver.from_arrays(arrays, c.nt)
    try:
        with pytest.raises(_lib.OidtError) as e:
            sol.set_launch(8, 0, 3)
        assert e.value.code == _lib.EINVAL
    finally:
        sol.close()
    d = arrays.desc()
    d.nf = 2
    assert _lib.lib().oidt_run_footprint(C.byref(d), C.byref(C.c_uint64())) == _lib.EINVAL


@pytest.mark.gpu
@pytest.mark.parametrize('scheme', ['tm', 'te'])
def test_study_realisations_batched_equal_alone(scheme):
    """Roughness realisations of a scattering-loss study in one batch (and a
    wave-impedance run beside them) give bit-identical arrays to solo runs."""
    cs = [variants.config(scheme, 'scattering_loss', profile_number=p, **SMALL[scheme]) for p in (0, 1)]
    cs.append(variants.config(scheme, 'wave_impedance', **SMALL[scheme]))
    together = variants.run_batch(cs)
    for c, got in zip(cs, together):
        alone = variants.run(c, (2, 4, 0))
        assert set(got) == set(alone)
        for k in alone:
            np.testing.assert_array_equal(got[k], alone[k], err_msg=(c.study, c.profile_number, k))
    assert not np.array_equal(together[0]['ez' if scheme == 'tm' else 'hz'],
                              together[1]['ez' if scheme == 'tm' else 'hz'])


@pytest.mark.gpu
def test_fp64_monitors_match_oracle_fields():
    """Monitors in a double-precision run: a line record (2-point average)
    and a whole-grid DFT (weights of the caller's choosing) against the FP64
    S-parameter oracle's own fields, step by step, to 1e-10."""
    from oidt_amd import configs
    from helpers import device_inputs
    from oracle import fdtd as ofdtd, geometry as ogeo, setup as osetup
    p = configs.small('te', 'float64', nt=96)
    s, up, lo, fg = device_inputs(p)
    nt, cx = s.nt, s.nx // 2
    w = np.linspace(0.0, 3.0, nt)
    cs = np.stack([np.cos(w), np.sin(w)], 1)
    sol = Solver(s, fg)
    try:
        i_line = sol.add_monitor(Monitor('hz', [(0, 0), (1, 0)], 0.5, [(cx, cx + 1), (0, s.ny)], nt))
        i_dft = sol.add_monitor(Monitor('ey', [(0, 0)], 1.0, [(0, s.nx), (0, s.ny)], nt, cs=cs))
        sol.run()
        line = sol.read_monitor(i_line)[:, 0, :, 0]
        re, im = sol.read_monitor(i_dft)
    finally:
        sol.close()
    c = osetup.explicit_setup(p)
    o = ofdtd.make(c, ogeo.fg_mask(c, up, lo))
    ref_line = np.zeros((nt, s.ny))
    ref = np.zeros((s.nx, s.ny), complex)
    for n in range(nt):
        o.step(n)
        ref_line[n] = 0.5 * (o.Hz[cx] + o.Hz[cx + 1])
        ref += o.Ey * (cs[n, 0] + 1j * cs[n, 1])
    assert np.abs(ref_line).max() > 0
    assert rel_l2(line, ref_line) < 1e-10
    assert rel_l2(re[:, :, 0] + 1j * im[:, :, 0], ref) < 1e-10
