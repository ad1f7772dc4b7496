"""On-device port power and S normalisation equal the host FP64 extraction
from the same phasors (relative 1e-12: only the summation order differs)."""
import numpy as np
import pytest

from oidt_amd import extract, params, simulation_launcher

from cli_cases import small_user


@pytest.mark.gpu
@pytest.mark.parametrize('scheme,lines', [('tm', 1), ('te', 1), ('3d', 1), ('tm', 2)])
def test_device_s_matrix_matches_host(tmp_path, scheme, lines):
    u = small_user(scheme, num_lines=lines)
    b, slots = simulation_launcher.run_batched(scheme, dict(u), workdir=str(tmp_path), keep=True)
    try:
        s = params.auto_setup(u, scheme, 'r')
        nf, L = u['sparam_num_freqs'], u['num_lines']
        host = extract.s_matrix(scheme, s, extract.npy_loader(str(tmp_path / 'Results')), L, nf)
        dev = extract.s_matrix_device(scheme, s, b, slots, L, nf)
        print(scheme, 'max |S_dev - S_host|', np.abs(dev - host).max(), flush=True)
        np.testing.assert_allclose(dev, host, rtol=1e-12, atol=1e-14)
        # the subtraction path (S_ii) and a plain incident power are exercised
        assert np.all(np.abs(dev[:, 0, 0]) < 0.5) and np.all(np.abs(dev[:, 1, 0]) > 0.2)
        # a window outside the DFT window is rejected, not read
        with pytest.raises(Exception):
            b.port_power(slots[(0, 'f', 'i')], 0, 0, 10 ** 6)
    finally:
        b.close()
