"""End-to-end drop-in: the reference protocol (4 runs per line + extraction)
driven from a user_config-style dictionary writes a Touchstone file that reads
back, with passive, reciprocal-looking S-parameters, sequentially and batched."""
import os

import numpy as np
import pytest

from oidt_amd import simulation_launcher, touchstone

from cli_cases import small_user


@pytest.mark.gpu
@pytest.mark.parametrize('scheme', ['tm', 'te', '3d'])
def test_protocol_writes_touchstone(tmp_path, scheme):
    u = small_user(scheme)
    seq = tmp_path / 'seq'
    bat = tmp_path / 'bat'
    seq.mkdir()
    bat.mkdir()
    p1 = simulation_launcher.main(scheme, workdir=str(seq), usr=dict(u))
    p2 = simulation_launcher.main(scheme, workdir=str(bat), batched=True, usr=dict(u))
    assert os.path.basename(p1) == 'small.s2p'
    f, S = touchstone.read_touchstone(p1)
    f2, S2 = touchstone.read_touchstone(p2)
    assert len(f) == 5
    print(scheme, 'S21', np.abs(S[:, 1, 0]), 'S11', np.abs(S[:, 0, 0]), flush=True)
    assert np.all(np.abs(S[:, 1, 0]) < 1.05) and np.all(np.abs(S[:, 1, 0]) > 0.2)
    assert np.all(np.abs(S[:, 0, 0]) < 0.5)
    np.testing.assert_allclose(S[:, 0, 1], S[:, 1, 0])
    # the batched protocol gives the same phasors as the sequential one
    np.testing.assert_array_equal(S, S2)
    names = sorted(os.listdir(seq / 'Results'))
    assert 'NEXT_STEPS_INSTRUCTIONS.txt' in names
    # fi and bi runs save port 1 / port 2, the two rough runs both ports:
    # 6 (port, tag) pairs x 2 components (2D) or 4 components (3D)
    assert len([n for n in names if n.endswith('.npy')]) == (12 if scheme != '3d' else 24)
