"""Drop-in driver host logic (no GPU): run kinds, roughness profile modes and
their file side effects (fdtd_main.py:67-79, fdtd_funcs.py:165-192)."""
import os

import numpy as np
import pytest

from oidt_amd import fdtd_main, params

from cli_cases import small_user


def test_profile_modes_follow_the_reference_protocol():
    u = small_user()
    assert fdtd_main.profile_mode(u, 0, 'f', 'i') == 'smooth'
    assert fdtd_main.profile_mode(u, 0, 'f', 'r') == 'gen'
    assert fdtd_main.profile_mode(u, 0, 'b', 'r') == 'load'
    assert fdtd_main.profile_mode(u, 1, 'f', 'r') == 'load'
    assert fdtd_main.profile_mode(dict(u, rough_toggle=False), 0, 'f', 'r') == 'smooth'


def test_generated_profiles_are_saved_and_reloaded(tmp_path):
    u = small_user()
    s, fg_gen = fdtd_main.prepare('tm', u, 0, 'f', 'r', workdir=str(tmp_path))
    files = sorted(os.listdir(tmp_path / 'rough_profiles'))
    assert files == ['profile_lower_s20nm_lc200nm_r0.npy', 'profile_upper_s20nm_lc200nm_r0.npy']
    up = np.load(tmp_path / 'rough_profiles' / files[1])
    assert len(up) == s.p2x - s.p1x and np.std(up) > 0
    _, fg_load = fdtd_main.prepare('tm', u, 0, 'b', 'r', workdir=str(tmp_path))
    assert np.array_equal(fg_gen, fg_load)
    s_i, fg_i = fdtd_main.prepare('tm', u, 0, 'f', 'i', workdir=str(tmp_path))
    assert s_i.nx < s.nx and s_i.nt == s.nt                 # incident runs: short guide, same nt
    rows = np.nonzero(fg_i[s_i.nx // 2])[0]
    assert list(rows) == list(range(s_i.by, s_i.by + s_i.ny_swg))   # smooth


def test_missing_run_arguments_raise(tmp_path, monkeypatch):
    import pickle
    monkeypatch.chdir(tmp_path)
    with open(tmp_path / 'user_dict_0.pkl', 'wb') as fh:
        pickle.dump(small_user(), fh)
    with pytest.raises(Exception, match='No S-parameter sim type selected'):
        fdtd_main.main('tm', ['fdtd_main.py', '0'], workdir=str(tmp_path))


def test_adjacent_profile_numbers_share_no_profile(tmp_path):
    """'gen' mode draws the two sidewalls of realisation k from independent
    child streams of its seed: no profile of realisation k reappears in
    realisation k + 1 (the reference's draws are independent)."""
    got = {}
    for k in (0, 1, 2):
        u = small_user()
        u['profile_number'] = k
        d = tmp_path / ('r%d' % k)
        d.mkdir()
        fdtd_main.prepare('tm', u, 0, 'f', 'r', workdir=str(d))
        files = sorted(os.listdir(d / 'rough_profiles'))
        got[k] = [np.load(d / 'rough_profiles' / f) for f in files]     # lower, upper
    for k in (0, 1):
        for a in got[k]:
            for b in got[k + 1]:
                assert not np.array_equal(a, b)
        lo, up = got[k]
        assert not np.array_equal(lo, up)
    # seeded: the same profile number gives the same profiles again
    u = small_user()
    d = tmp_path / 'again'
    d.mkdir()
    fdtd_main.prepare('tm', u, 0, 'f', 'r', workdir=str(d))
    again = [np.load(d / 'rough_profiles' / f) for f in sorted(os.listdir(d / 'rough_profiles'))]
    assert all(np.array_equal(a, b) for a, b in zip(again, got[0]))
