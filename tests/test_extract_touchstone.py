"""Extraction and Touchstone output."""
import os

import numpy as np
import pytest

from oidt_amd import configs, extract, params, touchstone
from oracle import extract as oextract
from oracle import setup as osetup

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize('name,ports', [('reference_example_excerpt.s4p', 4), ('reference_example_excerpt.s2p', 2)])
def test_touchstone_layout_matches_reference_files(name, ports):
    """Reading the reference's own output and writing it back is byte exact."""
    text = open(os.path.join(HERE, 'golden', name)).read()
    f, S = touchstone.read_touchstone(text, ports)
    assert touchstone.format_touchstone(f, S) == text


def test_touchstone_2port_column_order():
    S = np.zeros((1, 2, 2), complex)
    S[0, 1, 0] = 0.5 + 0.25j      # S21
    txt = touchstone.format_touchstone([1e14], S, acknowledgement=False)
    data = [l for l in txt.splitlines() if l and not l.startswith(('!', '#'))][0].split()
    assert float(data[3]) == 0.5 and float(data[4]) == 0.25      # freq, S11(2), S21(2)
    f, S2 = touchstone.read_touchstone(txt, 2)
    assert np.array_equal(S2, S)


def _synthetic(scheme, s, nf, rng, refl, trans):
    """Phasor files where every reflected/transmitted field is a known multiple
    of the incident one: S_ii = sqrt(|refl|^2 ...) closed form."""
    comps = extract.COMPS[scheme]
    shape = (nf, s.nx, s.ny) if scheme == '3d' else (nf, s.ny)
    inc = {c: rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for c in comps}
    files = {}
    for tag in ('fi', 'bi'):
        files.update({(c, tag, 1): inc[c] for c in comps})
    for tag, a in (('ft', 1 + refl), ('bt', 1 + refl), ('fr', trans), ('br', trans)):
        files.update({(c, tag, 1): a * inc[c] for c in comps})
    return files


@pytest.mark.parametrize('scheme', ['tm', 'te', '3d'])
def test_extraction_closed_form(scheme):
    """ref = (1 + r) inc on the same port and t inc on the other port gives
    S11 = sqrt(r^2) and S21 = sqrt(t^2) exactly (power ratio of scaled phasors)."""
    p = configs.small(scheme)
    s = params.explicit_setup(p)
    rng = np.random.default_rng(0)
    r, t = 0.1 - 0.05j, 0.8 + 0.3j
    files = _synthetic(scheme, s, 3, rng, r, t)
    S = extract.s_matrix(scheme, s, lambda c, tag, l: files[(c, tag, l)], 1, 3)
    np.testing.assert_allclose(S[:, 0, 0], np.sqrt(r * np.conj(r)), rtol=1e-12)
    np.testing.assert_allclose(S[:, 1, 0], np.sqrt(t * np.conj(t)), rtol=1e-12)
    np.testing.assert_allclose(S[:, 0, 1], S[:, 1, 0])
    # restated extraction agrees
    c = osetup.explicit_setup(p)
    comps = extract.COMPS[scheme]
    ofiles = {(tag, l): tuple(files[(cc, tag, l)] for cc in comps) for (cc, tag, l) in files}
    np.testing.assert_allclose(oextract.s_matrix(scheme, c, ofiles, 1, 3), S, rtol=1e-12)


def test_2d_upper_triangle_3d_lower_triangle_wins():
    """Asymmetric inputs: 2D keeps S[i,j] computed with i <= j, 3D the i >= j ones."""
    rng = np.random.default_rng(1)
    for scheme in ('tm', '3d'):
        p = configs.small(scheme)
        s = params.explicit_setup(p)
        files = _synthetic(scheme, s, 2, rng, 0.1, 0.7)
        for c in extract.COMPS[scheme]:
            files[(c, 'br', 1)] = 0.3 * files[(c, 'fi', 1)]     # S21 from fr, S12 from br differ
        S = extract.s_matrix(scheme, s, lambda c, tag, l: files[(c, tag, l)], 1, 2)
        want = 0.7 if scheme == 'tm' else 0.3
        np.testing.assert_allclose(np.abs(S[:, 0, 1]), want, rtol=1e-12)
        np.testing.assert_allclose(S[:, 0, 1], S[:, 1, 0])


def test_save_run_names(tmp_path):
    dft = {(k, p): np.full((2, 5), p + 0j) for k in ('e', 'h') for p in (1, 2)}
    files = extract.save_run(str(tmp_path), 'tm', dft, 'f', 'r', 0)
    names = sorted(os.path.basename(f) for f in files)
    assert names == ['ez_br_l1.npy', 'ez_ft_l1.npy', 'hy_br_l1.npy', 'hy_ft_l1.npy']
    assert np.load(tmp_path / 'ez_br_l1.npy')[0, 0] == 2
    files = extract.save_run(str(tmp_path), 'te', dft, 'b', 'i', 1)
    assert sorted(os.path.basename(f) for f in files) == ['ey_bi_l2.npy', 'hz_bi_l2.npy']
