"""S-parameter extraction from the port running-DFT phasors.

Restates FDTD/GPU/FDTD_2D_TM/s-parameter_extraction/extract_sparams.py:48-104
(TE :48-104 with the TE Poynting form, 3D FDTD_3D/.../extract_sparams.py:48-131)
in FP64: S = sqrt(Pref / Pinc) with Pinc / Pref the port-window sums of the
time-averaged Poynting vector of the incident and reflected/transmitted runs.

Phasor files follow the reference names ``{comp}_{tag}_l{line}.npy`` with tag
fi (forward incident, port 1), bi (backward incident, port 2), ft/br (forward
reflected run, ports 1/2) and fr/bt (backward reflected run, ports 1/2)
(fdtd_main.py:365-383).
"""
import os

import numpy as np

COMPS = {'tm': ('ez', 'hy'), 'te': ('ey', 'hz'), '3d': ('ex', 'ey', 'hx', 'hy')}


def run_tags(direction, kind):
    """Tags saved by one run: {port index 1|2: tag} (fdtd_main.py:365-383)."""
    if kind == 'i':
        return {1: 'fi'} if direction == 'f' else {2: 'bi'}
    return {1: 'ft', 2: 'br'} if direction == 'f' else {1: 'fr', 2: 'bt'}


def poynting(scheme, f):
    """Time-averaged normal Poynting density of phasor dict ``f``."""
    if scheme == 'tm':
        return -0.5 * (f['ez'] * np.conjugate(f['hy']))
    if scheme == 'te':
        return 0.5 * (f['ey'] * np.conjugate(f['hz']))
    return 0.5 * (f['ex'] * np.conjugate(f['hy']) - f['ey'] * np.conjugate(f['hx']))


def window_start(scheme, s, line):
    """First row (2D: y, 3D: x) of the port window of 1-based ``line``."""
    if scheme == '3d':
        return s.bx + s.cx_swg - s.ph + s.pp * (line - 1)
    return s.by + s.cy_swg - s.ph + s.pp * (line - 1)


def port_power(scheme, s, sav, line):
    start = window_start(scheme, s, line)
    if scheme == '3d':
        N = s.num_cpml
        return sav[:, start:start + s.pp, N:s.ny - N].sum(axis=(1, 2))
    return sav[:, start:start + s.pp].sum(axis=1)


def s_pairs(scheme, num_lines):
    """(i, j, iline, jline, inc_tag, ref_tag, cor_tag or None) of every S entry
    the reference computes, in its order (extract_sparams.py:52-104; 3D
    FDTD_3D extract_sparams.py:52-131)."""
    P = 2 * num_lines
    out = []
    for i in range(P):
        for j in (range(P) if scheme == '3d' else range(i, P)):
            inc_tag = 'fi' if j % 2 == 0 else 'bi'
            if i % 2 == j % 2:
                ref_tag = 'ft' if i % 2 == 0 else 'bt'
            else:
                ref_tag = 'fr' if i % 2 == 0 else 'br'
            cor_tag = ('fi' if i % 2 == 0 else 'bi') if i == j else None
            out.append((i, j, i // 2 + 1, j // 2 + 1, inc_tag, ref_tag, cor_tag))
    return out


def s_matrix(scheme, s, load, num_lines, nf):
    """S[nf, 2L, 2L] (complex128) from ``load(comp, tag, line)`` -> phasor array.

    2D fills the upper triangle and mirrors it (extract_sparams.py:52-104); 3D
    visits every (i, j) so the lower-triangle values are the ones left
    (FDTD_3D extract_sparams.py:52-131)."""
    P = 2 * num_lines
    S = np.zeros((nf, P, P), dtype=np.complex128)
    comps = COMPS[scheme]
    for i, j, iline, jline, inc_tag, ref_tag, cor_tag in s_pairs(scheme, num_lines):
        inc = {c: np.asarray(load(c, inc_tag, jline), dtype=np.complex128) for c in comps}
        ref = {c: np.array(load(c, ref_tag, jline), dtype=np.complex128) for c in comps}
        if cor_tag is not None:
            for c in comps:
                ref[c] -= np.asarray(load(c, cor_tag, iline), dtype=np.complex128)
        pinc = port_power(scheme, s, poynting(scheme, inc), jline)
        pref = port_power(scheme, s, poynting(scheme, ref), iline)
        S[:, i, j] = S[:, j, i] = np.sqrt(pref / pinc)
    return S


# run (direction, kind) and port index (0: port 1, 1: port 2) of each phasor tag
TAG_RUN = {'fi': (('f', 'i'), 0), 'bi': (('b', 'i'), 1), 'ft': (('f', 'r'), 0), 'br': (('f', 'r'), 1),
           'fr': (('b', 'r'), 0), 'bt': (('b', 'r'), 1)}


def s_matrix_device(scheme, s, batch, slots, num_lines, nf):
    """S[nf, 2L, 2L] as ``s_matrix``, with the port-window power sums done on
    the device from the FP64 accumulators (Batch.port_power); ``slots`` maps
    (line index 0.., direction, kind) to the batch slot of that run.  Only the
    per-frequency powers leave the device."""
    P = 2 * num_lines
    S = np.zeros((nf, P, P), dtype=np.complex128)
    N = s.num_cpml

    def where(tag, line):
        (d, k), port = TAG_RUN[tag]
        return slots[(line - 1, d, k)], port

    for i, j, iline, jline, inc_tag, ref_tag, cor_tag in s_pairs(scheme, num_lines):
        si, pi = where(inc_tag, jline)
        sr, pr = where(ref_tag, jline)
        sub = where(cor_tag, iline) if cor_tag is not None else None
        pinc = batch.port_power(si, pi, window_start(scheme, s, jline) - N, s.pp)
        pref = batch.port_power(sr, pr, window_start(scheme, s, iline) - N, s.pp, sub=sub)
        S[:, i, j] = S[:, j, i] = np.sqrt(pref / pinc)
    return S


def npy_loader(output_dir):
    def load(comp, tag, line):
        return np.load(os.path.join(output_dir, '{}_{}_l{}.npy'.format(comp, tag, line)))
    return load


def save_run(output_dir, scheme, dft, direction, kind, line):
    """Write the phasor files of one run (fdtd_main.py:365-383).

    ``dft``: {(component key, port): complex array}; component keys 'e'/'h' (2D)
    or 'ex'.. (3D) as returned by Solver.dft()."""
    os.makedirs(output_dir, exist_ok=True)
    names = {'tm': {'e': 'ez', 'h': 'hy'}, 'te': {'e': 'ey', 'h': 'hz'},
             '3d': {'ex': 'ex', 'ey': 'ey', 'hx': 'hx', 'hy': 'hy'}}[scheme]
    written = []
    for port, tag in run_tags(direction, kind).items():
        for key, comp in names.items():
            path = os.path.join(output_dir, '{}_{}_l{}'.format(comp, tag, line + 1))
            np.save(path, dft[(key, port)])
            written.append(path + '.npy')
    return written
