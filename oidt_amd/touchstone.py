"""Touchstone v1 output in the layout the reference produces.

The reference writes its ``.s{2L}p`` with scikit-rf 0.18.1
``Network.write_touchstone(form='ri', write_z0=True, skrf_comment=True)`` and
then prepends an acknowledgement comment line
(FDTD/GPU/FDTD_2D_TM/s-parameter_extraction/extract_sparams.py:106-117).
scikit-rf is not a dependency here; this module writes the same text layout
(observed in Results/example.s4p and FDTD_3D/.../Results/example.s2p) so that
``skrf.Network`` in the downstream SROPEE tool reads the file unchanged:

* option line ``# Hz S RI R 50.0 ``;
* one row of the S matrix per text line (2-port: S11 S21 S12 S22 on one line,
  the Touchstone v1 2-port order);
* a ``! Port Impedance`` comment after every frequency block.
"""
import numpy as np

ACK = ('!This S-parameter file was generated as part of the OIDT. This project was completed as part of '
       'research conducted with my major professor and advisor, Prof. Ata Zadehgol, at the Applied and '
       'Computational Electromagnetics Signal and Power Integrity (ACEM-SPI) Lab while working toward the '
       'Ph.D. in Electrical Engineering at the University of Idaho, Moscow, Idaho, USA. This project was '
       'funded, in part, by the National Science Foundation (NSF); award #1816542 [1].\n')


def _num(x):
    return repr(float(x))


def format_touchstone(freqs, S, z0=50.0, acknowledgement=True):
    S = np.asarray(S)
    nf, P, _ = S.shape
    out = [ACK] if acknowledgement else []
    out.append('!Created with skrf (http://scikit-rf.org).\n')
    out.append('# Hz S RI R {} \n'.format(float(z0)))
    if P == 2:
        out.append('!freq ReS11 ImS11 ReS21 ImS21 ReS12 ImS12 ReS22 ImS22\n')
    else:
        for r in range(P):
            names = ['ReS{0}{1} ImS{0}{1}'.format(r + 1, c + 1) for c in range(P)]
            for g in range(0, P, 3):
                prefix = '!freq ' if (r == 0 and g == 0) else '! '
                out.append(prefix + ' '.join(names[g:g + 3]) + '\n')
        out.append('!\n')
    zline = '! Port Impedance' + ''.join(' %.14f %.14f' % (z0, 0.0) for _ in range(P)) + '\n'
    for f in range(nf):
        if P == 2:
            vals = [S[f, 0, 0], S[f, 1, 0], S[f, 0, 1], S[f, 1, 1]]
            out.append(_num(freqs[f]) + ''.join(' %s %s' % (_num(v.real), _num(v.imag)) for v in vals) + '\n')
        else:
            for r in range(P):
                row = ''.join(' %s %s' % (_num(v.real), _num(v.imag)) for v in S[f, r])
                out.append((_num(freqs[f]) if r == 0 else '') + row + '\n')
        out.append(zline)
    return ''.join(out)


def write_touchstone(path, freqs, S, z0=50.0, acknowledgement=True):
    with open(path, 'w') as fh:
        fh.write(format_touchstone(freqs, S, z0, acknowledgement))
    return path


def read_touchstone(path_or_text, nports=None):
    """Minimal Touchstone v1 (RI, Hz) reader: (freqs, S[nf, P, P]).

    Follows the v1 rules: '!' starts a comment, '#' the option line, data are a
    stream of numbers with 1 + 2 P^2 per frequency (2-port in column order)."""
    text = path_or_text
    if '\n' not in path_or_text:
        with open(path_or_text) as fh:
            text = fh.read()
        if nports is None and path_or_text.lower().endswith('p') and '.s' in path_or_text.lower():
            nports = int(path_or_text.lower().rsplit('.s', 1)[1][:-1])
    nums = []
    for line in text.splitlines():
        line = line.split('!', 1)[0].strip()
        if not line or line.startswith('#'):
            continue
        nums.extend(float(t) for t in line.split())
    if nports is None:
        raise ValueError('port count unknown')
    P = nports
    per = 1 + 2 * P * P
    a = np.asarray(nums).reshape(-1, per)
    freqs = a[:, 0]
    z = a[:, 1::2] + 1j * a[:, 2::2]
    S = z.reshape(-1, P, P)
    if P == 2:
        S = S.transpose(0, 2, 1)
    return freqs, S
