"""Drop-in for the reference ``extract_sparams.py``.

Reads the run phasor files of every line, forms S = sqrt(Pref/Pinc) in FP64
(FDTD/GPU/FDTD_2D_TM/s-parameter_extraction/extract_sparams.py:48-104; 3D
FDTD_3D/.../extract_sparams.py:48-131) and writes
``./<output_dir>/<sparam_file>.s{2L}p`` in the Touchstone layout of the
reference (extract_sparams.py:106-117).  Shim for a reference directory::

    from oidt_amd.extract_sparams import main
    main('tm')
"""
import os
import sys

from . import extract, params, touchstone


def extract_to_file(scheme, usr, workdir='.'):
    s = params.auto_setup(usr, scheme, 'r')
    out_dir = os.path.join(workdir, usr.get('output_dir', 'Results'))
    S = extract.s_matrix(scheme, s, extract.npy_loader(out_dir), usr['num_lines'], usr['sparam_num_freqs'])
    path = os.path.join(out_dir, '{}.s{}p'.format(usr['sparam_file'], 2 * usr['num_lines']))
    touchstone.write_touchstone(path, s.sparam_freqs, S)
    return path, S


def main(scheme, argv=None, workdir='.'):
    argv = list(sys.argv if argv is None else argv)
    usr = params.load_user_dict(argv[1] if len(argv) > 1 else None, workdir=workdir)
    path, _ = extract_to_file(scheme, usr, workdir)
    print('Touchstone file written: {}'.format(path), flush=True)
    return path


if __name__ == '__main__':
    main(sys.argv[1], [sys.argv[0]] + sys.argv[2:])
