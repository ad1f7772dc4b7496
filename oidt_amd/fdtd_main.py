"""Drop-in for the reference simulation driver ``fdtd_main.py``.

Same command line, parameter files and outputs as
FDTD/GPU/FDTD_2D_TM/s-parameter_extraction/fdtd_main.py (and the TE / 3D
twins): ``python fdtd_main.py <line> <f|b> <i|r>`` in a directory holding
``user_config.py``; reads ``user_dict_<line>.pkl`` or re-pickles user_config
to ``user_dict.pkl`` (fdtd_auto_setup.py:26-53); generates, saves or loads the
roughness profiles as the reference does (fdtd_main.py:67-79); runs the time
loop on the GPU through liboidt_fdtd.so; writes the port phasors as
``./<output_dir>/{comp}_{tag}_l{line+1}.npy`` (fdtd_main.py:365-383).

A maintainer replaces the reference ``fdtd_main.py`` by a two-line shim::

    from oidt_amd.fdtd_main import main
    main('tm')        # 'te' or '3d' in the other directories

Phasors are complex128 (FP64 accumulators); the reference saved complex64
for FP32 runs, which its extraction script reads either way.
"""
import os
import pickle
import sys
import time

import numpy as np

from . import extract, materials, params, roughness
from .solver import Solver


def _usr(argv, workdir):
    usr = params.load_user_dict(argv[1] if len(argv) > 1 else None, workdir=workdir)
    if usr.get('auto_condition'):
        # fdtd_auto_setup.py:42-53 writes the conditioned values back to the pickle
        usr = dict(usr)
        usr.update(f0=0.5 * (usr['sparam_fmin'] + usr['sparam_fmax']), harmonics=2.0, points_per_wl=20,
                   num_flights=2.0)
        with open(os.path.join(workdir, 'user_dict.pkl'), 'wb') as fh:
            pickle.dump(usr, fh)
    return usr


def profile_mode(usr, line, direction, kind):
    """Roughness mode of one run (fdtd_main.py:67-79)."""
    if usr.get('sim_type', 's-param') != 's-param' or not usr.get('rough_toggle'):
        return 'smooth'
    if kind == 'i':
        return 'smooth'
    if line == 0 and direction == 'f' and kind == 'r':
        return 'gen'
    return 'load'


def prepare(scheme, usr, line, direction, kind, workdir='.'):
    """(Setup, core mask) of one run, with the reference's profile side effects."""
    s = params.auto_setup(usr, scheme, kind)
    mode = profile_mode(usr, line, direction, kind)
    length = (s.p2z - s.p1z) if scheme == '3d' else (s.p2x - s.p1x)
    up_rel, lo_rel = roughness.profile_names(usr.get('prof_dir', 'rough_profiles'), usr['rough_std'],
                                             usr['rough_acl'], usr.get('profile_number', 0))
    up_name = os.path.normpath(os.path.join(workdir, up_rel))
    lo_name = os.path.normpath(os.path.join(workdir, lo_rel))
    if mode == 'smooth':
        up, lo = roughness.generate_profiles('smooth', length, s, usr)
    elif mode == 'gen':
        up, lo = roughness.generate_profiles('gen', length, s, usr, up_name, lo_name,
                                             seed=int(usr.get('profile_seed', usr.get('profile_number', 0))))
    else:
        up, lo = roughness.generate_profiles('load', length, s, usr, up_name + '.npy', lo_name + '.npy')
    return s, materials.fg_mask(s, up, lo)


def run(scheme, usr, line, direction, kind, workdir='.', verbose=True):
    s, fg = prepare(scheme, usr, line, direction, kind, workdir)
    dims = 'x'.join(str(v) for v in s.shape)
    if verbose:
        print('The simulation size will be: {} cells over {} time steps\n'.format(dims, s.nt), flush=True)
    solver = Solver(s, fg, direction, line)
    t0 = time.time()
    pct = float(usr.get('feedback_at_n_percent', 10))
    every = max(1, int(s.nt * pct / 100.0))
    done = 0
    while done < s.nt:
        solver.advance(min(every, s.nt - done))
        solver.sync()
        done = solver.steps()
        if verbose:
            el = time.time() - t0
            print('Step {} of {} done, {:.1f} % complete, MC/sec: {:.1f}'.format(
                done, s.nt, 100.0 * done / s.nt, s.cells * done / max(el, 1e-9) / 1e6), flush=True)
    loop = time.time() - t0
    if verbose:
        print('Speed in MCells/sec: {}'.format(s.cells * s.nt / (1e6 * loop)), flush=True)
    out_dir = os.path.join(workdir, usr.get('output_dir', 'Results'))
    files = extract.save_run(out_dir, scheme, solver.dft(), direction, kind, line)
    solver.close()
    return s, files


def main(scheme, argv=None, workdir='.'):
    argv = list(sys.argv if argv is None else argv)
    usr = _usr(argv, workdir)
    if usr.get('sim_type', 's-param') == 's-param' and len(argv) < 4:
        raise Exception('No S-parameter sim type selected. Type 1, 2, 3, or 4 after the python launch command')
    line, direction, kind = int(argv[1]), argv[2], argv[3]
    params.kind_of(direction, kind)
    os.makedirs(os.path.join(workdir, usr.get('output_dir', 'Results')), exist_ok=True)
    run(scheme, usr, line, direction, kind, workdir)
    print('results saved!', flush=True)


if __name__ == '__main__':
    # python -m oidt_amd.fdtd_main <tm|te|3d> <line> <f|b> <i|r>
    main(sys.argv[1], [sys.argv[0]] + sys.argv[2:])
