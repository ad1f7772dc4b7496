"""Build the oracle's C restatement (test infrastructure): ``python -m oracle.build``.

gcc, FP64, OpenMP, no fused multiply-add (-ffp-contract=off) so every
operation rounds as the NumPy oracle's does.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, 'fdtd3d.c')
LIB = os.path.join(HERE, 'liboracle_fdtd3d.so')


def command(out=LIB):
    return ['gcc', '-O2', '-std=c11', '-fopenmp', '-ffp-contract=off', '-fno-fast-math', '-fPIC', '-shared',
            '-Wall', SRC, '-o', out, '-lm']


def build(verbose=False):
    tmp = LIB + '.tmp'
    cmd = command(tmp)
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError('gcc failed for the oracle C restatement')
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose=True))
