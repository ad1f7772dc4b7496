"""Build liboidt_fdtd.so for gfx950 with hipcc (in-tree, no JIT cache).

``python -m oidt_amd.build`` or ``oidt_amd.build.build()``.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = [os.path.join(HERE, 'csrc', 'fdtd_api.hip')]
DEPS = [os.path.join(HERE, 'csrc', f) for f in ('fdtd_kernels.h', 'fdtd_slot.h')] + \
       [os.path.join(ROOT, 'include', 'oidt_fdtd.h')]
LIB = os.path.join(HERE, 'liboidt_fdtd.so')
ARCH = os.environ.get('OIDT_OFFLOAD_ARCH', 'gfx950')


def hipcc():
    for c in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc'):
        if c and os.path.exists(c):
            return c
    return 'hipcc'


def command(out=LIB):
    return [hipcc(), '--offload-arch=' + ARCH, '-O3', '-std=c++17', '-fPIC', '-shared',
            # packed-FP32 instructions are kept out of the generated code
            '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
            '-Wall', '-Wno-unused-function', '-Wno-unused-variable',
            '-I' + os.path.join(ROOT, 'include'), *SRC, '-o', out]


def build(verbose=False):
    """Compile the library (always; a build is cheap and must reflect the sources)."""
    tmp = LIB + '.tmp'
    cmd = command(tmp)
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError('hipcc failed for liboidt_fdtd.so')
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose=True))
