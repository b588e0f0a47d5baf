"""Build recipe for the engine's shared library (gfx950, in-tree).

``python -m pyrdsim_amd.build`` compiles ``csrc/rds_engine.hip`` into
``pyrdsim_amd/librds_hip.so`` with hipcc.  No GPU is needed to build.
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, 'csrc', 'rds_engine.hip')
LIB = os.path.join(HERE, 'librds_hip.so')
INCLUDE = os.path.join(ROOT, 'include')

HIPCC_FLAGS = [
    '--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-shared',
    # packed FP32 instructions are kept out of the generated code
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-value', '-Wno-unused-result',
]


def hipcc():
    for cand in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', 'hipcc'):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError('hipcc not found')


def build(verbose=False):
    """Compile the engine library; returns its path."""
    tmp = LIB + '.tmp'
    cmd = [hipcc()] + HIPCC_FLAGS + ['-I' + INCLUDE, SRC, '-o', tmp]
    if verbose:
        print(' '.join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError('hipcc failed:\n' + res.stdout + res.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose='-v' in sys.argv))
