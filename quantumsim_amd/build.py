"""Compile libqsv.so (HIP kernels + C ABI) for gfx950, in-tree.

    python -m quantumsim_amd.build

Each ``csrc/*.hip`` file is compiled to an object in parallel, then linked
into ``quantumsim_amd/lib/libqsv.so``.
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
OUT_DIR = os.path.join(HERE, 'lib')
OBJ_DIR = os.path.join(HERE, 'lib', 'obj')
LIB = os.path.join(OUT_DIR, 'libqsv.so')

HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
FLAGS = [
    '--offload-arch=gfx950', '-O3', '-fPIC', '-std=c++17',
    # packed FP32 instructions are disabled on purpose (see DESIGN notes)
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-I' + os.path.join(ROOT, 'include'),
]


def _compile(src):
    obj = os.path.join(OBJ_DIR, os.path.basename(src) + '.o')
    cmd = [HIPCC] + FLAGS + ['-c', src, '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('hipcc failed on {}:\n{}'.format(src, r.stderr[-4000:]))
    return obj


def build(verbose=True):
    os.makedirs(OBJ_DIR, exist_ok=True)
    sources = sorted(glob.glob(os.path.join(CSRC, '*.hip')))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(sources))) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB + '.tmp'
    cmd = [HIPCC, '--offload-arch=gfx950', '-shared', '-fPIC', '-o', tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('link failed:\n' + r.stderr[-4000:])
    os.replace(tmp, LIB)
    if verbose:
        print('built', LIB, 'from', ', '.join(os.path.basename(s) for s in sources))
    return LIB


if __name__ == '__main__':
    build()
    sys.exit(0)
