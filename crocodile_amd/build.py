"""Build libcrocgpu.so (all HIP sources under csrc/) for gfx950 with hipcc.

Usage: ``python -m crocodile_amd.build`` (or ``build()`` from Python).  The
library is written next to this file, in-tree, so it travels with the source
tree to the machine that runs it.
"""

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OBJ = os.path.join(HERE, '_build')
LIB = os.path.join(HERE, 'libcrocgpu.so')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')

ARCH = 'gfx950'
CFLAGS = [
    '--offload-arch=' + ARCH, '-O3', '-fPIC', '-std=c++17',
    '-munsafe-fp-atomics',
    # packed FP32 VALU ops disabled (compiler-generated v_pk_*_f32 forms
    # are avoided on this hardware)
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wall', '-Wno-unused-result',
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith('.hip'))


def _compile(src, extra):
    obj = os.path.join(OBJ, os.path.basename(src) + '.o')
    cmd = [HIPCC] + CFLAGS + extra + ['-c', src, '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('hipcc failed for %s:\n%s\n%s' % (src, ' '.join(cmd), r.stderr))
    return obj, r.stderr


def build(verbose=False, extra=None):
    """Compile every csrc/*.hip and link libcrocgpu.so.  Always recompiles."""
    extra = list(extra or [])
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(lambda s: _compile(s, extra), srcs))
    objs = [o for o, _ in results]
    if verbose:
        for _, err in results:
            if err.strip():
                sys.stderr.write(err)
    tmp = LIB + '.tmp'
    cmd = [HIPCC, '--offload-arch=' + ARCH, '-shared', '-fPIC', '-o', tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('link failed:\n%s\n%s' % (' '.join(cmd), r.stderr))
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose='-v' in sys.argv))
