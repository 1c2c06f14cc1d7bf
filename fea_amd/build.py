"""Build the fea_amd HIP library (libfea_amd.so) for gfx950, in-tree.

Usage:  python -m fea_amd.build      (or fea_amd.build.build())

Every source under fea_amd/csrc is compiled with hipcc; the shared library is
written to fea_amd/_lib/libfea_amd.so, where fea_amd._native loads it.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OUT_DIR = os.path.join(HERE, '_lib')
LIB = os.path.join(OUT_DIR, 'libfea_amd.so')
ARCH = os.environ.get('FEA_AMD_ARCH', 'gfx950')

HIP_SOURCES = ['elements.hip', 'surfaces.hip', 'linalg.hip', 'pcg.hip', 'contact.hip', 'adjoint.hip', 'batched.hip', 'couple.hip', 'barrier.hip']
CPP_SOURCES = ['pattern.cpp', 'api.cpp']

# packed-FP32 instructions are disabled: the compiler may form them from scalar
# float code, and they are avoided on this target.
HIP_FLAGS = ['--offload-arch=' + ARCH, '-O3', '-std=c++17', '-fPIC',
             '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
             '-munsafe-fp-atomics']
CPP_FLAGS = ['-O3', '-std=c++17', '-fPIC']


def _hipcc():
    for cand in (shutil.which('hipcc'), '/opt/rocm/bin/hipcc'):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError('hipcc not found: the fea_amd HIP library cannot be built')


def _compile(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('compile failed: %s\n%s%s' % (' '.join(cmd), r.stdout, r.stderr))
    return cmd[-1]


def build(verbose=False):
    hipcc = _hipcc()
    obj_dir = os.path.join(OUT_DIR, 'obj')
    os.makedirs(obj_dir, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(obj_dir, src + '.o')
        jobs.append([hipcc] + HIP_FLAGS + ['-c', path, '-o', obj])
    for src in CPP_SOURCES:
        obj = os.path.join(obj_dir, src + '.o')
        jobs.append(['g++'] + CPP_FLAGS + ['-c', os.path.join(CSRC, src), '-o', obj])
    workers = min(len(jobs), max(1, min(8, os.cpu_count() or 1)))
    with concurrent.futures.ThreadPoolExecutor(workers) as ex:
        objs = list(ex.map(_compile, jobs))
    tmp = LIB + '.tmp'
    link = [hipcc, '--offload-arch=' + ARCH, '-shared', '-fPIC', '-o', tmp] + objs
    _compile(link)
    os.replace(tmp, LIB)
    if verbose:
        print('built', LIB)
    return LIB


if __name__ == '__main__':
    build(verbose=True)
    sys.exit(0)
