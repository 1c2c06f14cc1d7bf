"""Build the fea_amd HIP library (libfea_amd.so) for gfx950, in-tree.

Usage:  python -m fea_amd.build      (or fea_amd.build.build())

Every HIP / C++ source under fea_amd/csrc is compiled (hipcc, g++) into
fea_amd/_lib/libfea_amd.so, where fea_amd._native loads it; torch_ops.cpp is then
compiled with g++ against torch's headers into fea_amd/_lib/libfea_torch.so, the
TORCH_LIBRARY(fea, ...) operators over that C ABI.
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
TORCH_OPS_LIB = os.path.join(OUT_DIR, 'libfea_torch.so')
ARCH = os.environ.get('FEA_AMD_ARCH', 'gfx950')

HIP_SOURCES = ['elements.hip', 'surfaces.hip', 'linalg.hip', 'pcg.hip', 'contact.hip', 'adjoint.hip', 'batched.hip', 'couple.hip', 'barrier.hip']
CPP_SOURCES = ['pattern.cpp', 'api.cpp', 'amg_setup.cpp']

# packed-FP32 instructions are disabled: the compiler may form them from scalar
# float code, and they are avoided on this target.
HIP_FLAGS = ['--offload-arch=' + ARCH, '-O3', '-std=c++17', '-fPIC',
             '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
             '-munsafe-fp-atomics']
CPP_FLAGS = ['-O3', '-std=c++17', '-fPIC', '-pthread']


def _hipcc():
    for cand in (shutil.which('hipcc'), '/opt/rocm/bin/hipcc'):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError('hipcc not found: the fea_amd HIP library cannot be built')


def _compile(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('compile failed: %s\n%s%s' % (' '.join(cmd), r.stdout, r.stderr))
    return cmd[cmd.index('-o') + 1] if '-o' in cmd else cmd[-1]


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
    link = [hipcc, '--offload-arch=' + ARCH, '-shared', '-fPIC', '-pthread', '-Wl,-soname,libfea_amd.so', '-o', tmp] + objs
    _compile(link)
    os.replace(tmp, LIB)
    if verbose:
        print('built', LIB)
    build_torch_ops(verbose)
    return LIB


def build_torch_ops(verbose=False):
    """The PyTorch-ROCm operator library (TORCH_LIBRARY(fea, ...)) over libfea_amd.so:
    host C++ only, compiled with g++ against torch's headers and linked to the C ABI."""
    import torch
    import torch.utils.cpp_extension as ce
    rocm = os.environ.get('ROCM_PATH', '/opt/rocm')
    flags = ['-O2', '-std=c++17', '-fPIC', '-shared', '-D__HIP_PLATFORM_AMD__=1', '-DUSE_ROCM=1',
             '-D_GLIBCXX_USE_CXX11_ABI=%d' % int(torch._C._GLIBCXX_USE_CXX11_ABI), '-w']
    inc = ['-I' + p for p in ce.include_paths()] + ['-I' + os.path.join(rocm, 'include')]
    libs = ['-L' + p for p in ce.library_paths()] + ['-lc10', '-lc10_hip', '-ltorch', '-ltorch_cpu',
                                                     '-L' + OUT_DIR, '-lfea_amd', "-Wl,-rpath,$ORIGIN"]
    tmp = TORCH_OPS_LIB + '.tmp'
    _compile(['g++'] + flags + inc + [os.path.join(CSRC, 'torch_ops.cpp')] + libs + ['-o', tmp])
    os.replace(tmp, TORCH_OPS_LIB)
    if verbose:
        print('built', TORCH_OPS_LIB)
    return TORCH_OPS_LIB


if __name__ == '__main__':
    build(verbose=True)
    sys.exit(0)
