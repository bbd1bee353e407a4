"""Build the gfx950 shared library powder6_amd/_lib/libpowder6_pt.so with hipcc.

Flags: -O3, gfx950 only, packed-FP32 instructions disabled (-Xclang -target-feature
-Xclang -packed-fp32-ops), FP32 division and square root by faster sequences that are not
correctly rounded (-fno-hip-fp32-correctly-rounded-divide-sqrt; the profile constants they
feed are FP32 anyway), FP contraction left on for the device (FMA)
and off for the host plan compiler's reflection enumeration (bit-for-bit with the reference
ordering).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIBDIR = os.path.join(HERE, '_lib')
LIB = os.path.join(LIBDIR, 'libpowder6_pt.so')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
ARCH = '--offload-arch=gfx950'
DEVICE_FLAGS = ['-O3', '-std=c++17', '-fPIC', ARCH, '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
                '-munsafe-fp-atomics', '-fno-hip-fp32-correctly-rounded-divide-sqrt', '-Wall', '-Wno-unused-function']


def run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError('build failed: ' + ' '.join(cmd))
    return r.stdout


def build(verbose=False):
    os.makedirs(LIBDIR, exist_ok=True)
    objs = []
    for src in ('pt_kernels.hip', 'pt_mlp.hip', 'pt_jacobian.hip'):
        obj = os.path.join(LIBDIR, src.replace('.hip', '.o'))
        run([HIPCC] + DEVICE_FLAGS + ['-c', os.path.join(CSRC, src), '-o', obj])
        objs.append(obj)
    for src in ('pt_host.cpp', 'pt_mlp_host.cpp'):
        obj = os.path.join(LIBDIR, src.replace('.cpp', '.o'))
        run([HIPCC, '-O2', '-std=c++17', '-fPIC', '-ffp-contract=off', ARCH, '-Wall', '-x', 'hip',
             '-c', os.path.join(CSRC, src), '-o', obj])
        objs.append(obj)
    tmp = LIB + '.tmp'
    run([HIPCC, '-shared', '-fPIC', ARCH, '-o', tmp] + objs)
    os.replace(tmp, LIB)
    if verbose:
        print('built', LIB)
    build_addon(verbose)
    return LIB


NODE_INCLUDE = os.environ.get('NODE_INCLUDE', '/usr/include/node')
ADDON = os.path.join(LIBDIR, 'pt_addon.node')


def build_addon(verbose=False):
    """The Node N-API addon powder6_amd/_lib/pt_addon.node (links libpowder6_pt.so, found
    next to it through an $ORIGIN rpath).  Skipped when the Node headers are absent."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, 'node_api.h')):
        if verbose:
            print('node headers not found; addon not built')
        return None
    tmp = ADDON + '.tmp'
    run(['gcc', '-O2', '-fPIC', '-shared', '-std=c11', '-Wall', '-DNODE_GYP_MODULE_NAME=pt_addon', '-I' + NODE_INCLUDE,
         os.path.join(HERE, 'node', 'pt_addon.c'), '-L' + LIBDIR, '-lpowder6_pt', '-Wl,-rpath,$ORIGIN', '-o', tmp])
    os.replace(tmp, ADDON)
    if verbose:
        print('built', ADDON)
    return ADDON


if __name__ == '__main__':
    build(verbose=True)
