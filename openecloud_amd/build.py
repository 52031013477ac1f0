"""Builds liboecl.so (all HIP sources of csrc/) for gfx950 with hipcc.

Flags:
  -ffp-contract=off           no fused multiply-add, so FP64 arithmetic rounds exactly like
                              the reference's scalar C code
  -packed-fp32-ops disabled   no packed FP32 instructions are emitted
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'liboecl.so')
SOURCES = ['particles.hip', 'boundary.hip', 'emitter.hip', 'poisson.hip', 'fastpoisson.hip',
           'sort.hip', 'manage.hip', 'step.hip', 'unet.hip']
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
FLAGS = ['-O3', '-std=c++17', '--offload-arch=gfx950', '-fPIC', '-shared',
         '-ffp-contract=off', '-fvisibility=hidden',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-Wno-unused-result']


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + '.tmp'
    cmd = [HIPCC] + FLAGS + ['-o', tmp] + srcs + ['-L/opt/rocm/lib', '-lrocblas', '-lrocsolver',
                                                   '-Wl,-rpath,/opt/rocm/lib']
    if verbose:
        print(' '.join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    build(verbose='-v' in sys.argv)
