"""Build libpmhd.so (HIP, gfx950) in-tree, and the host test build of the same
kernel bodies used by tests/ (tests/hostbuild/libpmhd_host.so).

    python -m pythonmhd_amd.build            # both
    python -m pythonmhd_amd.build --gpu-only
"""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, 'pythonmhd_amd', 'csrc')
INC = os.path.join(ROOT, 'include')
GPU_LIB = os.path.join(ROOT, 'pythonmhd_amd', 'libpmhd.so')
HOST_SRC = os.path.join(ROOT, 'tests', 'hostbuild', 'pmhd_host.cpp')
HOST_LIB = os.path.join(ROOT, 'tests', 'hostbuild', 'libpmhd_host.so')
ARCH = os.environ.get('PYTORCH_ROCM_ARCH', 'gfx950')

HIPCC_FLAGS = [
    '--offload-arch=' + ARCH, '-O3', '-std=c++17', '-fPIC', '-shared',
    # keep the reference's operation order: no contraction into FMA
    '-ffp-contract=off',
    # no packed FP32 instructions
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-result', '-Wno-unused-value',
]


def _run(cmd):
    print('+', ' '.join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_gpu():
    hipcc = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
    tmp = GPU_LIB + '.tmp'
    _run([hipcc] + HIPCC_FLAGS + ['-I' + INC, '-I' + CSRC, os.path.join(CSRC, 'pmhd_hip.hip'), '-o', tmp])
    os.replace(tmp, GPU_LIB)
    return GPU_LIB


def build_host():
    tmp = HOST_LIB + '.tmp'
    _run(['g++', '-O2', '-std=c++17', '-fPIC', '-shared', '-ffp-contract=off', '-fopenmp',
          '-Wno-unknown-pragmas', '-I' + INC, '-I' + CSRC, HOST_SRC, '-o', tmp])
    os.replace(tmp, HOST_LIB)
    return HOST_LIB


def main(argv):
    build_gpu()
    if '--gpu-only' not in argv:
        build_host()


if __name__ == '__main__':
    main(sys.argv[1:])
