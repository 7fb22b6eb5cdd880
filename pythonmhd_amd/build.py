"""Build libpmhd.so (HIP, gfx950) in-tree, and the host test build of the same
kernel bodies used by tests/ (tests/hostbuild/libpmhd_host.so).

The library's translation units are compiled in parallel and linked into one
shared object: the FP64 step (IEEE arithmetic in the reference's operation order,
no contraction), the FP32-bulk step (contracted multiply-adds, hardware FP32
division and square root; its double-precision islands stay IEEE), the same FP32
step with non-temporal result stores for large grids, the adjoint (IEEE), the
allocation-free CFL entry point and the C ABI (host code).

    python -m pythonmhd_amd.build            # both
    python -m pythonmhd_amd.build --gpu-only [--only f32,abi]
"""
import concurrent.futures
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, 'pythonmhd_amd', 'csrc')
INC = os.path.join(ROOT, 'include')
GPU_LIB = os.path.join(ROOT, 'pythonmhd_amd', 'libpmhd.so')
GPU_OBJ = os.path.join(ROOT, 'pythonmhd_amd', '_build')
HOST_DIR = os.path.join(ROOT, 'tests', 'hostbuild')
HOST_LIB = os.path.join(HOST_DIR, 'libpmhd_host.so')
HOST_DUAL_LIB = os.path.join(HOST_DIR, 'libpmhd_host_dualc.so')
HOST_OBJ = os.path.join(HOST_DIR, '_build')
ARCH = os.environ.get('PYTORCH_ROCM_ARCH', 'gfx950')

HIPCC_COMMON = [
    '--offload-arch=' + ARCH, '-O3', '-std=c++17', '-fPIC',
    # no packed FP32 instructions
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-result', '-Wno-unused-value',
]
# keep the reference's operation order: no contraction into FMA, IEEE division/sqrt
IEEE = ['-ffp-contract=off']
FAST32 = ['-ffp-contract=fast', '-fno-hip-fp32-correctly-rounded-divide-sqrt', '-fno-signed-zeros',
          '-fgpu-flush-denormals-to-zero']

GPU_UNITS = [('pmhd_hip_f64.hip', IEEE), ('pmhd_hip_f32.hip', FAST32), ('pmhd_hip_f32nt.hip', FAST32),
             ('pmhd_hip_adj.hip', IEEE), ('pmhd_noalloc.hip', IEEE)]
HOST_UNITS = [(os.path.join(HOST_DIR, 'pmhd_host_f64.cpp'), []), (os.path.join(HOST_DIR, 'pmhd_host_f32.cpp'), []),
              (os.path.join(HOST_DIR, 'pmhd_host_adj.cpp'), []), (os.path.join(HOST_DIR, 'pmhd_host_noalloc.cpp'), []),
              (os.path.join(CSRC, 'pmhd_abi.cpp'), [])]
GXX_COMMON = ['-O2', '-std=c++17', '-fPIC', '-ffp-contract=off', '-Wno-unknown-pragmas']


def _run(cmd):
    print('+', ' '.join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _parallel(cmds):
    with concurrent.futures.ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        for f in [ex.submit(_run, c) for c in cmds]:
            f.result()


def _obj(dirname, src):
    return os.path.join(dirname, os.path.splitext(os.path.basename(src))[0] + '.o')


def build_gpu(only=None):
    """Compile every unit (or only the units whose file names contain one of the
    strings in `only`, reusing the other objects) and link libpmhd.so."""
    hipcc = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
    os.makedirs(GPU_OBJ, exist_ok=True)
    inc = ['-I' + INC, '-I' + CSRC]
    cmds, objs = [], []
    for src, flags in GPU_UNITS:
        o = _obj(GPU_OBJ, src)
        cmds.append([hipcc] + HIPCC_COMMON + flags + inc + ['-c', os.path.join(CSRC, src), '-o', o])
        objs.append(o)
    abi = _obj(GPU_OBJ, 'pmhd_abi.cpp')
    cmds.append(['g++'] + GXX_COMMON + inc + ['-c', os.path.join(CSRC, 'pmhd_abi.cpp'), '-o', abi])
    objs.append(abi)
    if only:
        cmds = [c for c, o in zip(cmds, objs) if any(u in os.path.basename(o) for u in only) or not os.path.exists(o)]
    _parallel(cmds)
    tmp = GPU_LIB + '.tmp'
    _run([hipcc, '--offload-arch=' + ARCH, '-shared', '-fPIC', '--hip-link'] + objs + ['-o', tmp])
    os.replace(tmp, GPU_LIB)
    return GPU_LIB


def build_host():
    """The host test library, and beside it the same library with every adjoint
    phase through the per-slot dual-number passes (libpmhd_host_dualc.so, the
    reference the hand-written reverses of the corrector and the update are tested
    against)."""
    os.makedirs(HOST_OBJ, exist_ok=True)
    inc = ['-I' + INC, '-I' + CSRC]
    cmds, objs = [], []
    for src, flags in HOST_UNITS:
        o = _obj(HOST_OBJ, src)
        cmds.append(['g++'] + GXX_COMMON + ['-fopenmp'] + flags + inc + ['-c', src, '-o', o])
        objs.append(o)
    adj = os.path.join(HOST_DIR, 'pmhd_host_adj.cpp')
    dual = os.path.join(HOST_OBJ, 'pmhd_host_adj_dualc.o')
    cmds.append(['g++'] + GXX_COMMON + ['-fopenmp', '-DPMHD_ADJ_DUAL_CORRECTOR'] + inc + ['-c', adj, '-o', dual])
    _parallel(cmds)
    for lib, lobjs in ((HOST_LIB, objs), (HOST_DUAL_LIB, [dual if o == _obj(HOST_OBJ, adj) else o for o in objs])):
        tmp = lib + '.tmp'
        _run(['g++', '-shared', '-fPIC', '-fopenmp'] + lobjs + ['-o', tmp])
        os.replace(tmp, lib)
    return HOST_LIB


def main(argv):
    only = argv[argv.index('--only') + 1].split(',') if '--only' in argv else None
    if '--gpu-only' in argv:
        build_gpu(only)
    else:
        with concurrent.futures.ThreadPoolExecutor(max_workers=2) as ex:
            for f in [ex.submit(build_gpu), ex.submit(build_host)]:
                f.result()


if __name__ == '__main__':
    main(sys.argv[1:])
