"""AddressSanitizer + UndefinedBehaviorSanitizer run of the host build (CPU only).

Builds the host test library (the same kernel bodies and C ABI as libpmhd.so,
compiled with g++ for the host) with ASan and UBSan into
tests/hostbuild/libpmhd_host_asan.so, then runs the host parity suite against it
in a child process with the sanitizer runtime preloaded.  Any invalid access or
undefined behaviour aborts the run.

    python tests/hostbuild/asan.py [pytest -k expression]
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from pythonmhd_amd import build as B  # noqa: E402

LIB = os.path.join(HERE, 'libpmhd_host_asan.so')
OBJ = os.path.join(HERE, '_build_asan')
SAN = ['-fsanitize=address,undefined', '-fno-omit-frame-pointer', '-fno-sanitize-recover=undefined', '-g']


def build():
    os.makedirs(OBJ, exist_ok=True)
    inc = ['-I' + B.INC, '-I' + B.CSRC]
    cmds, objs = [], []
    for src, flags in B.HOST_UNITS:
        o = B._obj(OBJ, src)
        cmds.append(['g++'] + B.GXX_COMMON + ['-O1'] + SAN + flags + inc + ['-c', src, '-o', o])
        objs.append(o)
    B._parallel(cmds)
    B._run(['g++', '-shared', '-fPIC'] + SAN + objs + ['-o', LIB])
    return LIB


def run(kexpr):
    asan = subprocess.run(['g++', '-print-file-name=libasan.so'], capture_output=True, text=True,
                          check=True).stdout.strip()
    ubsan = subprocess.run(['g++', '-print-file-name=libubsan.so'], capture_output=True, text=True,
                           check=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=asan + ' ' + ubsan, PMHD_HOST_LIB=LIB,
               ASAN_OPTIONS='detect_leaks=0:abort_on_error=1', UBSAN_OPTIONS='print_stacktrace=1')
    cmd = [sys.executable, '-m', 'pytest', os.path.join(ROOT, 'tests', 'test_host_parity.py'), '-x', '-q',
           '-p', 'no:cacheprovider']
    if kexpr:
        cmd += ['-k', kexpr]
    return subprocess.run(cmd, env=env, cwd=ROOT).returncode


if __name__ == '__main__':
    build()
    sys.exit(run(sys.argv[1] if len(sys.argv) > 1 else ''))
