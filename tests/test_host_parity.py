"""The parity suite on the CPU through the host build of the same kernel bodies
(tests/hostbuild/libpmhd_host.so) -- checks the kernels' arithmetic without a GPU."""
import os

import pytest

from tests.parity_suite import *  # noqa: F401,F403

# PMHD_HOST_LIB: another host build of the same sources (the sanitizer build of
# tests/hostbuild/asan.py)
HOST_LIB = os.environ.get('PMHD_HOST_LIB', os.path.join(os.path.dirname(__file__), 'hostbuild', 'libpmhd_host.so'))


@pytest.fixture(scope='module')
def lib():
    if not os.path.exists(HOST_LIB):
        from pythonmhd_amd import build
        build.build_host()
    from pythonmhd_amd._lib import load
    return load(HOST_LIB)
