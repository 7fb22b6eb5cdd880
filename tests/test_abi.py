"""The C ABI: the built libraries load and export every entry point include/pmhd.h
declares; argument validation of pmhd_create (host build, no GPU needed)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'pmhd.h')
GPU_LIB = os.path.join(ROOT, 'pythonmhd_amd', 'libpmhd.so')
HOST_LIB = os.path.join(ROOT, 'tests', 'hostbuild', 'libpmhd_host.so')


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r'^\s*(?:int|const char\*)\s+(pmhd_\w+)\s*\(', txt, re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert 'pmhd_create' in names and 'pmhd_step' in names and len(names) >= 10


@pytest.mark.parametrize('path', [GPU_LIB, HOST_LIB])
def test_library_exports_every_declared_symbol(path):
    if not os.path.exists(path):
        from pythonmhd_amd import build
        build.build_gpu() if path == GPU_LIB else build.build_host()
    lib = ctypes.CDLL(path)
    for name in declared():
        assert hasattr(lib, name), name


def test_python_binding_lists_all_exports():
    from pythonmhd_amd import _lib
    assert sorted(_lib.EXPORTS) == declared()


def test_product_load_has_no_fallback(tmp_path):
    from pythonmhd_amd import _lib
    with pytest.raises(_lib.PmhdError):
        _lib.load(str(tmp_path / 'missing.so'))


@pytest.fixture(scope='module')
def host():
    if not os.path.exists(HOST_LIB):
        from pythonmhd_amd import build
        build.build_host()
    from pythonmhd_amd import _lib
    return _lib.load(HOST_LIB)


@pytest.mark.parametrize('kw,msg', [
    (dict(order=3), 'reconstruct_order'),
    (dict(precision=16), 'precision'),
])
def test_create_rejects_bad_params(host, kw, msg):
    from pythonmhd_amd._lib import PmhdError, Solver
    with pytest.raises(PmhdError, match=msg):
        Solver((8, 8), (0.1, 0.1), (1, 1), 5 / 3, 0.8, lib=host, **kw)


def test_create_rejects_bad_grid(host):
    from pythonmhd_amd._lib import PmhdError, Solver
    with pytest.raises(PmhdError, match='boundary'):
        Solver((8, 8), (0.1, 0.1), (1, 2), 5 / 3, 0.8, lib=host)
    with pytest.raises(PmhdError, match='widths'):
        Solver((8, 8), (0.1, -0.1), (1, 1), 5 / 3, 0.8, lib=host)


def test_set_state_checks_shapes(host):
    import numpy as np
    from pythonmhd_amd._lib import Solver
    s = Solver((4, 5), (0.2, 0.25), (1, 1), 5 / 3, 0.8, lib=host)
    with pytest.raises(ValueError):
        s.set_state(np.ones((8, 5, 4)), np.ones((8, 5, 4)), (np.ones((4, 6)), np.ones((5, 5))))
    with pytest.raises(ValueError):
        s.set_state(np.ones((8, 4, 5)), np.ones((8, 4, 5)), (np.ones((4, 5)), np.ones((5, 5))))
