"""The parity suite on the MI355X through libpmhd.so (HIP)."""
import pytest

from tests.parity_suite import *  # noqa: F401,F403

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def lib():
    from pythonmhd_amd._lib import load
    return load()
