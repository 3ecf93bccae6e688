import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
OPS = os.path.join(ROOT, "ddfmm_amd", "_ops")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oplib():
    """(ldpath, hdpath) of the e^{ikr}/r operator library (ACCU=2, NPQ=4)."""
    from ddfmm_amd import mlib_gen
    return mlib_gen.ensure_library(OPS, ktype=0, accu=2, NPQ=4, hf_max_W=4, lf_min_W=2.0 ** -3)


@pytest.fixture(scope="session")
def oplib_expr():
    from ddfmm_amd import mlib_gen
    return mlib_gen.ensure_library(OPS, ktype=1, accu=2, NPQ=4, hf_max_W=4, lf_min_W=2.0 ** -3)


@pytest.fixture(scope="session")
def ops(oplib):
    from ddfmm_amd import serialize
    return serialize.read_ld(oplib[0]), serialize.read_hd(oplib[1])


@pytest.fixture(scope="session")
def ops_expr(oplib_expr):
    from ddfmm_amd import serialize
    return serialize.read_ld(oplib_expr[0]), serialize.read_hd(oplib_expr[1])
