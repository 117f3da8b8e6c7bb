import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def cm():
    """The HIP library binding; built here only if nothing is built yet."""
    import cumodp_amd
    if not os.path.exists(cumodp_amd.LIB_PATH):
        from cumodp_amd import build
        build.build()
    cumodp_amd.lib()
    return cumodp_amd


@pytest.fixture(scope="session")
def modpn():
    """modpn's CPU multiply compiled from the reference's C sources
    (oracle/_ref); skipped where the reference tree was absent at build time."""
    from oracle import modpn_ref
    if not modpn_ref.available():
        pytest.skip("oracle/_ref/libmodpn_fmul.so not built (reference tree absent)")
    modpn_ref.lib()
    return modpn_ref
