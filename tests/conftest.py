import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libmcx.so)")


@pytest.fixture(scope="session")
def workdir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("mcx"))


@pytest.fixture(scope="session")
def tiny_model(workdir):
    from mcmax_amd import configs
    return configs.load_config("tiny", workdir)


@pytest.fixture(scope="session")
def small_model(workdir):
    from mcmax_amd import configs
    return configs.load_config("small", workdir)


@pytest.fixture(scope="session")
def oracle_lib():
    import pyoracle
    pyoracle.build()
    return pyoracle


@pytest.fixture(scope="session")
def gpu():
    from mcmax_amd import transport
    tr = transport.Transport(0)
    yield tr
    tr.close()
