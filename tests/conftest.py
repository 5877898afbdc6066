import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


def load_package():
    """Import mwa-vcsbeam_amd/ (hyphenated directory) as `vcsbeam_amd`."""
    if "vcsbeam_amd" in sys.modules:
        return sys.modules["vcsbeam_amd"]
    d = os.path.join(ROOT, "mwa-vcsbeam_amd")
    spec = importlib.util.spec_from_file_location("vcsbeam_amd", os.path.join(d, "__init__.py"),
                                                  submodule_search_locations=[d])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["vcsbeam_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def pkg():
    return load_package()


@pytest.fixture(scope="session")
def orc():
    import oracle_lib
    return oracle_lib


@pytest.fixture(scope="session")
def gpu_lib(pkg):
    lib = pkg.load()
    if lib.vmDeviceCount() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    return lib
