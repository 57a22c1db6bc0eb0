import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "praca-magisterska_amd")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def load_package():
    """Imports the package directory praca-magisterska_amd/ as module `wiremom`."""
    if "wiremom" in sys.modules:
        return sys.modules["wiremom"]
    spec = importlib.util.spec_from_file_location("wiremom", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["wiremom"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def wm():
    return load_package()


@pytest.fixture(scope="session")
def geometries(wm):
    import wiremom.geometries as G
    return G


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def ctx(wm):
    c = wm.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def geo_dir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("geo"))
