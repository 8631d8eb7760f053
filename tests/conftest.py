"""Shared test setup: the `gpu` marker, loaders for the library host module
(its package directory name is not a Python identifier) and the oracle."""
import importlib.util
import pathlib
import sys

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
PKG = ROOT / "gpu-randomized-pseudoskeleton_amd"
sys.path.insert(0, str(ROOT / "oracle"))


def load_farcur():
    name = "farcur"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, PKG / "farcur.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")


@pytest.fixture(scope="session")
def farcur():
    return load_farcur()


@pytest.fixture(scope="session")
def oracle():
    import oracle as orc  # noqa: E402

    orc.lib()
    return orc


@pytest.fixture(scope="session")
def ctx(farcur):
    c = farcur.Context(0)
    yield c
    c.close()
