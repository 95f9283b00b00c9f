import importlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PKG_NAME = "hpc-reading-course_amd"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a gfx950 GPU (MI355X)")
    config.addinivalue_line("markers", "slow: longer CPU test")


def load_pkg():
    return importlib.import_module(PKG_NAME)


def load_problems():
    return importlib.import_module(PKG_NAME + ".problems")


@pytest.fixture(scope="session")
def problems():
    return load_problems()


@pytest.fixture(scope="session")
def orc():
    from oracle import pyoracle
    pyoracle.lib()
    pyoracle.lib(fast=True)
    return pyoracle
