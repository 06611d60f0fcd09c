"""Test configuration: import paths for the package and the oracle, the
``gpu`` marker, and shared helpers."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "tealeaf-kokkos_amd")
for p in (PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

DECKS = os.path.join(ROOT, "tests", "golden", "decks")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def read_deck(name: str) -> str:
    with open(os.path.join(DECKS, name)) as f:
        return f.read()


@pytest.fixture(scope="session")
def hip():
    import tealeaf_hip
    return tealeaf_hip


@pytest.fixture(scope="session")
def oracle():
    import oracle as o
    return o
