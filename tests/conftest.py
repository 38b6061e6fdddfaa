"""Shared fixtures: import paths, the oracle loaders, and the `gpu` marker.

Tests marked ``gpu`` need a gfx950 device and the built backend library; everything else runs on
the CPU (oracle, fixtures, host logic, and that the C-ABI library loads and exports its symbols).
"""
from __future__ import annotations

import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT / "optical-pathtracer_amd", ROOT / "oracle", ROOT / "tests"):
    if os.fspath(p) not in sys.path:
        sys.path.insert(0, os.fspath(p))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built hip backend")


@pytest.fixture(scope="session")
def oracle():
    from pyoracle import Oracle

    return Oracle()


@pytest.fixture(scope="session")
def reference():
    from pyoracle import Reference, reference_available

    if not reference_available():
        pytest.skip("oracle/_ref (the reference engine's own sources) is not built here")
    return Reference()


@pytest.fixture(scope="session")
def cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("a gpu-marked test ran without a visible GPU")
    return torch.device("cuda:0")
