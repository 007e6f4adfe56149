"""Shared fixtures.  The package directory is `vlasov.jl_amd/`, which is not a
valid module name, so it is loaded under the name `vlasov_amd`."""
from __future__ import annotations

import importlib.util
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def load_package():
    if "vlasov_amd" in sys.modules:
        return sys.modules["vlasov_amd"]
    pkg = ROOT / "vlasov.jl_amd"
    spec = importlib.util.spec_from_file_location("vlasov_amd", pkg / "__init__.py",
                                                  submodule_search_locations=[str(pkg)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["vlasov_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def va():
    return load_package()


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def config1(va):
    """Config 1 geometry: nfine = 28 (58^3), rcluster 78, rbox 235."""
    p = va.xenon_parameters(100_000, 28)
    fine, coarse = va.simulation.build_meshes(p)
    sm = va.grids.GaussianSmoothing.of(fine.axis)
    return p, fine, coarse, sm
