import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libmqgpu.so")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def fixtures():
    with open(os.path.join(GOLDEN, "fixtures.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def shapes():
    with open(os.path.join(GOLDEN, "shapes_and_wigner.json")) as f:
        return json.load(f)


def ham_from_fixture(fx):
    from qcmaquis_amd.sweep import Hamiltonian
    return Hamiltonian(fx["L"], fx["nelec"], fx["spin"], [(tuple(k), v) for k, v in fx["integrals"]])


@pytest.fixture(scope="session")
def oracle_engine():
    from pyoracle import OracleEngine
    return OracleEngine()


@pytest.fixture(scope="session")
def gpu_engine():
    from qcmaquis_amd.gpu import GpuEngine
    eng = GpuEngine(0)
    yield eng
    eng.close()
