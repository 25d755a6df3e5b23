import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "simple-dftd3_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP library)")


def load_golden(name):
    with open(os.path.join(GOLDEN, name)) as fh:
        return json.load(fh)


@pytest.fixture(scope="session")
def goldens():
    return {f["name"]: f for f in load_golden("reference_goldens.json")}


@pytest.fixture(scope="session")
def pairwise_goldens():
    return {f["name"]: f for f in load_golden("pairwise_goldens.json")}


@pytest.fixture(scope="session")
def family_goldens():
    return {f["name"]: f for f in load_golden("family_goldens.json")}


@pytest.fixture(scope="session")
def python_model():
    return load_golden("python_interface_model.json")
