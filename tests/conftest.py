import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__ as graft  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def clm():
    return graft.load_package()


@pytest.fixture(scope="session")
def O():
    return graft.load_oracle()


def read_pdb(path):
    xs = []
    with open(path) as f:
        for line in f:
            if line.startswith(("ATOM", "HETATM")):
                xs.append([float(line[30:38]), float(line[38:46]), float(line[46:54])])
    return np.array(xs)


@pytest.fixture(scope="session")
def argon():
    return read_pdb(os.path.join(ROOT, "tests", "golden", "argon_cubic.pdb"))


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
