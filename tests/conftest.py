import glob
import importlib
import os
import sys

import numpy
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def package():
    """The product package (its directory name is not a Python identifier)."""
    return importlib.import_module("pyscf-forge_amd")


def golden_cases():
    return sorted(os.path.basename(p)[:-4]
                  for p in glob.glob(os.path.join(GOLDEN, "*.npz")))


def load_golden(name):
    with numpy.load(os.path.join(GOLDEN, name + ".npz")) as f:
        d = {k: f[k] for k in f.files}
    d["norb"] = int(d["norb"])
    d["nelec"] = tuple(int(x) for x in d["nelec"])
    return d


def has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def pkg():
    return package()
