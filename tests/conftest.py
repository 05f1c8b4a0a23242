import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")


class Golden:
    """Access to a golden .npz as name/key groups."""

    def __init__(self, path):
        self.z = np.load(path, allow_pickle=False)
        self.keys = list(self.z.keys())

    def cases(self):
        return sorted({k.split("/")[0] for k in self.keys})

    def get(self, case, key):
        return self.z[f"{case}/{key}"]

    def has(self, case, key):
        return f"{case}/{key}" in self.z

    def tensor(self, case, prefix="ind", vals="vals"):
        d = self.get(case, "ndims")
        inds = [self.get(case, f"{prefix}{m}") for m in range(len(d))]
        return d, inds, self.get(case, vals)


@pytest.fixture(scope="session")
def golden_coo():
    return Golden(os.path.join(GOLDEN, "coo_mttkrp.npz"))


@pytest.fixture(scope="session")
def golden_hicoo():
    return Golden(os.path.join(GOLDEN, "hicoo.npz"))


@pytest.fixture(scope="session")
def golden_omp_reduce():
    return Golden(os.path.join(GOLDEN, "coo_omp_reduce.npz"))


@pytest.fixture(scope="session")
def golden_cpd():
    return Golden(os.path.join(GOLDEN, "cpd.npz"))


@pytest.fixture(scope="session")
def lib():
    from sptol_amd.parti import load_library
    return load_library()


@pytest.fixture(scope="session")
def gpu_lib(lib):
    import ctypes
    n = ctypes.c_int(0)
    if lib.sptolHipDeviceCount(ctypes.byref(n)) != 0 or n.value < 1:
        pytest.skip("no GPU visible to HIP")
    return lib


def rel_fro(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))
