import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")

# CH4 as the headline workload states it: C 21874.66, 4 x H 1837.15 electron
# masses, every charge +1.
CH4_MASSES = [21874.66] + [1837.15] * 4
CH4_CHARGES = [1.0] * 5


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import ffi
    return ffi.Oracle()


@pytest.fixture(scope="session")
def ce():
    from __graft_entry__ import load_package
    return load_package()


@pytest.fixture(scope="session")
def gpu(ce):
    if ce.device_count() < 1:
        pytest.fail("GPU test run without a visible HIP device")
    return ce


def per_atom_rel(p, ref):
    """max over atoms of |p_i - ref_i| / |ref_i| (per sim), p/ref [..., n, 3]."""
    p = np.asarray(p).reshape(ref.shape)
    return (np.linalg.norm(p - ref, axis=-1) / np.linalg.norm(ref, axis=-1)).max(axis=-1)


def config_rel(p, ref):
    """Reference's config-norm metric |dP|/|P| per sim (bench_batched_integrator.cpp:89-99)."""
    p = np.asarray(p).reshape(ref.shape)
    d = (p - ref).reshape(ref.shape[0], -1)
    return np.linalg.norm(d, axis=1) / np.linalg.norm(ref.reshape(ref.shape[0], -1), axis=1)
