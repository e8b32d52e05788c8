"""Shared test setup: markers, import paths and the config-1 workload."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "multi-vector-retrieval-gpu_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

# torch ships its own HIP runtime under the soname libamdhip64.so.7.  Load it
# before any test opens a HIP library of this tree (libmvr, the MFMA probe), so
# the process has one runtime and torch still sees the device.
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover
    pass


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")


@pytest.fixture(scope="session")
def c1():
    """Config-1 corpus, index and queries built by the CPU oracle."""
    from oracle import synth
    return synth.config1()


@pytest.fixture(scope="session")
def small():
    """A small workload (1,024 docs, 128 centroids) for fast checks."""
    from oracle import synth
    return synth.config1(n_doc=1024, n_comp=256, n_c=128, n_q=16, n_iter=4)
