import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def golden():
    """Inputs and outputs recorded from the compiled reference."""
    path = os.path.join(ROOT, "tests", "golden", "reference_small.npz")
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def golden_exact():
    """Rerank / filtering searches recorded from the compiled reference."""
    path = os.path.join(ROOT, "tests", "golden", "reference_exact.npz")
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def golden_repair():
    """Isolated-node repairs recorded from the compiled reference."""
    path = os.path.join(ROOT, "tests", "golden", "reference_repair.npz")
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def golden_mips():
    """Inner-product (--use_mips) outputs recorded from the compiled reference."""
    path = os.path.join(ROOT, "tests", "golden", "reference_mips.npz")
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def gpu():
    """Initialised nbpg_amd on device 0 (GPU tests only)."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import nbpg_amd as N
    N.init(0)
    return N


@pytest.fixture(scope="session")
def golden_branches():
    """Angle prune, cosine / top_degree builds and limit / degree_limit / wide-L
    searches recorded from the compiled reference."""
    path = os.path.join(ROOT, "tests", "golden", "reference_branches.npz")
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}
