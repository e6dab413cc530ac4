import os
import sys

import numpy as np
import pytest

# Hardware queues the HIP runtime spreads streams over, read when it starts:
# 16, so tests that run simulations concurrently overlap their kernels on the
# device as the serving loop does.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD Instinct MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def ref_vectors():
    return np.load(os.path.join(GOLDEN, "ref_vectors.npz"))


@pytest.fixture(scope="session")
def known_answers():
    import json
    with open(os.path.join(GOLDEN, "known_answers.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def orc():
    from oracle.orc import Oracle
    return Oracle("orc")


def block_row_rel_err(A, B):
    """max over 3x3 block rows of max|A-B| / max|B| on that block row."""
    n = B.shape[0]
    worst = 0.0
    for r in range(0, n, 3):
        den = np.abs(B[r:r + 3]).max()
        num = np.abs(A[r:r + 3] - B[r:r + 3]).max()
        worst = max(worst, num / den if den > 0 else num)
    return worst
