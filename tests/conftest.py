"""Shared test setup: markers, import paths, library fixtures."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tensorkraft-design_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")


@pytest.fixture(scope="session")
def ctx():
    import tk_hip

    c = tk_hip.Context(0)
    yield c
    c.close()


def lcg_next(state):
    """LCG of the reference tests (tk-linalg/tests/block_gemm_realistic.rs:12-15):
    s = s*6364136223846793005 + 1 (wrapping); x = ((s >> 33) as i32) / 2^30."""
    state = (state * 6364136223846793005 + 1) & 0xFFFFFFFFFFFFFFFF
    v = state >> 33
    if v >= 1 << 31:
        v -= 1 << 32
    return state, v / float(1 << 30)
