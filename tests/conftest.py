"""Shared test setup.

Markers: ``gpu`` tests need an MI355X (they call the HIP path through the
C ABI); everything else runs on the CPU.  The library and the oracle are
built in-tree when missing (hipcc cross-compiles gfx950 without a GPU).
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (HIP path through the C ABI)")


def _ensure_built():
    import mscresearch_amd as M
    import oracle as O

    if not os.path.exists(M.LIB_PATH) or not os.path.exists(M.CLI_PATH):
        from mscresearch_amd import build

        build.build()
    if not os.path.exists(O.LIB_PATH):
        O.build()


_ensure_built()


@pytest.fixture(scope="session")
def golden():
    import json

    with open(os.path.join(ROOT, "tests", "golden", "maxflow_golden.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def has_gpu():
    import mscresearch_amd as M

    return M.device_count() > 0
