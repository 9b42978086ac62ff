"""Shared fixtures. GPU tests are marked `gpu` and need a visible MI355X."""
from __future__ import annotations

import json
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (run with -m gpu on the MI355X box)")


def _gpu_count() -> int:
    try:
        from gary_amd import _lib
        return _lib.device_count()
    except Exception:
        return 0


def pytest_collection_modifyitems(config, items):
    if _gpu_count() > 0:
        return
    skip = pytest.mark.skip(reason="no GPU visible")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def golden_graphs():
    return json.loads((GOLDEN / "graphs.json").read_text())


@pytest.fixture(autouse=True)
def _announce(request):
    # name the case on the real stderr (bypassing capture) before any GPU work, so a run that
    # ends the process still says which case was running
    if "gpu" in request.keywords:
        sys.__stderr__.write(f"\n[case] {request.node.nodeid}\n")
        sys.__stderr__.flush()
    yield
