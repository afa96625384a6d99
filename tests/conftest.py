"""Shared pytest setup: import paths, the ``gpu`` marker, artifact fixtures."""

from __future__ import annotations

import sys
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parents[1]
for p in (REPO / "grg-spmv_amd", REPO):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def _gpu_available() -> bool:
    try:
        import torch

        return bool(torch.cuda.is_available())
    except Exception:
        return False


HAS_GPU = _gpu_available()


def pytest_collection_modifyitems(config, items):
    if HAS_GPU:
        return
    skip = pytest.mark.skip(reason="no ROCm GPU visible")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def artifact_dir(tmp_path_factory):
    return tmp_path_factory.mktemp("artifacts")


@pytest.fixture
def fake_hip_device(monkeypatch):
    """Planner tests on the CPU: one visible device, every stream on device 0."""
    from pygrgl_spmv.backends import _hip_stream

    class _FakeLib:
        def device_count(self):
            return 1

        def stream_device(self, ptr):
            return 0

    monkeypatch.setattr(_hip_stream, "_lib", lambda: _FakeLib())
    return _FakeLib()
