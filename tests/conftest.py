import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT, ROOT / "tt-toolbox_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libttgpu.so")


@pytest.fixture(scope="session")
def gpu_ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test collected but no GPU is visible")
    import ttgpu
    return ttgpu.default_context(0)


_session = {}


def pytest_sessionfinish(session, exitstatus):
    _session["status"] = int(exitstatus)


@pytest.hookimpl(trylast=True)
def pytest_unconfigure(config):
    """End a session that used the GPU without the HIP runtime's exit-time
    teardown.  Processes that ran the batched rounder (host worker threads,
    each with its own HIP stream) have crashed intermittently inside that
    teardown after every test had passed and all output was written; the
    session's status is kept and the OS releases the device resources."""
    torch = sys.modules.get("torch")
    if "status" in _session and torch is not None and torch.cuda.is_initialized():
        sys.stdout.flush()
        sys.stderr.flush()
        os._exit(_session["status"])
