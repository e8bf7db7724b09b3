import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT, ROOT / "argus-sdr_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libargus_hip.so)")


@pytest.fixture(scope="session")
def gpu():
    """Fails loudly (no fallback) if the HIP library or the GPU is missing."""
    import argus
    from argus import gpudemod
    argus.lib()
    assert gpudemod.Available(), "no HIP device visible"
    return gpudemod
