import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libnsamd.so")
    config.addinivalue_line("markers", "slow: long-running parity case")


def _oracle_libs_present():
    return (ROOT / "oracle" / "libns_oracle.so").exists()


@pytest.fixture(scope="session", autouse=True)
def _build_oracle():
    """The oracle is plain C: build it on the fly if it is missing."""
    if not _oracle_libs_present():
        from oracle import oracle as orc
        orc.build(with_ref=False)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def rel_max(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    s = np.abs(b).max()
    return np.abs(a - b).max() / (s if s > 0 else 1.0)


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False
