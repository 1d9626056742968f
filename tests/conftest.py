import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X")


def _ensure_oracle():
    so = os.path.join(ROOT, "oracle", "_build", "libzj_oracle.so")
    if not os.path.exists(so):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "-s", "oracle"], check=True)


_ensure_oracle()


@pytest.fixture(scope="session")
def golden_ops():
    return dict(np.load(os.path.join(GOLDEN, "operators.npz")))


@pytest.fixture(scope="session")
def golden_meshes():
    return dict(np.load(os.path.join(GOLDEN, "meshes.npz")))


@pytest.fixture(scope="session")
def golden_physics():
    return dict(np.load(os.path.join(GOLDEN, "physics.npz")))


def mesh_from_golden(g, name):
    pre = name + "__"
    return {k[len(pre):]: v for k, v in g.items() if k.startswith(pre)}
