import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "oracle"))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _ensure_built():
    import blamm_amd
    need = [blamm_amd.HOST_LIB, blamm_amd.BLAMM_BIN, blamm_amd.HIP_LIB]
    if not all(os.path.exists(p) for p in need):
        subprocess.run(["make", "-s", "-j8", "-C", os.path.join(REPO, "blamm_amd", "csrc")], check=True)
    if not os.path.exists(os.path.join(REPO, "oracle", "_build", "liboracle.so")) or \
            not os.path.exists(os.path.join(REPO, "oracle", "_build", "blamm_oracle")):
        subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle")], check=True)


@pytest.fixture(scope="session", autouse=True)
def built():
    _ensure_built()


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle.default()
