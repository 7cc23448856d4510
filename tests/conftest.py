import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    # the CPU oracle is test infrastructure: build it if this tree has no copy yet
    so = os.path.join(ROOT, "oracle", "liblh_oracle.so")
    if not os.path.exists(so):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True, capture_output=True)


@pytest.fixture(scope="session")
def root():
    return ROOT


def config_file(name):
    return os.path.join(ROOT, "liquidhairinteration_amd", "configs", name)
