import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libemc.so")


def pytest_collection_modifyitems(config, items):
    # print what is about to run on the GPU so a failing box names its case
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(pytest.mark.timeout(300))
