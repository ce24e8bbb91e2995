import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
# the product's Python binding (i-emic_amd/python/iemic_amd)
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "i-emic_amd", "python"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "ref: needs the reference build in oracle/_ref")


def pytest_collection_modifyitems(config, items):
    from thcm_cases import ref_available
    if ref_available():
        return
    skip = pytest.mark.skip(reason="oracle/_ref not built (no /root/reference here)")
    for it in items:
        if "ref" in it.keywords:
            it.add_marker(skip)
