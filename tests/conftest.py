import os
import sys

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
PKG = os.path.join(ROOT, "imau-ice_amd")
for p in (PKG, os.path.dirname(__file__)):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")
    # torch's HIP runtime must initialise before libimauvel.so's: a process
    # in which the library touched the GPU first cannot initialise torch's
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
