import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _built() -> bool:
    libs = [os.path.join(ROOT, "nexusbvh_amd", "lib", "libnxb.so"),
            os.path.join(ROOT, "nexusbvh_amd", "lib", "libnxb_scenes.so"),
            os.path.join(ROOT, "oracle", "lib", "libnxb_oracle.so")]
    return all(os.path.exists(p) for p in libs)


def pytest_sessionstart(session):
    # Build only when nothing is built (build() always rebuilds; a machine
    # that runs the GPU tests uses the libraries built beforehand).
    if not _built():
        from nexusbvh_amd import build as b
        b.build_all(verbose=False)


@pytest.fixture(scope="session")
def gpu_available():
    from nexusbvh_amd import native
    return native.device_count() > 0
