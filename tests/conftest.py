import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "scale-track_amd"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)

_LIB = os.path.join(ROOT, "scale-track_amd", "scaletrack", "libscaletrack.so")
_ORACLE = os.path.join(ROOT, "oracle", "libst_oracle.so")
if not (os.path.exists(_LIB) and os.path.exists(_ORACLE)):
    import __graft_entry__

    __graft_entry__.build()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libscaletrack.so)")


def _has_gpu():
    try:
        import scaletrack

        return scaletrack.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def ctx():
    import scaletrack

    if not _has_gpu():
        pytest.fail("GPU test run without a visible GPU")
    c = scaletrack.Context(0)
    yield c
    c.close()
