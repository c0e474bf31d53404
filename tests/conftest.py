import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import __graft_entry__  # noqa: E402

__graft_entry__.load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def srt():
    import snapshot_raytracer_amd as m
    return m


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o
    if not os.path.exists(o.LIB_PATH):
        o.build()
    return o


@pytest.fixture(scope="session")
def fields():
    from snapshot_raytracer_amd import fields as f
    return f


@pytest.fixture(scope="session")
def progress(request):
    """progress(msg): a line on the terminal now, even while pytest captures output (long GPU steps)."""
    rep = request.config.pluginmanager.get_plugin("terminalreporter")

    def emit(msg: str) -> None:
        if rep is not None:
            rep.write_line(msg)
            rep.flush()
        else:
            print(msg, flush=True)
    return emit
