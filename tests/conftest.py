import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def _print_case(item):
    # name the case before any GPU work so a fault can be attributed
    if item.get_closest_marker("gpu"):
        print(f"\n[gpu-case] {item.nodeid}", flush=True)


@pytest.hookimpl(tryfirst=True)
def pytest_runtest_setup(item):
    _print_case(item)


@pytest.fixture(autouse=True)
def _restore_process_defaults(request):
    """A GPU test that autotunes or loads a strategy table changes process-wide
    defaults (the default schedule, the learned table); every test after it
    starts from the built-in ones again."""
    yield
    if request.node.get_closest_marker("gpu"):
        from custingeralg_amd import learned
        learned.set_default_schedule(None)
        learned.unload()
