import contextlib
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def ctx():
    """Session-wide GPU context; GPU tests fail (not skip) if it cannot be made."""
    import mesh_amd

    c = mesh_amd.Context(0)
    yield c
    c.close()


def pytest_runtest_setup(item):
    if item.get_closest_marker("gpu"):
        print(f"\n[gpu] {item.nodeid}", flush=True)


# library defaults of the tuning options (mesh_set_option)
OPTION_DEFAULTS = {"force_variant": -1, "xcd_remap": 3, "rb_tile": 4, "fuse_md": 1, "compact_momenta": 1,
                   "u1_general": 0, "float_sums": 0}


@contextlib.contextmanager
def options(ctx, **kw):
    """set tuning options for a block and restore the library defaults"""
    try:
        for k, v in kw.items():
            ctx.set_option(k, v)
        yield
    finally:
        for k in kw:
            ctx.set_option(k, OPTION_DEFAULTS[k])
