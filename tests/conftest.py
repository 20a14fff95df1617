import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def _built():
    lib = os.path.join(ROOT, "unitigcoarsening_amd", "_lib")
    need = [os.path.join(lib, "libuc.so"), os.path.join(lib, "libuc_synth.so"), os.path.join(lib, "uc_main"),
            os.path.join(ROOT, "oracle", "liboracle.so")]
    return all(os.path.exists(p) for p in need)


def pytest_sessionstart(session):
    # build only when nothing is built (a machine without the toolchain runs on prebuilt files)
    if not _built():
        from unitigcoarsening_amd.build import build_all

        build_all()
