"""Shared pytest fixtures.

Markers:

``gpu``
    Needs a HIP device and the built ``libslc_hip.so``; these are the parity
    tests proper.  Run with ``-m gpu``.
``slow``
    Minutes of CPU work (full-length oracle runs).

The ``check`` fixture records every measured quantity and reprints it as a
``[PASS]/[FAIL] name detail`` table in the terminal summary, failing the test
if any check is false (the same soft-assertion style as the reference suite,
tests/conftest.py:44-76 there).
"""

import os
import sys

import pytest

_REPO = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
if _REPO not in sys.path:
    sys.path.insert(0, _REPO)

_ALL_CHECKS = []


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device and libslc_hip.so")
    config.addinivalue_line("markers", "slow: minutes of CPU work")


class Checker:
    def __init__(self, nodeid):
        self.nodeid = nodeid
        self.results = []

    def __call__(self, name, ok, detail=""):
        ok = bool(ok)
        self.results.append((name, ok, detail))
        return ok

    @property
    def failures(self):
        return [(n, d) for n, ok, d in self.results if not ok]


@pytest.fixture
def check(request):
    checker = Checker(request.node.nodeid)
    yield checker
    _ALL_CHECKS.extend((checker.nodeid, n, ok, d) for n, ok, d in checker.results)
    if not checker.results:
        pytest.fail("test recorded no checks")
    if checker.failures:
        pytest.fail("failed checks:\n  " + "\n  ".join(f"{n}  {d}" for n, d in checker.failures))


def pytest_collection_modifyitems(config, items):
    """Skip (never silently pass) `gpu` tests when no HIP device is visible."""
    import torch

    if torch.cuda.is_available():
        return
    skip = pytest.mark.skip(reason="needs a HIP device (run with -m gpu on an MI355X)")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def golden():
    """Loader for the committed golden fixtures (tests/golden/*.npz)."""
    import numpy as np

    def load(name):
        return dict(np.load(os.path.join(_REPO, "tests", "golden", name + ".npz")))

    return load


def pytest_terminal_summary(terminalreporter, exitstatus, config):
    if not _ALL_CHECKS:
        return
    tr = terminalreporter
    tr.write_sep("=", "measured quantities")
    last = None
    for nodeid, name, ok, detail in _ALL_CHECKS:
        head = nodeid.split("::")[0]
        if head != last:
            tr.write_line(head)
            last = head
        tr.write_line(f"  [{'PASS' if ok else 'FAIL'}] {name}  {detail}", green=ok, red=not ok)
