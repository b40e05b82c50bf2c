import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# (test id, label, error, bound) of the parity gates that reported one
_MARGINS = []


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (runs the HIP path through the C-ABI)")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture
def margin(request):
    """record(label, err, bound): a parity error and the bound it was held
    to, listed at the end of the run (worst ratios first)."""
    def record(label, err, bound):
        _MARGINS.append((request.node.nodeid.split("::")[-1], label, float(err), float(bound)))
    return record


def pytest_terminal_summary(terminalreporter):
    if not _MARGINS:
        return
    tr = terminalreporter
    tr.write_sep("-", "parity margins: error / bound (worst 12 of %d gates)" % len(_MARGINS))
    rows = sorted(_MARGINS, key=lambda m: -(m[2] / m[3] if m[3] > 0 else 0.0))
    for test, label, err, bound in rows[:12]:
        tr.write_line(f"{err / bound:6.3f}  {err:.2e} / {bound:.1e}  {test} {label}")
    full = [m for m in _MARGINS if m[1].startswith("full run")]
    if full:
        tr.write_sep("-", "full runs at the workloads' sizes")
        for test, label, err, bound in full:
            tr.write_line(f"{err:.2e} (gate {bound:.0e})  {label}")
