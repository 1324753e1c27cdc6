import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# unit roundoff of FP64
U64 = 2.0**-53


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD MI355X GPU (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU check (full golden runs)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle

    pyoracle.lib()
    return pyoracle


def stop_margin_bound(n):
    """Largest relative difference between the residual of the StopList
    computed with the device's fixed-order tree sums and the oracle's
    sequential sums: each of the two norms is a sum of n non-negative terms
    (each sum within (n - 1 + log2 n) u of the exact one, the square root
    halves it), plus the division; 4 n u covers both with room."""
    return 4.0 * max(n, 2) * U64


@pytest.fixture(autouse=True)
def stop_decision_margin(request):
    """In every GPU test, each minimise-type call of the oracle
    (minimise, minimise_truncate, timeStepsUntilEvent, and so every event of
    the quasistatic protocol) must decide its stop step with a margin above
    the rounding bound of the residual norms: the device sums the norms in
    another order, so a near-tie could otherwise flip the stop step without
    any test noticing."""
    if request.node.get_closest_marker("gpu") is None:
        yield
        return
    from oracle import pyoracle

    pyoracle.STOP_MARGINS.clear()
    yield
    bad = [(n, m) for n, m in pyoracle.STOP_MARGINS if not (m > stop_margin_bound(n))]
    assert not bad, f"StopList decisions within the rounding bound of the residual norm: {bad[:5]}"
