"""The reference's closed-form unit tests, run against the CPU oracle."""

import pytest

from tests import reference_cases as RC


@pytest.mark.parametrize("case", RC.ALL, ids=[f.__name__ for f in RC.ALL])
def test_reference_case(oracle, case):
    case(oracle)
