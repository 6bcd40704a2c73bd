"""The oracle's pins must reject plausible mistakes.

Each ORACLE_MUTANT build (oracle/nabo_oracle.cpp) plants one mistake on the
CPU: 1 = off[cd] not restored after a far child, 2 = wrong sign in the
incremental rectangle distance, 3 = self match not excluded, 4 = maxRadius
made exclusive, 5 = leftCount rule replaced by count/2, 6 = argMax keeps the
last widest dimension, 7 = a per-query radius (maxRadii overload) left
unsquared. Every mutant must fail at least one check of
tests/test_oracle.py.
"""
import numpy as np
import pytest

import oracle as orc
from tests import test_oracle as T


def checks(o):
    yield lambda: T.test_hand_worked_tree(o)
    yield lambda: T.test_first_widest_dimension(o)
    yield lambda: T.test_radius_is_inclusive(o)
    yield lambda: T.test_self_match_rule(o)
    yield lambda: T.test_knnvalidate_cases(o, "scan.3d", 10000, np.inf)
    yield lambda: T.test_knnvalidate_cases(o, "scan.3d.large", 1000, 0.5)
    yield lambda: T.test_tree_invariants(o, 8)
    yield lambda: T.test_pruning_visits_exactly_cells_within_radius(o)
    yield lambda: T.test_per_query_radii_equal_scalar_calls(o)


@pytest.mark.parametrize("mutant", [1, 2, 3, 4, 5, 6, 7])
def test_mutant_is_caught(mutant):
    o = orc.Oracle(orc.build(mutant))
    assert o.mutant == mutant
    caught = []
    for i, check in enumerate(checks(o)):
        try:
            check()
        except (AssertionError, RuntimeError, IndexError):
            caught.append(i)
    assert caught, f"mutant {mutant} passed every check"


def test_real_oracle_passes_all():
    o = orc.default()
    assert o.mutant == 0
    for check in checks(o):
        check()
