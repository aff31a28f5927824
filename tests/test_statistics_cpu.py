"""The ensemble comparison used by tests/test_gpu_statistics.py has power:
two ensembles of the oracle that differ only in the reaction selection
(the reference's biased rand1*srrate against the textbook rand1*totrate,
simulate_rdme.cpp:255-261) are told apart, and two independent ensembles
of the same algorithm are not."""
import numpy as np
import pytest

from spatialpy_amd import _abi, workloads

from test_gpu_statistics import _check


def _runs(oracle, sa, seeds):
    o = oracle.Oracle(sa)
    return np.stack([o.run(seed=s)["xx"].astype(np.int64) for s in seeds])


def test_comparison_accepts_independent_ensembles_of_one_algorithm(oracle):
    sa = workloads.birth_death_2d(seed=1, t_end=2, num_points=3)
    a = _runs(oracle, sa, range(1000, 1300))
    b = _runs(oracle, sa, range(2000, 2300))
    _check(sa, a, b, "same algorithm")


def test_comparison_detects_the_selection_rule(oracle):
    sa = workloads.birth_death_2d(seed=1, t_end=2, num_points=3)
    a = _runs(oracle, sa, range(1000, 1300))
    sa.flags |= _abi.FLAG_TEXTBOOK_SELECTION
    b = _runs(oracle, sa, range(2000, 2300))
    with pytest.raises(AssertionError):
        _check(sa, a, b, "biased vs textbook selection")
