"""The tie rules of tests/picard_ties.py on residual sequences: the two kinds
of parting they accept, and partings they must reject (a missed contraction,
a straddle outside the band, disagreeing floors)."""

import pytest

from picard_ties import parting_kind

ATOL = 1e-10


def test_same_count_is_no_tie():
    assert parting_kind([1e-3, 1e-7, 5e-11], [1e-3, 1e-7, 6e-11], ATOL) is None


def test_threshold_tie_both_orders():
    assert parting_kind([1e-3, 0.995e-10], [1e-3, 1.004e-10, 2e-14], ATOL) == 'threshold'
    assert parting_kind([1e-3, 1.004e-10, 2e-14], [1e-3, 0.995e-10], ATOL) == 'threshold'


def test_floor_tie_from_a_stalled_run():
    # a 512 x 1024 ensemble instance: both sides stall at ~1.01e-10, the GPU
    # run's residual rises at iteration 6 (revert), the C run's falls
    rc = [6.62e-3, 4.41e-6, 3.113e-9, 1.0266e-10, 1.0126e-10, 1.0105e-10, 1.0110e-10]
    rg = [6.62e-3, 4.41e-6, 3.111e-9, 1.0188e-10, 1.0085e-10, 1.0219e-10]
    assert parting_kind(rc, rg, ATOL) == 'floor'


def test_rejects_real_differences():
    with pytest.raises(AssertionError):          # one side still contracting
        parting_kind([1e-3, 1e-6, 1e-9, 5e-11], [1e-3, 1e-6, 1.2e-9], ATOL)
    with pytest.raises(AssertionError):          # straddle too wide
        parting_kind([1e-3, 0.9e-10], [1e-3, 1.2e-10, 1e-13], ATOL)
    with pytest.raises(AssertionError):          # both stalled, floors disagree
        parting_kind([1e-3, 2e-10, 1.9e-10, 1.95e-10], [1e-3, 3e-10, 3.1e-10], ATOL)
