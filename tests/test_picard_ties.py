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
    # an ensemble instance: 9.84e-11 (stopped) against 1.00001e-10 (went on), 1.6 % apart
    rc = [5.988e-2, 3.8628e-05, 2.5515e-08, 9.8436e-11]
    rg = [5.988e-2, 3.8628e-05, 2.5512e-08, 1.00001e-10, 8.318e-11]
    assert parting_kind(rc, rg, ATOL) == 'threshold'


def test_floor_tie_from_a_stalled_run():
    # a 512 x 1024 ensemble instance: both sides stall at ~1.01e-10, the GPU
    # run's residual rises at iteration 6 (revert), the C run's falls
    rc = [6.62e-3, 4.41e-6, 3.113e-9, 1.0266e-10, 1.0126e-10, 1.0105e-10, 1.0110e-10]
    rg = [6.62e-3, 4.41e-6, 3.111e-9, 1.0188e-10, 1.0085e-10, 1.0219e-10]
    assert parting_kind(rc, rg, ATOL) == 'floor'


def test_floor_tie_where_the_floor_is_reached():
    # the same instance on another run: the C side lands at 9.7e-11 (below
    # snes_atol), the GPU side at 1.02e-10; both are ~40x above what the
    # contraction predicts ((3.1e-9)^2 / 4.4e-6 = 2.2e-12)
    rc = [6.62e-3, 4.4128e-06, 3.1085e-09, 9.706e-11]
    rg = [6.62e-3, 4.4128e-06, 3.1114e-09, 1.0188e-10, 1.0085e-10, 1.0219e-10]
    assert parting_kind(rc, rg, ATOL) in ('threshold', 'floor')


def test_floor_tie_across_the_threshold():
    # an ensemble instance: 8.90e-11 (stopped) against 1.02e-10 (went on)
    rc = [5.2823e-2, 3.39748e-05, 2.23938e-08, 8.9019e-11]
    rg = [5.2823e-2, 3.39748e-05, 2.24016e-08, 1.02087e-10, 7.967e-11]
    assert parting_kind(rc, rg, ATOL) == 'floor'


def test_rejects_real_differences():
    with pytest.raises(AssertionError):          # one side still contracting
        parting_kind([1e-3, 1e-6, 1e-9, 5e-11], [1e-3, 1e-6, 1.2e-9], ATOL)
    with pytest.raises(AssertionError):          # far apart near the threshold
        parting_kind([1e-3, 0.4e-10], [1e-3, 1.2e-10, 1e-13], ATOL)
    with pytest.raises(AssertionError):          # one side's residual rose, the other's fell
        parting_kind([1e-3, 1e-6, 2e-6], [1e-3, 1e-6, 1e-9, 1e-12], ATOL)
    with pytest.raises(AssertionError):          # both stalled, floors disagree
        parting_kind([1e-3, 2e-10, 1.9e-10, 1.95e-10], [1e-3, 3e-10, 3.1e-10], ATOL)
    with pytest.raises(AssertionError):          # one side still on its contraction line
        parting_kind([6.6e-3, 4.4e-6, 3.1e-9, 2.5e-12], [6.6e-3, 4.4e-6, 3.1e-9, 1.0e-10, 1.01e-10], ATOL)
    with pytest.raises(AssertionError):          # both well above the threshold, still contracting
        parting_kind([1e-3, 1e-6, 5e-9], [1e-3, 1e-6, 6e-9, 1e-11], ATOL)
