"""The batch autotuner's decision rule (host logic, no GPU): timings within the
tolerance count as equal, so noise cannot flip the choice between settings of
the same speed, and among those the lower-SAH one wins."""

from nexusbvh_amd.batch import better_setting, faster_than


def test_faster_only_beyond_the_tolerance():
    assert faster_than(4.80, 5.00, 0.01)
    assert not faster_than(4.96, 5.00, 0.01)  # 0.8% faster: noise
    assert not faster_than(5.00, 5.00, 0.01)
    assert not faster_than(5.20, 5.00, 0.01)


def test_equal_speed_goes_to_the_lower_sah():
    # radius 4 (lower SAH) and radius 2 within timing noise of each other, in
    # either order of trial
    assert better_setting(4.76, 29.90, 4.75, 30.02, 0.01)
    assert not better_setting(4.75, 30.02, 4.76, 29.90, 0.01)


def test_clearly_faster_wins_over_lower_sah():
    assert better_setting(4.50, 30.20, 4.75, 29.90, 0.01)


def test_slower_beyond_the_tolerance_loses_even_with_lower_sah():
    assert not better_setting(4.90, 29.00, 4.75, 30.00, 0.01)
