"""CPU tests of the Nu/Re time statistics (the criteria of calNuRe,
RaNu.F90:59-132), on constructed series whose answers follow from the rules."""
import math

from mglc_amd.statistics import NuReStatistics


def test_constant_series_is_stationary_at_200_and_converged_at_400(tmp_path):
    s = NuReStatistics(output_dir=str(tmp_path))
    for _ in range(399):
        s.add(4.5, 120.0)
    # at t = 100 the previous mean is the empty mean 0: no; at t = 200 the mean
    # has not moved over 100 units: stationary
    assert s.stationary and s.stationary_time == 200 and not s.converged
    s.add(4.5, 120.0)
    assert s.converged and s.time == 400  # 200 units after stationarity, halves agree
    assert (tmp_path / "Nu_Bulk.dat").read_text().count("\n") == 400


def test_drifting_series_never_stationary():
    s = NuReStatistics()
    for t in range(1, 1001):
        s.add(1.0 + 0.01 * t, 50.0)
    assert not s.stationary  # the running mean moves 8 % per 100 units at t = 1000


def test_oscillation_converges_only_after_the_flow_reversal_time():
    s = NuReStatistics(flow_reversal=True, flow_reversal_time=1000)
    for t in range(1, 1801):
        s.add(4.0 + 0.01 * math.sin(t / 7.0), 100.0 + 0.5 * math.cos(t / 11.0))
        if s.stationary and s.time < s.stationary_time + 1000:
            assert not s.converged
    assert s.stationary and s.converged
    assert s.time >= s.stationary_time + 1000


def test_half_windows_include_the_middle_sample_twice():
    # stationary at 200, then a jump: the halves over [200, 300] and [300, 400]
    s = NuReStatistics()
    for _ in range(200):
        s.add(2.0, 10.0)
    assert s.stationary_time == 200
    for t in range(201, 401):
        s.add(2.0 if t <= 300 else 2.1, 10.0)
    # first half mean 2.0, last half (t = 300 .. 400): (2.0 + 100 * 2.1) / 101
    assert not s.converged
