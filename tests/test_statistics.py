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


# ---- against the reference's own calNuRe ------------------------------------
# tests/golden/ranu_series.npz: RaNu.F90 built unmodified (oracle/_ref/ref_ranu)
# over 960 time units of given fields: its Nu and Re values, running means and
# states after every call, and the files it wrote (tests/golden/make_ranu.py).
import os  # noqa: E402

import numpy as np  # noqa: E402
import pytest  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ranu_series.npz")


def _ranu():
    return np.load(GOLDEN)


def test_python_statistics_follow_the_reference_calNuRe():
    """Fed the reference's own Nu and Re values, the running means equal its
    bit for bit (both sum in time order) and the stationary / converged
    states agree after every time unit (stationary at 500, the convergence
    test failing at 700 and passing at 900)."""
    d = _ranu()
    rec = d["rec"]
    s = NuReStatistics()
    for t in range(rec.shape[0]):
        s.add(rec[t, 0], rec[t, 1])
        assert s.nu_mean[-1] == rec[t, 2] and s.re_mean[-1] == rec[t, 3], t
        assert (int(s.stationary), s.stationary_time, int(s.converged)) == tuple(int(x) for x in rec[t, 4:7]), t
    assert s.stationary_time == 500 and s.converged
    assert int(np.argmax(rec[:, 6] > 0)) + 1 == 900


def test_volume_nusselt_and_reynolds_formulas_equal_calNuRe(oracle_lib):
    """The oracle's Nu_vol (the formula the GPU's lbm_nusselt is held to) and
    the Reynolds formula the GPU's lbm_reynolds is held to, on the three
    random fields the reference evaluated: <w T> nz / kappa + 1 and
    sqrt(<u^2 + v^2 + w^2>) nz / nu (RaNu.F90:13-23, 36-44)."""
    O = oracle_lib
    d = _ranu()
    kappa, visc = d["consts"][0], d["consts"][1]
    fields, rec = d["fields"], d["rec"]
    nz, ny, nx = fields.shape[2:]
    p = O.thermal_params(nx, ny, nz, 0.6, 0.0, 0.0, tconst_axis=1)
    for q in range(3):
        u, v, w, T = fields[q]
        nu_vol = O.thermal_nusselt(p, w, T, kappa)[0]
        assert abs(nu_vol - rec[q, 0]) <= 1e-14 * abs(rec[q, 0]), q
        re = float(np.sqrt(np.mean(u * u + v * v + w * w)) * nz / visc)
        assert abs(re - rec[q, 1]) <= 1e-14 * rec[q, 1], q


FC = "/opt/rocm/bin/amdflang"


@pytest.mark.skipif(not os.path.exists(FC), reason="amdflang not available")
def test_driver_statistics_files_equal_the_reference_files(tmp_path):
    """The Fortran driver's calNuRe (mglc_amd/fortran/buoyancy_data.f90),
    given the reference's Nu and Re values, the same itc and timeUnit, writes
    Nu_Bulk.dat, Re_Bulk.dat and statistically.txt byte for byte as the
    reference's calNuRe wrote them."""
    import subprocess
    d = _ranu()
    rec, consts, itc_step = d["rec"], d["consts"], int(d["itc_step"])
    src = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mglc_amd", "fortran",
                       "buoyancy_data.f90")
    harness = tmp_path / "stats_main.f90"
    harness.write_text(f"""program stats_main
    use buoyancy_data
    implicit none
    integer :: q, n
    real(8), allocatable :: vals(:, :)
    n = {rec.shape[0]}
    allocate(vals(2, n))
    open(10, file='series.bin', form='unformatted', access='stream', status='old')
    read(10) timeUnit, vals
    close(10)
    allocate(NuVolAvg(0:1000), ReVolAvg(0:1000), NuVolAvg_mean(0:1000), ReVolAvg_mean(0:1000))
    NuVolAvg = 0.0d0; ReVolAvg = 0.0d0; NuVolAvg_mean = 0.0d0; ReVolAvg_mean = 0.0d0
    output_dir = '.'
    do q = 1, n
        itc = q * {itc_step}
        call calNuRe(vals(1, q), vals(2, q))
    enddo
end program stats_main
""")
    mod = tmp_path / "mod"
    mod.mkdir()
    exe = tmp_path / "stats"
    subprocess.run([FC, "-O2", "-J", str(mod), src, str(harness), "-o", str(exe)], check=True, cwd=tmp_path)
    np.concatenate([consts[2:3], rec[:, :2].ravel()]).tofile(tmp_path / "series.bin")
    subprocess.run([str(exe)], check=True, cwd=tmp_path)
    assert (tmp_path / "Nu_Bulk.dat").read_bytes() == bytes(d["nu_bulk"])
    assert (tmp_path / "Re_Bulk.dat").read_bytes() == bytes(d["re_bulk"])
    assert (tmp_path / "statistically.txt").read_bytes() == bytes(d["statistically"])
