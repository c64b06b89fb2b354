"""Time statistics of the buoyancy-driven cavity: the bulk Nusselt and
Reynolds series, their running means, and the stationarity and convergence
tests of the reference's calNuRe
(Buoyancy_driven_cavity/fortran/3d/mpi_blocked/RaNu.F90:1-137).

The reference calls calNuRe once per time unit with Nu_vol and Re_vol of the
current fields; here the caller passes those values (from
``Cavity.nusselt()`` and ``Cavity.reynolds()``) to :meth:`NuReStatistics.add`.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field


@dataclass
class NuReStatistics:
    """Running statistics with the reference's criteria:

    * stationary: every 100 time units, the running means of Nu and Re have
      both moved by at most 1 % over the last 100 units (RaNu.F90:107-132);
    * converged: once stationary, every 200 units the means of the first and
      the second half of the stationary window agree within 1 % for both
      (RaNu.F90:59-105); with flow reversal, not before
      ``flow_reversal_time`` units after stationarity.
    """

    flow_reversal: bool = False
    flow_reversal_time: int = 0
    output_dir: str | None = None   # append "Nu_Bulk.dat" / "Re_Bulk.dat" as the reference does
    nu: list = field(default_factory=list)
    re: list = field(default_factory=list)
    nu_mean: list = field(default_factory=list)
    re_mean: list = field(default_factory=list)
    stationary: bool = False
    stationary_time: int = 0
    converged: bool = False
    log: list = field(default_factory=list)

    @property
    def time(self) -> int:
        """dimensionlessTime: the number of samples so far."""
        return len(self.nu)

    def _mean_at(self, series, t):
        return series[t - 1] if t >= 1 else 0.0

    def add(self, nu: float, re: float, itc_over_time_unit: float | None = None):
        self.nu.append(float(nu))
        self.re.append(float(re))
        t = self.time
        self.nu_mean.append(sum(self.nu) / t)
        self.re_mean.append(sum(self.re) / t)
        if self.output_dir is not None:
            stamp = t if itc_over_time_unit is None else itc_over_time_unit
            with open(os.path.join(self.output_dir, "Nu_Bulk.dat"), "a") as fh:
                fh.write(f" {stamp} {nu!r}\n")
            with open(os.path.join(self.output_dir, "Re_Bulk.dat"), "a") as fh:
                fh.write(f" {stamp} {re!r}\n")
        if self.stationary and (t - self.stationary_time) % 200 == 0:
            half = (self.stationary_time + t) // 2
            first = range(self.stationary_time, half + 1)
            last = range(half, t + 1)
            nf = sum(self.nu[i - 1] for i in first) / len(first)
            rf = sum(self.re[i - 1] for i in first) / len(first)
            nl = sum(self.nu[i - 1] for i in last) / len(last)
            rl = sum(self.re[i - 1] for i in last) / len(last)
            if abs((nf - nl) / nl) <= 0.01 and abs((rf - rl) / rl) <= 0.01:
                if not self.flow_reversal or t >= self.stationary_time + self.flow_reversal_time:
                    if not self.converged:
                        self.log.append(f"statistically converged at time {t}")
                    self.converged = True
        if not self.stationary and t % 100 == 0:
            nm, rm = self.nu_mean[t - 1], self.re_mean[t - 1]
            nm0, rm0 = self._mean_at(self.nu_mean, t - 100), self._mean_at(self.re_mean, t - 100)
            if abs((nm - nm0) / nm) <= 0.01 and abs((rm - rm0) / rm) <= 0.01:
                self.stationary = True
                self.stationary_time = t
                self.log.append(f"statistically stationary at time {t}")
