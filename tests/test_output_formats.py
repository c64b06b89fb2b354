"""The cavity driver's result files against the reference's own writers:
mglc_amd/fortran/output.f90 (output_Tecplot, getVelocity, output_binary),
given the fields the reference's writers were given
(tests/golden/make_output.py: the reference's output_binary, output_Tecplot
and getVelocity, lid_driven_cavity_3d.f90:583-723, built unmodified, on its
65^3 grid), writes every file byte for byte as they did: the Tecplot
#!TDV101 binary, the Fortran sequential binary records and the two centre
lines."""
import hashlib
import json
import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
FC = "/opt/rocm/bin/amdflang"

sys.path.insert(0, os.path.join(HERE, "golden"))
import make_output  # noqa: E402  (the fixture's field recipe)


@pytest.mark.skipif(not os.path.exists(FC), reason="amdflang not available")
def test_cavity_output_files_equal_the_reference_writers(tmp_path):
    ref = json.load(open(os.path.join(HERE, "golden", "output_files.json")))
    n, itc, u0 = ref["n"], ref["itc"], ref["U0"]
    harness = tmp_path / "out_main.f90"
    harness.write_text(f"""program out_main
    use cavity_output
    implicit none
    integer, parameter :: n = {n}
    real(8) :: xp(0:n+1), yp(0:n+1), zp(0:n+1)
    real(8), allocatable :: rho(:, :, :), u(:, :, :), v(:, :, :), w(:, :, :)
    integer :: i
    allocate(rho(n, n, n), u(n, n, n), v(n, n, n), w(n, n, n))
    open(10, file='fields.bin', form='unformatted', access='stream', status='old')
    read(10) rho, u, v, w
    close(10)
    ! node coordinates as the cavity driver sets them (mglc_amd/fortran/main.f90)
    xp(0) = 0.0d0; xp(n+1) = dble(n)
    do i = 1, n
        xp(i) = dble(i) - 0.5d0
    enddo
    yp = xp
    zp = xp
    call output_binary('.', u, v, rho, n, n, n, {itc})
    call output_Tecplot('.', xp, yp, zp, u, v, w, rho, n, n, n, {itc})
    call getVelocity('.', xp, zp, u, w, n, n, n, {u0!r}d0, {itc})
end program out_main
""")
    mod = tmp_path / "mod"
    mod.mkdir()
    exe = tmp_path / "out"
    src = os.path.join(ROOT, "mglc_amd", "fortran", "output.f90")
    subprocess.run([FC, "-O2", "-J", str(mod), src, str(harness), "-o", str(exe)], check=True, cwd=tmp_path)
    np.concatenate(make_output.fields()).tofile(tmp_path / "fields.bin")
    subprocess.run([str(exe)], check=True, cwd=tmp_path)
    for name, want in ref["files"].items():
        data = (tmp_path / name).read_bytes()
        assert len(data) == want["bytes"], name
        assert hashlib.sha256(data).hexdigest() == want["sha256"], name


@pytest.mark.skipif(not os.path.exists(FC), reason="amdflang not available")
def test_buoyancy_output_files_equal_the_reference_writers(tmp_path):
    """The buoyancy driver's Tecplot (X, Y, Z, U, V, W, T) and sequential
    binary (u, v, w, T) against the reference's own writers
    (bouyancy3d.F90:984-1164, 51^3)."""
    ref = json.load(open(os.path.join(HERE, "golden", "output_files.json")))
    n, itc = ref["n_buoyancy"], ref["itc"]
    harness = tmp_path / "out_main.f90"
    harness.write_text(f"""program out_main
    use cavity_output
    implicit none
    integer, parameter :: n = {n}
    real(8) :: xp(0:n+1), yp(0:n+1), zp(0:n+1)
    real(8), allocatable :: u(:, :, :), v(:, :, :), w(:, :, :), T(:, :, :)
    integer :: i
    allocate(u(n, n, n), v(n, n, n), w(n, n, n), T(n, n, n))
    open(10, file='fields.bin', form='unformatted', access='stream', status='old')
    read(10) u, v, w, T
    close(10)
    ! node coordinates as the buoyancy driver sets them (mglc_amd/fortran/buoyancy_main.f90)
    xp(0) = 0.0d0; xp(n+1) = dble(n)
    do i = 1, n
        xp(i) = dble(i) - 0.5d0
    enddo
    yp = xp
    zp = xp
    call output_binary_buoyancy('.', u, v, w, T, n, n, n, {itc})
    call output_Tecplot_buoyancy('.', xp, yp, zp, u, v, w, T, n, n, n, {itc})
end program out_main
""")
    mod = tmp_path / "mod"
    mod.mkdir()
    exe = tmp_path / "out"
    src = os.path.join(ROOT, "mglc_amd", "fortran", "output.f90")
    subprocess.run([FC, "-O2", "-J", str(mod), src, str(harness), "-o", str(exe)], check=True, cwd=tmp_path)
    np.concatenate(make_output.fields_buoyancy()).tofile(tmp_path / "fields.bin")
    subprocess.run([str(exe)], check=True, cwd=tmp_path)
    for name, want in ref["files_buoyancy"].items():
        data = (tmp_path / name).read_bytes()
        assert len(data) == want["bytes"], name
        assert hashlib.sha256(data).hexdigest() == want["sha256"], name
