"""ctypes loader for the CPU oracle (oracle/liblbm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker or the timed CPU
baseline -- never the thing measured or shipped.  The product path
(mglc_amd) does not import this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbm_oracle.so")
REF_CAVITY = os.path.join(HERE, "_ref", "ref_cavity")

MRT, BGK = 0, 1


class CavityParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int),
        ("ny", ctypes.c_int),
        ("nz", ctypes.c_int),
        ("collision", ctypes.c_int),
        ("tau", ctypes.c_double),
        ("u_lid", ctypes.c_double),
        ("rho0", ctypes.c_double),
        ("meq12_rho", ctypes.c_int),
    ]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-f", os.path.join(HERE, "Makefile")], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH} (run make -f oracle/Makefile)")
        _lib = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        _lib.oracle_cavity_run.argtypes = [ctypes.POINTER(CavityParams), ctypes.c_int, ctypes.c_int,
                                           dp, dp, dp, dp, dp, ctypes.c_int, ctypes.c_int]
        _lib.oracle_cavity_run.restype = ctypes.c_int
        _lib.oracle_cavity_run_from.argtypes = [ctypes.POINTER(CavityParams), ctypes.c_int,
                                                dp, dp, dp, dp, ctypes.c_int]
        _lib.oracle_cavity_run_from.restype = ctypes.c_int
        _lib.oracle_d3q19_basis.argtypes = [dp, dp]
        _lib.oracle_last_loop_seconds.restype = ctypes.c_double
        for name in ("oracle_cavity_collision",):
            getattr(_lib, name).restype = None
        _lib.oracle_cavity_check.restype = ctypes.c_double
    return _lib


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def cavity_params(nx, ny, nz, tau, u_lid=0.1, collision=MRT, rho0=1.0, meq12_rho=0) -> CavityParams:
    return CavityParams(nx, ny, nz, collision, tau, u_lid, rho0, meq12_rho)


def fields(nx, ny, nz):
    """Four (nz, ny, nx) arrays -- C order with x fastest, i.e. Fortran (nx,ny,nz)."""
    return [np.zeros((nz, ny, nx), dtype=np.float64) for _ in range(4)]


def cavity_run(p: CavityParams, nsteps: int, check_every: int = 0, nthreads: int = 0):
    """Reference initial + nsteps steps.  Returns (rho, u, v, w, errs)."""
    rho, u, v, w = fields(p.nx, p.ny, p.nz)
    nmax = max(1, nsteps // check_every) if check_every > 0 else 1
    errs = np.zeros(nmax, dtype=np.float64)
    r = lib().oracle_cavity_run(ctypes.byref(p), nsteps, check_every, _dp(rho), _dp(u), _dp(v), _dp(w),
                                _dp(errs), nmax, nthreads)
    if r < 0:
        raise MemoryError("oracle_cavity_run failed to allocate")
    return rho, u, v, w, errs[:r]


def cavity_run_from(p: CavityParams, nsteps: int, rho, u, v, w, nthreads: int = 0):
    rho, u, v, w = (np.ascontiguousarray(a, dtype=np.float64).copy() for a in (rho, u, v, w))
    r = lib().oracle_cavity_run_from(ctypes.byref(p), nsteps, _dp(rho), _dp(u), _dp(v), _dp(w), nthreads)
    if r < 0:
        raise MemoryError("oracle_cavity_run_from failed to allocate")
    return rho, u, v, w


def last_loop_seconds() -> float:
    return lib().oracle_last_loop_seconds()


def d3q19_basis():
    M = np.zeros((19, 19), dtype=np.float64)
    Minv = np.zeros((19, 19), dtype=np.float64)
    lib().oracle_d3q19_basis(_dp(M), _dp(Minv))
    return M, Minv


def run_reference_cavity(nx, ny, nz, nsteps, outfile):
    """Run the reference Fortran (oracle/_ref/ref_cavity, built from the
    reference's own sources) and read back (tauf, U0, rho0), rho, u, v, w."""
    subprocess.run([REF_CAVITY, str(nx), str(ny), str(nz), str(nsteps), outfile], check=True,
                   stdout=subprocess.DEVNULL)
    raw = np.fromfile(outfile, dtype=np.float64)
    head, body = raw[:3], raw[3:]
    n = nx * ny * nz
    assert body.size == 4 * n, (body.size, n)
    rho, u, v, w = (body[i * n:(i + 1) * n].reshape(nz, ny, nx) for i in range(4))
    return dict(tauf=head[0], U0=head[1], rho0=head[2]), rho, u, v, w
