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
        ("periodic_mask", ctypes.c_int),
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


def cavity_params(nx, ny, nz, tau, u_lid=0.1, collision=MRT, rho0=1.0, meq12_rho=0,
                  periodic_mask=0) -> CavityParams:
    return CavityParams(nx, ny, nz, collision, tau, u_lid, rho0, meq12_rho, periodic_mask)


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


def reference_loop_seconds(kind, nx, ny, nz, nsteps):
    """Wall time of the step loop of the reference's own subroutines
    (oracle/_ref/ref_cavity or ref_thermal, one process, initial and output
    excluded) over nsteps steps: the CPU baseline of the reference's code."""
    import tempfile
    exe = REF_CAVITY if kind == "cavity" else REF_THERMAL
    with tempfile.TemporaryDirectory() as d:
        out = subprocess.run([exe, str(nx), str(ny), str(nz), str(nsteps), os.path.join(d, "fields.bin")],
                             check=True, capture_output=True, text=True).stdout
    for line in out.splitlines():
        parts = line.split()
        if len(parts) == 2 and parts[0] == "ref_loop_seconds":
            return float(parts[1])
    raise RuntimeError(f"{exe} printed no ref_loop_seconds line")


def reference_loop_seconds_slabs(kind, nx, ny, nzs, nsteps):
    """The reference's own step loop (as reference_loop_seconds) run as
    len(nzs) processes at once, process p on an nx x ny x nzs[p] slab: the
    subdomains of a z decomposition of the reference's MPI run, each a whole
    walled cavity, without the halo exchange between them.  Returns each
    process's loop seconds."""
    import tempfile
    exe = REF_CAVITY if kind == "cavity" else REF_THERMAL
    with tempfile.TemporaryDirectory() as d:
        procs = [subprocess.Popen([exe, str(nx), str(ny), str(nz), str(nsteps), os.path.join(d, f"fields{p}.bin")],
                                  stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
                 for p, nz in enumerate(nzs)]
        outs = [pr.communicate() for pr in procs]
    secs = []
    for pr, (out, err) in zip(procs, outs):
        if pr.returncode != 0:
            raise RuntimeError(f"{exe} failed ({pr.returncode}): {err[-500:]}")
        vals = [float(l.split()[1]) for l in out.splitlines() if l.startswith("ref_loop_seconds")]
        if not vals:
            raise RuntimeError(f"{exe} printed no ref_loop_seconds line")
        secs.append(vals[0])
    return secs


# ---- buoyancy-driven cavity (D3Q19 + Guo forcing, D3Q7 temperature) ----------
REF_THERMAL = os.path.join(HERE, "_ref", "ref_thermal")


class ThermalParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("tconst_axis", ctypes.c_int),
        ("tau", ctypes.c_double), ("para_a", ctypes.c_double), ("g_beta", ctypes.c_double),
        ("omega_rot", ctypes.c_double), ("t_hot", ctypes.c_double), ("t_cold", ctypes.c_double),
        ("t_ref", ctypes.c_double),
    ]


def thermal_derived(N, Rayleigh, Prandtl, Mach=0.1, Ekman=None):
    """The parameter statements of module.F90:19-29, computed in the same order."""
    import math
    tauf = 0.5 + Mach * float(N) * math.sqrt(3.0 * Prandtl / Rayleigh)
    viscosity = (tauf - 0.5) / 3.0
    diffusivity = viscosity / Prandtl
    omega_rot = 0.0 if Ekman is None else viscosity / 2.0 / Ekman / float(N * N)
    para_a = 42.0 * math.sqrt(3.0) * diffusivity - 6.0
    g_beta1 = Rayleigh * viscosity * diffusivity / float(N)
    g_beta = g_beta1 / float(N * N)
    return dict(tau=tauf, viscosity=viscosity, diffusivity=diffusivity, omega_rot=omega_rot, para_a=para_a,
                g_beta=g_beta)


def thermal_params(nx, ny, nz, tau, para_a, g_beta, omega_rot=0.0, t_hot=0.5, t_cold=-0.5, t_ref=0.0,
                   tconst_axis=1) -> ThermalParams:
    return ThermalParams(nx, ny, nz, tconst_axis, tau, para_a, g_beta, omega_rot, t_hot, t_cold, t_ref)


def _thermal_lib():
    L = lib()
    if not getattr(L, "_thermal_ready", False):
        dp = ctypes.POINTER(ctypes.c_double)
        L.oracle_thermal_run.argtypes = [ctypes.POINTER(ThermalParams), ctypes.c_int, ctypes.c_int, dp, dp, dp,
                                         dp, dp, ctypes.c_int, dp, ctypes.c_int, ctypes.c_int]
        L.oracle_thermal_run.restype = ctypes.c_int
        L.oracle_guo_source_moments.argtypes = [ctypes.c_double] * 6 + [dp]
        L.oracle_d3q7_basis.argtypes = [dp, dp]
        L.oracle_thermal_nusselt.argtypes = [ctypes.POINTER(ThermalParams), dp, dp, ctypes.c_double, dp, dp]
        L._thermal_ready = True
    return L


def thermal_run(p: ThermalParams, nsteps: int, fields=None, check_every: int = 0, nthreads: int = 0):
    """Reference initial (or `fields` = (rho, u, v, w, T)) + nsteps steps.
    Returns (rho, u, v, w, T, errs) with errs an (n, 2) array of (errorU, errorT)."""
    shape = (p.nz, p.ny, p.nx)
    if fields is None:
        arrs = [np.zeros(shape) for _ in range(5)]
        from_fields = 0
    else:
        arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(shape).copy() for a in fields]
        from_fields = 1
    nmax = max(1, nsteps // check_every) if check_every > 0 else 1
    errs = np.zeros(2 * nmax)
    r = _thermal_lib().oracle_thermal_run(ctypes.byref(p), nsteps, check_every, *[_dp(a) for a in arrs],
                                          from_fields, _dp(errs), nmax, nthreads)
    if r < 0:
        raise MemoryError("oracle_thermal_run failed to allocate")
    return (*arrs, errs[:2 * r].reshape(r, 2))


def guo_source_moments(u, v, w, Fx, Fy, Fz):
    S = np.zeros(19)
    _thermal_lib().oracle_guo_source_moments(u, v, w, Fx, Fy, Fz, _dp(S))
    return S


def d3q7_basis():
    N = np.zeros((7, 7)); Ninv = np.zeros((7, 7))
    _thermal_lib().oracle_d3q7_basis(_dp(N), _dp(Ninv))
    return N, Ninv


def thermal_nusselt(p: ThermalParams, w, T, kappa):
    """(nu_vol, nu_wall): calNuRe's <w T> nz / kappa + 1 (RaNu.F90:13-23) and
    the conductive flux through the hot wall."""
    a = np.ascontiguousarray(w, dtype=np.float64)
    t = np.ascontiguousarray(T, dtype=np.float64)
    nv, nw = ctypes.c_double(), ctypes.c_double()
    _thermal_lib().oracle_thermal_nusselt(ctypes.byref(p), _dp(a), _dp(t), kappa, ctypes.byref(nv), ctypes.byref(nw))
    return nv.value, nw.value


def run_reference_thermal(nx, ny, nz, nsteps, outfile):
    """Run oracle/_ref/ref_thermal (the reference's own buoyancy subroutines) and
    read back its parameters and rho, u, v, w, T."""
    subprocess.run([REF_THERMAL, str(nx), str(ny), str(nz), str(nsteps), outfile], check=True,
                   stdout=subprocess.DEVNULL)
    raw = np.fromfile(outfile, dtype=np.float64)
    head, body = raw[:8], raw[8:]
    n = nx * ny * nz
    assert body.size == 5 * n, (body.size, n)
    keys = ("tauf", "paraA", "gBeta", "omegaRatating", "Thot", "Tcold", "Tref", "diffusivity")
    fields = [body[i * n:(i + 1) * n].reshape(nz, ny, nx) for i in range(5)]
    return dict(zip(keys, head)), fields


# ---- moving spheres ------------------------------------------------------------
class ParticleParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("np", ctypes.c_int),
        ("collision", ctypes.c_int), ("meq12_rho", ctypes.c_int),
        ("tau", ctypes.c_double), ("rho0", ctypes.c_double), ("rho_solid", ctypes.c_double),
        ("gravity", ctypes.c_double), ("threshold_wall", ctypes.c_double), ("stiff_wall", ctypes.c_double),
        ("threshold_particle", ctypes.c_double), ("stiff_particle", ctypes.c_double),
    ]


def particle_params(nx, ny, nz, np_, tau, rho_solid=1.5, gravity=1e-5, rho0=1.0, collision=MRT, meq12_rho=1,
                    threshold_wall=4.0, stiff_wall=0.10, threshold_particle=4.0, stiff_particle=0.10):
    return ParticleParams(nx, ny, nz, np_, collision, meq12_rho, tau, rho0, rho_solid, gravity, threshold_wall,
                          stiff_wall, threshold_particle, stiff_particle)


def particle_library(path: str) -> ctypes.CDLL:
    """Another build of the oracle library (a copy compiled elsewhere, e.g.
    with a formula altered on purpose to show that a test catches it), set
    up for particle_run(..., L=)."""
    return _particle_lib(ctypes.CDLL(path))


def _particle_lib(L=None):
    L = lib() if L is None else L
    if not getattr(L, "_particle_ready", False):
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.oracle_particle_map.argtypes = [ctypes.POINTER(ParticleParams), dp, dp, ip]
        L.oracle_calq.argtypes = [dp, dp, ctypes.c_int, dp, ctypes.c_int, dp, dp]
        L.oracle_calq.restype = ctypes.c_int
        L.oracle_calq_closed.argtypes = [dp, dp, ctypes.c_int, dp, ctypes.c_int]
        L.oracle_calq_closed.restype = ctypes.c_double
        L.oracle_particle_run.argtypes = [ctypes.POINTER(ParticleParams), ctypes.c_int, dp, dp, dp, dp, dp, dp, ip,
                                          dp, dp, dp, dp, dp]
        L.oracle_particle_run.restype = ctypes.c_int
        L._particle_ready = True
    return L


def _ip(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def particle_map(p: ParticleParams, xc, r):
    owner = np.zeros((p.nz, p.ny, p.nx), dtype=np.int32)
    xc = np.ascontiguousarray(xc, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    _particle_lib().oracle_particle_map(ctypes.byref(p), _dp(xc), _dp(r), _ip(owner))
    return owner


def calq(xc, r, pid, x, a):
    xc = np.ascontiguousarray(xc, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    x0 = np.zeros(3)
    q = ctypes.c_double()
    rc = _particle_lib().oracle_calq(_dp(xc), _dp(r), pid, _dp(x), a, _dp(x0), ctypes.byref(q))
    closed = _particle_lib().oracle_calq_closed(_dp(xc), _dp(r), pid, _dp(x), a)
    return rc, q.value, x0, closed


def set_threads(n: int, L=None) -> None:
    """Threads of the oracle's OpenMP loops that take no thread count of
    their own (the particle oracles' per-node loops)."""
    L = lib() if L is None else L
    L.oracle_set_threads.argtypes = [ctypes.c_int]
    L.oracle_set_threads.restype = None
    L.oracle_set_threads(int(n))


def particle_run(p: ParticleParams, nsteps: int, xc, r, U=None, Om=None, L=None):
    """Spheres at xc (np x 3, 1-based node coordinates) with radii r, fluid at
    rest; returns a dict of the state after nsteps (rc != 0 on a stop).  L:
    a library from particle_library (default: this directory's build)."""
    n = p.np
    xc = np.ascontiguousarray(xc, dtype=np.float64).reshape(n, 3).copy()
    r = np.ascontiguousarray(r, dtype=np.float64)
    U = np.zeros((n, 3)) if U is None else np.ascontiguousarray(U, dtype=np.float64).copy()
    Om = np.zeros((n, 3)) if Om is None else np.ascontiguousarray(Om, dtype=np.float64).copy()
    F = np.zeros((n, 3)); T = np.zeros((n, 3))
    shape = (p.nz, p.ny, p.nx)
    owner = np.zeros(shape, dtype=np.int32)
    rho, u, v, w = (np.zeros(shape) for _ in range(4))
    ra = ctypes.c_double()
    rc = _particle_lib(L).oracle_particle_run(ctypes.byref(p), nsteps, _dp(xc), _dp(U), _dp(Om), _dp(r), _dp(F), _dp(T),
                                             _ip(owner), _dp(rho), _dp(u), _dp(v), _dp(w), ctypes.byref(ra))
    return dict(rc=rc, xc=xc, U=U, Om=Om, F=F, T=T, owner=owner, rho=rho, u=u, v=v, w=w, rho_avg=ra.value)


# ---- the reference's two-dimensional two-disc program (CaseIV_DKT.F90) --------
REF_DKT = os.path.join(HERE, "_ref", "ref_dkt")
DKT_SCALARS = ("itc", "shift", "rho_avg", "xc", "yc", "uc", "vc", "omega", "fx", "fy", "torque")


def _dkt_unpack(scal):
    out = {"itc": int(scal[0]), "shift": int(scal[1]), "rho_avg": float(scal[2])}
    for q, name in enumerate(DKT_SCALARS[3:]):
        out[name] = np.array(scal[3 + 2 * q: 5 + 2 * q])
    return out


def dkt_dims():
    nx, ny = ctypes.c_int(), ctypes.c_int()
    lib().oracle_dkt_dims(ctypes.byref(nx), ctypes.byref(ny))
    return nx.value, ny.value


def dkt_run(nsteps: int, with_f: bool = False):
    """The restatement of CaseIV_DKT.F90 after nsteps steps: a dict of the
    disc state (see DKT_SCALARS) and the fields rho, u, v, obst (ny, nx) and,
    with with_f, f (ny, nx, 9).  Raises on a stop of the reference's."""
    L = lib()
    if not getattr(L, "_dkt_ready", False):
        dp = ctypes.POINTER(ctypes.c_double)
        L.oracle_dkt_run.argtypes = [ctypes.c_int, dp, dp, dp, dp, dp, dp]
        L.oracle_dkt_run.restype = ctypes.c_int
        L._dkt_ready = True
    nx, ny = dkt_dims()
    scal = np.zeros(19)
    rho, u, v, ob = (np.zeros((ny, nx)) for _ in range(4))
    f = np.zeros((ny, nx, 9)) if with_f else None
    rc = L.oracle_dkt_run(int(nsteps), _dp(scal), _dp(rho), _dp(u), _dp(v), _dp(ob), _dp(f) if with_f else None)
    if rc != 0:
        raise RuntimeError(f"oracle_dkt_run stopped with code {rc}")
    out = _dkt_unpack(scal)
    out.update(rho=rho, u=u, v=v, obst=ob)
    if with_f:
        out["f"] = f
    return out


def run_reference_dkt(nsteps: int):
    """Run oracle/_ref/ref_dkt (the reference's own CaseIV_DKT.F90
    subroutines) for nsteps steps in a scratch directory and read its state
    in the layout of dkt_run(with_f=True)."""
    import tempfile
    nx, ny = dkt_dims()
    with tempfile.TemporaryDirectory() as d:
        subprocess.run([REF_DKT, str(int(nsteps))], cwd=d, check=True, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL)
        raw = np.fromfile(os.path.join(d, "dkt_state.bin"), dtype=np.float64)
    n = nx * ny
    assert raw.size == 19 + 4 * n + 9 * n, raw.size
    out = _dkt_unpack(raw[:19])
    body = raw[19:]
    for q, name in enumerate(("rho", "u", "v", "obst")):
        out[name] = body[q * n:(q + 1) * n].reshape(ny, nx)
    out["f"] = body[4 * n:].reshape(ny, nx, 9)
    return out
