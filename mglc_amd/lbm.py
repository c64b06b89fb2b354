"""Python binding of libmglc_lbm.so (the C-ABI of include/mglc_lbm.h).

The reference drives its step as Fortran subroutines on module globals
(MPI/Lid_driven_cavity/fortran/3d/mpi_3d_blocked/main.f90:75-103):
``initial``, then ``collision / streaming / bounceback / macro`` per step,
``check`` every 2000 steps and ``output`` at the end.  :class:`Cavity` exposes
the same operations -- ``initial()``, ``step(n)``, ``check()``,
``macro()`` -- backed only by the HIP library.  There is no CPU fallback: if
the library is missing or fails to load, every call raises.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmglc_lbm.so")

LBM_MRT, LBM_BGK = 0, 1
STORE_F64, STORE_F32, STORE_F16 = 0, 1, 2
STORAGE_NAMES = {"f64": STORE_F64, "f32": STORE_F32, "f16": STORE_F16}
COLLISION_NAMES = {"mrt": LBM_MRT, "bgk": LBM_BGK}
TRANSFORM_NAMES = {"auto": 0, "valu": 1, "mfma": 2}

LBM_OK, LBM_ERR_NAN = 0, 3

# Every entry point include/mglc_lbm.h declares.
EXPORTED_SYMBOLS = (
    "lbm_create", "lbm_destroy", "lbm_set_stream", "lbm_initial", "lbm_load_fields", "lbm_step",
    "lbm_check", "lbm_download_macro", "lbm_synchronize", "lbm_get_info", "lbm_status_string",
    "lbm_last_error", "lbm_check_sums", "lbm_halo_layout", "lbm_halo_buffer_bytes", "lbm_halo_pack",
    "lbm_halo_unpack", "lbm_check_thermal", "lbm_nusselt", "lbm_download_temperature",
    "lbm_batch_create", "lbm_batch_destroy", "lbm_batch_initial", "lbm_batch_step", "lbm_batch_step_each",
    "lbm_batch_autotune", "lbm_batch_set_launch", "lbm_batch_synchronize", "lbm_batch_download_macro",
    "lbm_batch_get_info", "lbm_particles_init", "lbm_particles_get", "lbm_particles_map",
    "lbm_state_bytes", "lbm_save_state", "lbm_load_state", "lbm_reynolds", "lbm_bandwidth_probe",
    "lbm_step_begin", "lbm_step_end", "lbm_check_sums_noalloc", "lbm_batch_set_cells_per_thread",
)

FACE_XM, FACE_XP, FACE_YM, FACE_YP, FACE_ZM, FACE_ZP = range(6)
NUM_REGIONS = 18


class LbmParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("collision", ctypes.c_int), ("storage", ctypes.c_int), ("meq12_rho", ctypes.c_int),
        ("f16_scale_log2", ctypes.c_int), ("thermal", ctypes.c_int),
        ("block_x", ctypes.c_int), ("block_y", ctypes.c_int),
        ("halo_faces", ctypes.c_int), ("halo_loopback", ctypes.c_int), ("thermal_axis", ctypes.c_int),
        ("halo_overlap", ctypes.c_int), ("waves_per_cu", ctypes.c_int), ("transform", ctypes.c_int),
        ("tau", ctypes.c_double), ("u_lid", ctypes.c_double), ("rho0", ctypes.c_double),
        ("para_a", ctypes.c_double), ("g_beta", ctypes.c_double), ("omega_rot", ctypes.c_double),
        ("t_hot", ctypes.c_double), ("t_cold", ctypes.c_double), ("t_ref", ctypes.c_double),
        ("reserved_d", ctypes.c_double * 2),
    ]


class LbmInfo(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_longlong), ("ny", ctypes.c_longlong), ("nz", ctypes.c_longlong),
        ("pitch_x", ctypes.c_longlong), ("pop_stride", ctypes.c_longlong),
        ("steps_done", ctypes.c_longlong), ("device_bytes", ctypes.c_longlong),
        ("bytes_per_lup", ctypes.c_double), ("storage", ctypes.c_int), ("collision", ctypes.c_int),
    ]


class LbmBatchInfo(ctypes.Structure):
    _fields_ = [
        ("instances", ctypes.c_longlong), ("total_cells", ctypes.c_longlong), ("total_tiles", ctypes.c_longlong),
        ("steps_done", ctypes.c_longlong), ("device_bytes", ctypes.c_longlong), ("bytes_per_lup", ctypes.c_double),
        ("threads", ctypes.c_int), ("blocks", ctypes.c_int), ("graph_steps", ctypes.c_int),
        ("cells_per_thread", ctypes.c_int), ("stream", ctypes.c_void_p),
    ]


class LbmParticleParams(ctypes.Structure):
    _fields_ = [
        ("np", ctypes.c_int), ("reserved_i", ctypes.c_int),
        ("rho_solid", ctypes.c_double), ("gravity", ctypes.c_double),
        ("threshold_wall", ctypes.c_double), ("stiff_wall", ctypes.c_double),
        ("threshold_particle", ctypes.c_double), ("stiff_particle", ctypes.c_double),
        ("reserved_d", ctypes.c_double * 4),
    ]


class LbmError(RuntimeError):
    def __init__(self, status: int, where: str, message: str):
        super().__init__(f"{where}: status {status}: {message}")
        self.status = status


_lib = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load the HIP library.  Raises if it is not built -- no fallback."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libmglc_lbm.so is not built ({path}); run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int
    lib.lbm_create.argtypes = [ctypes.POINTER(LbmParams), ctypes.POINTER(vp)]
    lib.lbm_destroy.argtypes = [vp]
    lib.lbm_set_stream.argtypes = [vp, vp]
    lib.lbm_initial.argtypes = [vp]
    lib.lbm_load_fields.argtypes = [vp, dp, dp, dp, dp, dp]
    lib.lbm_check_thermal.argtypes = [vp, dp, dp]
    lib.lbm_nusselt.argtypes = [vp, dp, dp]
    lib.lbm_download_temperature.argtypes = [vp, dp]
    lib.lbm_step.argtypes = [vp, ip]
    lib.lbm_step_begin.argtypes = [vp]
    lib.lbm_step_end.argtypes = [vp]
    lib.lbm_check.argtypes = [vp, dp]
    lib.lbm_download_macro.argtypes = [vp, dp, dp, dp, dp]
    lib.lbm_synchronize.argtypes = [vp]
    lib.lbm_get_info.argtypes = [vp, ctypes.POINTER(LbmInfo)]
    lib.lbm_check_sums.argtypes = [vp, dp, dp]
    lib.lbm_check_sums_noalloc.argtypes = [vp, vp, vp, vp, vp, ctypes.c_longlong, vp, vp]
    lib.lbm_halo_layout.argtypes = [vp, ip, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]
    lib.lbm_halo_buffer_bytes.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t)]
    lib.lbm_halo_pack.argtypes = [vp, vp]
    lib.lbm_halo_unpack.argtypes = [vp, vp]
    lib.lbm_batch_create.argtypes = [ctypes.POINTER(LbmParams), ip, ctypes.POINTER(vp)]
    lib.lbm_batch_destroy.argtypes = [vp]
    lib.lbm_batch_initial.argtypes = [vp]
    lib.lbm_batch_step.argtypes = [vp, ip]
    lib.lbm_batch_step_each.argtypes = [vp, ip]
    lib.lbm_batch_autotune.argtypes = [vp, ip, ctypes.POINTER(ip), ctypes.POINTER(ip)]
    lib.lbm_batch_set_launch.argtypes = [vp, ip, ip]
    lib.lbm_batch_set_cells_per_thread.argtypes = [vp, ip]
    lib.lbm_batch_synchronize.argtypes = [vp]
    lib.lbm_batch_download_macro.argtypes = [vp, ip, dp, dp, dp, dp]
    lib.lbm_batch_get_info.argtypes = [vp, ctypes.POINTER(LbmBatchInfo)]
    lib.lbm_particles_init.argtypes = [vp, ctypes.POINTER(LbmParticleParams), dp, dp, dp, dp]
    lib.lbm_particles_get.argtypes = [vp, dp, dp, dp, dp, dp]
    lib.lbm_particles_map.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    lib.lbm_reynolds.argtypes = [vp, dp]
    lib.lbm_state_bytes.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t)]
    lib.lbm_save_state.argtypes = [vp, vp, ctypes.c_size_t]
    lib.lbm_load_state.argtypes = [vp, vp, ctypes.c_size_t]
    lib.lbm_bandwidth_probe.argtypes = [ip, ip, ip, ip, dp]
    lib.lbm_status_string.argtypes = [ip]
    lib.lbm_status_string.restype = ctypes.c_char_p
    lib.lbm_last_error.restype = ctypes.c_char_p
    for name in EXPORTED_SYMBOLS:
        if name not in ("lbm_status_string", "lbm_last_error"):
            getattr(lib, name).restype = ctypes.c_int
    _lib = lib
    return lib


def bandwidth_probe(pattern: int, storage: str = "f64", n: int = 256, reps: int = 20) -> float:
    """GB/s (read + write) the device delivers for the step's traffic with no
    arithmetic: pattern 0 = one-stream copy, 1 = the D3Q19 pull of an n^3
    lattice in the library's layout (lbm_bandwidth_probe)."""
    lib = load_library()
    g = ctypes.c_double()
    st = {"f64": STORE_F64, "f32": STORE_F32, "f16": STORE_F16}[storage]
    rc = lib.lbm_bandwidth_probe(int(pattern), st, int(n), int(reps), ctypes.byref(g))
    if rc != 0:
        raise LbmError(rc, "lbm_bandwidth_probe", lib.lbm_last_error().decode())
    return g.value


def _dp(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@dataclass
class CavityConfig:
    """Parameters named as in the reference's module commondata
    (mpi_3d_blocked/commondata.f90:4-42)."""
    nx: int
    ny: int
    nz: int
    Reynolds: float = 1000.0
    U0: float = 0.1
    rho0: float = 1.0
    collision: str = "mrt"
    storage: str = "f64"
    tauf: float | None = None      # default: commondata.f90:9, 3 U0 N / Re + 0.5 with N = nx
    meq12_rho: int = 0
    f16_scale_log2: int = 0
    block_x: int = 0
    block_y: int = 0
    halo_faces: int = 0        # bit f: face f reads a halo (FACE_XM..FACE_ZP) instead of a wall
    halo_loopback: int = 0     # 1: exchange looped back onto this subdomain before every step
    halo_overlap: int = 0      # 1: boundary shell + interior per step, exchange overlapped with the interior
    waves_per_cu: int = 0      # occupancy cap of the step kernels (0 = the library's choice)
    transform: str = "auto"    # FP16 MRT moment transforms: "auto", "valu" or "mfma"

    def tau(self) -> float:
        if self.tauf is not None:
            return float(self.tauf)
        return self.U0 * float(self.nx) / self.Reynolds * 3.0 + 0.5

    def params(self) -> LbmParams:
        p = LbmParams()
        p.nx, p.ny, p.nz = self.nx, self.ny, self.nz
        p.collision = COLLISION_NAMES[self.collision]
        p.storage = STORAGE_NAMES[self.storage]
        p.meq12_rho = self.meq12_rho
        p.f16_scale_log2 = self.f16_scale_log2
        p.block_x, p.block_y = self.block_x, self.block_y
        p.halo_faces, p.halo_loopback, p.halo_overlap = self.halo_faces, self.halo_loopback, self.halo_overlap
        p.waves_per_cu = self.waves_per_cu
        p.transform = TRANSFORM_NAMES[self.transform]
        p.tau, p.u_lid, p.rho0 = self.tau(), self.U0, self.rho0
        return p


def splitmix_uniform(shape, seed: int, amp: float = 1e-4):
    """Uniform numbers in [-amp, amp) from SplitMix64 of the linear cell index
    (x fastest) plus `seed`: the seeded noise of the buoyancy case's initial
    temperature, reproducible on any host (mglc_amd/fortran/buoyancy_data.f90 uses
    the same generator)."""
    n = int(np.prod(shape))
    with np.errstate(over="ignore"):
        z = (np.arange(n, dtype=np.uint64) + np.uint64(seed)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u01 = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return (amp * (2.0 * u01 - 1.0)).reshape(shape)


def buoyancy_initial_temperature(nx, ny, nz, t_hot=0.5, t_cold=-0.5, seed=2024, amp=1e-4):
    """Initial T of the buoyancy case: the reference's held layers next to the
    y walls (initial.F90:126-144) and T = 0 plus seeded uniform noise in
    [-amp, amp) on every other cell."""
    T = np.zeros((nz, ny, nx))
    T[:, 1:-1, :] = splitmix_uniform((nz, ny - 2, nx), seed, amp)
    T[:, 0, :] = t_hot
    T[:, -1, :] = t_cold
    return T


def random_spheres(nx: int, ny: int, nz: int, count: int = 2000, rmin: float = 4.0, rmax: float = 8.0,
                   seed: int = 2025, gap: float = 1.0, min_gap: float | None = 2.5, max_tries: int = 200000,
                   relax_iters: int = 5000):
    """Seeded random non-overlapping spheres in an nx x ny x nz box.

    Radii are uniform in [rmin, rmax]; random sequential addition places them
    largest first at uniform random centres (1-based node coordinates)
    keeping `gap` lattice units from every other sphere and from the walls
    (the low wall of axis d lies at coordinate 1, the high one at N, as
    particle_force.F90:128-186 places them).  With min_gap set, a
    deterministic relaxation then pushes overlapping pairs (and spheres too
    close to a wall) apart until every gap is at least min_gap, so that the
    spring forces of the particle scheme (active below 4 lattice units) start
    weak.  Returns (centres (count, 3), radii (count,))."""
    rng = np.random.default_rng(seed)
    radii = np.sort(rng.uniform(rmin, rmax, count))[::-1].copy()
    N = np.array([nx, ny, nz], dtype=np.float64)
    cs = 2.0 * rmax + gap  # grid cell: no two touching candidates are two cells apart
    ncell = np.maximum((N // cs).astype(int), 1)
    grid: dict = {}
    xc = np.zeros((count, 3))
    offsets = [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1)]
    for q, r in enumerate(radii):
        lo, hi = 1.0 + r + gap, N - r - gap
        if np.any(hi <= lo):
            raise ValueError(f"sphere of radius {r} does not fit the box")
        for _ in range(max_tries):
            p = lo + rng.random(3) * (hi - lo)
            cell = tuple(int(v) for v in np.minimum((p // cs).astype(int), ncell - 1))
            clear = True
            for o3 in offsets:
                for o in grid.get((cell[0] + o3[0], cell[1] + o3[1], cell[2] + o3[2]), ()):
                    d = xc[o] - p
                    if d @ d < (r + radii[o] + gap) ** 2:
                        clear = False
                        break
                if not clear:
                    break
            if clear:
                xc[q] = p
                grid.setdefault(cell, []).append(q)
                break
        else:
            raise RuntimeError(f"could not place sphere {q} of {count} (packing too dense)")
    if min_gap is not None and min_gap > gap:
        xc = _relax_spheres(xc, radii, N, min_gap, relax_iters)
    return xc, radii


def jittered_grid_spheres(nx: int, ny: int, nz: int, count: int, radius: float = 10.0, pitch: float = 50.0,
                          jitter: float = 20.0, seed: int = 1):
    """Centres on a grid of `pitch` spacing, each moved by a uniform offset in
    [-jitter/2, jitter/2) per axis, as the reference places its discs
    (mpi_particle/initial.F90:48-84: start at pitch/2 = 25, step 50 along x,
    and when x passes nx - 1 (200 for its 201-wide box) restart at 25 one
    row up; offsets (rand - 0.5) * 20).  Carried to 3-D, a full row moves y
    and a full layer moves z.  The offsets come from a seeded NumPy generator,
    not the compiler's random_number (whose stream is its own), so the
    layout, not the exact centres, is the reference's.  All spheres get
    `radius` (radius = radius0 in the reference).  Returns (centres (count,
    3), radii (count,)) in 1-based node coordinates."""
    rng = np.random.default_rng(seed)
    start = pitch / 2.0
    x, y, z = start, start, start
    xc = np.zeros((count, 3))
    for q in range(count):
        off = (rng.random(3) - 0.5) * jitter
        xc[q] = (x + off[0], y + off[1], z + off[2])
        x += pitch
        if x > nx - 1:
            x = start
            y += pitch
            if y > ny - 1:
                y = start
                z += pitch
    if np.any(xc[:, 2] + radius > nz):
        raise ValueError(f"{count} spheres at pitch {pitch} do not fit a {nx} x {ny} x {nz} box")
    return xc, np.full(count, float(radius))


def _relax_spheres(xc, r, N, min_gap, iters, step=0.6):
    """Push pairs closer than min_gap (and spheres closer than min_gap to a
    wall) apart, each by a fraction of its overlap, until none is left.
    Pairs are visited in sorted order, so the result is reproducible."""
    from scipy.spatial import cKDTree

    slack = 0.05  # pushes aim this far beyond min_gap, so the iteration ends in finitely many steps
    target = min_gap + slack
    xc = xc.copy()
    for _ in range(iters):
        pairs = cKDTree(xc).query_pairs(2.0 * r.max() + target, output_type="ndarray")
        pairs = pairs[np.lexsort((pairs[:, 1], pairs[:, 0]))] if len(pairs) else pairs.reshape(0, 2)
        i, j = pairs[:, 0], pairs[:, 1]
        d = xc[j] - xc[i]
        dist = np.sqrt((d * d).sum(axis=1))
        over = r[i] + r[j] + target - dist
        hit = over > 0
        disp = np.zeros_like(xc)
        if hit.any():
            push = d[hit] / dist[hit, None] * (0.5 * step * over[hit])[:, None]
            np.add.at(disp, i[hit], -push)
            np.add.at(disp, j[hit], push)
        lo = (1.0 + r[:, None] + target) - xc
        hi = xc - (N - r[:, None] - target)
        disp += step * (np.maximum(lo, 0.0) - np.maximum(hi, 0.0))
        if (not hit.any() or over[hit].max() <= slack) and max(lo.max(), hi.max()) <= slack:
            return xc
        xc += disp
    raise RuntimeError(f"could not relax the packing to gaps of {min_gap}")


@dataclass
class ThermalConfig:
    """Buoyancy-driven cavity, parameters named as in the reference's module
    (buoyancy .../3d/mpi_blocked/module.F90:11-33); derived constants follow
    its parameter statements.  thermal_axis: 1 = y walls held (LeftRightWallsConstT)."""
    nx: int
    ny: int
    nz: int
    Rayleigh: float = 1e5
    Prandtl: float = 0.71
    Mach: float = 0.1
    Thot: float = 0.5
    Tcold: float = -0.5
    Tref: float = 0.0
    Ekman: float | None = None     # None: no Coriolis force
    thermal_axis: int = 1
    N: int | None = None           # length used in the parameter statements (default nz)
    block_x: int = 0
    block_y: int = 0
    storage: str = "f64"           # f64, or f32 / f16: FP32 arithmetic on shifted FP32 / scaled FP16 populations
    waves_per_cu: int = 0
    halo_faces: int = 0            # as CavityConfig: faces bordering neighbour subdomains
    halo_loopback: int = 0
    halo_overlap: int = 0

    def derived(self):
        import math
        N = float(self.N if self.N is not None else self.nz)
        tauf = 0.5 + self.Mach * N * math.sqrt(3.0 * self.Prandtl / self.Rayleigh)
        viscosity = (tauf - 0.5) / 3.0
        diffusivity = viscosity / self.Prandtl
        omega = 0.0 if self.Ekman is None else viscosity / 2.0 / self.Ekman / (N * N)
        para_a = 42.0 * math.sqrt(3.0) * diffusivity - 6.0
        g_beta = self.Rayleigh * viscosity * diffusivity / N / (N * N)
        return dict(tau=tauf, viscosity=viscosity, diffusivity=diffusivity, omega_rot=omega, para_a=para_a,
                    g_beta=g_beta)

    def params(self) -> LbmParams:
        d = self.derived()
        p = LbmParams()
        p.nx, p.ny, p.nz = self.nx, self.ny, self.nz
        p.collision, p.thermal = LBM_MRT, 1
        p.storage = {"f64": STORE_F64, "f32": STORE_F32, "f16": STORE_F16}[self.storage]
        p.block_x, p.block_y = self.block_x, self.block_y
        p.waves_per_cu = self.waves_per_cu
        p.halo_faces, p.halo_loopback, p.halo_overlap = self.halo_faces, self.halo_loopback, self.halo_overlap
        p.thermal_axis = self.thermal_axis
        p.tau, p.u_lid, p.rho0 = d["tau"], 0.0, 1.0
        p.para_a, p.g_beta, p.omega_rot = d["para_a"], d["g_beta"], d["omega_rot"]
        p.t_hot, p.t_cold, p.t_ref = self.Thot, self.Tcold, self.Tref
        return p


class Cavity:
    """One GPU's lid-driven cavity.  Arrays are returned as (nz, ny, nx) C-order
    numpy arrays, i.e. the reference's Fortran (nx, ny, nz) layout."""

    def __init__(self, cfg, stream: int | None = None):
        """cfg: CavityConfig (lid-driven) or ThermalConfig (buoyancy-driven),
        or an LbmParams filled by the caller."""
        self.cfg = cfg
        self.lib = load_library()
        self._ctx = ctypes.c_void_p()
        p = cfg if isinstance(cfg, LbmParams) else cfg.params()
        self._call("lbm_create", ctypes.byref(p), ctypes.byref(self._ctx))
        if stream is not None:
            self._call("lbm_set_stream", self._ctx, ctypes.c_void_p(stream))

    def _call(self, name, *args):
        rc = getattr(self.lib, name)(*args)
        if rc != LBM_OK:
            raise LbmError(rc, name, self.lib.lbm_last_error().decode())
        return rc

    @property
    def shape(self):
        return (self.cfg.nz, self.cfg.ny, self.cfg.nx)

    def initial(self):
        self._call("lbm_initial", self._ctx)

    def load_fields(self, rho, u, v, w, T=None):
        arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(self.shape) for a in (rho, u, v, w)]
        t = None if T is None else np.ascontiguousarray(T, dtype=np.float64).reshape(self.shape)
        self._call("lbm_load_fields", self._ctx, *[_dp(a) for a in arrs], _dp(t))

    def step(self, nsteps: int = 1):
        self._call("lbm_step", self._ctx, int(nsteps))

    def step_begin(self):
        """Boundary shell of the next step (halo_overlap contexts); halo_pack
        now packs it, step_end completes the step."""
        self._call("lbm_step_begin", self._ctx)

    def step_end(self):
        self._call("lbm_step_end", self._ctx)

    def check(self) -> float:
        e = ctypes.c_double()
        self._call("lbm_check", self._ctx, ctypes.byref(e))
        return e.value

    def check_thermal(self):
        """(errorU, errorT) of the buoyancy check.F90."""
        eu, et = ctypes.c_double(), ctypes.c_double()
        self._call("lbm_check_thermal", self._ctx, ctypes.byref(eu), ctypes.byref(et))
        return eu.value, et.value

    def nusselt(self):
        """(nu_vol, nu_wall): calNuRe's <w T> nz / kappa + 1 (RaNu.F90:13-23) and
        the conductive flux through the hot wall."""
        a, b = ctypes.c_double(), ctypes.c_double()
        self._call("lbm_nusselt", self._ctx, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def reynolds(self) -> float:
        """Volume Reynolds number sqrt(<|u|^2>) nz / nu (calNuRe, RaNu.F90:36-44)."""
        r = ctypes.c_double()
        self._call("lbm_reynolds", self._ctx, ctypes.byref(r))
        return r.value

    def temperature(self):
        T = np.empty(self.shape, dtype=np.float64)
        self._call("lbm_download_temperature", self._ctx, _dp(T))
        return T

    def check_sums(self):
        """(e1, e2) of check.f90 on this subdomain, before the Allreduce."""
        e1, e2 = ctypes.c_double(), ctypes.c_double()
        self._call("lbm_check_sums", self._ctx, ctypes.byref(e1), ctypes.byref(e2))
        return e1.value, e2.value

    def halo_layout(self, region: int):
        off, cnt = ctypes.c_longlong(), ctypes.c_longlong()
        self._call("lbm_halo_layout", self._ctx, int(region), ctypes.byref(off), ctypes.byref(cnt))
        return off.value, cnt.value

    def halo_buffer_bytes(self) -> int:
        n = ctypes.c_size_t()
        self._call("lbm_halo_buffer_bytes", self._ctx, ctypes.byref(n))
        return n.value

    def halo_pack(self, dev_ptr: int):
        self._call("lbm_halo_pack", self._ctx, ctypes.c_void_p(dev_ptr))

    def halo_unpack(self, dev_ptr: int):
        self._call("lbm_halo_unpack", self._ctx, ctypes.c_void_p(dev_ptr))

    def macro(self):
        out = [np.empty(self.shape, dtype=np.float64) for _ in range(4)]
        self._call("lbm_download_macro", self._ctx, *[_dp(a) for a in out])
        return tuple(out)

    def synchronize(self):
        self._call("lbm_synchronize", self._ctx)

    def save_state(self) -> bytes:
        """The exact-restart state (lbm_save_state) as bytes."""
        n = ctypes.c_size_t()
        self._call("lbm_state_bytes", self._ctx, ctypes.byref(n))
        buf = ctypes.create_string_buffer(n.value)
        self._call("lbm_save_state", self._ctx, buf, n)
        return buf.raw

    def load_state(self, blob: bytes):
        buf = ctypes.create_string_buffer(bytes(blob), len(blob))
        self._call("lbm_load_state", self._ctx, buf, ctypes.c_size_t(len(blob)))

    def particles_init(self, centers, radii, rho_solid=1.5, gravity=1e-5, threshold_wall=4.0, stiff_wall=0.10,
                       threshold_particle=4.0, stiff_particle=0.10, U=None, omega=None):
        """Attach spheres (centres in 1-based node coordinates) to a fluid at rest."""
        xc = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1, 3)
        r = np.ascontiguousarray(radii, dtype=np.float64).reshape(-1)
        if r.shape[0] != xc.shape[0]:
            raise ValueError("one radius per centre")
        pp = LbmParticleParams(np=xc.shape[0], rho_solid=rho_solid, gravity=gravity, threshold_wall=threshold_wall,
                               stiff_wall=stiff_wall, threshold_particle=threshold_particle,
                               stiff_particle=stiff_particle)
        vel = None if U is None else np.ascontiguousarray(U, dtype=np.float64).reshape(xc.shape)
        om = None if omega is None else np.ascontiguousarray(omega, dtype=np.float64).reshape(xc.shape)
        self._np = xc.shape[0]
        self._call("lbm_particles_init", self._ctx, ctypes.byref(pp), _dp(xc), _dp(r), _dp(vel), _dp(om))

    def particles(self):
        """dict of np x 3 arrays: xc, U, Om, F (total force), T (hydrodynamic torque)."""
        out = {k: np.empty((self._np, 3)) for k in ("xc", "U", "Om", "F", "T")}
        self._call("lbm_particles_get", self._ctx, *[_dp(out[k]) for k in ("xc", "U", "Om", "F", "T")])
        return out

    def particle_map(self):
        owner = np.empty(self.shape, dtype=np.int32)
        self._call("lbm_particles_map", self._ctx, owner.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return owner

    def info(self) -> LbmInfo:
        inf = LbmInfo()
        self._call("lbm_get_info", self._ctx, ctypes.byref(inf))
        return inf

    def close(self):
        if self._ctx:
            self.lib.lbm_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_configs(count: int = 64, seed: int = 12345, sizes=(32, 48, 64, 80, 96), storage: str = "f32",
                  collision: str = "mrt"):
    """The batched-cavity case: `count` cavities with N drawn from `sizes`, Re
    uniform in [100, 1000] and U0 uniform in [0.05, 0.1], from a seeded numpy
    generator (PCG64)."""
    rng = np.random.default_rng(seed)
    ns = rng.choice(np.asarray(sizes), size=count)
    res = rng.uniform(100.0, 1000.0, size=count)
    us = rng.uniform(0.05, 0.1, size=count)
    return [CavityConfig(int(n), int(n), int(n), Reynolds=float(r), U0=float(u), storage=storage, collision=collision)
            for n, r, u in zip(ns, res, us)]


class Batch:
    """Independent lid-driven cavities stepped together (lbm_batch_*)."""

    def __init__(self, cfgs):
        self.cfgs = list(cfgs)
        self.lib = load_library()
        arr = (LbmParams * len(self.cfgs))(*[c.params() for c in self.cfgs])
        self._b = ctypes.c_void_p()
        self._call("lbm_batch_create", arr, len(self.cfgs), ctypes.byref(self._b))

    def _call(self, name, *args):
        rc = getattr(self.lib, name)(*args)
        if rc != LBM_OK:
            raise LbmError(rc, name, self.lib.lbm_last_error().decode())
        return rc

    def initial(self):
        self._call("lbm_batch_initial", self._b)

    def step(self, nsteps: int):
        self._call("lbm_batch_step", self._b, int(nsteps))

    def step_each(self, nsteps: int):
        self._call("lbm_batch_step_each", self._b, int(nsteps))

    def autotune(self, trial_steps: int = 4):
        t, b = ctypes.c_int(), ctypes.c_int()
        self._call("lbm_batch_autotune", self._b, int(trial_steps), ctypes.byref(t), ctypes.byref(b))
        return t.value, b.value

    def set_launch(self, threads: int, blocks: int):
        self._call("lbm_batch_set_launch", self._b, int(threads), int(blocks))

    def set_cells_per_thread(self, cpt: int):
        self._call("lbm_batch_set_cells_per_thread", self._b, int(cpt))

    def synchronize(self):
        self._call("lbm_batch_synchronize", self._b)

    def macro(self, m: int):
        c = self.cfgs[m]
        out = [np.empty((c.nz, c.ny, c.nx)) for _ in range(4)]
        self._call("lbm_batch_download_macro", self._b, int(m), *[_dp(a) for a in out])
        return tuple(out)

    def info(self) -> LbmBatchInfo:
        inf = LbmBatchInfo()
        self._call("lbm_batch_get_info", self._b, ctypes.byref(inf))
        return inf

    def close(self):
        if self._b:
            self.lib.lbm_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
