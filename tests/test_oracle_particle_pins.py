"""Pins of the moving-sphere oracle's three-dimensional assembly (the parts
with no reference output: the reference's particle scheme is two-dimensional,
MPI/Micro_particles/fortran/case4/mpi), each by what mathematics or physics
fixes independently of the oracle's code:

  body kinematics   U' = U + F/m, x' = x + U + F/(2m), Omega' = Omega + T/I
                    with m = rho_s 4/3 pi r^3 and I = 2/5 m r^2, the solid
                    sphere's mass and moment of inertia (the 3-D form of
                    particle_update.F90:26-42);
  buoyancy          a sphere at rest in fluid at rest feels only
                    -(rho_s - rho_avg) 4/3 pi r^3 g along z (particle_force.F90:191;
                    the momentum exchange of a closed lattice surface in a
                    uniform rest state sums to zero);
  wall springs      near each of the six walls, (rho_s - rho0) V g / stiff
                    ((d - t) / t)^2 pushing away from that wall, d the gap
                    d = x - r - 1 to a low wall and N - x - r to a high wall
                    (particle_force.F90:126-186, which has the low walls and
                    the high x wall);
  refill            every node a sphere uncovers gets rho by 3-point
                    extrapolation along the lattice direction of largest
                    projection on the outward normal, and momentum
                    rho_avg' (U + Omega x r) (particle_update.F90:110-181);
  spin-down         a heavy sphere spinning in a closed box decays at
                    15 mu / (rho_s r^2) / (1 - (r/R)^3), Stokes' torque
                    -8 pi mu r^3 Omega in a concentric spherical container of
                    radius R (Lamb) over I = 2/5 m r^2;
  settling          the terminal velocity lies between Stokes' velocity
                    corrected for the concentric spherical containers that fit
                    inside and around the box (Happel & Brenner's wall factor;
                    Stokes drag grows as the container shrinks, so the box's
                    drag lies between the two).

test_mutations_are_caught builds copies of the oracle with one of these
formulas altered and shows that the named checks fail on each.
"""
import concurrent.futures
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

EX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
EY = np.array([0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1])
EZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1])


def sphere_mass(rho_s, r):
    return rho_s * 4.0 / 3.0 * math.pi * r ** 3


def nu_of(tau):
    return (tau - 0.5) / 3.0


# ---- checks: each takes (O, L): the oracle module and a library (None = the
# oracle as built) and raises AssertionError when the pin does not hold ------
def check_kinematics(O, L):
    """One step of a spinning, moving sphere: the new velocity, centre and
    angular velocity follow from the step's force and torque through the
    solid sphere's m and I."""
    n = (22, 20, 24)
    xc, r = np.array([[11.3, 10.2, 12.6]]), np.array([4.4])
    rho_s = 1.7
    U0, Om0 = np.array([[0.02, -0.015, 0.01]]), np.array([[0.003, -0.002, 0.004]])
    p = O.particle_params(*n, 1, 0.8, rho_solid=rho_s, gravity=3e-5)
    a = O.particle_run(p, 5, xc, r, U=U0, Om=Om0, L=L)
    b = O.particle_run(p, 6, xc, r, U=U0, Om=Om0, L=L)
    assert a["rc"] == 0 and b["rc"] == 0
    m = sphere_mass(rho_s, r[0])
    I = 0.4 * m * r[0] ** 2
    dU, dOm = b["U"] - a["U"], b["Om"] - a["Om"]
    assert np.abs(dU * m - b["F"]).max() <= 1e-12 * np.abs(b["F"]).max()
    assert np.abs(dOm * I - b["T"]).max() <= 1e-12 * np.abs(b["T"]).max()
    assert np.abs(b["xc"] - (a["xc"] + a["U"] + 0.5 * dU)).max() <= 1e-13


def check_buoyancy(O, L):
    """First step from rest, far from every wall: the force is the buoyancy
    of the sphere's volume alone."""
    n = (24, 24, 24)
    xc, r = np.array([[12.4, 12.7, 12.2]]), np.array([4.6])
    rho_s, g = 1.5, 1e-4
    out = O.particle_run(O.particle_params(*n, 1, 0.8, rho_solid=rho_s, gravity=g), 1, xc, r, L=L)
    assert out["rc"] == 0
    want = -(rho_s - out["rho_avg"]) * 4.0 / 3.0 * math.pi * r[0] ** 3 * g
    assert abs(out["rho_avg"] - 1.0) < 1e-15
    assert abs(out["F"][0, 2] - want) <= 1e-12 * abs(want)
    assert np.abs(out["F"][0, :2]).max() <= 1e-12 * abs(want)


def check_wall_springs(O, L):
    """A sphere at rest 1.7 lattice units from each wall in turn (inside the
    spring threshold 4): after one step the force is the buoyancy plus one
    spring along that wall's axis, pointing into the box."""
    N = 26
    r, rho_s, g, stiff, thr, gap = 4.2, 1.5, 1e-4, 0.1, 4.0, 1.7
    V = 4.0 / 3.0 * math.pi * r ** 3
    spring = (rho_s - 1.0) * V * g / stiff * ((gap - thr) / thr) ** 2
    buoy = -(rho_s - 1.0) * V * g
    mid = N / 2.0 + 0.3
    for axis in range(3):
        for high in (False, True):
            c = [mid, mid - 0.2, mid + 0.1]
            c[axis] = N - r - gap if high else 1.0 + r + gap
            p = O.particle_params(N, N, N, 1, 0.8, rho_solid=rho_s, gravity=g, threshold_wall=thr, stiff_wall=stiff)
            out = O.particle_run(p, 1, np.array([c]), np.array([r]), L=L)
            assert out["rc"] == 0
            want = np.zeros(3)
            want[2] += buoy
            want[axis] += -spring if high else spring
            assert np.abs(out["F"][0] - want).max() <= 1e-12 * spring, (axis, high, out["F"][0], want)


def check_refill(O, L):
    """Every node a moving, spinning sphere uncovers over 14 steps: density
    from 3 f1 - 3 f2 + f3 along the direction of largest e . (x - x_c) (a
    stencil point that is not fluid in both maps counts as rest density 1),
    momentum rho_avg' (U + Omega x (x - x_c)), with the new centre and
    velocities; rho_avg' is the mean density over the new map's fluid nodes
    at refill time, when uncovered nodes still hold rho_s."""
    n = (22, 20, 24)
    xc0, r = np.array([[11.3, 10.2, 12.6]]), np.array([4.4])
    U0, Om0 = np.array([[0.05, -0.035, 0.025]]), np.array([[0.004, -0.006, 0.008]])
    rho_s = 1.5
    p = O.particle_params(*n, 1, 0.8, rho_solid=rho_s, gravity=0.0)
    prev = O.particle_run(p, 0, xc0, r, U=U0, Om=Om0, L=L)
    checked = 0
    for k in range(1, 15):
        cur = O.particle_run(p, k, xc0, r, U=U0, Om=Om0, L=L)
        assert cur["rc"] == 0
        uncovered = (prev["owner"] >= 0) & (cur["owner"] < 0)
        fluid_new = cur["owner"] < 0
        at_refill = np.where(uncovered, rho_s, cur["rho"])
        rho_avg2 = at_refill[fluid_new].mean()
        for kk, jj, ii in zip(*np.nonzero(uncovered)):
            x = np.array([ii + 1.0, jj + 1.0, kk + 1.0])
            rv = x - cur["xc"][0]
            ec = 1 + int(np.argmax((rv[0] * EX + rv[1] * EY + rv[2] * EZ)[1:]))

            def rho_at(m):
                i2, j2, k2 = ii + m * EX[ec], jj + m * EY[ec], kk + m * EZ[ec]
                if not (0 <= i2 < n[0] and 0 <= j2 < n[1] and 0 <= k2 < n[2]):
                    return 1.0
                if prev["owner"][k2, j2, i2] >= 0 or cur["owner"][k2, j2, i2] >= 0:
                    return 1.0
                return cur["rho"][k2, j2, i2]

            rho = cur["rho"][kk, jj, ii]
            assert abs(rho - (3 * rho_at(1) - 3 * rho_at(2) + rho_at(3))) <= 1e-13
            j = rho * np.array([cur["u"][kk, jj, ii], cur["v"][kk, jj, ii], cur["w"][kk, jj, ii]])
            want = rho_avg2 * (cur["U"][0] + np.cross(cur["Om"][0], rv))
            assert np.abs(j - want).max() <= 1e-12 * np.abs(want).max()
            checked += 1
        prev = cur
    assert checked >= 20


def container_radii(c, N):
    """Radii of the largest sphere about c inside the box of nodes 1..N
    (halfway walls at 0.5 and N + 0.5) and of the smallest one around it."""
    c = np.asarray(c, dtype=float)
    lo, hi = c - 0.5, N + 0.5 - c
    return float(np.minimum(lo, hi).min()), float(np.sqrt((np.maximum(lo, hi) ** 2).sum()))


def check_spin_down(O, L):
    """A sphere of radius 5 and density 200 spinning about z in a closed
    32^3 box: the decay rate of Omega between steps 300 and 1000 against
    15 mu / (rho_s r^2) / (1 - (r/R)^3) for R between the inscribed and
    circumscribed containers, with 8 % allowed for the lattice sphere (a
    lattice sphere of radius 5 turns about 4 % slower than Stokes' sphere).
    I = 1/2 m r^2 would decay 20 % slower."""
    N, r, tau, rho_s = 32, 5.0, 1.1, 200.0
    xc = np.array([[16.63, 16.29, 16.57]])
    Om0 = np.array([[0.0, 0.0, 1e-3]])
    p = O.particle_params(N, N, N, 1, tau, rho_solid=rho_s, gravity=0.0)
    a = O.particle_run(p, 300, xc, np.array([r]), Om=Om0, L=L)
    b = O.particle_run(p, 1000, xc, np.array([r]), Om=Om0, L=L)
    assert a["rc"] == 0 and b["rc"] == 0
    rate = math.log(a["Om"][0, 2] / b["Om"][0, 2]) / 700.0
    mu = nu_of(tau)
    r_in, r_out = container_radii(xc[0], N)
    slow = 15.0 * mu / (rho_s * r * r) / (1.0 - (r / r_out) ** 3)
    fast = 15.0 * mu / (rho_s * r * r) / (1.0 - (r / r_in) ** 3)
    assert 0.92 * slow <= rate <= 1.05 * fast, (rate, slow, fast)


def wall_factor(lam):
    """Drag of a sphere translating at the centre of a spherical container,
    over Stokes' drag, lam = r / R (Happel & Brenner, Low Reynolds Number
    Hydrodynamics, 4-22)."""
    return (1.0 - lam ** 5) / (1.0 - 2.25 * lam + 2.5 * lam ** 3 - 2.25 * lam ** 5 + lam ** 6)


def check_settling(O, L):
    """A sphere of radius 4, rho_s 1.5, settling under g = 1e-4 in a closed
    32^3 box (Re = 0.03): its velocity after 600 steps (terminal to 1 %)
    against Stokes' (rho_s - rho_f) V g / (6 pi mu r) divided by the wall
    factors of the containers inside and around the box along the path,
    5 % allowed for the lattice sphere."""
    N, r, tau, rho_s, g = 32, 4.0, 1.1, 1.5, 1e-4
    xc = np.array([[16.63, 16.29, 16.87]])
    out = O.particle_run(O.particle_params(N, N, N, 1, tau, rho_solid=rho_s, gravity=g), 600, xc, np.array([r]), L=L)
    assert out["rc"] == 0
    mu = nu_of(tau)
    stokes = (rho_s - 1.0) * 4.0 / 3.0 * math.pi * r ** 3 * g / (6.0 * math.pi * mu * r)
    ins = min(container_radii(xc[0], N)[0], container_radii(out["xc"][0], N)[0])
    outs = max(container_radii(xc[0], N)[1], container_radii(out["xc"][0], N)[1])
    u = -out["U"][0, 2]
    assert 0.95 * stokes / wall_factor(r / ins) <= u <= 1.05 * stokes / wall_factor(r / outs), \
        (u, stokes / wall_factor(r / ins), stokes / wall_factor(r / outs))


CHECKS = {"kinematics": check_kinematics, "buoyancy": check_buoyancy, "wall_springs": check_wall_springs,
          "refill": check_refill, "spin_down": check_spin_down, "settling": check_settling}


@pytest.mark.parametrize("name", list(CHECKS))
def test_pin(oracle_lib, name):
    CHECKS[name](oracle_lib, None)


# ---- the same checks against copies of the oracle with one formula altered --
# (substring of particle_oracle.c, its replacement, the checks that must fail)
MUTATIONS = {
    "inertia_half": ("I = 0.4 * m", "I = 0.5 * m", ["kinematics", "spin_down"]),
    "mass_excess": ("m = p->rho_solid * V", "m = (p->rho_solid - p->rho0) * V", ["kinematics"]),
    "buoyancy_volume": ("- (p->rho_solid - s->rho_avg) * V * p->gravity",
                        "- (p->rho_solid - s->rho_avg) * (V * 0.75) * p->gravity", ["buoyancy", "settling"]),
    "wall_high_sign": ("fw[d] -= ps_spring(scale_w, hi", "fw[d] += ps_spring(scale_w, hi", ["wall_springs"]),
    "wall_low_place": ("- s->r[a] - 1.0;", "- s->r[a] - 0.5;", ["wall_springs"]),
    "refill_cosine": ("double tn = (rx * EX[a] + ry * EY[a] + rz * EZ[a]) / rn;",
                      "double tn = (rx * EX[a] + ry * EY[a] + rz * EZ[a]) / rn / "
                      "sqrt((double)(EX[a] * EX[a] + EY[a] * EY[a] + EZ[a] * EZ[a]));", ["refill"]),
    "refill_old_mean": ("m[3] = rho_avg2 *", "m[3] = s->rho_avg *", ["refill"]),
    "refill_no_spin": ("m[7] = rho_avg2 * (ub[2] + wb[2]);", "m[7] = rho_avg2 * ub[2];", ["refill"]),
}


def _build_mutant(src_dir, work, name, old, new):
    d = os.path.join(work, name)
    os.makedirs(d)
    for f in os.listdir(src_dir):
        if f.endswith((".c", ".h")):
            shutil.copy(os.path.join(src_dir, f), d)
    path = os.path.join(d, "particle_oracle.c")
    text = open(path).read()
    assert text.count(old) == 1, f"{name}: '{old}' occurs {text.count(old)} times"
    open(path, "w").write(text.replace(old, new))
    lib = os.path.join(d, "liboracle_mutant.so")
    srcs = sorted(os.path.join(d, f) for f in os.listdir(d) if f.endswith(".c"))
    subprocess.run(["gcc", "-O2", "-fPIC", "-fopenmp", "-std=c11", "-shared", "-o", lib, *srcs, "-lm"], check=True)
    return lib


def test_mutations_are_caught(oracle_lib, tmp_path):
    O = oracle_lib
    src = os.path.dirname(O.LIB_PATH)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        libs = {name: ex.submit(_build_mutant, src, str(tmp_path), name, old, new)
                for name, (old, new, _) in MUTATIONS.items()}
        libs = {name: fut.result() for name, fut in libs.items()}
    missed = []
    for name, (_, _, checks) in MUTATIONS.items():
        L = O.particle_library(libs[name])
        O.set_threads(os.cpu_count() or 1, L)
        for c in checks:
            try:
                CHECKS[c](O, L)
                missed.append(f"{name}: {c}")
            except AssertionError:
                pass
    assert not missed, f"mutations the checks did not catch: {missed}"
