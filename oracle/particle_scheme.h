/* particle_scheme.h -- TEST INFRASTRUCTURE: the per-link formulas of the
 * reference's moving-particle scheme, written once and used by both oracles:
 * the two-dimensional restatement of the reference's own program
 * (particle2d_oracle.c, checked bit for bit against states that program
 * wrote) and the three-dimensional generalisation the GPU path is held to
 * (particle_oracle.c).  A disc of the two-dimensional program is a sphere at
 * z = 0 spinning about z; every function below then performs exactly the
 * reference's operations in the reference's order (x + 0.0 and 0.0 - x are
 * exact), so the two-dimensional bit-for-bit agreement pins these functions
 * for the three-dimensional oracle too.  Products are rounded one by one and
 * sums taken left to right, as the reference evaluates them.
 */
#ifndef MGLC_PARTICLE_SCHEME_H
#define MGLC_PARTICLE_SCHEME_H

#include <math.h>

/* The bisection's stop tolerance: `real(8), parameter :: epsRadius=1e-9` is a
 * default-real literal, so it holds single-precision 1e-9 widened to double
 * (case4/seq/CaseIV_DKT.F90:559, case4/mpi/particle_bounceback.F90:109). */
#define PS_CALQ_EPS ((double)1e-9f)

static inline double ps_dist2(double dx, double dy, double dz) { return (dx * dx + dy * dy) + dz * dz; }

/* calQ (CaseIV_DKT.F90:551-591, particle_bounceback.F90:101-141): wall
 * fraction q of the link from node X along e into the sphere (c, rad) by
 * bisection from q = 1/2 until |d - rad| < epsRadius; the last point goes to
 * p.  Returns 0, or 1 if q leaves [0, 1] (the reference's "error q!"). */
static inline int ps_calq(const double c[3], double rad, const double X[3], const int e[3], double p[3],
                          double *qout) {
    double q = 0.5, qt = 0.5;
    double px = X[0] + qt * (double)e[0], py = X[1] + qt * (double)e[1], pz = X[2] + qt * (double)e[2];
    for (int it = 0; it < 200; ++it) {
        const double d = sqrt(ps_dist2(px - c[0], py - c[1], pz - c[2]));
        if (!(fabs(d - rad) >= PS_CALQ_EPS)) break;
        qt = qt / 2.0;
        if (d > rad) {
            px = px + qt * (double)e[0];
            py = py + qt * (double)e[1];
            pz = pz + qt * (double)e[2];
            q = q + qt;
        } else {
            px = px - qt * (double)e[0];
            py = py - qt * (double)e[1];
            pz = pz - qt * (double)e[2];
            q = q - qt;
        }
    }
    p[0] = px;
    p[1] = py;
    p[2] = pz;
    *qout = q;
    return (q > 1.0 || q < 0.0) ? 1 : 0;
}

/* The rotation part of the body velocity at offset r from the centre,
 * Omega x r (CaseIV_DKT.F90:509-510: temp1 = -(y0 - yc) Omega,
 * temp2 = (x0 - xc) Omega). */
static inline void ps_spin(const double om[3], const double r[3], double w[3]) {
    w[0] = om[1] * r[2] - om[2] * r[1];
    w[1] = om[2] * r[0] - om[0] * r[2];
    w[2] = om[0] * r[1] - om[1] * r[0];
}

/* e . (U + W), the moving-wall velocity along e (CaseIV_DKT.F90:515, 519:
 * ex(r)*(Uc + temp1) + ey(r)*(Vc + temp2)). */
static inline double ps_wall_dot(const int e[3], const double U[3], const double W[3]) {
    return (double)e[0] * (U[0] + W[0]) + (double)e[1] * (U[1] + W[1]) + (double)e[2] * (U[2] + W[2]);
}

/* Quadratic interpolated bounce-back (CaseIV_DKT.F90:512-520,
 * particle_bounceback.F90:65-75): the population entering the fluid node
 * along -e from a wall at fraction q.  fa, fa1, fa2 = f_post(e) at x, x - e,
 * x - 2e; fb, fb1 = f_post(-e) at x, x - e; wa = weight of e; wall = the
 * wall velocity along -e (ps_wall_dot). */
static inline double ps_interp_bb(double q, double fa, double fa1, double fa2, double fb, double fb1, double wa,
                                  double rho_avg, double wall) {
    if (q < 0.5)
        return q * (1.0 + 2.0 * q) * fa + (1.0 - 4.0 * q * q) * fa1 - q * (1.0 - 2.0 * q) * fa2 +
               6.0 * wa * rho_avg * wall;
    return fa / q / (1.0 + 2.0 * q) + fb * (2.0 * q - 1.0) / q - fb1 * (2.0 * q - 1.0) / (2.0 * q + 1.0) +
           6.0 * wa * rho_avg / q / (1.0 + 2.0 * q) * wall;
}

/* One component of the momentum exchange with the Galilean correction of
 * Wen et al. JCP 2014 (CaseIV_DKT.F90:712-713, particle_force.F90:60-62):
 * (e_a - U - W) f_post(a) - (e_b - U - W) f(b), b = -a. */
static inline double ps_exchange(int ea, int eb, double U, double W, double fpost_a, double f_b) {
    return ((double)ea - U - W) * fpost_a - ((double)eb - U - W) * f_b;
}

/* Three-point extrapolation of a refilled population along the outward
 * direction (CaseIV_DKT.F90:1136, particle_update.F90:126-148). */
static inline double ps_extrapolate(double f1, double f2, double f3) { return 3.0 * f1 - 3.0 * f2 + f3; }

/* The spring magnitude scale (s / threshold)^2, s = distance - threshold,
 * the distance being what each spring measures: the centre distance less
 * both radii between particles, the centre-to-wall or gap-to-wall distance at
 * a wall (CaseIV_DKT.F90:754-757, 776, 788; particle_force.F90:108-111,
 * 143, 159, 175).  (x/t)**2.0d0 is the product (x/t)*(x/t). */
static inline double ps_spring(double scale, double s, double threshold) {
    const double g = s / threshold;
    return scale * (g * g);
}

#endif
