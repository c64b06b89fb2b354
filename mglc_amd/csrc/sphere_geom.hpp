// sphere_geom.hpp -- node/sphere geometry of the particle coupling, shared by
// the device kernels and host-side checks (host and device builds give the
// same bits: no contraction, correctly rounded sqrt).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace lbm {

// The bisection's stop tolerance: the reference declares
// `real(8), parameter :: epsRadius=1e-9` (particle_bounceback.F90:109, and
// CaseIV_DKT.F90:559), a default-real literal, so its value is the single
// precision 1e-9 widened to double: 9.99999971718068537e-10.
constexpr double CALQ_EPS = (double)1e-9f;

// Squared distance with each product rounded and the sum taken left to
// right, as (x - xc)**2 + (y - yc)**2 + (z - zc)**2 of initial.F90:78-89 and
// particle_update.F90:69-79 evaluates: the solid map built from it is exact.
__host__ __device__ inline double sphere_dist2(double dx, double dy, double dz) {
#pragma clang fp contract(off)
    return (dx * dx + dy * dy) + dz * dz;
}

// calQ (particle_bounceback.F90:101-141): wall fraction q of the link from
// node X along e = (ex, ey, ez) into the sphere (cx, cy, cz, rad) by
// bisection from q = 1/2, stopping when |d - rad| < CALQ_EPS (at most 200
// halvings); (x0, y0, z0) receives the last point.  Returns false when q
// leaves [0, 1].
// Same halvings, same stop rule and same branches as the reference's loop
// (bit for bit), at a fraction of its cost: along the link
// g(q) = |X + q e - c|^2 - rad^2 = (a q + 2 b) q + g0 is a quadratic, so each
// halving costs two fused multiply-adds.  Where |g| > T the point is more
// than 2e-9 away from the surface (|d - rad| = |g| / (d + rad), d <= rad + 3
// on a boundary link): the reference's stop test cannot fire and its branch
// d > rad is g > 0.  T also covers the rounding of g (below 1e-14 (rad + 2)^2,
// allowed 1e-12 (rad + 2)^2).  So the halvings run in two loops: while
// |g| > T, on g alone (no point, no square root); from the first halving
// inside the band on, the reference's own test on the point, accumulated
// step by step as the reference accumulates it.  The point at the switch is
// X + q e exactly, as the reference's sum is: q then has at most
// CALQ_FAR_MAX + 1 bits after the binary point, so both are exact for node
// coordinates below 2^(52 - CALQ_FAR_MAX) = 2^20 (a link enters the band
// after about 30 halvings; one that has not after CALQ_FAR_MAX goes on in
// the second loop).
constexpr int CALQ_FAR_MAX = 32;

__host__ __device__ inline bool calq(double cx, double cy, double cz, double rad, double X, double Y, double Z, int ex,
                                     int ey, int ez, double& x0, double& y0, double& z0, double& qout) {
#pragma clang fp contract(off)
    const double Dx = X - cx, Dy = Y - cy, Dz = Z - cz;
    const double ga = (double)(ex * ex + ey * ey + ez * ez);
    const double gb2 = 2.0 * (ex * Dx + ey * Dy + ez * Dz);
    const double g0 = sphere_dist2(Dx, Dy, Dz) - rad * rad;
    const double band = (2.0 * rad + 3.0) * 2e-9 + 1e-12 * (rad + 2.0) * (rad + 2.0);
    double q = 0.5, qt = 0.5;
    int it = 0;
#pragma unroll 1  // unrolled, the loop holds many more registers for no gain
    for (; it < CALQ_FAR_MAX; ++it) {
        const double g = __builtin_fma(__builtin_fma(ga, q, gb2), q, g0);
        if (fabs(g) <= band) break;
        qt = qt / 2.0;
        q += g > 0.0 ? qt : -qt;
    }
    double px = X + q * ex, py = Y + q * ey, pz = Z + q * ez;
#pragma unroll 1
    for (; it < 200; ++it) {
        const double g = __builtin_fma(__builtin_fma(ga, q, gb2), q, g0);
        bool outside;
        if (fabs(g) <= band) {
            const double d = sqrt(sphere_dist2(px - cx, py - cy, pz - cz));
            if (fabs(d - rad) < CALQ_EPS) break;
            outside = d > rad;
        } else {
            outside = g > 0.0;
        }
        qt = qt / 2.0;
        const double s = outside ? qt : -qt;
        px += s * ex;
        py += s * ey;
        pz += s * ez;
        q += s;
    }
    x0 = px;
    y0 = py;
    z0 = pz;
    qout = q;
    return !(q > 1.0 || q < 0.0);
}

}  // namespace lbm
