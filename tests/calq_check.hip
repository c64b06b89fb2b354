// calq_check.hip -- host-only check (built with --cuda-host-only, run on the
// CPU) that lbm::calq, the two-loop bisection the particle kernels use
// (halvings on the quadratic alone far from the surface, the reference's own
// test near it), returns bit for bit what the reference's plain loop
// (particle_bounceback.F90:101-141, restated below) returns, on every
// boundary link of many random spheres.  Test infrastructure.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>

#include "../mglc_amd/csrc/sphere_geom.hpp"

static const int EX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
static const int EY[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
static const int EZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};

// the reference's loop: a square root every halving
static bool calq_plain(double cx, double cy, double cz, double rad, double X, double Y, double Z, int ex, int ey,
                       int ez, double& x0, double& y0, double& z0, double& qout) {
#pragma clang fp contract(off)
    double q = 0.5, qt = 0.5;
    double px = X + qt * ex, py = Y + qt * ey, pz = Z + qt * ez;
    for (int it = 0; it < 200; ++it) {
        const double d = sqrt(lbm::sphere_dist2(px - cx, py - cy, pz - cz));
        if (fabs(d - rad) < (double)1e-9f) break;  // the reference's epsRadius
        qt = qt / 2.0;
        if (d > rad) {
            px += qt * ex; py += qt * ey; pz += qt * ez; q += qt;
        } else {
            px -= qt * ex; py -= qt * ey; pz -= qt * ez; q -= qt;
        }
    }
    x0 = px; y0 = py; z0 = pz; qout = q;
    return !(q > 1.0 || q < 0.0);
}

static unsigned long long rng = 0x9E3779B97F4A7C15ull;
static double uni() {  // splitmix64 -> [0, 1)
    unsigned long long z = (rng += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (double)((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
}

int main(int argc, char** argv) {
    const int spheres = argc > 1 ? atoi(argv[1]) : 300;
    long long links = 0, mismatches = 0, band = 0;
    for (int s = 0; s < spheres; ++s) {
        const double rad = s % 7 == 0 ? (double)(2 + s % 5) : 2.0 + 8.0 * uni();  // some integer radii (ties)
        // centres near the origin and far from it (node coordinates up to ~30000)
        const double base = s % 3 == 0 ? 20.0 : (s % 3 == 1 ? 400.0 : 30000.0);
        const double cx = base + (s % 11 == 0 ? 0.0 : uni()), cy = base + uni(), cz = base + (s % 13 == 0 ? 0.5 : uni());
        const int b = (int)base;
        for (int k = b - 12; k <= b + 12; ++k)
            for (int j = b - 12; j <= b + 12; ++j)
                for (int i = b - 12; i <= b + 12; ++i) {
                    if (lbm::sphere_dist2(i - cx, j - cy, k - cz) <= rad * rad) continue;
                    for (int a = 1; a < 19; ++a) {
                        if (lbm::sphere_dist2(i + EX[a] - cx, j + EY[a] - cy, k + EZ[a] - cz) > rad * rad) continue;
                        double x1, y1, z1, q1, x2, y2, z2, q2;
                        const bool ok1 = lbm::calq(cx, cy, cz, rad, i, j, k, EX[a], EY[a], EZ[a], x1, y1, z1, q1);
                        const bool ok2 = calq_plain(cx, cy, cz, rad, i, j, k, EX[a], EY[a], EZ[a], x2, y2, z2, q2);
                        ++links;
                        if (ok1 != ok2 || memcmp(&q1, &q2, 8) || memcmp(&x1, &x2, 8) || memcmp(&y1, &y2, 8) ||
                            memcmp(&z1, &z2, 8))
                            ++mismatches;
                        const double d = sqrt(lbm::sphere_dist2(x2 - cx, y2 - cy, z2 - cz));
                        if (fabs(d - rad) < (double)1e-9f) ++band;
                    }
                }
    }
    printf("links %lld mismatches %lld converged %lld\n", links, mismatches, band);
    return mismatches != 0;
}
