/* particle2d_oracle.c -- TEST INFRASTRUCTURE: a CPU restatement, in C, of the
 * reference's own two-dimensional moving-particle program
 * (MPI/Micro_particles/fortran/case4/seq/CaseIV_DKT.F90: two discs settling
 * side by side -- drafting, kissing and tumbling -- in a 201 x 801 D2Q9
 * channel whose window moves with them).  It is the oracle that ties the
 * particle scheme (calQ bisection, quadratic interpolated bounce-back with
 * the moving-wall term, Wen et al. momentum exchange, springs, kinematics,
 * new map, refill by extrapolation and moment replacement) to the
 * reference's own numbers: tests/golden holds states that the reference
 * program itself wrote (oracle/ref/ref_dkt_main.f90), and this restatement
 * must reproduce them.  Only tests load it; nothing in the product does.
 *
 * Every sum runs in the reference's order and every expression is evaluated
 * in its order (left to right, products rounded; x**2.0d0 is x*x as the
 * reference's -O3 build computes it), so with contraction off the results
 * are the reference's bits; the per-node loops whose writes never overlap
 * (collision, push streaming, macro) run in parallel, which changes no value.
 * The per-link and per-particle formulas are the functions of
 * particle_scheme.h, shared with the three-dimensional oracle.  Parameters
 * are the reference's compile-time ones (CaseIV_DKT.F90:1-62).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "lbm_oracle.h"
#include "particle_scheme.h"

#define NX 201
#define NY 801
#define NCP 2 /* cNumMax */

static const int EX2[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1}; /* CaseIV_DKT.F90:44-45 */
static const int EY2[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
static const int R2[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6}; /* r(1:8), :49-50 */

/* f, f_post: (0:8, -1:nx+2, -1:ny+2); obst: (0:nx+1, 0:ny+1); fields (1:nx, 1:ny) */
#define FI(a, i, j) ((size_t)(a) + 9 * ((size_t)((i) + 1) + (size_t)(NX + 4) * (size_t)((j) + 1)))
#define OI(i, j) ((size_t)(i) + (size_t)(NX + 2) * (size_t)(j))
#define MI(i, j) ((size_t)((i) - 1) + (size_t)NX * (size_t)((j) - 1))

struct dkt {
    double *f, *fp, *rho, *u, *v, *up, *vp;
    int *obst, *obst_new;
    double xc[NCP], yc[NCP], xco[NCP], yco[NCP], om[NCP], omo[NCP], uc[NCP], vc[NCP], uco[NCP], vco[NCP];
    double radius[NCP], rho_avg, fx[NCP], fy[NCP], torque[NCP];
    double fwx[NCP], fwy[NCP]; /* calForce's wall springs: kept between calls (see oracle/ref/Makefile) */
    int itc, shift;
    /* parameters */
    double pi, l0, t0, radius0, rho0, rho_solid, tauf, snu, sq, gravity, th_wall, st_wall, th_part, st_part;
    double omega[9];
    int ny_half;
};

static double sq2(double x) { return x * x; }

/* CaseIV_DKT.F90:134-268 */
static void initial(struct dkt *s) {
    s->pi = 4.0 * atan(1.0);
    s->l0 = 1.0 / 100.0;
    s->t0 = 5.0 / 10000.0;
    s->radius0 = 20.0 / 2.0;
    s->rho0 = 1.0;
    s->rho_solid = 1.01;
    s->tauf = 3.0 * 0.05 + 0.5;
    s->snu = 1.0 / s->tauf;
    s->sq = 8.0 * (2.0 * s->tauf - 1.0) / (8.0 * s->tauf - 1.0);
    s->gravity = 980.0 * sq2(s->t0) / s->l0;
    s->th_wall = 4.0;
    s->st_wall = 0.25;
    s->th_part = 4.0;
    s->st_part = 0.25;
    s->ny_half = (NY - 1) / 2 + 1;
    s->itc = 0;
    s->xc[0] = 101.01;
    s->yc[0] = (double)s->ny_half + 40.0;
    s->xc[1] = 101.0;
    s->yc[1] = (double)s->ny_half;
    for (int c = 0; c < NCP; ++c) {
        s->xco[c] = s->xc[c];
        s->yco[c] = s->yc[c];
        s->radius[c] = s->radius0;
        s->uc[c] = s->vc[c] = s->om[c] = 0.0;
        s->uco[c] = s->vco[c] = s->omo[c] = 0.0;
        s->fwx[c] = s->fwy[c] = 0.0;
    }
    for (int j = 1; j <= NY; ++j)
        for (int i = 1; i <= NX; ++i) {
            s->rho[MI(i, j)] = s->rho0;
            for (int c = 0; c < NCP; ++c)
                if (sq2(i - s->xc[c]) + sq2(j - s->yc[c]) <= sq2(s->radius[c])) {
                    s->obst[OI(i, j)] = 1;
                    s->rho[MI(i, j)] = s->rho_solid;
                }
        }
    s->omega[0] = 4.0 / 9.0;
    for (int a = 1; a <= 4; ++a) s->omega[a] = 1.0 / 9.0;
    for (int a = 5; a <= 8; ++a) s->omega[a] = 1.0 / 36.0;
    for (int j = 1; j <= NY; ++j)
        for (int i = 1; i <= NX; ++i)
            if (s->obst[OI(i, j)] == 0) {
                const double uu = s->u[MI(i, j)], vv = s->v[MI(i, j)];
                const double us2 = uu * uu + vv * vv;
                for (int a = 0; a <= 8; ++a) {
                    const double un = uu * EX2[a] + vv * EY2[a];
                    s->f[FI(a, i, j)] = s->omega[a] * (1.0 + 3.0 * un + 4.5 * un * un - 1.5 * us2);
                }
            }
    s->shift = 0;
}

/* the D2Q9 moment transform and its inverse (CaseIV_DKT.F90:356-394, 1140-1166) */
static void moments(const double *f, double m[9]) {
    m[0] = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
    m[1] = -4.0 * f[0] - f[1] - f[2] - f[3] - f[4] + 2.0 * f[5] + 2.0 * f[6] + 2.0 * f[7] + 2.0 * f[8];
    m[2] = 4.0 * f[0] - 2.0 * f[1] - 2.0 * f[2] - 2.0 * f[3] - 2.0 * f[4] + f[5] + f[6] + f[7] + f[8];
    m[3] = f[1] - f[3] + f[5] - f[6] - f[7] + f[8];
    m[4] = -2.0 * f[1] + 2.0 * f[3] + f[5] - f[6] - f[7] + f[8];
    m[5] = f[2] - f[4] + f[5] + f[6] - f[7] - f[8];
    m[6] = -2.0 * f[2] + 2.0 * f[4] + f[5] + f[6] - f[7] - f[8];
    m[7] = f[1] - f[2] + f[3] - f[4];
    m[8] = f[5] - f[6] + f[7] - f[8];
}

static void inverse(const double m[9], double *f) {
    f[0] = (m[0] - m[1] + m[2]) / 9.0;
    f[1] = m[0] / 9.0 - m[1] / 36.0 - m[2] / 18.0 + m[3] / 6.0 - m[4] / 6.0 + m[7] * 0.25;
    f[2] = m[0] / 9.0 - m[1] / 36.0 - m[2] / 18.0 + m[5] / 6.0 - m[6] / 6.0 - m[7] * 0.25;
    f[3] = m[0] / 9.0 - m[1] / 36.0 - m[2] / 18.0 - m[3] / 6.0 + m[4] / 6.0 + m[7] * 0.25;
    f[4] = m[0] / 9.0 - m[1] / 36.0 - m[2] / 18.0 - m[5] / 6.0 + m[6] / 6.0 - m[7] * 0.25;
    f[5] = m[0] / 9.0 + m[1] / 18.0 + m[2] / 36.0 + m[3] / 6.0 + m[4] / 12.0 + m[5] / 6.0 + m[6] / 12.0 + m[8] * 0.25;
    f[6] = m[0] / 9.0 + m[1] / 18.0 + m[2] / 36.0 - m[3] / 6.0 - m[4] / 12.0 + m[5] / 6.0 + m[6] / 12.0 - m[8] * 0.25;
    f[7] = m[0] / 9.0 + m[1] / 18.0 + m[2] / 36.0 - m[3] / 6.0 - m[4] / 12.0 - m[5] / 6.0 - m[6] / 12.0 + m[8] * 0.25;
    f[8] = m[0] / 9.0 + m[1] / 18.0 + m[2] / 36.0 + m[3] / 6.0 + m[4] / 12.0 - m[5] / 6.0 - m[6] / 12.0 - m[8] * 0.25;
}

/* collision with the moving window (CaseIV_DKT.F90:272-410) */
static int collision(struct dkt *s) {
    const double shift = (double)(int)((double)s->ny_half - s->yc[1]);
    if (fabs(shift) >= 2.0) return 5;
    if (shift == 1.0 || shift == -1.0) {
        const int up = shift == 1.0;
        for (int c = 0; c < NCP; ++c) s->yc[c] = up ? s->yc[c] + 1.0 : s->yc[c] - 1.0;
        s->shift += up ? 1 : -1;
        if (up) {
            for (int j = NY - 1; j >= 1; --j)
                for (int i = 1; i <= NX; ++i) {
                    for (int a = 0; a <= 8; ++a) s->f[FI(a, i, j + 1)] = s->f[FI(a, i, j)];
                    s->u[MI(i, j + 1)] = s->u[MI(i, j)];
                    s->v[MI(i, j + 1)] = s->v[MI(i, j)];
                    s->rho[MI(i, j + 1)] = s->rho[MI(i, j)];
                    s->obst[OI(i, j + 1)] = s->obst[OI(i, j)];
                }
        } else {
            for (int j = 2; j <= NY; ++j)
                for (int i = 1; i <= NX; ++i) {
                    for (int a = 0; a <= 8; ++a) s->f[FI(a, i, j - 1)] = s->f[FI(a, i, j)];
                    s->u[MI(i, j - 1)] = s->u[MI(i, j)];
                    s->v[MI(i, j - 1)] = s->v[MI(i, j)];
                    s->rho[MI(i, j - 1)] = s->rho[MI(i, j)];
                    s->obst[OI(i, j - 1)] = s->obst[OI(i, j)];
                }
        }
        const int j = up ? 1 : NY;
        for (int i = 1; i <= NX; ++i) {
            s->u[MI(i, j)] = 0.0;
            s->v[MI(i, j)] = 0.0;
            s->rho[MI(i, j)] = s->rho0;
            s->obst[OI(i, j)] = 0;
            for (int a = 0; a <= 8; ++a) s->f[FI(a, i, j)] = s->omega[a] * s->rho0;
        }
    }
    const double rates[9] = {0.0, s->snu, s->snu, 0.0, s->sq, 0.0, s->sq, s->snu, s->snu};
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= NY; ++j)
        for (int i = 1; i <= NX; ++i) {
            if (s->obst[OI(i, j)] != 0) continue;
            double m[9], meq[9], mp[9], fl[9];
            for (int a = 0; a <= 8; ++a) fl[a] = s->f[FI(a, i, j)];
            moments(fl, m);
            const double r = s->rho[MI(i, j)], uu = s->u[MI(i, j)], vv = s->v[MI(i, j)];
            meq[0] = r;
            meq[1] = r * (-2.0 + 3.0 * (uu * uu + vv * vv));
            meq[2] = r * (1.0 - 3.0 * (uu * uu + vv * vv));
            meq[3] = r * uu;
            meq[4] = -meq[3];
            meq[5] = r * vv;
            meq[6] = -meq[5];
            meq[7] = r * (uu * uu - vv * vv);
            meq[8] = r * (uu * vv);
            for (int a = 0; a <= 8; ++a) mp[a] = m[a] - rates[a] * (m[a] - meq[a]);
            inverse(mp, fl);
            for (int a = 0; a <= 8; ++a) s->fp[FI(a, i, j)] = fl[a];
        }
    return 0;
}

/* calQ (CaseIV_DKT.F90:551-591) through the shared bisection; a disc is a
 * sphere at z = 0 and the link stays in that plane.  Returns 3 if q leaves
 * [0, 1]. */
static int calq(const struct dkt *s, int c, double i, double j, int a, double *x0, double *y0, double *qout) {
    const double ctr[3] = {s->xc[c], s->yc[c], 0.0}, X[3] = {i, j, 0.0};
    const int e[3] = {EX2[a], EY2[a], 0};
    double p[3];
    const int bad = ps_calq(ctr, s->radius[c], X, e, p, qout);
    *x0 = p[0];
    *y0 = p[1];
    return bad ? 3 : 0;
}

/* the disc's body velocity split as the reference uses it: U and the spin
 * part W = Omega x (x0 - c) (CaseIV_DKT.F90:509-510, 701-702) */
static void disc_motion(const struct dkt *s, int c, double x0, double y0, double U[3], double W[3]) {
    const double om[3] = {0.0, 0.0, s->om[c]}, r[3] = {x0 - s->xc[c], y0 - s->yc[c], 0.0};
    U[0] = s->uc[c];
    U[1] = s->vc[c];
    U[2] = 0.0;
    ps_spin(om, r, W);
}

static int inside(const struct dkt *s, int c, int i, int j, const double *xc, const double *yc) {
    return sq2(i - xc[c]) + sq2(j - yc[c]) <= sq2(s->radius[c]);
}

/* bounceback at the channel walls (CaseIV_DKT.F90:594-633, moving window) */
static void bounceback(struct dkt *s) {
    for (int j = 1; j <= NY; ++j) {
        s->f[FI(1, 1, j)] = s->fp[FI(3, 1, j)];
        s->f[FI(5, 1, j)] = s->fp[FI(7, 1, j)];
        s->f[FI(8, 1, j)] = s->fp[FI(6, 1, j)];
        s->f[FI(3, NX, j)] = s->fp[FI(1, NX, j)];
        s->f[FI(6, NX, j)] = s->fp[FI(8, NX, j)];
        s->f[FI(7, NX, j)] = s->fp[FI(5, NX, j)];
    }
    for (int i = 1; i <= NX; ++i) {
        s->f[FI(4, i, NY)] = s->f[FI(4, i, NY - 1)];
        s->f[FI(7, i, NY)] = s->f[FI(7, i, NY - 1)];
        s->f[FI(8, i, NY)] = s->f[FI(8, i, NY - 1)];
        s->f[FI(2, i, 1)] = s->fp[FI(4, i, 1)];
        s->f[FI(5, i, 1)] = s->fp[FI(7, i, 1)];
        s->f[FI(6, i, 1)] = s->fp[FI(8, i, 1)];
    }
}

/* streaming with the interpolated bounce-back at the discs
 * (CaseIV_DKT.F90:413-548) */
static int streaming(struct dkt *s) {
    /* every (node, direction) writes its own target: order-free */
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= NY; ++j)
        for (int i = 1; i <= NX; ++i)
            if (s->obst[OI(i, j)] == 0)
                for (int a = 0; a <= 8; ++a) s->f[FI(a, i + EX2[a], j + EY2[a])] = s->fp[FI(a, i, j)];
    bounceback(s);
    double avg = 0.0;
    int n = 0;
    for (int j = 1; j <= NY; ++j)
        for (int i = 1; i <= NX; ++i)
            if (s->obst[OI(i, j)] == 0) {
                avg = avg + s->rho[MI(i, j)];
                n = n + 1;
            }
    s->rho_avg = avg / (double)n;
    for (int j = -1; j <= NY + 2; ++j)
        for (int a = 0; a <= 8; ++a) {
            const double w = s->omega[a] * s->rho0;
            s->f[FI(a, 0, j)] = w;
            s->f[FI(a, -1, j)] = w;
            s->f[FI(a, NX + 1, j)] = w;
            s->f[FI(a, NX + 2, j)] = w;
            s->fp[FI(a, 0, j)] = w;
            s->fp[FI(a, -1, j)] = w;
            s->fp[FI(a, NX + 1, j)] = w;
            s->fp[FI(a, NX + 2, j)] = w;
        }
    for (int i = -1; i <= NX + 2; ++i)
        for (int a = 0; a <= 8; ++a) {
            const double w = s->omega[a] * s->rho0;
            s->f[FI(a, i, 0)] = w;
            s->f[FI(a, i, -1)] = w;
            s->f[FI(a, i, NY + 1)] = w;
            s->f[FI(a, i, NY + 2)] = w;
            s->fp[FI(a, i, 0)] = w;
            s->fp[FI(a, i, -1)] = w;
            s->fp[FI(a, i, NY + 1)] = w;
            s->fp[FI(a, i, NY + 2)] = w;
        }
    for (int j = 1; j <= NY; ++j)
        for (int i = 1; i <= NX; ++i) {
            if (s->obst[OI(i, j)] != 0) continue;
            for (int a = 0; a <= 8; ++a) {
                const int ip = i + EX2[a], jp = j + EY2[a];
                if (s->obst[OI(ip, jp)] != 1) continue;
                int found = 0;
                for (int c = 0; c < NCP; ++c) {
                    if (!inside(s, c, ip, jp, s->xc, s->yc)) continue;
                    found = 1;
                    double x0, y0, q;
                    const int rc = calq(s, c, (double)i, (double)j, a, &x0, &y0, &q);
                    if (rc) return rc;
                    double U[3], W[3];
                    disc_motion(s, c, x0, y0, U, W);
                    const int ra = R2[a];
                    const int er[3] = {EX2[ra], EY2[ra], 0};
                    const double wall = ps_wall_dot(er, U, W);
                    const double val =
                        ps_interp_bb(q, s->fp[FI(a, i, j)], s->fp[FI(a, i - EX2[a], j - EY2[a])],
                                     s->fp[FI(a, i - 2 * EX2[a], j - 2 * EY2[a])], s->fp[FI(ra, i, j)],
                                     s->fp[FI(ra, i - EX2[a], j - EY2[a])], s->omega[a], s->rho_avg, wall);
                    s->f[FI(ra, i, j)] = val;
                }
                if (!found) return 7;
            }
        }
    return 0;
}

/* CaseIV_DKT.F90:636-656 */
static void macro(struct dkt *s) {
#pragma omp parallel for schedule(static)
    for (int j = 1; j <= NY; ++j)
        for (int i = 1; i <= NX; ++i)
            if (s->obst[OI(i, j)] == 0) {
                const double *f = &s->f[FI(0, i, j)];
                const double r = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
                s->rho[MI(i, j)] = r;
                s->u[MI(i, j)] = (f[1] - f[3] + f[5] - f[6] - f[7] + f[8]) / r;
                s->v[MI(i, j)] = (f[2] - f[4] + f[5] + f[6] - f[7] - f[8]) / r;
            }
}

/* CaseIV_DKT.F90:659-836 */
static int cal_force(struct dkt *s) {
    for (int c = 0; c < NCP; ++c) s->fx[c] = s->fy[c] = s->torque[c] = 0.0;
    for (int j = 1; j <= NY; ++j)
        for (int i = 1; i <= NX; ++i) {
            if (s->obst[OI(i, j)] != 0) continue;
            for (int a = 1; a <= 8; ++a) {
                const int ip = i + EX2[a], jp = j + EY2[a];
                if (s->obst[OI(ip, jp)] != 1) continue;
                int found = 0;
                for (int c = 0; c < NCP; ++c) {
                    if (!inside(s, c, ip, jp, s->xc, s->yc)) continue;
                    found = 1;
                    double x0, y0, q;
                    const int rc = calq(s, c, (double)i, (double)j, a, &x0, &y0, &q);
                    if (rc) return rc;
                    double U[3], W[3];
                    disc_motion(s, c, x0, y0, U, W);
                    const int ra = R2[a];
                    const double fpa = s->fp[FI(a, i, j)], fr = s->f[FI(ra, i, j)];
                    const double tfx = ps_exchange(EX2[a], EX2[ra], U[0], W[0], fpa, fr);
                    const double tfy = ps_exchange(EY2[a], EY2[ra], U[1], W[1], fpa, fr);
                    const double tt = (x0 - s->xc[c]) * tfy - (y0 - s->yc[c]) * tfx;
                    s->fx[c] = s->fx[c] + tfx;
                    s->fy[c] = s->fy[c] + tfy;
                    s->torque[c] = s->torque[c] + tt;
                }
                if (!found) return 7;
            }
        }
    double fxij[NCP], fyij[NCP];
    for (int c = 0; c < NCP; ++c) {
        double scale = s->pi * sq2(s->radius[c]) * (s->rho_solid - s->rho_avg) * s->gravity / s->st_part;
        fxij[c] = 0.0;
        fyij[c] = 0.0;
        for (int c2 = 0; c2 < NCP; ++c2) {
            if (c2 == c) continue;
            const double dij = sqrt(sq2(s->xc[c] - s->xc[c2]) + sq2(s->yc[c] - s->yc[c2]));
            if (dij >= s->radius[c] + s->radius[c2] + s->th_part) {
                /* no force */
            } else if (dij < s->radius[c] + s->radius[c2] + s->th_part && dij >= s->radius[c] + s->radius[c2]) {
                const double mag = ps_spring(scale, dij - s->radius[c] - s->radius[c2] - s->th_part, s->th_part);
                fxij[c] = fxij[c] + mag * (s->xc[c] - s->xc[c2]) / dij;
                fyij[c] = fyij[c] + mag * (s->yc[c] - s->yc[c2]) / dij;
            } else {
                return 1;
            }
        }
        scale = s->pi * sq2(s->radius[c]) * (s->rho_solid - s->rho_avg) * s->gravity / s->st_wall;
        if (s->xc[c] < 1.5 * s->radius0) {
            const double dw = s->xc[c] - 1.0;
            if (dw >= s->radius[c] + s->th_wall) {
                s->fwx[c] = 0.0;
                s->fwy[c] = 0.0;
            } else if (dw < s->radius[c] + s->th_wall && dw >= s->radius[c]) {
                s->fwx[c] = ps_spring(scale, dw - s->th_wall, s->th_wall);
                s->fwy[c] = 0.0;
            } else {
                return 2;
            }
        } else if (s->xc[c] > (double)NX - 1.5 * s->radius0) {
            const double dw = (double)NX - s->xc[c];
            if (dw >= s->radius[c] + s->th_wall) {
                s->fwx[c] = 0.0;
                s->fwy[c] = 0.0;
            } else if (dw < s->radius[c] + s->th_wall && dw >= s->radius[c]) {
                s->fwx[c] = -ps_spring(scale, dw - s->th_wall, s->th_wall);
                s->fwy[c] = 0.0;
            } else {
                return 2;
            }
        }
    }
    for (int c = 0; c < NCP; ++c) {
        s->fx[c] = s->fx[c] + fxij[c] + s->fwx[c];
        s->fy[c] = s->fy[c] - (s->rho_solid - s->rho_avg) * s->pi * sq2(s->radius0) * s->gravity + fyij[c] + s->fwy[c];
        s->xco[c] = s->xc[c];
        s->yco[c] = s->yc[c];
        s->uco[c] = s->uc[c];
        s->vco[c] = s->vc[c];
        s->omo[c] = s->om[c];
    }
    return 0;
}

/* CaseIV_DKT.F90:1004-1193 */
static int update_center(struct dkt *s) {
    for (int c = 0; c < NCP; ++c) {
        const double ax = s->fx[c] / s->pi / sq2(s->radius0) / s->rho_solid;
        const double ay = s->fy[c] / s->pi / sq2(s->radius0) / s->rho_solid;
        const double r2 = sq2(s->radius[c]);
        const double aom = s->torque[c] / (0.5 * s->rho_solid * s->pi * (r2 * r2));
        s->uc[c] = s->uco[c] + ax;
        s->vc[c] = s->vco[c] + ay;
        s->om[c] = s->omo[c] + aom;
        s->xc[c] = s->xco[c] + s->uco[c] + 0.5 * ax;
        s->yc[c] = s->yco[c] + s->vco[c] + 0.5 * ay;
    }
    for (int c = 0; c < NCP; ++c)
        if (s->yc[c] <= 2.0 * s->radius0) return 6;
    for (int j = 1; j <= NY; ++j)
        for (int i = 1; i <= NX; ++i) {
            s->obst_new[OI(i, j)] = 0;
            for (int c = 0; c < NCP; ++c)
                if (inside(s, c, i, j, s->xc, s->yc)) {
                    s->obst_new[OI(i, j)] = 1;
                    s->rho[MI(i, j)] = s->rho_solid;
                    s->u[MI(i, j)] = 0.0;
                    s->v[MI(i, j)] = 0.0;
                }
        }
    double avg = 0.0;
    int n = 0;
    for (int j = 1; j <= NY; ++j)
        for (int i = 1; i <= NX; ++i)
            if (s->obst_new[OI(i, j)] == 0) {
                avg = avg + s->rho[MI(i, j)];
                n = n + 1;
            }
    s->rho_avg = avg / (double)n;
    for (int j = 1; j <= NY; ++j)
        for (int i = 1; i <= NX; ++i) {
            if (!(s->obst[OI(i, j)] == 1 && s->obst_new[OI(i, j)] == 0)) continue;
            int found = 0;
            for (int c = 0; c < NCP; ++c) {
                if (!inside(s, c, i, j, s->xco, s->yco)) continue;
                found = 1;
                double out = 0.0;
                int ec = 0;
                for (int a = 1; a <= 8; ++a) {
                    const double tn = ((i - s->xc[c]) * EX2[a] + (j - s->yc[c]) * EY2[a]) /
                                      sqrt(sq2(i - s->xc[c]) + sq2(j - s->yc[c]));
                    if (tn > out) {
                        out = tn;
                        ec = a;
                    }
                }
                if (ec == 0) return 4;
                double *f = &s->f[FI(0, i, j)];
                for (int a = 0; a <= 8; ++a)
                    f[a] = ps_extrapolate(s->f[FI(a, i + EX2[ec], j + EY2[ec])],
                                          s->f[FI(a, i + 2 * EX2[ec], j + 2 * EY2[ec])],
                                          s->f[FI(a, i + 3 * EX2[ec], j + 3 * EY2[ec])]);
                double m[9];
                moments(f, m);
                double U[3], W[3];
                disc_motion(s, c, (double)i, (double)j, U, W);
                m[3] = s->rho_avg * (U[0] + W[0]);
                m[5] = s->rho_avg * (U[1] + W[1]);
                inverse(m, f);
                const double r = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
                s->rho[MI(i, j)] = r;
                s->u[MI(i, j)] = (f[1] - f[3] + f[5] - f[6] - f[7] + f[8]) / r;
                s->v[MI(i, j)] = (f[2] - f[4] + f[5] + f[6] - f[7] - f[8]) / r;
            }
            if (!found) return 7;
        }
    memcpy(s->obst, s->obst_new, sizeof(int) * (NX + 2) * (NY + 2));
    return 0;
}

/* CaseIV_DKT.F90:861-867 check(): the residual history, kept for fidelity */
static void check(struct dkt *s) {
    for (int j = 1; j <= NY; ++j)
        for (int i = 1; i <= NX; ++i)
            if (s->obst[OI(i, j)] == 0) {
                s->up[MI(i, j)] = s->u[MI(i, j)];
                s->vp[MI(i, j)] = s->v[MI(i, j)];
            }
}

int oracle_dkt_dims(int *nx, int *ny) {
    *nx = NX;
    *ny = NY;
    return 0;
}

int oracle_dkt_run(int nsteps, double *scal, double *rho, double *u, double *v, double *obst, double *f) {
    struct dkt s;
    memset(&s, 0, sizeof s);
    const size_t nf = 9 * (size_t)(NX + 4) * (size_t)(NY + 4), nm = (size_t)NX * NY, no = (size_t)(NX + 2) * (NY + 2);
    s.f = calloc(nf, sizeof(double));
    s.fp = calloc(nf, sizeof(double));
    s.rho = calloc(nm, sizeof(double));
    s.u = calloc(nm, sizeof(double));
    s.v = calloc(nm, sizeof(double));
    s.up = calloc(nm, sizeof(double));
    s.vp = calloc(nm, sizeof(double));
    s.obst = calloc(no, sizeof(int));
    s.obst_new = calloc(no, sizeof(int));
    int rc = -1;
    if (s.f && s.fp && s.rho && s.u && s.v && s.up && s.vp && s.obst && s.obst_new) {
        rc = 0;
        initial(&s);
        for (int step = 0; step < nsteps && rc == 0; ++step) {
            rc = collision(&s);
            if (!rc) rc = streaming(&s);
            if (rc) break;
            macro(&s);
            rc = cal_force(&s);
            if (rc) break;
            s.itc = s.itc + 1;
            if (s.itc % 2000 == 0) check(&s);
            rc = update_center(&s);
        }
        scal[0] = s.itc;
        scal[1] = s.shift;
        scal[2] = s.rho_avg;
        for (int c = 0; c < NCP; ++c) {
            scal[3 + c] = s.xc[c];
            scal[5 + c] = s.yc[c];
            scal[7 + c] = s.uc[c];
            scal[9 + c] = s.vc[c];
            scal[11 + c] = s.om[c];
            scal[13 + c] = s.fx[c];
            scal[15 + c] = s.fy[c];
            scal[17 + c] = s.torque[c];
        }
        for (int j = 1; j <= NY; ++j)
            for (int i = 1; i <= NX; ++i) {
                const size_t m = MI(i, j);
                rho[m] = s.rho[m];
                u[m] = s.u[m];
                v[m] = s.v[m];
                obst[m] = (double)s.obst[OI(i, j)];
                if (f)
                    for (int a = 0; a <= 8; ++a) f[9 * m + a] = s.f[FI(a, i, j)];
            }
    }
    free(s.f);
    free(s.fp);
    free(s.rho);
    free(s.u);
    free(s.v);
    free(s.up);
    free(s.vp);
    free(s.obst);
    free(s.obst_new);
    return rc;
}
