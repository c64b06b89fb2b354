/*
 * particle_oracle.c -- CPU restatement of the reference's moving-particle
 * coupling (MPI/Micro_particles/fortran/case4/mpi: fluid.F90,
 * particle_bounceback.F90, particle_force.F90, particle_update.F90,
 * initial.F90), generalised from D2Q9 discs to D3Q19 spheres.
 *
 * TEST INFRASTRUCTURE ONLY (parity checker for the GPU particle path).
 *
 * The reference's particle scheme exists only in 2-D.  Its per-link and
 * per-particle formulas (calQ, the interpolated bounce-back with the moving
 * wall, the momentum exchange, the springs, the refill extrapolation, the body
 * velocity) are computed here by the functions of particle_scheme.h, which
 * the 2-D restatement of the reference's sequential program
 * (particle2d_oracle.c) also uses and which that restatement checks bit for
 * bit against states the reference program itself wrote.  The 3-D assembly
 * around them has no reference output to compare with; it is pinned by
 * closed forms and physics (tests/test_oracle_particle_pins.py,
 * tests/test_oracle_particles.py): the particle map against brute-force
 * enumeration, the bisection wall distance against the reference's own closed
 * form calQ2, rest-state invariance, the solid sphere's m and I through the
 * kinematic update, the buoyancy volume, the six wall springs, the refill's
 * direction, extrapolation and momentum on every uncovered node, Stokes
 * spin-down and settling in the closed box; altering any of those formulas
 * makes a named check fail.
 *
 * Step (main.F90:35-69): collision on fluid nodes -> masked pull streaming ->
 * box-wall bounce-back -> interpolated bounce-back on particle links with the
 * moving-wall term -> macro on fluid nodes -> momentum-exchange force and
 * torque + spring and buoyancy forces -> kinematics -> new solid map ->
 * refill of uncovered nodes.
 *
 * 3-D generalisation choices (made identically in the GPU path):
 *  - node coordinates are 1-based integers (X = i, as initial.F90:34-39 with
 *    i + i_start_global); the box walls are halfway bounce-back;
 *  - volume V = 4/3 pi r^3, moment of inertia I = 2/5 rho_s V r^2 per
 *    particle radius (the 2-D code uses pi r0^2 and 1/2 rho_s pi r^4);
 *  - spring forces against all six walls (low walls at coordinate 1, high
 *    walls at N, as particle_force.F90:128-186 places them);
 *  - interpolation and refill stencil points that are not fluid nodes (solid,
 *    or outside the box) contribute the rest equilibrium w_a rho0
 *    (particle_bounceback.F90 reads ghost layers initialised to w_a rho0 for
 *    points outside; stale values for solid points are replaced);
 *  - refill reads the streamed populations before any refill of the step.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "lbm_oracle.h"
#include "particle_scheme.h"

static const int EX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
static const int EY[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
static const int EZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
static const int OPP[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
static const double PI = 3.14159265358979323846;

static double wgt(int a) { return a == 0 ? 1.0 / 3.0 : (a <= 6 ? 1.0 / 18.0 : 1.0 / 36.0); }

#define PCELL(p, i, j, k) ((size_t)(i) + (size_t)(p)->nx * ((size_t)(j) + (size_t)(p)->ny * (size_t)(k)))

static int inside(const oracle_particle_params *p, int i, int j, int k) {
    return i >= 0 && i < p->nx && j >= 0 && j < p->ny && k >= 0 && k < p->nz;
}

/* Squared distance with each product rounded, as x**2.0d0 compiles at -O3
 * (no fused multiply-add, left-to-right sum). */
static double dist2(double dx, double dy, double dz) {
    volatile double a = dx * dx, b = dy * dy, c = dz * dz;
    volatile double s = a + b;
    return s + c;
}

/* particle_update.F90:69-79 / initial.F90:78-89: node (i,j,k) (1-based
 * coordinates i+1, j+1, k+1) is solid iff inside some sphere; owner is that
 * sphere's index (the last one when several claim it), -1 for fluid. */
void oracle_particle_map(const oracle_particle_params *p, const double *xc, const double *r, int *owner) {
    size_t n = (size_t)p->nx * p->ny * p->nz;
    for (size_t c = 0; c < n; ++c) owner[c] = -1;
    for (int q = 0; q < p->np; ++q) {
        const double cx = xc[3 * q], cy = xc[3 * q + 1], cz = xc[3 * q + 2], rr = r[q] * r[q];
        int i0 = (int)floor(cx - r[q]) - 1, i1 = (int)ceil(cx + r[q]) + 1;
        int j0 = (int)floor(cy - r[q]) - 1, j1 = (int)ceil(cy + r[q]) + 1;
        int k0 = (int)floor(cz - r[q]) - 1, k1 = (int)ceil(cz + r[q]) + 1;
        for (int k = k0; k <= k1; ++k)
            for (int j = j0; j <= j1; ++j)
                for (int i = i0; i <= i1; ++i) {
                    if (!inside(p, i - 1, j - 1, k - 1)) continue;
                    if (dist2((double)i - cx, (double)j - cy, (double)k - cz) <= rr)
                        owner[PCELL(p, i - 1, j - 1, k - 1)] = q;
                }
    }
}

/* calQ (particle_bounceback.F90:101-141) through the shared bisection: wall
 * fraction q of the link from node x (1-based coordinates) along e_a into
 * sphere q_id; x0 receives the boundary point.  Returns 0, or -1 if q leaves
 * [0, 1]. */
int oracle_calq(const double *xc, const double *r, int id, const double x[3], int a, double x0[3],
                double *qout) {
    const int e[3] = {EX[a], EY[a], EZ[a]};
    return ps_calq(&xc[3 * id], r[id], x, e, x0, qout) ? -1 : 0;
}

/* calQ2 (particle_bounceback.F90:144-185): the closed form of the same q, the
 * smaller root of |x + q e - c|^2 = r^2. */
double oracle_calq_closed(const double *xc, const double *r, int id, const double x[3], int a) {
    double dx = x[0] - xc[3 * id], dy = x[1] - xc[3 * id + 1], dz = x[2] - xc[3 * id + 2];
    double ca = (double)(EX[a] * EX[a] + EY[a] * EY[a] + EZ[a] * EZ[a]);
    double cb = 2.0 * (EX[a] * dx + EY[a] * dy + EZ[a] * dz);
    double cc = dx * dx + dy * dy + dz * dz - r[id] * r[id];
    double delta = cb * cb - 4.0 * ca * cc;
    return (-cb - sqrt(delta)) / (2.0 * ca);
}

typedef struct {
    double M[19][19], Minv[19][19];
} pbases;

static void collide_cell(const oracle_particle_params *p, const pbases *b, const double *f, double rho,
                         double u, double v, double w, double *fo) {
    const double snu = 1.0 / p->tau, sq = 8.0 * (2.0 * p->tau - 1.0) / (8.0 * p->tau - 1.0);
    static const int kind[19] = {0, 1, 1, 0, 2, 0, 2, 0, 2, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2};
    if (p->collision == ORACLE_BGK) {
        double us2 = u * u + v * v + w * w;
        for (int a = 0; a < 19; ++a) {
            double un = u * EX[a] + v * EY[a] + w * EZ[a];
            double feq = rho * wgt(a) * (1.0 + 3.0 * un + 4.5 * un * un - 1.5 * us2);
            fo[a] = f[a] - snu * (f[a] - feq);
        }
        return;
    }
    double m[19], meq[19], ms[19];
    for (int q = 0; q < 19; ++q) {
        double acc = 0.0;
        for (int a = 0; a < 19; ++a) acc += b->M[q][a] * f[a];
        m[q] = acc;
    }
    double us2 = u * u + v * v + w * w;
    meq[0] = rho;
    meq[1] = rho * (-11.0 + 19.0 * us2);
    meq[2] = rho * (3.0 - 11.0 / 2.0 * us2);
    meq[3] = rho * u;
    meq[4] = -2.0 / 3.0 * rho * u;
    meq[5] = rho * v;
    meq[6] = -2.0 / 3.0 * rho * v;
    meq[7] = rho * w;
    meq[8] = -2.0 / 3.0 * rho * w;
    meq[9] = rho * (2.0 * u * u - v * v - w * w);
    meq[10] = -0.5 * rho * (2.0 * u * u - v * v - w * w);
    meq[11] = rho * (v * v - w * w);
    meq[12] = p->meq12_rho ? -0.5 * rho * (v * v - w * w) : -0.5 * (v * v - w * w);
    meq[13] = rho * u * v;
    meq[14] = rho * v * w;
    meq[15] = rho * u * w;
    meq[16] = meq[17] = meq[18] = 0.0;
    for (int q = 0; q < 19; ++q) {
        double s = kind[q] == 0 ? 0.0 : (kind[q] == 1 ? snu : sq);
        ms[q] = m[q] - s * (m[q] - meq[q]);
    }
    for (int a = 0; a < 19; ++a) {
        double acc = 0.0;
        for (int q = 0; q < 19; ++q) acc += b->Minv[a][q] * ms[q];
        fo[a] = acc;
    }
}

/* The motion of rigid body `id` seen at point x, split as the scheme uses
 * it: the translation U and the spin part W = Omega x (x - c). */
static void body_motion(const double *xc, const double *U, const double *Om, int id, const double x[3], double u[3],
                        double w[3]) {
    const double r[3] = {x[0] - xc[3 * id], x[1] - xc[3 * id + 1], x[2] - xc[3 * id + 2]};
    u[0] = U[3 * id];
    u[1] = U[3 * id + 1];
    u[2] = U[3 * id + 2];
    ps_spin(&Om[3 * id], r, w);
}

/* f_post of node (i,j,k) direction a for interpolation stencils: the rest
 * equilibrium for points that are not fluid nodes. */
static double fpost_at(const oracle_particle_params *p, const double *fpost, const int *owner, int i, int j, int k,
                       int a) {
    if (!inside(p, i, j, k) || owner[PCELL(p, i, j, k)] >= 0) return wgt(a) * p->rho0;
    return fpost[19 * PCELL(p, i, j, k) + a];
}

struct pstate {
    oracle_particle_params p;
    pbases b;
    double *f, *fpost, *rho, *u, *v, *w;
    int *owner, *owner_new;
    double *xc, *U, *Om, *r;
    double *F, *T;   /* per-particle force, torque of the last step */
    double rho_avg;
};

static double fluid_mean_rho(const struct pstate *s, const int *owner) {
    size_t n = (size_t)s->p.nx * s->p.ny * s->p.nz;
    double sum = 0.0;
    long long cnt = 0;
    for (size_t c = 0; c < n; ++c)
        if (owner[c] < 0) {
            sum += s->rho[c];
            ++cnt;
        }
    return sum / (double)cnt;
}

/* One step.  Returns 0, or a negative code: -2 interpenetration, -3 wall
 * penetration, -4 bad q. */
static int pstep(struct pstate *s) {
    const oracle_particle_params *p = &s->p;
    const int nx = p->nx, ny = p->ny, nz = p->nz, np = p->np;
    const size_t n = (size_t)nx * ny * nz;
    /* collision on fluid nodes (fluid.F90:1-72); nodes are independent, so
     * the loop runs in parallel without changing any value */
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < (long long)n; ++c)
        if (s->owner[c] < 0) collide_cell(p, &s->b, &s->f[19 * c], s->rho[c], s->u[c], s->v[c], s->w[c], &s->fpost[19 * c]);
    /* rhoAvg over fluid nodes of the last macro (particle_bounceback.F90:15-35) */
    s->rho_avg = fluid_mean_rho(s, s->owner);
    /* streaming + walls + particle bounce-back (fluid.F90:75-161,
     * particle_bounceback.F90:37-96); each node writes only its own
     * populations, so the loop runs in parallel */
    int bad_q = 0;
#pragma omp parallel for schedule(static) reduction(| : bad_q)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t c = PCELL(p, i, j, k);
                if (s->owner[c] >= 0) continue;
                for (int bdir = 0; bdir < 19; ++bdir) {
                    int si = i - EX[bdir], sj = j - EY[bdir], sk = k - EZ[bdir];
                    if (!inside(p, si, sj, sk)) {
                        s->f[19 * c + bdir] = s->fpost[19 * c + OPP[bdir]];
                        continue;
                    }
                    int id = s->owner[PCELL(p, si, sj, sk)];
                    if (id < 0) {
                        s->f[19 * c + bdir] = s->fpost[19 * PCELL(p, si, sj, sk) + bdir];
                        continue;
                    }
                    const int a = OPP[bdir]; /* link direction from x into the solid */
                    double x[3] = {(double)(i + 1), (double)(j + 1), (double)(k + 1)}, x0[3], q, ub[3], wb[3];
                    if (oracle_calq(s->xc, s->r, id, x, a, x0, &q)) {
                        bad_q = 1;
                        continue;
                    }
                    body_motion(s->xc, s->U, s->Om, id, x0, ub, wb);
                    const int eb[3] = {EX[bdir], EY[bdir], EZ[bdir]};
                    const double eu = ps_wall_dot(eb, ub, wb);
                    const double fa = s->fpost[19 * c + a];
                    /* q < 1/2 reads f_post(a) at x - e_a, x - 2 e_a; q >= 1/2
                     * reads f_post(-a) at x and x - e_a */
                    const double f1 = fpost_at(p, s->fpost, s->owner, i - EX[a], j - EY[a], k - EZ[a], a);
                    const double f2 = fpost_at(p, s->fpost, s->owner, i - 2 * EX[a], j - 2 * EY[a], k - 2 * EZ[a], a);
                    const double fb = s->fpost[19 * c + bdir];
                    const double fb1 = fpost_at(p, s->fpost, s->owner, i - EX[a], j - EY[a], k - EZ[a], bdir);
                    const double val = ps_interp_bb(q, fa, f1, f2, fb, fb1, wgt(a), s->rho_avg, eu);
                    s->f[19 * c + bdir] = val;
                }
            }
    if (bad_q) return -4;
    /* macro on fluid nodes (fluid.F90:164-184) */
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < (long long)n; ++c) {
        if (s->owner[c] >= 0) continue;
        double r = 0, x = 0, y = 0, z = 0;
        for (int a = 0; a < 19; ++a) {
            double fa = s->f[19 * c + a];
            r += fa;
            x += fa * EX[a];
            y += fa * EY[a];
            z += fa * EZ[a];
        }
        s->rho[c] = r;
        s->u[c] = x / r;
        s->v[c] = y / r;
        s->w[c] = z / r;
    }
    /* momentum exchange (particle_force.F90:31-92), per particle in FP64 */
    for (int q = 0; q < 3 * np; ++q) s->F[q] = s->T[q] = 0.0;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t c = PCELL(p, i, j, k);
                if (s->owner[c] >= 0) continue;
                for (int a = 1; a < 19; ++a) {
                    int ti = i + EX[a], tj = j + EY[a], tk = k + EZ[a];
                    if (!inside(p, ti, tj, tk)) continue;
                    int id = s->owner[PCELL(p, ti, tj, tk)];
                    if (id < 0) continue;
                    double x[3] = {(double)(i + 1), (double)(j + 1), (double)(k + 1)}, x0[3], qq, ub[3], wb[3];
                    if (oracle_calq(s->xc, s->r, id, x, a, x0, &qq)) return -4;
                    body_motion(s->xc, s->U, s->Om, id, x0, ub, wb);
                    const int b = OPP[a];
                    double fp = s->fpost[19 * c + a], fb = s->f[19 * c + b];
                    double dF[3] = {ps_exchange(EX[a], EX[b], ub[0], wb[0], fp, fb),
                                    ps_exchange(EY[a], EY[b], ub[1], wb[1], fp, fb),
                                    ps_exchange(EZ[a], EZ[b], ub[2], wb[2], fp, fb)};
                    double rx = x0[0] - s->xc[3 * id], ry = x0[1] - s->xc[3 * id + 1], rz = x0[2] - s->xc[3 * id + 2];
                    s->F[3 * id] += dF[0];
                    s->F[3 * id + 1] += dF[1];
                    s->F[3 * id + 2] += dF[2];
                    s->T[3 * id] += ry * dF[2] - rz * dF[1];
                    s->T[3 * id + 1] += rz * dF[0] - rx * dF[2];
                    s->T[3 * id + 2] += rx * dF[1] - ry * dF[0];
                }
            }
    /* springs, buoyancy (particle_force.F90:96-191) and kinematics
     * (particle_update.F90:26-42) */
    const double N[3] = {(double)nx, (double)ny, (double)nz};
    double *Fn = (double *)malloc(sizeof(double) * 3 * np);
    for (int a = 0; a < np; ++a) {
        const double V = 4.0 / 3.0 * PI * s->r[a] * s->r[a] * s->r[a];
        /* particle and wall springs kept apart and added last, as Fxij and
         * Fwxij are (particle_force.F90:192-193) */
        double fs[3] = {0, 0, 0}, fw[3] = {0, 0, 0};
        const double scale_p = V * (p->rho_solid - s->rho_avg) * p->gravity / p->stiff_particle;
        for (int b2 = 0; b2 < np; ++b2) {
            if (b2 == a) continue;
            double dx = s->xc[3 * a] - s->xc[3 * b2], dy = s->xc[3 * a + 1] - s->xc[3 * b2 + 1],
                   dz = s->xc[3 * a + 2] - s->xc[3 * b2 + 2];
            double dij = sqrt(dist2(dx, dy, dz));
            if (dij >= s->r[a] + s->r[b2] + p->threshold_particle) continue;
            if (dij < s->r[a] + s->r[b2]) { free(Fn); return -2; }
            const double mag = ps_spring(scale_p, dij - s->r[a] - s->r[b2] - p->threshold_particle,
                                         p->threshold_particle);
            fs[0] += mag * dx / dij;
            fs[1] += mag * dy / dij;
            fs[2] += mag * dz / dij;
        }
        const double scale_w = V * (p->rho_solid - p->rho0) * p->gravity / p->stiff_wall;
        for (int d = 0; d < 3; ++d) {
            double lo = s->xc[3 * a + d] - s->r[a] - 1.0;
            double hi = N[d] - s->xc[3 * a + d] - s->r[a];
            if (lo < 0.0 || hi < 0.0) { free(Fn); return -3; }
            if (lo < p->threshold_wall) fw[d] += ps_spring(scale_w, lo - p->threshold_wall, p->threshold_wall);
            if (hi < p->threshold_wall) fw[d] -= ps_spring(scale_w, hi - p->threshold_wall, p->threshold_wall);
        }
        Fn[3 * a] = s->F[3 * a] + fs[0] + fw[0];
        Fn[3 * a + 1] = s->F[3 * a + 1] + fs[1] + fw[1];
        Fn[3 * a + 2] = s->F[3 * a + 2] - (p->rho_solid - s->rho_avg) * V * p->gravity + fs[2] + fw[2];
    }
    for (int a = 0; a < np; ++a) {
        const double V = 4.0 / 3.0 * PI * s->r[a] * s->r[a] * s->r[a];
        const double m = p->rho_solid * V, I = 0.4 * m * s->r[a] * s->r[a];
        for (int d = 0; d < 3; ++d) {
            double acc = Fn[3 * a + d] / m;
            double Uold = s->U[3 * a + d];
            s->U[3 * a + d] = Uold + acc;
            s->xc[3 * a + d] = s->xc[3 * a + d] + Uold + 0.5 * acc;
            s->Om[3 * a + d] += s->T[3 * a + d] / I;
            s->F[3 * a + d] = Fn[3 * a + d];
        }
    }
    free(Fn);
    /* new map, solid nodes at rest with rho_s (particle_update.F90:69-91) */
    oracle_particle_map(p, s->xc, s->r, s->owner_new);
    for (size_t c = 0; c < n; ++c)
        if (s->owner_new[c] >= 0) {
            s->rho[c] = p->rho_solid;
            s->u[c] = s->v[c] = s->w[c] = 0.0;
        }
    const double rho_avg2 = fluid_mean_rho(s, s->owner_new);
    /* refill of uncovered nodes (particle_update.F90:110-186) */
    double *fnew = NULL;
    size_t *cells = NULL;
    int nref = 0;
    for (size_t c = 0; c < n; ++c)
        if (s->owner[c] >= 0 && s->owner_new[c] < 0) ++nref;
    if (nref) {
        fnew = (double *)malloc(sizeof(double) * 19 * nref);
        cells = (size_t *)malloc(sizeof(size_t) * nref);
    }
    int t = 0;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t c = PCELL(p, i, j, k);
                if (!(s->owner[c] >= 0 && s->owner_new[c] < 0)) continue;
                const int id = s->owner[c];
                double rx = (double)(i + 1) - s->xc[3 * id], ry = (double)(j + 1) - s->xc[3 * id + 1],
                       rz = (double)(k + 1) - s->xc[3 * id + 2];
                double rn = sqrt(dist2(rx, ry, rz));
                double best = 0.0;
                int ec = 0;
                for (int a = 1; a < 19; ++a) {
                    double tn = (rx * EX[a] + ry * EY[a] + rz * EZ[a]) / rn;
                    if (tn > best) {
                        best = tn;
                        ec = a;
                    }
                }
                double fr[19];
                for (int a = 0; a < 19; ++a) {
                    double pt[3];
                    for (int m = 1; m <= 3; ++m) {
                        int pi_ = i + m * EX[ec], pj = j + m * EY[ec], pk = k + m * EZ[ec];
                        int ok = inside(p, pi_, pj, pk) && s->owner[PCELL(p, pi_, pj, pk)] < 0 &&
                                 s->owner_new[PCELL(p, pi_, pj, pk)] < 0;
                        pt[m - 1] = ok ? s->f[19 * PCELL(p, pi_, pj, pk) + a] : wgt(a) * p->rho0;
                    }
                    fr[a] = ps_extrapolate(pt[0], pt[1], pt[2]);
                }
                /* replace the momentum moments with rhoAvg times the body velocity */
                double m[19];
                for (int q = 0; q < 19; ++q) {
                    double acc = 0.0;
                    for (int a = 0; a < 19; ++a) acc += s->b.M[q][a] * fr[a];
                    m[q] = acc;
                }
                double xn[3] = {(double)(i + 1), (double)(j + 1), (double)(k + 1)}, ub[3], wb[3];
                body_motion(s->xc, s->U, s->Om, id, xn, ub, wb);
                m[3] = rho_avg2 * (ub[0] + wb[0]);
                m[5] = rho_avg2 * (ub[1] + wb[1]);
                m[7] = rho_avg2 * (ub[2] + wb[2]);
                for (int a = 0; a < 19; ++a) {
                    double acc = 0.0;
                    for (int q = 0; q < 19; ++q) acc += s->b.Minv[a][q] * m[q];
                    fnew[19 * t + a] = acc;
                }
                cells[t++] = c;
            }
    for (int q = 0; q < nref; ++q) {
        size_t c = cells[q];
        double r = 0, x = 0, y = 0, z = 0;
        for (int a = 0; a < 19; ++a) {
            double fa = fnew[19 * q + a];
            s->f[19 * c + a] = fa;
            r += fa;
            x += fa * EX[a];
            y += fa * EY[a];
            z += fa * EZ[a];
        }
        s->rho[c] = r;
        s->u[c] = x / r;
        s->v[c] = y / r;
        s->w[c] = z / r;
    }
    free(fnew);
    free(cells);
    memcpy(s->owner, s->owner_new, sizeof(int) * n);
    return 0;
}

int oracle_particle_run(const oracle_particle_params *p, int nsteps, double *xc, double *U, double *Om,
                        const double *r, double *F, double *T, int *owner, double *rho, double *u, double *v,
                        double *w, double *rho_avg) {
    struct pstate s;
    memset(&s, 0, sizeof s);
    s.p = *p;
    oracle_d3q19_basis(s.b.M, s.b.Minv);
    const size_t n = (size_t)p->nx * p->ny * p->nz;
    s.f = (double *)calloc(19 * n, sizeof(double));
    s.fpost = (double *)calloc(19 * n, sizeof(double));
    s.owner_new = (int *)malloc(sizeof(int) * n);
    if (!s.f || !s.fpost || !s.owner_new) {
        free(s.f); free(s.fpost); free(s.owner_new);
        return -1;
    }
    s.rho = rho; s.u = u; s.v = v; s.w = w; s.owner = owner;
    s.xc = xc; s.U = U; s.Om = Om; s.r = (double *)r; s.F = F; s.T = T;
    /* initial.F90:78-135: map, rho0 in the fluid, rho_s in solids, u = 0,
     * f = feq(rho0, 0) = w rho0 at fluid nodes */
    oracle_particle_map(p, xc, r, owner);
    for (size_t c = 0; c < n; ++c) {
        rho[c] = owner[c] >= 0 ? p->rho_solid : p->rho0;
        u[c] = v[c] = w[c] = 0.0;
        if (owner[c] < 0)
            for (int a = 0; a < 19; ++a) s.f[19 * c + a] = wgt(a) * p->rho0;
    }
    int rc = 0;
    for (int it = 0; it < nsteps && rc == 0; ++it) rc = pstep(&s);
    if (rho_avg) *rho_avg = s.rho_avg;
    free(s.f); free(s.fpost); free(s.owner_new);
    return rc;
}
