/*
 * lbm_oracle.c -- CPU restatement of the reference's D3Q19 lid-driven cavity
 * step (MPI/Lid_driven_cavity/fortran/3d/mpi_3d_blocked/ sources).
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker and the CPU baseline).  The
 * product library never links it.
 *
 * Parity status of each function (how it is pinned):
 *   initial, collision (MRT), streaming, bounceback, macro: compared with the
 *     reference Fortran itself, compiled unmodified from its sources by
 *     oracle/ref/Makefile, at several grid shapes and step counts (golden
 *     fixtures in tests/golden/).
 *   collision (BGK): the reference's BGK lines are commented out
 *     (collision.f90:192-198) and cannot be compiled unedited.  Pinned by
 *     closed forms instead: BGK equals MRT with every rate set to 1/tau and
 *     meq(12) carrying rho; feq reproduces the reference meq through M; rest
 *     state f = w is a fixed point.
 *   check: closed form (hand-computable residual on constructed fields).
 *
 * The moment basis is built from the d'Humieres (2002) polynomial definition
 * rather than from the reference's written-out rows, so that comparing with
 * the reference pins both.
 */
#include "lbm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* D3Q19 velocity set, weights (commondata.f90:29-40). */
static const int EX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
static const int EY[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
static const int EZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};

static double weight(int a) {
    if (a == 0) return 1.0 / 3.0;
    return a <= 6 ? 1.0 / 18.0 : 1.0 / 36.0;
}

static double g_loop_seconds = 0.0;
double oracle_last_loop_seconds(void) { return g_loop_seconds; }

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ---- moment basis ------------------------------------------------------ */

/* Row k of the D3Q19 MRT basis evaluated at velocity e (integer valued). */
static int basis_row(int k, int ex, int ey, int ez) {
    int e2 = ex * ex + ey * ey + ez * ez;
    switch (k) {
    case 0: return 1;
    case 1: return 19 * e2 - 30;
    case 2: return (21 * e2 * e2 - 53 * e2 + 24) / 2;
    case 3: return ex;
    case 4: return (5 * e2 - 9) * ex;
    case 5: return ey;
    case 6: return (5 * e2 - 9) * ey;
    case 7: return ez;
    case 8: return (5 * e2 - 9) * ez;
    case 9: return 3 * ex * ex - e2;
    case 10: return (3 * e2 - 5) * (3 * ex * ex - e2);
    case 11: return ey * ey - ez * ez;
    case 12: return (3 * e2 - 5) * (ey * ey - ez * ez);
    case 13: return ex * ey;
    case 14: return ey * ez;
    case 15: return ex * ez;
    case 16: return (ey * ey - ez * ez) * ex;
    case 17: return (ez * ez - ex * ex) * ey;
    case 18: return (ex * ex - ey * ey) * ez;
    }
    return 0;
}

void oracle_d3q19_basis(double M[19][19], double Minv[19][19]) {
    for (int k = 0; k < 19; ++k) {
        double norm2 = 0.0;
        for (int a = 0; a < 19; ++a) {
            M[k][a] = (double)basis_row(k, EX[a], EY[a], EZ[a]);
            norm2 += M[k][a] * M[k][a];
        }
        for (int a = 0; a < 19; ++a) Minv[a][k] = M[k][a] / norm2;
    }
}

/* Sparse copies of M and M^-1 (zeros dropped) for the collision loops. */
typedef struct {
    int n[19];
    int idx[19][19];
    double c[19][19];
} sparse19;

static sparse19 g_M, g_Minv;
static int g_basis_ready = 0;

static void build_sparse(void) {
    if (g_basis_ready) return;
    double M[19][19], Minv[19][19];
    oracle_d3q19_basis(M, Minv);
    for (int r = 0; r < 19; ++r) {
        g_M.n[r] = g_Minv.n[r] = 0;
        for (int c = 0; c < 19; ++c) {
            if (M[r][c] != 0.0) {
                g_M.idx[r][g_M.n[r]] = c;
                g_M.c[r][g_M.n[r]++] = M[r][c];
            }
            if (Minv[r][c] != 0.0) {
                g_Minv.idx[r][g_Minv.n[r]] = c;
                g_Minv.c[r][g_Minv.n[r]++] = Minv[r][c];
            }
        }
    }
    g_basis_ready = 1;
}

/* ---- stages ------------------------------------------------------------ */

#define CELL(p, i, j, k) ((size_t)(i) + (size_t)(p)->nx * ((size_t)(j) + (size_t)(p)->ny * (size_t)(k)))
/* f_post with halo: 0..nx+1 etc. */
#define HCELL(p, i, j, k) \
    ((size_t)(i) + (size_t)((p)->nx + 2) * ((size_t)(j) + (size_t)((p)->ny + 2) * (size_t)(k)))

/* Equilibrium (initial.f90:63-73). */
static void feq19(double rho, double u, double v, double w, double *feq) {
    double us2 = u * u + v * v + w * w;
    for (int a = 0; a < 19; ++a) {
        double un = u * EX[a] + v * EY[a] + w * EZ[a];
        feq[a] = rho * weight(a) * (1.0 + 3.0 * un + 4.5 * un * un - 1.5 * us2);
    }
}

void oracle_feq_init(const oracle_cavity_params *p, double *f, const double *rho,
                     const double *u, const double *v, const double *w) {
    for (int k = 0; k < p->nz; ++k)
        for (int j = 0; j < p->ny; ++j)
            for (int i = 0; i < p->nx; ++i) {
                size_t c = CELL(p, i, j, k);
                feq19(rho[c], u[c], v[c], w[c], &f[19 * c]);
            }
}

/* Reference initial state (initial.f90:46-73): rho = rho0, u = v = w = 0,
 * except u = U0 on the top fluid layer k = nz (initial.f90:54-61); then
 * f = feq(rho, u). */
void oracle_cavity_initial(const oracle_cavity_params *p, double *f, double *rho,
                           double *u, double *v, double *w) {
    size_t n = (size_t)p->nx * p->ny * p->nz;
    for (size_t c = 0; c < n; ++c) {
        rho[c] = p->rho0;
        u[c] = v[c] = w[c] = 0.0;
    }
    for (int j = 0; j < p->ny; ++j)
        for (int i = 0; i < p->nx; ++i) u[CELL(p, i, j, p->nz - 1)] = p->u_lid;
    oracle_feq_init(p, f, rho, u, v, w);
}

/* Equilibrium moments (collision.f90:73-91).  meq(12) has no rho in the
 * cavity code (collision.f90:85); the thermal code has it (evolution_f.F90:75). */
static void meq19(double rho, double u, double v, double w, int meq12_rho, double *meq) {
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
    meq[12] = meq12_rho ? -0.5 * rho * (v * v - w * w) : -0.5 * (v * v - w * w);
    meq[13] = rho * u * v;
    meq[14] = rho * v * w;
    meq[15] = rho * u * w;
    meq[16] = meq[17] = meq[18] = 0.0;
}

/* Relaxation rates s(0:18) (collision.f90:94-112; commondata.f90:42). */
static void rates19(double tau, double *s) {
    double snu = 1.0 / tau;
    double sq = 8.0 * (2.0 * tau - 1.0) / (8.0 * tau - 1.0);
    static const int kind[19] = {0, 1, 1, 0, 2, 0, 2, 0, 2, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2};
    for (int a = 0; a < 19; ++a) s[a] = kind[a] == 0 ? 0.0 : (kind[a] == 1 ? snu : sq);
}

/* collision.f90:15-202 (MRT: m = M f, m* = m - s (m - meq), f_post = M^-1 m*)
 * and the BGK form of collision.f90:192-198 (f_post = f - (f - feq)/tau).
 * Writes the interior of the haloed f_post. */
void oracle_cavity_collision(const oracle_cavity_params *p, const double *f,
                             const double *rho, const double *u, const double *v,
                             const double *w, double *fpost) {
    build_sparse();
    double s[19];
    rates19(p->tau, s);
    const double snu = 1.0 / p->tau;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < p->nz; ++k)
        for (int j = 0; j < p->ny; ++j)
            for (int i = 0; i < p->nx; ++i) {
                size_t c = CELL(p, i, j, k);
                const double *fc = &f[19 * c];
                double *fo = &fpost[19 * HCELL(p, i + 1, j + 1, k + 1)];
                if (p->collision == ORACLE_BGK) {
                    double feq[19];
                    feq19(rho[c], u[c], v[c], w[c], feq);
                    for (int a = 0; a < 19; ++a) fo[a] = fc[a] - snu * (fc[a] - feq[a]);
                } else {
                    double m[19], meq[19], ms[19];
                    for (int r = 0; r < 19; ++r) {
                        double acc = 0.0;
                        for (int q = 0; q < g_M.n[r]; ++q) acc += g_M.c[r][q] * fc[g_M.idx[r][q]];
                        m[r] = acc;
                    }
                    meq19(rho[c], u[c], v[c], w[c], p->meq12_rho, meq);
                    for (int r = 0; r < 19; ++r) ms[r] = m[r] - s[r] * (m[r] - meq[r]);
                    for (int a = 0; a < 19; ++a) {
                        double acc = 0.0;
                        for (int q = 0; q < g_Minv.n[a]; ++q)
                            acc += g_Minv.c[a][q] * ms[g_Minv.idx[a][q]];
                        fo[a] = acc;
                    }
                }
            }
}

/* streaming.f90:8-20: pull, f(a, x) = f_post(a, x - e_a).  Sources in the
 * halo hold nothing meaningful on wall faces; bounceback overwrites every such
 * value.  On periodic axes the source wraps to the opposite side, which is what
 * the halo exchange of ex_sendrecv.f90 delivers when the neighbour across the
 * face is the opposite face. */
static int wrap(int s, int n, int periodic) {
    if (!periodic) return s;
    if (s < 0) return s + n;
    if (s >= n) return s - n;
    return s;
}

void oracle_cavity_streaming(const oracle_cavity_params *p, const double *fpost,
                             double *f) {
    const int px = p->periodic_mask & 1, py = (p->periodic_mask >> 1) & 1, pz = (p->periodic_mask >> 2) & 1;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < p->nz; ++k)
        for (int j = 0; j < p->ny; ++j)
            for (int i = 0; i < p->nx; ++i) {
                size_t c = CELL(p, i, j, k);
                for (int a = 0; a < 19; ++a) {
                    int si = wrap(i - EX[a], p->nx, px), sj = wrap(j - EY[a], p->ny, py),
                        sk = wrap(k - EZ[a], p->nz, pz);
                    f[19 * c + a] = fpost[19 * HCELL(p, si + 1, sj + 1, sk + 1) + a];
                }
            }
}

/* Opposite directions: 1<->2, 3<->4, 5<->6, 7<->10, 8<->9, 11<->14, 12<->13,
 * 15<->18, 16<->17. */
static const int OPP[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};

/* bounce_back.f90:6-83: halfway bounce-back on the six walls, in the
 * reference order (x-, x+, y-, y+, z-, then the moving lid z+ last so it wins
 * on pops 13/14 at the lid edges).  Lid: f(14) = f_post(11) - rho/6*U0,
 * f(13) = f_post(12) - rho/6*(-U0), rho = the previous step's macro value
 * (bounce_back.f90:72-83). */
void oracle_cavity_bounceback(const oracle_cavity_params *p, const double *fpost,
                              const double *rho, double *f) {
    const int nx = p->nx, ny = p->ny, nz = p->nz;
    /* incoming populations per wall: those whose e points into the domain */
    for (int wall = 0; wall < 6; ++wall) {
        int axis = wall / 2, hi = wall % 2;
        if ((p->periodic_mask >> axis) & 1) continue;  /* no wall on a periodic axis */
        int n0 = axis == 0 ? ny : nx, n1 = axis == 2 ? ny : nz;
        for (int b = 0; b < n1; ++b)
            for (int a0 = 0; a0 < n0; ++a0) {
                int i, j, k;
                if (axis == 0) { i = hi ? nx - 1 : 0; j = a0; k = b; }
                else if (axis == 1) { i = a0; j = hi ? ny - 1 : 0; k = b; }
                else { i = a0; j = b; k = hi ? nz - 1 : 0; }
                size_t c = CELL(p, i, j, k);
                const double *fp = &fpost[19 * HCELL(p, i + 1, j + 1, k + 1)];
                for (int a = 1; a < 19; ++a) {
                    int e = axis == 0 ? EX[a] : (axis == 1 ? EY[a] : EZ[a]);
                    if (e != (hi ? -1 : 1)) continue;
                    double val = fp[OPP[a]];
                    if (axis == 2 && hi && (a == 13 || a == 14)) {
                        double uw = a == 14 ? p->u_lid : -p->u_lid;
                        val = fp[OPP[a]] - rho[c] / 6.0 * uw;
                    }
                    f[19 * c + a] = val;
                }
            }
    }
}

/* macro.f90:6-25: rho = sum f, u = sum f e / rho, accumulated a = 0..18. */
void oracle_cavity_macro(const oracle_cavity_params *p, const double *f,
                         double *rho, double *u, double *v, double *w) {
    size_t n = (size_t)p->nx * p->ny * p->nz;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < n; ++c) {
        double r = 0.0, x = 0.0, y = 0.0, z = 0.0;
        for (int a = 0; a < 19; ++a) {
            double fa = f[19 * c + a];
            r += fa;
            x += fa * EX[a];
            y += fa * EY[a];
            z += fa * EZ[a];
        }
        rho[c] = r;
        u[c] = x / r;
        v[c] = y / r;
        w[c] = z / r;
    }
}

/* check.f90:9-34: errorU = sqrt(sum (u-up)^2 + (v-vp)^2) / sqrt(sum |u|^2);
 * the numerator has no w term (check.f90:15).  Then up = u, vp = v. */
double oracle_cavity_check(const oracle_cavity_params *p, const double *u,
                           const double *v, const double *w, double *up,
                           double *vp) {
    size_t n = (size_t)p->nx * p->ny * p->nz;
    double e1 = 0.0, e2 = 0.0;
    for (size_t c = 0; c < n; ++c) {
        e1 += (u[c] - up[c]) * (u[c] - up[c]) + (v[c] - vp[c]) * (v[c] - vp[c]);
        e2 += u[c] * u[c] + v[c] * v[c] + w[c] * w[c];
    }
    memcpy(up, u, n * sizeof(double));
    memcpy(vp, v, n * sizeof(double));
    return sqrt(e1) / sqrt(e2);
}

static int run_loop(const oracle_cavity_params *p, int nsteps, int check_every,
                    double *f, double *rho, double *u, double *v, double *w,
                    double *errs, int nerr_max) {
    size_t n = (size_t)p->nx * p->ny * p->nz;
    size_t nh = (size_t)(p->nx + 2) * (p->ny + 2) * (p->nz + 2);
    double *fpost = (double *)calloc(19 * nh, sizeof(double));
    double *up = (double *)calloc(n, sizeof(double));
    double *vp = (double *)calloc(n, sizeof(double));
    if (!fpost || !up || !vp) {
        free(fpost); free(up); free(vp);
        return -1;
    }
    int nchk = 0;
    double t0 = now_s();
    /* main.f90:85-103 loop body */
    for (int itc = 1; itc <= nsteps; ++itc) {
        oracle_cavity_collision(p, f, rho, u, v, w, fpost);
        oracle_cavity_streaming(p, fpost, f);
        oracle_cavity_bounceback(p, fpost, rho, f);
        oracle_cavity_macro(p, f, rho, u, v, w);
        if (check_every > 0 && itc % check_every == 0) {
            double e = oracle_cavity_check(p, u, v, w, up, vp);
            if (errs && nchk < nerr_max) errs[nchk] = e;
            ++nchk;
        }
    }
    g_loop_seconds = now_s() - t0;
    free(fpost); free(up); free(vp);
    return nchk;
}

static void set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

/* Threads of the OpenMP loops that follow (the particle oracles' per-node
 * loops); n <= 0 leaves the setting as it is. */
void oracle_set_threads(int n) { set_threads(n); }

int oracle_cavity_run(const oracle_cavity_params *p, int nsteps, int check_every,
                      double *rho, double *u, double *v, double *w,
                      double *errs, int nerr_max, int nthreads) {
    set_threads(nthreads);
    size_t n = (size_t)p->nx * p->ny * p->nz;
    double *f = (double *)malloc(19 * n * sizeof(double));
    if (!f) return -1;
    oracle_cavity_initial(p, f, rho, u, v, w);
    int r = run_loop(p, nsteps, check_every, f, rho, u, v, w, errs, nerr_max);
    free(f);
    return r;
}

int oracle_cavity_run_from(const oracle_cavity_params *p, int nsteps,
                           double *rho, double *u, double *v, double *w,
                           int nthreads) {
    set_threads(nthreads);
    size_t n = (size_t)p->nx * p->ny * p->nz;
    double *f = (double *)malloc(19 * n * sizeof(double));
    if (!f) return -1;
    oracle_feq_init(p, f, rho, u, v, w);
    int r = run_loop(p, nsteps, 0, f, rho, u, v, w, NULL, 0);
    free(f);
    return r;
}
