/*
 * thermal_oracle.c -- CPU restatement of the reference's 3-D buoyancy-driven
 * cavity step: D3Q19 MRT flow with Guo forcing (Boussinesq + Coriolis) and
 * D3Q7 MRT temperature (MPI/Buoyancy_driven_cavity/fortran/3d/mpi_blocked/
 * module.F90, initial.F90, evolution_f.F90, evolution_g.F90, check.F90), in the
 * step order of bouyancy3d_mpi.F90:222-248.
 *
 * TEST INFRASTRUCTURE ONLY (parity checker and CPU baseline).
 *
 * Pinned by: golden fixtures written by the reference's own subroutines
 * (oracle/ref/Makefile, thermal harness), the Guo source moments against
 * M (Guo 2002 velocity-space forcing), D3Q7 N N^-1 = I, and the conductive
 * steady state (pure diffusion between the held walls is linear in y).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "lbm_oracle.h"

static const int EX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
static const int EY[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
static const int EZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
static const int OPP[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};

#define CELL(p, i, j, k) ((size_t)(i) + (size_t)(p)->nx * ((size_t)(j) + (size_t)(p)->ny * (size_t)(k)))
#define HCELL(p, i, j, k) \
    ((size_t)(i) + (size_t)((p)->nx + 2) * ((size_t)(j) + (size_t)((p)->ny + 2) * (size_t)(k)))

static double wf(int a) { return a == 0 ? 1.0 / 3.0 : (a <= 6 ? 1.0 / 18.0 : 1.0 / 36.0); }

/* D3Q7 MRT basis (evolution_g.F90:15-21) and the relaxation rates
 * Qd = 3 - sqrt(3), Qnu = 4 sqrt(3) - 6 (module.F90:55). */
static int n7(int k, int a) {
    const int ex = EX[a], ey = EY[a], ez = EZ[a];
    switch (k) {
    case 0: return 1;
    case 1: return ex;
    case 2: return ey;
    case 3: return ez;
    case 4: return a == 0 ? -6 : 1;
    case 5: return 3 * ex * ex - (ex * ex + ey * ey + ez * ez);
    case 6: return ey * ey - ez * ez;
    }
    return 0;
}

void oracle_d3q7_basis(double N[7][7], double Ninv[7][7]) {
    for (int k = 0; k < 7; ++k) {
        double norm2 = 0.0;
        for (int a = 0; a < 7; ++a) {
            N[k][a] = n7(k, a);
            norm2 += N[k][a] * N[k][a];
        }
        for (int a = 0; a < 7; ++a) Ninv[a][k] = N[k][a] / norm2;
    }
}

/* Guo source in moment space (evolution_f.F90:107-125). */
void oracle_guo_source_moments(double u, double v, double w, double Fx, double Fy, double Fz,
                               double *S) {
    double uF = u * Fx + v * Fy + w * Fz;
    S[0] = 0.0;
    S[1] = 38.0 * uF;
    S[2] = -11.0 * uF;
    S[3] = Fx;
    S[4] = -2.0 / 3.0 * Fx;
    S[5] = Fy;
    S[6] = -2.0 / 3.0 * Fy;
    S[7] = Fz;
    S[8] = -2.0 / 3.0 * Fz;
    S[9] = 4.0 * u * Fx - 2.0 * v * Fy - 2.0 * w * Fz;
    S[10] = -2.0 * u * Fx + v * Fy + w * Fz;
    S[11] = 2.0 * v * Fy - 2.0 * w * Fz;
    S[12] = -v * Fy + w * Fz;
    S[13] = u * Fy + v * Fx;
    S[14] = v * Fz + w * Fy;
    S[15] = u * Fz + w * Fx;
    S[16] = S[17] = S[18] = 0.0;
}

typedef struct {
    double M[19][19], Minv[19][19];
    double N[7][7], Ninv[7][7];
} bases;

static void get_bases(bases *b) {
    oracle_d3q19_basis(b->M, b->Minv);
    oracle_d3q7_basis(b->N, b->Ninv);
}

/* Initial state (initial.F90:86-186): rho = 1, u = 0, T = 0 except the held
 * wall layers (Thot at y = 1 and Tcold at y = ny for LeftRightWallsConstT,
 * bottom / top for TopBottomPlatesConstT); f = feq(rho, u),
 * g = omegaT T (1 + 21/(6 + paraA) e.u) with omegaT = (1 - paraA)/7, (paraA + 6)/42. */
void oracle_thermal_initial_T(const oracle_thermal_params *p, double *T) {
    for (int k = 0; k < p->nz; ++k)
        for (int j = 0; j < p->ny; ++j)
            for (int i = 0; i < p->nx; ++i) {
                int q = p->tconst_axis == 0 ? i : (p->tconst_axis == 1 ? j : k);
                int nq = p->tconst_axis == 0 ? p->nx : (p->tconst_axis == 1 ? p->ny : p->nz);
                double t = 0.0;
                if (p->tconst_axis >= 0 && q == 0) t = p->t_hot;
                if (p->tconst_axis >= 0 && q == nq - 1) t = p->t_cold;
                T[CELL(p, i, j, k)] = t;
            }
}

static void init_populations(const oracle_thermal_params *p, double *f, double *g, const double *rho,
                             const double *u, const double *v, const double *w, const double *T) {
    size_t n = (size_t)p->nx * p->ny * p->nz;
    const double wt0 = (1.0 - p->para_a) / 7.0, wt1 = (p->para_a + 6.0) / 42.0;
    for (size_t c = 0; c < n; ++c) {
        double us2 = u[c] * u[c] + v[c] * v[c] + w[c] * w[c];
        for (int a = 0; a < 19; ++a) {
            double un = u[c] * EX[a] + v[c] * EY[a] + w[c] * EZ[a];
            f[19 * c + a] = rho[c] * wf(a) * (1.0 + 3.0 * un + 4.5 * un * un - 1.5 * us2);
        }
        for (int a = 0; a < 7; ++a) {
            double un = u[c] * EX[a] + v[c] * EY[a] + w[c] * EZ[a];
            g[7 * c + a] = (a == 0 ? wt0 : wt1) * T[c] * (1.0 + 21.0 / (6.0 + p->para_a) * un);
        }
    }
}

/* collision (evolution_f.F90:1-224): forces from the old fields, MRT with the
 * Guo source; collisionT (evolution_g.F90:1-57): D3Q7 MRT with the old u, T. */
static void collide_all(const oracle_thermal_params *p, const bases *b, const double *f, const double *g,
                        const double *rho, const double *u, const double *v, const double *w,
                        const double *T, double *Fx, double *Fy, double *Fz, double *fpost,
                        double *gpost) {
    const double snu = 1.0 / p->tau, sq = 8.0 * (2.0 * p->tau - 1.0) / (8.0 * p->tau - 1.0);
    static const int kind[19] = {0, 1, 1, 0, 2, 0, 2, 0, 2, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2};
    const double qd = 3.0 - sqrt(3.0), qnu = 4.0 * sqrt(3.0) - 6.0;
    const double q7[7] = {0.0, qd, qd, qd, qnu, qnu, qnu};
#pragma omp parallel for schedule(static)
    for (int k = 0; k < p->nz; ++k)
        for (int j = 0; j < p->ny; ++j)
            for (int i = 0; i < p->nx; ++i) {
                size_t c = CELL(p, i, j, k);
                const double r = rho[c], uu = u[c], vv = v[c], ww = w[c];
                Fx[c] = -2.0 * r * vv * p->omega_rot;
                Fy[c] = 2.0 * r * uu * p->omega_rot;
                Fz[c] = r * p->g_beta * (T[c] - p->t_ref);
                double m[19], meq[19], S[19];
                for (int q = 0; q < 19; ++q) {
                    double acc = 0.0;
                    for (int a = 0; a < 19; ++a) acc += b->M[q][a] * f[19 * c + a];
                    m[q] = acc;
                }
                double us2 = uu * uu + vv * vv + ww * ww;
                meq[0] = r;
                meq[1] = -11.0 * r + 19.0 * r * us2;
                meq[2] = 3.0 * r - 11.0 / 2.0 * r * us2;
                meq[3] = r * uu;
                meq[4] = -2.0 / 3.0 * meq[3];
                meq[5] = r * vv;
                meq[6] = -2.0 / 3.0 * meq[5];
                meq[7] = r * ww;
                meq[8] = -2.0 / 3.0 * meq[7];
                meq[9] = r * (2.0 * uu * uu - vv * vv - ww * ww);
                meq[10] = -0.5 * meq[9];
                meq[11] = r * (vv * vv - ww * ww);
                meq[12] = -0.5 * meq[11];
                meq[13] = r * (uu * vv);
                meq[14] = r * (vv * ww);
                meq[15] = r * (ww * uu);
                meq[16] = meq[17] = meq[18] = 0.0;
                oracle_guo_source_moments(uu, vv, ww, Fx[c], Fy[c], Fz[c], S);
                double *fo = &fpost[19 * HCELL(p, i + 1, j + 1, k + 1)];
                double ms[19];
                for (int q = 0; q < 19; ++q) {
                    double s = kind[q] == 0 ? 0.0 : (kind[q] == 1 ? snu : sq);
                    ms[q] = m[q] - s * (m[q] - meq[q]) + (1.0 - 0.5 * s) * S[q];
                }
                for (int a = 0; a < 19; ++a) {
                    double acc = 0.0;
                    for (int q = 0; q < 19; ++q) acc += b->Minv[a][q] * ms[q];
                    fo[a] = acc;
                }
                /* D3Q7 */
                double nn[7], neq[7], ns[7];
                for (int q = 0; q < 7; ++q) {
                    double acc = 0.0;
                    for (int a = 0; a < 7; ++a) acc += b->N[q][a] * g[7 * c + a];
                    nn[q] = acc;
                }
                const double t = T[c];
                neq[0] = t;
                neq[1] = t * uu;
                neq[2] = t * vv;
                neq[3] = t * ww;
                neq[4] = t * p->para_a;
                neq[5] = neq[6] = 0.0;
                for (int q = 0; q < 7; ++q) ns[q] = nn[q] - q7[q] * (nn[q] - neq[q]);
                double *go = &gpost[7 * HCELL(p, i + 1, j + 1, k + 1)];
                for (int a = 0; a < 7; ++a) {
                    double acc = 0.0;
                    for (int q = 0; q < 7; ++q) acc += b->Ninv[a][q] * ns[q];
                    go[a] = acc;
                }
            }
}

/* streaming + bounceback of f (no-slip walls everywhere, evolution_f.F90:227-343)
 * and of g (evolution_g.F90:60-197): adiabatic walls bounce back; held walls
 * use anti-bounce-back g(a) = -g_post(abar) + (6 + paraA)/21 Twall. */
static void stream_all(const oracle_thermal_params *p, const double *fpost, const double *gpost, double *f,
                       double *g) {
    const double cw = (6.0 + p->para_a) / 21.0;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < p->nz; ++k)
        for (int j = 0; j < p->ny; ++j)
            for (int i = 0; i < p->nx; ++i) {
                size_t c = CELL(p, i, j, k);
                const size_t hc = HCELL(p, i + 1, j + 1, k + 1);
                for (int a = 0; a < 19; ++a) {
                    int si = i - EX[a], sj = j - EY[a], sk = k - EZ[a];
                    int out = si < 0 || si >= p->nx || sj < 0 || sj >= p->ny || sk < 0 || sk >= p->nz;
                    f[19 * c + a] = out ? fpost[19 * hc + OPP[a]]
                                        : fpost[19 * HCELL(p, si + 1, sj + 1, sk + 1) + a];
                }
                for (int a = 0; a < 7; ++a) {
                    int si = i - EX[a], sj = j - EY[a], sk = k - EZ[a];
                    int out = si < 0 || si >= p->nx || sj < 0 || sj >= p->ny || sk < 0 || sk >= p->nz;
                    if (!out) {
                        g[7 * c + a] = gpost[7 * HCELL(p, si + 1, sj + 1, sk + 1) + a];
                        continue;
                    }
                    int axis = EX[a] != 0 ? 0 : (EY[a] != 0 ? 1 : 2);
                    double bb = gpost[7 * hc + OPP[a]];
                    if (axis == p->tconst_axis) {
                        /* the pull leaves through the low wall when e > 0 */
                        int low = (axis == 0 ? EX[a] : (axis == 1 ? EY[a] : EZ[a])) > 0;
                        g[7 * c + a] = -bb + cw * (low ? p->t_hot : p->t_cold);
                    } else {
                        g[7 * c + a] = bb;
                    }
                }
            }
}

/* macro (evolution_f.F90:346-370): u = (sum f e + F/2) / rho; macroT: T = sum g. */
static void macro_all(const oracle_thermal_params *p, const double *f, const double *g, const double *Fx,
                      const double *Fy, const double *Fz, double *rho, double *u, double *v, double *w,
                      double *T) {
    size_t n = (size_t)p->nx * p->ny * p->nz;
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < n; ++c) {
        const double *fc = &f[19 * c];
        double r = 0.0, x = 0.0, y = 0.0, z = 0.0, t = 0.0;
        for (int a = 0; a < 19; ++a) {
            r += fc[a];
            x += fc[a] * EX[a];
            y += fc[a] * EY[a];
            z += fc[a] * EZ[a];
        }
        rho[c] = r;
        u[c] = (x + 0.5 * Fx[c]) / r;
        v[c] = (y + 0.5 * Fy[c]) / r;
        w[c] = (z + 0.5 * Fz[c]) / r;
        for (int a = 0; a < 7; ++a) t += g[7 * c + a];
        T[c] = t;
    }
}

int oracle_thermal_run(const oracle_thermal_params *p, int nsteps, int check_every, double *rho, double *u,
                       double *v, double *w, double *T, int from_fields, double *errs, int nerr_max,
                       int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    size_t n = (size_t)p->nx * p->ny * p->nz;
    size_t nh = (size_t)(p->nx + 2) * (p->ny + 2) * (p->nz + 2);
    bases *b = (bases *)malloc(sizeof(bases));
    double *f = (double *)malloc(19 * n * sizeof(double));
    double *g = (double *)malloc(7 * n * sizeof(double));
    double *fpost = (double *)calloc(19 * nh, sizeof(double));
    double *gpost = (double *)calloc(7 * nh, sizeof(double));
    double *F = (double *)calloc(3 * n, sizeof(double));
    double *prev = (double *)calloc(4 * n, sizeof(double));
    if (!b || !f || !g || !fpost || !gpost || !F || !prev) {
        free(b); free(f); free(g); free(fpost); free(gpost); free(F); free(prev);
        return -1;
    }
    get_bases(b);
    if (!from_fields) {
        for (size_t c = 0; c < n; ++c) {
            rho[c] = 1.0;
            u[c] = v[c] = w[c] = 0.0;
        }
        oracle_thermal_initial_T(p, T);
    }
    init_populations(p, f, g, rho, u, v, w, T);
    int nchk = 0;
    for (int itc = 1; itc <= nsteps; ++itc) {
        collide_all(p, b, f, g, rho, u, v, w, T, F, F + n, F + 2 * n, fpost, gpost);
        stream_all(p, fpost, gpost, f, g);
        macro_all(p, f, g, F, F + n, F + 2 * n, rho, u, v, w, T);
        if (check_every > 0 && itc % check_every == 0) {
            /* check.F90:16-41: errorU with w, errorT = sum|T - Tp| / sum|T| */
            double e1 = 0, e2 = 0, e5 = 0, e6 = 0;
            for (size_t c = 0; c < n; ++c) {
                double du = u[c] - prev[c], dv = v[c] - prev[n + c], dw = w[c] - prev[2 * n + c];
                e1 += du * du + dv * dv + dw * dw;
                e2 += u[c] * u[c] + v[c] * v[c] + w[c] * w[c];
                e5 += fabs(T[c] - prev[3 * n + c]);
                e6 += fabs(T[c]);
                prev[c] = u[c];
                prev[n + c] = v[c];
                prev[2 * n + c] = w[c];
                prev[3 * n + c] = T[c];
            }
            if (errs && 2 * nchk + 1 < 2 * nerr_max) {
                errs[2 * nchk] = sqrt(e1) / sqrt(e2);
                errs[2 * nchk + 1] = e5 / e6;
            }
            ++nchk;
        }
    }
    free(b); free(f); free(g); free(fpost); free(gpost); free(F); free(prev);
    return nchk;
}

/* Nusselt numbers of the buoyancy cavity, FP64:
 *   nu_vol  = <w T> nz / kappa + 1 -- calNuRe's volume Nusselt number
 *             (RaNu.F90:13-23): the velocity along gravity (z) and the z
 *             extent, whichever walls are held;
 *   nu_wall = mean over the hot wall (axis a = tconst_axis, hot at the low
 *             side) of (Thot - T(first layer)) / (1/2) * N_a / (Thot - Tcold):
 *             the conductive flux through the wall at the half-link.
 * `w` is the z velocity; the volume sum runs over the cells in storage
 * order (x fastest), as RaNu.F90's k, j, i loops do. */
void oracle_thermal_nusselt(const oracle_thermal_params *p, const double *w, const double *T, double kappa,
                            double *nu_vol, double *nu_wall) {
    size_t n = (size_t)p->nx * p->ny * p->nz;
    const int ax = p->tconst_axis;
    const int na = ax == 0 ? p->nx : (ax == 1 ? p->ny : p->nz);
    double s = 0.0;
    for (size_t c = 0; c < n; ++c) s += w[c] * T[c];
    *nu_vol = s / (double)n * (double)p->nz / kappa + 1.0;
    double sw = 0.0;
    size_t nwall = 0;
    for (int k = 0; k < p->nz; ++k)
        for (int j = 0; j < p->ny; ++j)
            for (int i = 0; i < p->nx; ++i) {
                int q = ax == 0 ? i : (ax == 1 ? j : k);
                if (q != 0) continue;
                sw += (p->t_hot - T[CELL(p, i, j, k)]) * 2.0;
                ++nwall;
            }
    *nu_wall = sw / (double)nwall * (double)na / (p->t_hot - p->t_cold);
}
