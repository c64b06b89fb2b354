/*
 * lbm_oracle.h -- CPU restatement of the reference lattice Boltzmann step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * path: only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py may load it.  The product (libmglc_lbm.so) never links or calls it.
 *
 * Layout follows the reference's Fortran arrays, 0-based:
 *   f[a + 19*(i + nx*(j + ny*k))]        f(0:18, nx, ny, nz)
 *   fpost with a one-cell halo:           f_post(0:18, 0:nx+1, 0:ny+1, 0:nz+1)
 *   rho/u/v/w[i + nx*(j + ny*k)]          rho(nx, ny, nz)   (x fastest)
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORACLE_MRT = 0, ORACLE_BGK = 1 };

typedef struct {
    int nx, ny, nz;
    int collision;       /* ORACLE_MRT (reference default) or ORACLE_BGK */
    double tau;          /* tauf */
    double u_lid;        /* U0 */
    double rho0;
    int meq12_rho;       /* 0: cavity meq(12) without rho (collision.f90:85) */
    int periodic_mask;   /* bit 0/1/2: x/y/z faces exchange with the opposite face
                            (the halo a periodic Cartesian exchange would deliver)
                            instead of bouncing back; 0 = the reference cavity */
} oracle_cavity_params;

/* Whole-run driver: reference initial, then nsteps of
 * collision -> streaming -> bounceback -> macro, with check() every
 * check_every steps (0 = never).  Outputs rho,u,v,w after the last step; errs
 * (may be NULL) receives errorU at each check, nerr_max entries at most.
 * Returns the number of checks done, or -1 on allocation failure. */
/* Threads of the OpenMP loops that follow (the particle oracles' per-node
 * loops); n <= 0 leaves the setting as it is. */
void oracle_set_threads(int n);

int oracle_cavity_run(const oracle_cavity_params *p, int nsteps, int check_every,
                      double *rho, double *u, double *v, double *w,
                      double *errs, int nerr_max, int nthreads);

/* Same run from given initial fields (rho,u,v,w in, overwritten with the
 * result).  f starts at equilibrium of those fields. */
int oracle_cavity_run_from(const oracle_cavity_params *p, int nsteps,
                           double *rho, double *u, double *v, double *w,
                           int nthreads);

/* Per-stage entry points, for tests that pin single stages. */
void oracle_cavity_initial(const oracle_cavity_params *p, double *f, double *rho,
                           double *u, double *v, double *w);
void oracle_feq_init(const oracle_cavity_params *p, double *f, const double *rho,
                     const double *u, const double *v, const double *w);
void oracle_cavity_collision(const oracle_cavity_params *p, const double *f,
                             const double *rho, const double *u, const double *v,
                             const double *w, double *fpost);
void oracle_cavity_streaming(const oracle_cavity_params *p, const double *fpost,
                             double *f);
void oracle_cavity_bounceback(const oracle_cavity_params *p, const double *fpost,
                              const double *rho, double *f);
void oracle_cavity_macro(const oracle_cavity_params *p, const double *f,
                         double *rho, double *u, double *v, double *w);
double oracle_cavity_check(const oracle_cavity_params *p, const double *u,
                           const double *v, const double *w, double *up,
                           double *vp);

/* The D3Q19 moment basis built from its polynomial definition, and its
 * inverse via row norms (M^-1 = M^T diag(1/|row|^2)).  Exposed so tests can
 * check M M^-1 = I and the equilibrium moments M feq = meq. */
void oracle_d3q19_basis(double M[19][19], double Minv[19][19]);

/* ---- buoyancy-driven cavity: D3Q19 + Guo forcing, D3Q7 temperature ---- */
typedef struct {
    int nx, ny, nz;
    int tconst_axis;     /* walls held at t_hot (low side) / t_cold (high side):
                            1 = y (LeftRightWallsConstT), 2 = z (TopBottomPlatesConstT),
                            0 = x, -1 = none; the other walls are adiabatic */
    double tau;          /* tauf (module.F90:19) */
    double para_a;       /* paraA = 42 sqrt(3) kappa - 6 (module.F90:27) */
    double g_beta;       /* gBeta (module.F90:28-29) */
    double omega_rot;    /* Coriolis rate omegaRatating (module.F90:25); 0 = off */
    double t_hot, t_cold, t_ref;
} oracle_thermal_params;

/* Reference initial T (or the given fields when from_fields), then nsteps of
 * collision -> streaming -> bounceback -> collisionT -> streamingT ->
 * bouncebackT -> macro -> macroT (bouyancy3d_mpi.F90:222-248).  errs receives
 * (errorU, errorT) pairs of check.F90 every check_every steps. */
int oracle_thermal_run(const oracle_thermal_params *p, int nsteps, int check_every, double *rho,
                       double *u, double *v, double *w, double *T, int from_fields, double *errs,
                       int nerr_max, int nthreads);
void oracle_thermal_initial_T(const oracle_thermal_params *p, double *T);
void oracle_guo_source_moments(double u, double v, double w, double Fx, double Fy, double Fz,
                               double *S);
void oracle_d3q7_basis(double N[7][7], double Ninv[7][7]);
void oracle_thermal_nusselt(const oracle_thermal_params *p, const double *w, const double *T,
                            double kappa, double *nu_vol, double *nu_wall);

/* ---- moving spheres (D3Q19 generalisation of case4/mpi) ------------------- */
typedef struct {
    int nx, ny, nz;
    int np;              /* number of spheres */
    int collision;       /* ORACLE_MRT or ORACLE_BGK */
    int meq12_rho;
    double tau;
    double rho0, rho_solid;
    double gravity;      /* along -z */
    double threshold_wall, stiff_wall, threshold_particle, stiff_particle;
} oracle_particle_params;

/* owner[c] = index of the sphere containing node c (1-based coordinates
 * i+1, j+1, k+1), -1 for fluid. */
void oracle_particle_map(const oracle_particle_params *p, const double *xc, const double *r, int *owner);
int oracle_calq(const double *xc, const double *r, int id, const double x[3], int a, double x0[3],
                double *q);
double oracle_calq_closed(const double *xc, const double *r, int id, const double x[3], int a);
/* nsteps steps from rest; xc, U, Om (3 per sphere) are updated in place;
 * F, T receive the last step's total force and the hydrodynamic torque.
 * Returns 0, -1 (allocation), -2 (interpenetration), -3 (wall penetration),
 * -4 (wall distance q outside [0, 1]). */
int oracle_particle_run(const oracle_particle_params *p, int nsteps, double *xc, double *U, double *Om,
                        const double *r, double *F, double *T, int *owner, double *rho, double *u,
                        double *v, double *w, double *rho_avg);

/* Wall-clock seconds spent inside the step loop of the last run. */
double oracle_last_loop_seconds(void);

/* The reference's two-dimensional two-disc particle program
 * (Micro_particles/fortran/case4/seq/CaseIV_DKT.F90) restated
 * (particle2d_oracle.c).  After nsteps steps: scal[19] = itc, shiftTimes,
 * rhoAvg, xCenter[2], yCenter[2], Uc[2], Vc[2], rationalOmega[2],
 * wallTotalForceX[2], wallTotalForceY[2], totalTorque[2]; rho, u, v, obst
 * (nx*ny, x fastest) and, if f is not NULL, f (9*nx*ny, population fastest).
 * Returns 0, -1 (allocation) or the reference's stop: 1 interpenetration,
 * 2 wall penetration, 3 q outside [0, 1], 4 no extrapolation direction,
 * 5 window moving too fast, 6 a disc near the bottom, 7 no disc owns a node. */
int oracle_dkt_dims(int *nx, int *ny);
int oracle_dkt_run(int nsteps, double *scal, double *rho, double *u, double *v, double *obst, double *f);

#ifdef __cplusplus
}
#endif
#endif
