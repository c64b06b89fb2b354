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
} oracle_cavity_params;

/* Whole-run driver: reference initial, then nsteps of
 * collision -> streaming -> bounceback -> macro, with check() every
 * check_every steps (0 = never).  Outputs rho,u,v,w after the last step; errs
 * (may be NULL) receives errorU at each check, nerr_max entries at most.
 * Returns the number of checks done, or -1 on allocation failure. */
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

/* Wall-clock seconds spent inside the step loop of the last run. */
double oracle_last_loop_seconds(void);

#ifdef __cplusplus
}
#endif
#endif
