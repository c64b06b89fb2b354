/*
 * mglc_lbm.h -- C-ABI of the MI355X lattice Boltzmann step library
 * (libmglc_lbm.so).
 *
 * The reference (cheryli/MGLC) is a set of Fortran programs whose hot path is
 * a chain of no-argument subroutines on module globals:
 *
 *   initial -> do while: collision -> exchange -> streaming -> bounceback
 *              -> macro -> (check every 2000 steps) -> output
 *   (MPI/Lid_driven_cavity/fortran/3d/mpi_3d_blocked/main.f90:75-103)
 *
 * This library replaces that chain for one GPU's subdomain.  A Fortran driver
 * binds these entry points with iso_c_binding (mglc_amd/fortran/mglc_lbm_capi.f90).
 * Conventions:
 *   - every call returns an int status (LBM_OK = 0, else an lbm_status);
 *     lbm_last_error() gives a message for the last failure on this thread;
 *   - host arrays are Fortran column-major (nx, ny, nz), x fastest, FP64,
 *     owned by the caller; no pointer is kept across calls;
 *   - all device memory is owned by the context; lbm_destroy frees it;
 *   - a context is used from one host thread; work runs on one HIP stream
 *     (the null stream unless lbm_set_stream names another).
 */
#ifndef MGLC_LBM_H
#define MGLC_LBM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum lbm_status {
    LBM_OK = 0,
    LBM_ERR_ARG = 1,      /* invalid argument / parameter */
    LBM_ERR_HIP = 2,      /* HIP runtime error */
    LBM_ERR_NAN = 3,      /* residual is NaN (the flow blew up) */
    LBM_ERR_TAU = 4,      /* relaxation time out of range (tau <= 0.5) */
    LBM_ERR_NOMEM = 5,    /* device allocation failed */
    LBM_ERR_STATE = 6,    /* call out of order (e.g. step before initial) */
    LBM_ERR_UNSUPPORTED = 7
};

enum lbm_collision { LBM_MRT = 0, LBM_BGK = 1 };
enum lbm_storage { LBM_STORE_F64 = 0, LBM_STORE_F32 = 1, LBM_STORE_F16 = 2 };

typedef struct lbm_ctx lbm_ctx;

/* Run parameters.  Field names follow the reference's module commondata
 * (mpi_3d_blocked/commondata.f90:4-42, buoyancy module.F90:11-57); derived
 * constants (tauf, paraA, gBeta, ...) are computed by the caller, as the
 * reference computes them in its parameter statements. */
typedef struct lbm_params {
    int nx, ny, nz;       /* subdomain = whole cavity on one GPU */
    int collision;        /* enum lbm_collision; the reference's operator is MRT */
    int storage;          /* enum lbm_storage */
    int meq12_rho;        /* 0: cavity meq(12) = -(v^2-w^2)/2, no rho (collision.f90:85);
                             1: thermal form with rho (evolution_f.F90:75) */
    int f16_scale_log2;   /* FP16 storage holds (f - w) * 2^k; 0 selects k = 12 */
    int thermal;          /* 0 only in this version (D3Q7 coupling is a later row) */
    int block_x, block_y; /* launch shape; 0 = default 64 x 4 */
    int reserved_i[6];
    double tau;           /* tauf: commondata.f90:9 tauf = 3 U0 N / Re + 0.5 */
    double u_lid;         /* U0: lid velocity along +x on the top wall k = nz */
    double rho0;          /* initial density */
    double reserved_d[8];
} lbm_params;

typedef struct lbm_info {
    long long nx, ny, nz;
    long long pitch_x;          /* elements between rows in device storage */
    long long pop_stride;       /* elements between populations */
    long long steps_done;       /* itc */
    long long device_bytes;     /* device memory held by the context */
    double bytes_per_lup;       /* algorithmic HBM bytes per lattice update: 2*19*sizeof(store) */
    int storage;
    int collision;
} lbm_info;

/* Allocate device state for p.  Replaces the allocations of initial.f90:35-44. */
int lbm_create(const lbm_params *p, lbm_ctx **ctx);
int lbm_destroy(lbm_ctx *ctx);

/* Launch all work of ctx on this stream (a hipStream_t).  NULL = null stream. */
int lbm_set_stream(lbm_ctx *ctx, void *hip_stream);

/* Reference initial state (initial.f90:46-73): rho = rho0, u = v = w = 0,
 * u = U0 on the top fluid layer, f = feq(rho, u); up = vp = wp = 0; itc = 0. */
int lbm_initial(lbm_ctx *ctx);

/* Equilibrium start from caller fields (each nx*ny*nz FP64, x fastest):
 * f = feq(rho, u, v, w), macro fields = the given ones, itc = 0.  The hook
 * for seeded-noise starts and restarts. */
int lbm_load_fields(lbm_ctx *ctx, const double *rho, const double *u, const double *v,
                    const double *w);

/* nsteps iterations of collision -> streaming -> bounceback -> macro
 * (main.f90:87-97), fused into one pass over the lattice per step.  The
 * macroscopic fields of the last step are kept for lbm_check and
 * lbm_download_macro. */
int lbm_step(lbm_ctx *ctx, int nsteps);

/* Reference check (check.f90:9-34): errorU = sqrt(sum (u-up)^2 + (v-vp)^2) /
 * sqrt(sum u^2 + v^2 + w^2) over the last macro fields, then up = u, vp = v.
 * FP64 device reduction.  Returns LBM_ERR_NAN (with *error_u set) when the
 * residual is NaN. */
int lbm_check(lbm_ctx *ctx, double *error_u);

/* Copy rho, u, v, w of the last step to host arrays (nx*ny*nz each, x
 * fastest; any pointer may be NULL to skip that field). */
int lbm_download_macro(lbm_ctx *ctx, double *rho, double *u, double *v, double *w);

/* Wait for all work queued on ctx's stream. */
int lbm_synchronize(lbm_ctx *ctx);

int lbm_get_info(const lbm_ctx *ctx, lbm_info *info);
const char *lbm_status_string(int status);
const char *lbm_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* MGLC_LBM_H */
