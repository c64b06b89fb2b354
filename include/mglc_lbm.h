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
    LBM_ERR_UNSUPPORTED = 7,
    LBM_ERR_PARTICLE = 8  /* particle run stopped: interpenetration, wall penetration
                             or a wall distance q outside [0, 1] (particle_force.F90:117-186,
                             particle_bounceback.F90:128-138) */
};

enum lbm_collision { LBM_MRT = 0, LBM_BGK = 1 };
enum lbm_storage { LBM_STORE_F64 = 0, LBM_STORE_F32 = 1, LBM_STORE_F16 = 2 };

typedef struct lbm_ctx lbm_ctx;

/* Halo regions of a subdomain (mpi_3d_blocked/main.f90:50-72, ex_sendrecv.f90:12-123):
 * six faces carry the 5 populations that cross them over the whole face plane,
 * twelve edges carry the one diagonal population that crosses both faces along
 * the edge line; corners carry nothing in D3Q19.  Edge regions are numbered by
 * that population a = 7..18 as region 6 + (a - 7).  The outgoing region r of a
 * subdomain goes to the neighbour across r; what arrives through region r
 * comes from that neighbour's outgoing region opposite to r (faces) or its
 * outgoing region r (edges: the same diagonal population). */
enum lbm_region {
    LBM_FACE_XM = 0, LBM_FACE_XP = 1, LBM_FACE_YM = 2, LBM_FACE_YP = 3,
    LBM_FACE_ZM = 4, LBM_FACE_ZP = 5, LBM_EDGE_FIRST = 6, LBM_NUM_REGIONS = 18
};

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
    int thermal;          /* 1: buoyancy-driven cavity -- D3Q19 MRT with the Guo body force
                             (Boussinesq + Coriolis) coupled to a D3Q7 MRT temperature
                             (buoyancy .../3d/mpi_blocked/evolution_f.F90, evolution_g.F90);
                             FP64 storage, all walls no-slip, no lid */
    int block_x, block_y; /* launch shape; 0 = default (x: min(256, nx rounded to 64)) */
    int halo_faces;       /* bit f (enum lbm_region, faces 0..5) set: face f borders a
                             neighbour subdomain and reads its halo; clear: wall
                             (halfway bounce-back; the z+ wall is the moving lid) */
    int halo_loopback;    /* 1: the exchange looped back on one GPU: every step also
                             writes the populations leaving through each face into the
                             opposite face's incoming region of this same subdomain, and
                             each diagonal one leaving through an edge into that edge's
                             region -- what a pack followed by an unpack into the
                             opposite regions delivers -- for the regions the next step
                             reads (all of them when halo_faces is 0: the exchange is
                             then rehearsed, not read) */
    int thermal_axis;     /* thermal: axis of the walls held at t_hot (low side) and t_cold
                             (high side): 1 = y (LeftRightWallsConstT), 2 = z
                             (TopBottomPlatesConstT), 0 = x, -1 = none; other walls adiabatic */
    int halo_overlap;     /* 1: split every step into the boundary shell (cells on halo
                             faces) and the interior, as the reference's non-blocking code
                             does (lid3_mpi_nonblock.f90:1108-1231); with halo_loopback the
                             shell kernel also writes the populations it sends straight into
                             the next step's halo, so no pack / unpack runs between steps.
                             Halo faces must be y and z faces only (x, the contiguous axis,
                             stays whole); isothermal contexts only */
    int waves_per_cu;     /* occupancy cap of the step kernels (the fused step, the split
                             step's interior, the thermal step), in waves per CU (a multiple
                             of the waves of one block, at most 32), held by an LDS
                             reservation the kernels never touch; 0 = the library's choice:
                             16 for the isothermal FP64 / FP32 steps that read no halo, no
                             cap otherwise */
    int transform;        /* MRT moment transforms of FP16-stored steps: 0 = the library's
                             choice (VALU), 1 = VALU, 2 = the forward transform on the
                             matrix cores (MFMA; measured slower than VALU on MI355X, kept
                             as a selectable form); other storages and BGK use VALU */
    double tau;           /* tauf: commondata.f90:9 tauf = 3 U0 N / Re + 0.5
                             (thermal: module.F90:19 tauf = 0.5 + Ma N sqrt(3 Pr / Ra)) */
    double u_lid;         /* U0: lid velocity along +x on the top wall k = nz (0 when thermal) */
    double rho0;          /* initial density */
    /* thermal constants, computed by the caller as module.F90:19-33 does */
    double para_a;        /* paraA = 42 sqrt(3) kappa - 6 */
    double g_beta;        /* gBeta: the buoyancy force is rho gBeta (T - t_ref) along +z */
    double omega_rot;     /* rotation rate of the Coriolis force (0 = none) */
    double t_hot, t_cold, t_ref;
    double reserved_d[2];
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

/* Allocate device state for p.  Replaces the allocations of initial.f90:35-44.
 * Step kernels under an occupancy cap (waves_per_cu; by default the
 * isothermal FP64 and FP32 steps that read no halo, held to 16 waves per CU,
 * which streams faster) launch with an LDS reservation per block; kernels of
 * other work on the same device see that LDS in use meanwhile. */
int lbm_create(const lbm_params *p, lbm_ctx **ctx);
int lbm_destroy(lbm_ctx *ctx);

/* Launch all work of ctx on this stream (a hipStream_t).  NULL = null stream. */
int lbm_set_stream(lbm_ctx *ctx, void *hip_stream);

/* Reference initial state (initial.f90:46-73): rho = rho0, u = v = w = 0,
 * u = U0 on the top fluid layer when the z+ face is the lid wall (the
 * reference's coords(2) == dims(2) - 1 test), f = feq(rho, u); up = vp = 0;
 * itc = 0.  Thermal: T = 0 except the layers next to the held walls, which
 * start at t_hot / t_cold (initial.F90:126-162). */
int lbm_initial(lbm_ctx *ctx);

/* Equilibrium start from caller fields (each nx*ny*nz FP64, x fastest):
 * f = feq(rho, u, v, w), macro fields = the given ones, itc = 0.  The hook
 * for seeded-noise starts and restarts.  T is the temperature of a thermal
 * context (g = its equilibrium, initial.F90:176-179; NULL = 0) and is
 * ignored otherwise. */
int lbm_load_fields(lbm_ctx *ctx, const double *rho, const double *u, const double *v,
                    const double *w, const double *T);

/* nsteps iterations of collision -> streaming -> bounceback -> macro
 * (main.f90:87-97), fused into one pass over the lattice per step.  The
 * macroscopic fields of the last step are kept for lbm_check and
 * lbm_download_macro.  With halo_loopback each step writes the next step's
 * halo itself (the current state's halo is exchanged once first when it is
 * not already in place); with halo faces and no loopback the caller
 * exchanges between steps (lbm_halo_pack / transport / lbm_halo_unpack), so
 * nsteps must be 1. */
int lbm_step(lbm_ctx *ctx, int nsteps);

/* One step in two halves, for a caller that overlaps its own exchange with the
 * interior update (lid3_mpi_nonblock.f90:1108-1231: boundary collision, then
 * Isend/Irecv, then interior collision, then Waitall).  Requires halo_overlap.
 *   lbm_step_begin: update the boundary shell (every cell on a halo face) of
 *                   the next step; lbm_halo_pack called now packs that shell;
 *   lbm_step_end:   update the interior and complete the step (macro fields
 *                   written, as lbm_step(ctx, 1)).
 * The usual sequence per step: lbm_halo_unpack(previous receive), begin, pack,
 * start the transport on another stream, end, wait for the transport.  Results
 * equal lbm_step(ctx, 1) bit for bit. */
int lbm_step_begin(lbm_ctx *ctx);
int lbm_step_end(lbm_ctx *ctx);

/* Reference check (check.f90:9-34): errorU = sqrt(sum (u-up)^2 + (v-vp)^2) /
 * sqrt(sum u^2 + v^2 + w^2) over the last macro fields, then up = u, vp = v.
 * FP64 device reduction.  Returns LBM_ERR_NAN (with *error_u set) when the
 * residual is NaN. */
int lbm_check(lbm_ctx *ctx, double *error_u);

/* Thermal residuals (buoyancy check.F90:16-41): errorU with the w term,
 * errorT = sum |T - Tp| / sum |T|; then up, vp, wp, Tp = u, v, w, T. */
int lbm_check_thermal(lbm_ctx *ctx, double *error_u, double *error_t);

/* Nusselt numbers of a thermal context (FP64):
 * nu_vol  = <w T> nz / kappa + 1, calNuRe's volume Nusselt number
 *           (RaNu.F90:13-23): w is the velocity along gravity (z), whichever
 *           walls are held;
 * nu_wall = mean over the hot wall of 2 (t_hot - T_first_layer) N_a / (t_hot - t_cold),
 *           the conductive flux at the half-link wall.  kappa = (paraA + 6) / (42 sqrt 3). */
int lbm_nusselt(lbm_ctx *ctx, double *nu_vol, double *nu_wall);

/* Copy T of the last step (thermal contexts) to a host array of nx*ny*nz. */
int lbm_download_temperature(lbm_ctx *ctx, double *T);
/* Volume Reynolds number of the last macro fields, as calNuRe
 * (RaNu.F90:36-44): sqrt(<u^2 + v^2 + w^2>) * nz / nu with nu = (tau - 1/2)/3,
 * FP64, deterministic.  Any context without spheres. */
int lbm_reynolds(lbm_ctx *ctx, double *re);

/* The two global sums behind errorU (check.f90:12-17 before the Allreduce of
 * check.f90:25-28): e1 = sum (u-up)^2 + (v-vp)^2, e2 = sum u^2 + v^2 + w^2 over
 * this subdomain, then up = u, vp = v.  A decomposed run adds them over all
 * subdomains and takes errorU = sqrt(e1) / sqrt(e2). */
int lbm_check_sums(lbm_ctx *ctx, double *e1, double *e2);

/* The sums of lbm_check_sums over the caller's own device arrays, allocating
 * nothing.  No counterpart in the reference: for callers that cannot allocate
 * device memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc).  u, v, w, up, vp: distinct device arrays of n doubles;
 * e1 = sum (u-up)^2 + (v-vp)^2 and e2 = sum u^2 + v^2 + w^2 are written to
 * sums_dev[0], sums_dev[1] (device memory), then up = u, vp = v; the sums equal
 * lbm_check_sums's bit for bit on the same fields.  Queued on hip_stream
 * (NULL = null stream) without waiting.  The per-block partial sums live in a
 * fixed library array of 2 x 1024 doubles (16 KiB) per device, whatever n;
 * every call on a device shares it, so calls that overlap on one GPU (other
 * streams, other threads) race.  n in [1, 2^40], else LBM_ERR_ARG. */
int lbm_check_sums_noalloc(const double *u, const double *v, const double *w, double *up, double *vp,
                           long long n, double *sums_dev, void *hip_stream);

/* Packed halo buffer: all 18 regions concatenated in region order, elements of
 * the storage type (8, 4 or 2 bytes).  Faces: [slot 0..4][plane], slots the 5
 * populations in ascending order, plane index j + ny*k (x faces), i + nx*k
 * (y faces), i + nx*j (z faces).  Edges: the line index along the free axis. */
int lbm_halo_layout(const lbm_ctx *ctx, int region, long long *offset, long long *count);
int lbm_halo_buffer_bytes(const lbm_ctx *ctx, size_t *bytes);

/* Pack every outgoing region of the current state into the device buffer
 * dev_send (lbm_halo_buffer_bytes long): the post-collision populations that
 * leave through each face and edge, as the exchange of ex_sendrecv.f90 sends
 * them.  Call after lbm_initial / lbm_step and before the next lbm_step;
 * between lbm_step_begin and lbm_step_end it packs the shell of the step in
 * progress (regions of faces in halo_faces and edges between two of them). */
int lbm_halo_pack(lbm_ctx *ctx, void *dev_send);

/* Take the incoming regions from the device buffer dev_recv (same layout;
 * region r holds what arrives through r).  Only regions of faces in
 * halo_faces (and edges between two such faces) are read by the next step. */
int lbm_halo_unpack(lbm_ctx *ctx, const void *dev_recv);

/* Copy rho, u, v, w of the last step to host arrays (nx*ny*nz each, x
 * fastest; any pointer may be NULL to skip that field). */
int lbm_download_macro(lbm_ctx *ctx, double *rho, double *u, double *v, double *w);

/* Wait for all work queued on ctx's stream. */
int lbm_synchronize(lbm_ctx *ctx);

int lbm_get_info(const lbm_ctx *ctx, lbm_info *info);

/* ---- checkpoint / restart ---------------------------------------------------
 * Replaces backupData (Buoyancy_driven_cavity/fortran/3d/seq/bouyancy3d.F90:
 * 1011-1030) and the loadInitField restart switch
 * (3d/mpi_blocked/initial.F90:102, 195-206).  The state blob holds what the
 * next step reads (stored populations in the device layout, lid density,
 * macro fields, residual history, incoming halo, thermal populations, force
 * and temperature) behind a header naming the grid, storage, collision and
 * flags it belongs to; a run restarted from it continues bit for bit.  The
 * caller owns the host buffer and its file I/O.  Contexts with attached
 * spheres return LBM_ERR_UNSUPPORTED. */
int lbm_state_bytes(lbm_ctx *ctx, size_t *bytes);
int lbm_save_state(lbm_ctx *ctx, void *host_blob, size_t bytes);
/* The context must have the blob's grid, storage, collision, thermal and
 * halo settings (else LBM_ERR_ARG); the step counter is restored. */
int lbm_load_state(lbm_ctx *ctx, const void *host_blob, size_t bytes);

/* ---- moving spheres --------------------------------------------------------
 * Rigid spheres in a closed box (no-slip walls, no lid), coupled to the
 * D3Q19 fluid as the reference couples discs to D2Q9
 * (MPI/Micro_particles/fortran/case4/mpi: fluid.F90, particle_bounceback.F90,
 * particle_force.F90, particle_update.F90): collision on fluid nodes, masked
 * streaming, quadratic interpolated bounce-back with the moving-wall term
 * (wall distance q by bisection), momentum-exchange force and torque summed
 * per sphere in FP64 (deterministic block sums), spring and buoyancy forces,
 * kinematics, new solid map, refill of uncovered nodes.  Coordinates are
 * 1-based node coordinates (node (i,j,k) sits at (i+1, j+1, k+1)); gravity
 * acts along -z.  Storage FP64 or FP32. */
typedef struct lbm_particle_params {
    int np;
    int reserved_i;
    double rho_solid;        /* rhoSolid */
    double gravity;          /* |g| along -z, lattice units */
    double threshold_wall, stiff_wall;          /* wall spring (particle_force.F90:126-186) */
    double threshold_particle, stiff_particle;  /* particle spring (particle_force.F90:96-124) */
    double reserved_d[4];
} lbm_particle_params;

/* Attach np spheres (centres np*3, radii np; velocities U and angular
 * velocities omega np*3 each, NULL for rest) to a context and set the initial
 * state of initial.F90:78-135: solid map, rho = rho_s in solids and rho0 in
 * the fluid, u = 0, f = w rho0 on fluid nodes.  Replaces lbm_initial /
 * lbm_load_fields (which detach the spheres again).  Needs an isothermal
 * FP64/FP32 context with u_lid = 0 and no halo faces. */
int lbm_particles_init(lbm_ctx *ctx, const lbm_particle_params *pp, const double *centers,
                       const double *radii, const double *U, const double *omega);
/* A step call that meets a stop condition returns LBM_ERR_PARTICLE naming
 * the first condition met and its step; the context then refuses further
 * steps (the fields and sphere state are those of an unspecified point of the
 * stopping step) until lbm_particles_init is called again. */
/* Sphere state of the last step (each np*3; any pointer may be NULL):
 * centre, velocity, angular velocity, total force (hydrodynamic + springs +
 * buoyancy) and hydrodynamic torque. */
int lbm_particles_get(lbm_ctx *ctx, double *centers, double *U, double *omega, double *F, double *T);
/* Owner map of the current step: sphere index per node, -1 for fluid
 * (nx*ny*nz int32, x fastest). */
int lbm_particles_map(lbm_ctx *ctx, int *owner);

/* ---- batched cavities ----------------------------------------------------
 * Many independent lid-driven cavities of uneven sizes, Re and U0 stepped
 * together: one persistent kernel walks 64-cell tiles of every instance
 * through a per-instance offset table; whole 100-step blocks are replayed
 * from a captured hipGraph; the persistent launch shape is picked by
 * lbm_batch_autotune.  Each instance reproduces a single-context run of the
 * same parameters exactly.  No reference counterpart: the reference runs one
 * cavity per program (mpi_3d_blocked/main.f90). */
typedef struct lbm_batch lbm_batch;

typedef struct lbm_batch_info {
    long long instances, total_cells, total_tiles, steps_done, device_bytes;
    double bytes_per_lup;       /* 2*19*sizeof(store) */
    int threads, blocks;        /* persistent launch shape in use */
    int graph_steps;            /* steps per captured graph */
    int cells_per_thread;       /* cells each thread updates per step (1, 2 or 4) */
    void *stream;               /* the batch's own hipStream_t */
} lbm_batch_info;

/* params[m] gives instance m's nx, ny, nz, tau and u_lid; storage, collision
 * and meq12_rho must be the same for all (instance 0's). */
int lbm_batch_create(const lbm_params *params, int count, lbm_batch **batch);
int lbm_batch_destroy(lbm_batch *batch);
/* Reference initial state (initial.f90:46-73) in every instance. */
int lbm_batch_initial(lbm_batch *batch);
/* nsteps steps of every instance: the persistent kernel, graph-replayed in
 * blocks of 100 steps.  Macro fields of the last step are kept. */
int lbm_batch_step(lbm_batch *batch, int nsteps);
/* The same steps with one single-domain launch per instance per step (the
 * instances launched one at a time): the baseline the batch is measured
 * against. */
int lbm_batch_step_each(lbm_batch *batch, int nsteps);
/* Time trial_steps steps at each candidate launch shape (64/128/256 threads,
 * 4..64 waves per CU, 1/2/4 cells per thread), keep the fastest, and restart
 * from the initial state.  The cells per thread chosen are in lbm_batch_info. */
int lbm_batch_autotune(lbm_batch *batch, int trial_steps, int *threads, int *blocks);
int lbm_batch_set_launch(lbm_batch *batch, int threads, int blocks);
/* Each wave takes cells_per_thread consecutive 64-cell tiles per turn (1, 2
 * or 4): a thread updates that many cells per step.  Results do not depend on
 * it. */
int lbm_batch_set_cells_per_thread(lbm_batch *batch, int cells_per_thread);
int lbm_batch_synchronize(lbm_batch *batch);
int lbm_batch_download_macro(lbm_batch *batch, int instance, double *rho, double *u, double *v,
                             double *w);
int lbm_batch_get_info(const lbm_batch *batch, lbm_batch_info *info);
/* ---- diagnostics --------------------------------------------------------------
 * No counterpart in the reference.  Measures, on the current device, the HBM
 * bandwidth (GB/s, read + write) the memory system delivers for the step's
 * traffic with no arithmetic: pattern 0 = a one-stream copy of 19 * n^3
 * elements, pattern 1 = the D3Q19 pull of an n^3 lattice in the library's
 * layout (19 upstream loads and 19 nontemporal stores per cell), pattern 2 =
 * the buoyancy step's traffic (D3Q19 + D3Q7 pulls and one force plane, FP64:
 * 27 loads and 27 stores per cell); patterns 1 and 2 report the best of four
 * occupancies (all waves, 16, 12 and 8 waves per CU).  storage = element size
 * (LBM_STORE_*), 16 <= n <= 512, reps timed launches after 3 untimed ones.
 * Allocates and frees two lattices of its own. */
int lbm_bandwidth_probe(int pattern, int storage, int n, int reps, double *gbps);

const char *lbm_status_string(int status);
const char *lbm_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* MGLC_LBM_H */
