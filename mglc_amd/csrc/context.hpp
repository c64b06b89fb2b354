// context.hpp -- the library's internal state (struct lbm_ctx) and helpers
// shared by the translation units of libmglc_lbm.so.  Not part of the C-ABI.
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/mglc_lbm.h"
#include "halo.hpp"

// Record a message for lbm_last_error() and return `code`.
int lbm_fail(int code, const char* fmt, ...);
// LBM_OK, or LBM_ERR_HIP if the last kernel launch failed.
int lbm_check_launch();

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return lbm_fail(LBM_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                            __FILE__, __LINE__);                                               \
    } while (0)

// Deterministic FP64 reductions: a fixed grid of per-block partials, then one
// block sums the partials in a fixed order.
constexpr int RED_THREADS = 256;
constexpr int RED_BLOCKS = 1024;
constexpr int RED_MAX_SUMS = 4;

struct ParticleState;  // moving spheres (particles.hip)

struct lbm_ctx {
    lbm_params p{};
    int device = 0;
    hipStream_t stream = nullptr;
    int px = 0;
    long long pxy = 0, ps = 0;
    size_t elem = 8;
    void* buf[2] = {nullptr, nullptr};
    int cur = 0;                 // buffer holding the current f_post
    double* lid_rho = nullptr;
    double* macro = nullptr;     // rho, u, v, w (4 * n)
    double* prev = nullptr;      // up, vp (2 * n)
    double* red = nullptr;       // per-block partials + totals of the FP64 reductions
    long long n = 0;
    long long itc = 0;
    bool initialised = false;
    dim3 block{64, 4, 1};
    unsigned step_lds = 0;       // LDS bytes reserved per block of the step kernels: caps the
                                 // blocks resident per CU (0: no cap); see lbm_create
    unsigned step_lds_halo = 0;  // the same for the kernels that read a halo (and the thermal step)
    float scale = 1.0f, inv_scale = 1.0f;
    long long device_bytes = 0;
    long long hoff[lbm::NREGION_ALL] = {}, hcount[lbm::NREGION_ALL] = {}, htotal = 0;  // D3Q7 regions: thermal only
    void* hin = nullptr;         // incoming halo regions (packed layout)
    void* hsend = nullptr;       // loopback send buffer
    // buoyancy-driven cavity (thermal = 1): D3Q7 temperature populations,
    // the body force of the last collision, T and its previous check value
    void* gbuf[2] = {nullptr, nullptr};  // 7 * ps each
    void* force = nullptr;       // Fx, Fy, Fz of the last collision (3 * ps; FP64 with FP64 storage, else FP32)
    double* temp = nullptr;      // T (n)
    double* prev_t = nullptr;    // wp, Tp for the thermal check (2 * n)
    ParticleState* part = nullptr;  // moving spheres, when attached
    // split steps (halo_overlap): with the loopback, the halo the shell
    // writes for the next step (swapped with hin after each step) and whether
    // hin belongs to the current state; whether a step is between
    // lbm_step_begin and lbm_step_end
    void* hin_next = nullptr;
    bool halo_fresh = false;
    bool in_step = false;
};

// Thermal path (thermal.hip).
// The looped-back exchange of every region the next step reads (pack from the
// current state, unpack into the opposite faces' regions): the thermal path
// runs it between steps.
int halo_loopback_exchange(lbm_ctx* c);
int thermal_alloc(lbm_ctx* c);
size_t thermal_force_elem(const lbm_ctx* c);  // bytes of one stored force component
void thermal_release(lbm_ctx* c);
int thermal_load(lbm_ctx* c, const double* d_rho, const double* d_u, const double* d_v, const double* d_w,
                 const double* d_T);
int thermal_step(lbm_ctx* c, int nsteps);
int thermal_check(lbm_ctx* c, double* sums4);
int thermal_nusselt(lbm_ctx* c, double* nu_vol, double* nu_wall);
int flow_reynolds(lbm_ctx* c, double* re);

// Particle path (particles.hip).
int particles_step(lbm_ctx* c, int nsteps);
void particles_release(lbm_ctx* c);

