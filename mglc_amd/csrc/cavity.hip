// cavity.hip -- D3Q19 lid-driven cavity on one MI355X: the fused pull-scheme
// stream + bounce-back + macro + collide kernel, its initialisation, the FP64
// residual reduction, and the C-ABI of include/mglc_lbm.h.
//
// Device layout (structure of arrays, no halo):
//   population a of cell (i, j, k) lives at  buf[a * ps + i + px * (j + ny * k)]
//   px = nx rounded up to a multiple of 16 elements; ps = px * ny * nz plus a
//   stagger so the 19 population streams do not start on the same HBM
//   channel.  Two such buffers (A-B): each step reads one, writes the other.
// Walls are not stored: a thread knows from (i, j, k) which of its 19 pulls
// leave the domain and bounces those back from its own cell.
//
// The stored state between steps is the post-collision f_post of the
// reference (collision.f90), so one kernel does, per cell:
//   streaming.f90:8-20    pull f(a, x) = f_post(a, x - e_a)
//   bounce_back.f90:6-83  halfway bounce-back, moving lid on k = nz
//   macro.f90:6-25        rho, u, v, w
//   collision.f90:15-202  the next step's collision (MRT or BGK)
// which is the reference's step order with the collision of step n+1 moved
// to the end of step n.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>

#include "../../include/mglc_lbm.h"
#include "collide.hpp"

using namespace lbm;

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(LBM_ERR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                            \
    } while (0)

struct StepArgs {
    const void* src;
    void* dst;
    int nx, ny, nz, px;
    long long pxy;      // px * ny
    long long ps;       // population stride (elements)
    double u_lid;
    double snu;         // 1 / tau
    double sq;          // 8 (2 tau - 1) / (8 tau - 1)
    float scale, inv_scale;
    int meq12_rho;
    double* lid_rho;    // nx * ny: rho of the top layer from the last macro
    double* rho;        // compact nx*ny*nz macro output (written when MACRO)
    double* u;
    double* v;
    double* w;
};

enum { COLL_MRT = 0, COLL_BGK = 1 };

template <class S, int COLL>
__device__ __forceinline__ void collide(typename S::Real f[Q],
                                        const Moments<typename S::Real, S::SHIFT>& q,
                                        const StepArgs& a) {
    using R = typename S::Real;
    if constexpr (COLL == COLL_BGK)
        collide_bgk<R, S::SHIFT>(f, q, R(a.snu));
    else
        collide_mrt<R, S::SHIFT>(f, q, R(a.snu), R(a.sq), a.meq12_rho != 0);
}

// One lattice update per thread.  Block = (64, by): one wavefront per x-row
// segment of 64 cells, so every population load and store of a wave is 64
// consecutive elements.
template <class S, int COLL, bool MACRO>
__global__ __launch_bounds__(256) void k_cavity_step(const StepArgs a) {
    using T = typename S::T;
    using R = typename S::Real;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= a.nx || j >= a.ny) return;
    const int c = i + a.px * (j + a.ny * k);
    const T* __restrict__ src = static_cast<const T*>(a.src);
    T* __restrict__ dst = static_cast<T*>(a.dst);
    const bool top = (k == a.nz - 1);
    const int lid_idx = i + a.nx * j;
    double rho_lid = 0.0;
    if (top) rho_lid = a.lid_rho[lid_idx];

    R f[Q];
    sfor<Q>([&](auto A) {
        constexpr int al = decltype(A)::value;
        constexpr int ex = EX[al], ey = EY[al], ez = EZ[al];
        const T* base = src + (long long)al * a.ps;
        if constexpr (ex == 0 && ey == 0 && ez == 0) {
            f[al] = S::ld(base + c, a.inv_scale);
        } else {
            bool in = true;
            if constexpr (ex != 0) in = in && (unsigned)(i - ex) < (unsigned)a.nx;
            if constexpr (ey != 0) in = in && (unsigned)(j - ey) < (unsigned)a.ny;
            if constexpr (ez != 0) in = in && (unsigned)(k - ez) < (unsigned)a.nz;
            if (in) {
                const long long off = (long long)ex + (long long)ey * a.px + (long long)ez * a.pxy;
                f[al] = S::ld(base + (c - off), a.inv_scale);
            } else {
                // halfway bounce-back: f(a, x) = f_post(abar, x)
                R val = S::ld(src + (long long)OPP[al] * a.ps + c, a.inv_scale);
                if constexpr (ez == -1 && ex != 0) {
                    // moving lid (bounce_back.f90:77-78): f(14) = f_post(11) - rho/6 U0,
                    // f(13) = f_post(12) - rho/6 (-U0); the lid is applied last, so it
                    // wins over the side walls on these two populations.
                    if (top) {
                        const double t = rho_lid / 6.0 * a.u_lid;
                        val = (ex > 0) ? R(double(val) + t) : R(double(val) - t);
                    }
                }
                f[al] = val;
            }
        }
    });

    const Moments<R, S::SHIFT> q = moments<R, S::SHIFT>(f);
    if (top) a.lid_rho[lid_idx] = S::SHIFT ? 1.0 + double(q.drho) : double(q.rho);
    if constexpr (MACRO) {
        const long long m = i + (long long)a.nx * (j + (long long)a.ny * k);
        a.rho[m] = S::SHIFT ? 1.0 + double(q.drho) : double(q.rho);
        a.u[m] = double(q.u);
        a.v[m] = double(q.v);
        a.w[m] = double(q.w);
    }
    collide<S, COLL>(f, q, a);
    sfor<Q>([&](auto A) {
        constexpr int al = decltype(A)::value;
        S::st(dst + (long long)al * a.ps + c, f[al], a.scale);
    });
}

// Equilibrium start (initial.f90:63-73) followed by the first collision with
// the given fields (the reference's first collision uses the initial rho, u,
// not moments of feq).  Writes buffer dst, the lid density and the macro
// fields.
template <class S, int COLL>
__global__ __launch_bounds__(256) void k_cavity_init(const StepArgs a, const double* __restrict__ rho_in,
                                                     const double* __restrict__ u_in,
                                                     const double* __restrict__ v_in,
                                                     const double* __restrict__ w_in) {
    using T = typename S::T;
    using R = typename S::Real;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= a.nx || j >= a.ny) return;
    const int c = i + a.px * (j + a.ny * k);
    const long long m = i + (long long)a.nx * (j + (long long)a.ny * k);
    const double rho = rho_in[m], u = u_in[m], v = v_in[m], w = w_in[m];
    const double us2 = u * u + v * v + w * w;
    R f[Q];
    sfor<Q>([&](auto A) {
        constexpr int al = decltype(A)::value;
        double un = -0.0;
        if constexpr (EX[al] != 0) un = mac<EX[al]>(un, u);
        if constexpr (EY[al] != 0) un = mac<EY[al]>(un, v);
        if constexpr (EZ[al] != 0) un = mac<EZ[al]>(un, w);
        const double feq = rho * weight(al) * (1.0 + 3.0 * un + 4.5 * un * un - 1.5 * us2);
        f[al] = S::SHIFT ? R(feq - weight(al)) : R(feq);
    });
    Moments<R, S::SHIFT> q;
    q.rho = R(rho);
    q.drho = R(rho - 1.0);
    q.u = R(u);
    q.v = R(v);
    q.w = R(w);
    a.rho[m] = rho;
    a.u[m] = u;
    a.v[m] = v;
    a.w[m] = w;
    if (k == a.nz - 1) a.lid_rho[i + a.nx * j] = rho;
    collide<S, COLL>(f, q, a);
    T* __restrict__ dst = static_cast<T*>(a.dst);
    sfor<Q>([&](auto A) {
        constexpr int al = decltype(A)::value;
        S::st(dst + (long long)al * a.ps + c, f[al], a.scale);
    });
}

// ---- FP64 residual (check.f90:9-34) -----------------------------------------------
constexpr int RED_THREADS = 256;
constexpr int RED_BLOCKS = 1024;

__device__ __forceinline__ double wave_sum(double x) {
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    return x;
}

// Two-pass deterministic reduction: fixed grid, per-block partials in a
// fixed order, then one block sums the partials.
__global__ __launch_bounds__(RED_THREADS) void k_check_partial(const double* __restrict__ u,
                                                               const double* __restrict__ v,
                                                               const double* __restrict__ w,
                                                               double* __restrict__ up,
                                                               double* __restrict__ vp, long long n,
                                                               double* __restrict__ partial) {
    double e1 = 0.0, e2 = 0.0;
    for (long long c = blockIdx.x * (long long)RED_THREADS + threadIdx.x; c < n;
         c += (long long)gridDim.x * RED_THREADS) {
        const double uu = u[c], vv = v[c], ww = w[c];
        const double du = uu - up[c], dv = vv - vp[c];
        e1 += du * du + dv * dv;
        e2 += uu * uu + vv * vv + ww * ww;
        up[c] = uu;
        vp[c] = vv;
    }
    __shared__ double s1[RED_THREADS / 64], s2[RED_THREADS / 64];
    e1 = wave_sum(e1);
    e2 = wave_sum(e2);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        s1[wid] = e1;
        s2[wid] = e2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t1 = 0.0, t2 = 0.0;
        for (int q = 0; q < RED_THREADS / 64; ++q) {
            t1 += s1[q];
            t2 += s2[q];
        }
        partial[2 * blockIdx.x] = t1;
        partial[2 * blockIdx.x + 1] = t2;
    }
}

__global__ __launch_bounds__(RED_THREADS) void k_reduce_pairs(const double* __restrict__ partial,
                                                              int nblocks, double* __restrict__ out) {
    double e1 = 0.0, e2 = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += RED_THREADS) {
        e1 += partial[2 * b];
        e2 += partial[2 * b + 1];
    }
    __shared__ double s1[RED_THREADS / 64], s2[RED_THREADS / 64];
    e1 = wave_sum(e1);
    e2 = wave_sum(e2);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
        s1[wid] = e1;
        s2[wid] = e2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t1 = 0.0, t2 = 0.0;
        for (int q = 0; q < RED_THREADS / 64; ++q) {
            t1 += s1[q];
            t2 += s2[q];
        }
        out[0] = t1;
        out[1] = t2;
    }
}

__global__ void k_fill(double* __restrict__ p, long long n, double val) {
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n;
         c += (long long)gridDim.x * blockDim.x)
        p[c] = val;
}

using StepFn = void (*)(const StepArgs);
using InitFn = void (*)(const StepArgs, const double*, const double*, const double*, const double*);

template <class S>
struct Table {
    static constexpr StepFn step[2][2] = {
        {k_cavity_step<S, COLL_MRT, false>, k_cavity_step<S, COLL_MRT, true>},
        {k_cavity_step<S, COLL_BGK, false>, k_cavity_step<S, COLL_BGK, true>}};
    static constexpr InitFn init[2] = {k_cavity_init<S, COLL_MRT>, k_cavity_init<S, COLL_BGK>};
};

}  // namespace

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
    double* red = nullptr;       // 2 * RED_BLOCKS partials + 2 totals
    long long n = 0;
    long long itc = 0;
    bool initialised = false;
    dim3 block{64, 4, 1};
    float scale = 1.0f, inv_scale = 1.0f;
    long long device_bytes = 0;
};

namespace {

StepArgs make_args(const lbm_ctx* c) {
    StepArgs a{};
    a.nx = c->p.nx;
    a.ny = c->p.ny;
    a.nz = c->p.nz;
    a.px = c->px;
    a.pxy = c->pxy;
    a.ps = c->ps;
    a.u_lid = c->p.u_lid;
    a.snu = 1.0 / c->p.tau;
    a.sq = 8.0 * (2.0 * c->p.tau - 1.0) / (8.0 * c->p.tau - 1.0);
    a.scale = c->scale;
    a.inv_scale = c->inv_scale;
    a.meq12_rho = c->p.meq12_rho;
    a.lid_rho = c->lid_rho;
    a.rho = c->macro;
    a.u = c->macro + c->n;
    a.v = c->macro + 2 * c->n;
    a.w = c->macro + 3 * c->n;
    return a;
}

dim3 grid_of(const lbm_ctx* c) {
    return dim3((c->p.nx + c->block.x - 1) / c->block.x, (c->p.ny + c->block.y - 1) / c->block.y,
                c->p.nz);
}

StepFn step_fn(const lbm_ctx* c, bool macro) {
    const int coll = c->p.collision == LBM_BGK ? 1 : 0;
    switch (c->p.storage) {
    case LBM_STORE_F32: return Table<StoreF32>::step[coll][macro];
    case LBM_STORE_F16: return Table<StoreF16>::step[coll][macro];
    default: return Table<StoreF64>::step[coll][macro];
    }
}

InitFn init_fn(const lbm_ctx* c) {
    const int coll = c->p.collision == LBM_BGK ? 1 : 0;
    switch (c->p.storage) {
    case LBM_STORE_F32: return Table<StoreF32>::init[coll];
    case LBM_STORE_F16: return Table<StoreF16>::init[coll];
    default: return Table<StoreF64>::init[coll];
    }
}

int check_launch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(LBM_ERR_HIP, "kernel launch failed: %s", hipGetErrorString(e));
    return LBM_OK;
}

void release(lbm_ctx* c) {
    for (void* b : c->buf)
        if (b) (void)hipFree(b);
    if (c->lid_rho) (void)hipFree(c->lid_rho);
    if (c->macro) (void)hipFree(c->macro);
    if (c->prev) (void)hipFree(c->prev);
    if (c->red) (void)hipFree(c->red);
}

}  // namespace

extern "C" {

const char* lbm_last_error(void) { return g_err; }

const char* lbm_status_string(int status) {
    switch (status) {
    case LBM_OK: return "ok";
    case LBM_ERR_ARG: return "invalid argument";
    case LBM_ERR_HIP: return "HIP runtime error";
    case LBM_ERR_NAN: return "residual is NaN";
    case LBM_ERR_TAU: return "relaxation time out of range";
    case LBM_ERR_NOMEM: return "device allocation failed";
    case LBM_ERR_STATE: return "call out of order";
    case LBM_ERR_UNSUPPORTED: return "unsupported option";
    }
    return "unknown status";
}

int lbm_create(const lbm_params* p, lbm_ctx** out) {
    if (!p || !out) return fail(LBM_ERR_ARG, "lbm_create: null argument");
    *out = nullptr;
    if (p->nx < 2 || p->ny < 2 || p->nz < 2)
        return fail(LBM_ERR_ARG, "lbm_create: grid %d x %d x %d too small", p->nx, p->ny, p->nz);
    if (!(p->tau > 0.5) || !std::isfinite(p->tau))
        return fail(LBM_ERR_TAU, "lbm_create: tau = %g must be > 0.5", p->tau);
    if (p->collision != LBM_MRT && p->collision != LBM_BGK)
        return fail(LBM_ERR_ARG, "lbm_create: collision %d unknown", p->collision);
    if (p->storage < LBM_STORE_F64 || p->storage > LBM_STORE_F16)
        return fail(LBM_ERR_ARG, "lbm_create: storage %d unknown", p->storage);
    if (p->thermal != 0) return fail(LBM_ERR_UNSUPPORTED, "lbm_create: thermal coupling not built");
    if (p->f16_scale_log2 < 0 || p->f16_scale_log2 > 20)
        return fail(LBM_ERR_ARG, "lbm_create: f16_scale_log2 %d out of [0, 20]", p->f16_scale_log2);
    const int bx = p->block_x ? p->block_x : 64, by = p->block_y ? p->block_y : 4;
    if (bx % 64 != 0 || bx * by > 256 || by < 1)
        return fail(LBM_ERR_ARG, "lbm_create: block %d x %d (x must be a multiple of 64, <= 256 threads)",
                    bx, by);

    lbm_ctx* c = new (std::nothrow) lbm_ctx();
    if (!c) return fail(LBM_ERR_NOMEM, "lbm_create: host allocation failed");
    c->p = *p;
    if (c->p.rho0 == 0.0) c->p.rho0 = 1.0;
    c->block = dim3(bx, by, 1);
    c->elem = p->storage == LBM_STORE_F64 ? 8 : (p->storage == LBM_STORE_F32 ? 4 : 2);
    c->px = (p->nx + 15) / 16 * 16;
    c->pxy = (long long)c->px * p->ny;
    const long long cells = c->pxy * p->nz;
    if (cells >= (1LL << 31)) {
        delete c;
        return fail(LBM_ERR_ARG, "lbm_create: %lld padded cells exceed the 32-bit cell index", cells);
    }
    // population stride: whole 4 KiB pages plus a 1 KiB stagger per stream
    const long long page = 4096 / (long long)c->elem;
    c->ps = (cells + page - 1) / page * page + 1024 / (long long)c->elem;
    c->n = (long long)p->nx * p->ny * p->nz;
    const int k = p->f16_scale_log2 ? p->f16_scale_log2 : 12;
    c->scale = p->storage == LBM_STORE_F16 ? std::ldexp(1.0f, k) : 1.0f;
    c->inv_scale = 1.0f / c->scale;

    (void)hipGetDevice(&c->device);
    const size_t lat = (size_t)c->ps * Q * c->elem;
    hipError_t e = hipSuccess;
    for (auto& b : c->buf)
        if (e == hipSuccess) e = hipMalloc(&b, lat);
    if (e == hipSuccess) e = hipMalloc(&c->lid_rho, sizeof(double) * p->nx * p->ny);
    if (e == hipSuccess) e = hipMalloc(&c->macro, sizeof(double) * 4 * c->n);
    if (e == hipSuccess) e = hipMalloc(&c->prev, sizeof(double) * 2 * c->n);
    if (e == hipSuccess) e = hipMalloc(&c->red, sizeof(double) * (2 * RED_BLOCKS + 2));
    if (e != hipSuccess) {
        release(c);
        delete c;
        (void)hipGetLastError();
        return fail(LBM_ERR_NOMEM, "lbm_create: hipMalloc failed: %s", hipGetErrorString(e));
    }
    c->device_bytes = (long long)(2 * lat + sizeof(double) * (p->nx * p->ny + 6 * c->n + 2 * RED_BLOCKS + 2));
    *out = c;
    return LBM_OK;
}

int lbm_destroy(lbm_ctx* c) {
    if (!c) return LBM_OK;
    (void)hipStreamSynchronize(c->stream);
    release(c);
    delete c;
    return LBM_OK;
}

int lbm_set_stream(lbm_ctx* c, void* stream) {
    if (!c) return fail(LBM_ERR_ARG, "lbm_set_stream: null context");
    c->stream = static_cast<hipStream_t>(stream);
    return LBM_OK;
}

static int load_device_fields(lbm_ctx* c, const double* d_rho, const double* d_u, const double* d_v,
                              const double* d_w) {
    StepArgs a = make_args(c);
    a.dst = c->buf[0];
    hipLaunchKernelGGL(init_fn(c), grid_of(c), c->block, 0, c->stream, a, d_rho, d_u, d_v, d_w);
    int rc = check_launch();
    if (rc) return rc;
    // up = vp = 0 (initial.f90:51-53); a fill kernel rather than a memset so
    // the sequence stays valid under stream capture
    hipLaunchKernelGGL(k_fill, dim3(2048), dim3(256), 0, c->stream, c->prev, 2 * c->n, 0.0);
    rc = check_launch();
    if (rc) return rc;
    c->cur = 0;
    c->itc = 0;
    c->initialised = true;
    return LBM_OK;
}

int lbm_initial(lbm_ctx* c) {
    if (!c) return fail(LBM_ERR_ARG, "lbm_initial: null context");
    // initial fields into the macro arrays, then equilibrium + first collision
    // from them.  The init kernel reads and rewrites the same macro values
    // cell by cell, so the arrays can serve as both input and output.
    const long long n = c->n, layer = (long long)c->p.nx * c->p.ny;
    const int tb = 256, nb = 2048;
    hipLaunchKernelGGL(k_fill, dim3(nb), dim3(tb), 0, c->stream, c->macro, n, c->p.rho0);
    hipLaunchKernelGGL(k_fill, dim3(nb), dim3(tb), 0, c->stream, c->macro + n, n, 0.0);
    hipLaunchKernelGGL(k_fill, dim3(nb), dim3(tb), 0, c->stream, c->macro + 2 * n, 2 * n, 0.0);
    // u = U0 on the top fluid layer k = nz (initial.f90:54-61)
    hipLaunchKernelGGL(k_fill, dim3(64), dim3(tb), 0, c->stream, c->macro + n + (n - layer), layer,
                       c->p.u_lid);
    int rc = check_launch();
    if (rc) return rc;
    return load_device_fields(c, c->macro, c->macro + n, c->macro + 2 * n, c->macro + 3 * n);
}

int lbm_load_fields(lbm_ctx* c, const double* rho, const double* u, const double* v, const double* w) {
    if (!c || !rho || !u || !v || !w) return fail(LBM_ERR_ARG, "lbm_load_fields: null argument");
    const size_t bytes = sizeof(double) * c->n;
    HIP_TRY(hipMemcpyAsync(c->macro, rho, bytes, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->macro + c->n, u, bytes, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->macro + 2 * c->n, v, bytes, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->macro + 3 * c->n, w, bytes, hipMemcpyHostToDevice, c->stream));
    int rc = load_device_fields(c, c->macro, c->macro + c->n, c->macro + 2 * c->n, c->macro + 3 * c->n);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));  // caller may free the host arrays on return
    return LBM_OK;
}

int lbm_step(lbm_ctx* c, int nsteps) {
    if (!c) return fail(LBM_ERR_ARG, "lbm_step: null context");
    if (nsteps < 0) return fail(LBM_ERR_ARG, "lbm_step: nsteps = %d", nsteps);
    if (!c->initialised) return fail(LBM_ERR_STATE, "lbm_step before lbm_initial / lbm_load_fields");
    StepArgs a = make_args(c);
    const dim3 grid = grid_of(c);
    for (int s = 0; s < nsteps; ++s) {
        a.src = c->buf[c->cur];
        a.dst = c->buf[1 - c->cur];
        hipLaunchKernelGGL(step_fn(c, s == nsteps - 1), grid, c->block, 0, c->stream, a);
        c->cur = 1 - c->cur;
    }
    c->itc += nsteps;
    return check_launch();
}

int lbm_check(lbm_ctx* c, double* error_u) {
    if (!c || !error_u) return fail(LBM_ERR_ARG, "lbm_check: null argument");
    if (!c->initialised) return fail(LBM_ERR_STATE, "lbm_check before lbm_initial");
    const long long n = c->n;
    hipLaunchKernelGGL(k_check_partial, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, c->stream,
                       c->macro + n, c->macro + 2 * n, c->macro + 3 * n, c->prev, c->prev + n, n, c->red);
    hipLaunchKernelGGL(k_reduce_pairs, dim3(1), dim3(RED_THREADS), 0, c->stream, c->red, RED_BLOCKS,
                       c->red + 2 * RED_BLOCKS);
    int rc = check_launch();
    if (rc) return rc;
    double tot[2];
    HIP_TRY(hipMemcpyAsync(tot, c->red + 2 * RED_BLOCKS, sizeof tot, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    *error_u = std::sqrt(tot[0]) / std::sqrt(tot[1]);
    if (std::isnan(*error_u)) return fail(LBM_ERR_NAN, "lbm_check: residual is NaN at itc = %lld", c->itc);
    return LBM_OK;
}

int lbm_download_macro(lbm_ctx* c, double* rho, double* u, double* v, double* w) {
    if (!c) return fail(LBM_ERR_ARG, "lbm_download_macro: null context");
    if (!c->initialised) return fail(LBM_ERR_STATE, "lbm_download_macro before lbm_initial");
    double* dst[4] = {rho, u, v, w};
    const size_t bytes = sizeof(double) * c->n;
    for (int q = 0; q < 4; ++q)
        if (dst[q])
            HIP_TRY(hipMemcpyAsync(dst[q], c->macro + q * c->n, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_synchronize(lbm_ctx* c) {
    if (!c) return fail(LBM_ERR_ARG, "lbm_synchronize: null context");
    HIP_TRY(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_get_info(const lbm_ctx* c, lbm_info* info) {
    if (!c || !info) return fail(LBM_ERR_ARG, "lbm_get_info: null argument");
    info->nx = c->p.nx;
    info->ny = c->p.ny;
    info->nz = c->p.nz;
    info->pitch_x = c->px;
    info->pop_stride = c->ps;
    info->steps_done = c->itc;
    info->device_bytes = c->device_bytes;
    info->bytes_per_lup = 2.0 * Q * (double)c->elem;
    info->storage = c->p.storage;
    info->collision = c->p.collision;
    return LBM_OK;
}

}  // extern "C"
