// batch.hip -- many independent lid-driven cavities of uneven sizes stepped
// together: one persistent kernel walks a flat list of 64-cell tiles of every
// instance (per-instance offset table in device memory), steps are replayed
// from a hipGraph holding 100 step launches, and the persistent launch shape
// is chosen by timing candidates (autotune).  Each cell is updated by the
// same per-cell code as the single-domain kernel (step.hpp), so every
// instance reproduces a single-domain run of it exactly.
//
// Device layout: instance m's lattice is the single-domain SoA layout of its
// own size, placed at element offset off[m] inside two shared buffers (A-B);
// macro fields and lid densities are concatenated the same way.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <new>
#include <vector>

#include "collide.hpp"
#include "context.hpp"
#include "step.hpp"

using namespace lbm;

namespace {

constexpr int TILE = 64;          // cells per tile = one wavefront
constexpr int GRAPH_STEPS = 100;  // steps captured per graph replay (even)

struct BatchInst {
    long long off;         // element offset of this instance's lattice in each buffer
    long long pxy, ps;
    long long macro_off;   // offset of this instance's cells in the macro arrays
    long long lid_off;     // offset of its top layer in the lid density array
    long long ncells;      // nx * ny * nz
    int nx, ny, nz, px;
    double snu, sq, u_lid;
    int tile_begin;        // first tile of this instance in the flat tile list
    int pad;
};

struct BatchArgs {
    const BatchInst* inst;
    int ninst;
    int total_tiles;
    const void* src;       // buffer read this step (A or B)
    void* dst;
    double* lid_rho;
    double *rho, *u, *v, *w;   // concatenated macro arrays
    float scale, inv_scale;
    int meq12_rho;
    const unsigned* step_ctr;  // device step counter (seed of FP16 stochastic rounding)
    unsigned step_add;         // this launch's step within its sequence: seed = *step_ctr + step_add
};

// Per-instance view with the members cavity_cell reads.
struct InstView {
    const void* src;
    void* dst;
    int nx, ny, nz, px;
    long long pxy, ps;
    double u_lid, snu, sq;
    float scale, inv_scale;
    int meq12_rho;
    double* lid_rho;
    double *rho, *u, *v, *w;
    unsigned seed;
};

// CPT: cells per thread -- each wave takes CPT consecutive 64-cell tiles per
// turn (one cell per lane from each), so a thread updates CPT cells per step
// in a run of CPT * 64 consecutive cells.
template <class S, int COLL, bool MACRO, int CPT>
__global__ __launch_bounds__(256) void k_batch_step(const BatchArgs b) {
    using T = typename S::T;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
    const int nwaves = (int)((gridDim.x * blockDim.x) >> 6);
    const unsigned seed = *b.step_ctr + b.step_add;
    int m = 0;  // instance of the current tile; tiles only move forward, so m does too
    for (int t0 = wave * CPT; t0 < b.total_tiles; t0 += nwaves * CPT)
    for (int t = t0; t < t0 + CPT && t < b.total_tiles; ++t) {
        while (m + 1 < b.ninst && b.inst[m + 1].tile_begin <= t) ++m;
        const BatchInst& in = b.inst[m];
        // the tile's first cell, once per wave (uniform, 32-bit), then each
        // lane steps forward by its lane number, carrying across rows
        const unsigned nx = (unsigned)in.nx, ny = (unsigned)in.ny;
        const unsigned lc0 = (unsigned)(t - in.tile_begin) * TILE;
        unsigned k = lc0 / (nx * ny);
        const unsigned rem = lc0 - k * nx * ny;
        unsigned j = rem / nx;
        unsigned i = rem - j * nx + (unsigned)lane;
        while (i >= nx) {
            i -= nx;
            if (++j == ny) {
                j = 0;
                ++k;
            }
        }
        if (k >= (unsigned)in.nz) continue;  // past the instance's last cell
        InstView a;
        a.src = static_cast<const T*>(b.src) + in.off;
        a.dst = static_cast<T*>(b.dst) + in.off;
        a.nx = in.nx;
        a.ny = in.ny;
        a.nz = in.nz;
        a.px = in.px;
        a.pxy = in.pxy;
        a.ps = in.ps;
        a.u_lid = in.u_lid;
        a.snu = in.snu;
        a.sq = in.sq;
        a.scale = b.scale;
        a.inv_scale = b.inv_scale;
        a.meq12_rho = b.meq12_rho;
        a.lid_rho = b.lid_rho + in.lid_off;
        a.rho = b.rho + in.macro_off;
        a.u = b.u + in.macro_off;
        a.v = b.v + in.macro_off;
        a.w = b.w + in.macro_off;
        a.seed = seed;
        cavity_cell<S, COLL, MACRO, false, false, true>(a, (int)i, (int)j, (int)k);
    }
}

// Equilibrium start + first collision of every instance (the initial state of
// initial.f90), from the macro arrays (rho = rho0, u = U0 on each top layer).
template <class S, int COLL>
__global__ __launch_bounds__(256) void k_batch_init(const BatchArgs b) {
    using T = typename S::T;
    using R = typename S::Real;
    const int lane = threadIdx.x & 63;
    const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int nwaves = (int)((gridDim.x * blockDim.x) >> 6);
    for (int t = wave; t < b.total_tiles; t += nwaves) {
        int lo = 0, hi = b.ninst - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (b.inst[mid].tile_begin <= t) lo = mid;
            else hi = mid - 1;
        }
        const BatchInst& in = b.inst[lo];
        const long long lc = (long long)(t - in.tile_begin) * TILE + lane;
        if (lc >= in.ncells) continue;
        const int i = (int)(lc % in.nx);
        const int j = (int)((lc / in.nx) % in.ny);
        const int k = (int)(lc / ((long long)in.nx * in.ny));
        const long long mc = in.macro_off + lc;
        const double rho = b.rho[mc], u = b.u[mc], v = b.v[mc], w = b.w[mc];
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
        if (k == in.nz - 1) b.lid_rho[in.lid_off + i + (long long)in.nx * j] = rho;
        InstView a;
        a.snu = in.snu;
        a.sq = in.sq;
        a.meq12_rho = b.meq12_rho;
        collide<S, COLL>(f, q, a);
        T* dst = static_cast<T*>(b.dst) + in.off;
        const int c = i + in.px * (j + in.ny * k);
        const unsigned h = cell_random<S>(c, 0u);
        sfor<Q>([&](auto A) {
            constexpr int al = decltype(A)::value;
            S::st(dst + (long long)al * in.ps + c, f[al], b.scale, store_random<S, al>(h));
        });
    }
}

__global__ void k_batch_fill_initial(const BatchInst* __restrict__ inst, int ninst, double rho0,
                                     double* __restrict__ rho, double* __restrict__ u, double* __restrict__ v,
                                     double* __restrict__ w, long long total) {
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < total;
         c += (long long)gridDim.x * blockDim.x) {
        int lo = 0, hi = ninst - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (inst[mid].macro_off <= c) lo = mid;
            else hi = mid - 1;
        }
        const BatchInst& in = inst[lo];
        const long long lc = c - in.macro_off;
        const long long layer = (long long)in.nx * in.ny;
        rho[c] = rho0;
        u[c] = (lc >= in.ncells - layer) ? in.u_lid : 0.0;  // top fluid layer (initial.f90:54-61)
        v[c] = 0.0;
        w[c] = 0.0;
    }
}

__global__ void k_counter_set(unsigned* ctr, unsigned v) { *ctr = v; }
__global__ void k_counter_add(unsigned* ctr, unsigned n) { *ctr = *ctr + n; }

using BatchStepFn = void (*)(const BatchArgs);
using BatchInitFn = void (*)(const BatchArgs);

// cells per thread the kernels are built for (the autotuner's choices)
constexpr int CPT_CHOICES[3] = {1, 2, 4};

template <class S>
struct BatchTable {
    // step[collision][macro][cells per thread: 1, 2, 4]
    static constexpr BatchStepFn step[2][2][3] = {
        {{k_batch_step<S, COLL_MRT, false, 1>, k_batch_step<S, COLL_MRT, false, 2>, k_batch_step<S, COLL_MRT, false, 4>},
         {k_batch_step<S, COLL_MRT, true, 1>, k_batch_step<S, COLL_MRT, true, 2>, k_batch_step<S, COLL_MRT, true, 4>}},
        {{k_batch_step<S, COLL_BGK, false, 1>, k_batch_step<S, COLL_BGK, false, 2>, k_batch_step<S, COLL_BGK, false, 4>},
         {k_batch_step<S, COLL_BGK, true, 1>, k_batch_step<S, COLL_BGK, true, 2>, k_batch_step<S, COLL_BGK, true, 4>}}};
    static constexpr BatchInitFn init[2] = {k_batch_init<S, COLL_MRT>, k_batch_init<S, COLL_BGK>};
};

}  // namespace

struct lbm_batch {
    int ninst = 0;
    lbm_params p0{};
    std::vector<BatchInst> host;
    BatchInst* inst = nullptr;     // device copy of the offset table
    size_t elem = 8;
    long long total_elems = 0, total_cells = 0, total_lid = 0;
    int total_tiles = 0;
    void* buf[2] = {nullptr, nullptr};
    int cur = 0;
    double* lid_rho = nullptr;
    double* macro = nullptr;       // 4 * total_cells
    unsigned* ctr = nullptr;       // device step counter
    hipStream_t stream = nullptr;
    float scale = 1.0f, inv_scale = 1.0f;
    long long itc = 0;
    bool initialised = false;
    int threads = 256;             // launch shape of the persistent kernel
    int blocks = 0;
    int cpt = 1;                   // cells per thread (1, 2 or 4)
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_cur = -1;            // buffer parity the graph was captured for
    int graph_threads = 0, graph_blocks = 0, graph_cpt = 0;
    std::vector<StepArgs> single; // per-instance arguments of the single-domain kernel (baseline)
    long long device_bytes = 0;
};

namespace {

BatchArgs batch_args(const lbm_batch* b) {
    BatchArgs a{};
    a.inst = b->inst;
    a.ninst = b->ninst;
    a.total_tiles = b->total_tiles;
    a.lid_rho = b->lid_rho;
    a.rho = b->macro;
    a.u = b->macro + b->total_cells;
    a.v = b->macro + 2 * b->total_cells;
    a.w = b->macro + 3 * b->total_cells;
    a.scale = b->scale;
    a.inv_scale = b->inv_scale;
    a.meq12_rho = b->p0.meq12_rho;
    a.step_ctr = b->ctr;
    return a;
}

BatchStepFn batch_step_fn(const lbm_batch* b, bool macro) {
    const int coll = b->p0.collision == LBM_BGK ? 1 : 0;
    const int ci = b->cpt == 4 ? 2 : (b->cpt == 2 ? 1 : 0);
    switch (b->p0.storage) {
    case LBM_STORE_F32: return BatchTable<StoreF32>::step[coll][macro][ci];
    case LBM_STORE_F16: return BatchTable<StoreF16>::step[coll][macro][ci];
    default: return BatchTable<StoreF64>::step[coll][macro][ci];
    }
}

BatchInitFn batch_init_fn(const lbm_batch* b) {
    const int coll = b->p0.collision == LBM_BGK ? 1 : 0;
    switch (b->p0.storage) {
    case LBM_STORE_F32: return BatchTable<StoreF32>::init[coll];
    case LBM_STORE_F16: return BatchTable<StoreF16>::init[coll];
    default: return BatchTable<StoreF64>::init[coll];
    }
}

// One persistent step launch on b->stream (src = buf[cur]), step `step_add`
// of a sequence; the sequence ends with advance_counter.
void launch_batch_step(lbm_batch* b, int cur, bool macro, int step_add) {
    BatchArgs a = batch_args(b);
    a.src = b->buf[cur];
    a.dst = b->buf[1 - cur];
    a.step_add = (unsigned)step_add;
    hipLaunchKernelGGL(batch_step_fn(b, macro), dim3(b->blocks), dim3(b->threads), 0, b->stream, a);
}

// The device step counter moves past a sequence of n steps (one launch per
// sequence instead of one per step).
void advance_counter(lbm_batch* b, int n) {
    if (n > 0) hipLaunchKernelGGL(k_counter_add, dim3(1), dim3(1), 0, b->stream, b->ctr, (unsigned)n);
}

void drop_graph(lbm_batch* b) {
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
    b->graph_cur = -1;
}

// Capture GRAPH_STEPS step launches (starting from buffer `cur`) into a
// graph; the last one writes the macro fields.
int capture_graph(lbm_batch* b, int cur) {
    drop_graph(b);
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeGlobal));
    int c = cur;
    for (int s = 0; s < GRAPH_STEPS; ++s) {
        launch_batch_step(b, c, s == GRAPH_STEPS - 1, s);
        c = 1 - c;
    }
    advance_counter(b, GRAPH_STEPS);
    hipError_t e = hipStreamEndCapture(b->stream, &b->graph);
    if (e != hipSuccess) {
        b->graph = nullptr;
        return lbm_fail(LBM_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    }
    HIP_TRY(hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
    b->graph_cur = cur;
    b->graph_threads = b->threads;
    b->graph_blocks = b->blocks;
    b->graph_cpt = b->cpt;
    return LBM_OK;
}

int default_blocks(int threads) {
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
        cus = prop.multiProcessorCount;
    return cus * (2048 / threads);  // enough waves to fill every CU
}

void release(lbm_batch* b) {
    drop_graph(b);
    for (void* p : b->buf)
        if (p) (void)hipFree(p);
    if (b->inst) (void)hipFree(b->inst);
    if (b->lid_rho) (void)hipFree(b->lid_rho);
    if (b->macro) (void)hipFree(b->macro);
    if (b->ctr) (void)hipFree(b->ctr);
    if (b->stream) (void)hipStreamDestroy(b->stream);
}

}  // namespace

extern "C" {

int lbm_batch_create(const lbm_params* params, int count, lbm_batch** out) {
    if (!params || !out || count < 1) return lbm_fail(LBM_ERR_ARG, "lbm_batch_create: bad arguments");
    *out = nullptr;
    const lbm_params& p0 = params[0];
    for (int m = 0; m < count; ++m) {
        const lbm_params& p = params[m];
        // the tile walk indexes an instance's cells in 32 bits
        if (p.nx < 2 || p.ny < 2 || p.nz < 2 || (long long)p.nx * p.ny * p.nz > (1ll << 31) - 64)
            return lbm_fail(LBM_ERR_ARG, "lbm_batch_create: instance %d grid %d x %d x %d", m, p.nx, p.ny, p.nz);
        if (!(p.tau > 0.5) || !std::isfinite(p.tau))
            return lbm_fail(LBM_ERR_TAU, "lbm_batch_create: instance %d tau = %g", m, p.tau);
        if (p.storage != p0.storage || p.collision != p0.collision || p.meq12_rho != p0.meq12_rho ||
            p.f16_scale_log2 != p0.f16_scale_log2 || p.rho0 != p0.rho0)
            return lbm_fail(LBM_ERR_ARG, "lbm_batch_create: instance %d: storage, collision, meq12_rho, "
                                         "f16_scale_log2 and rho0 must match instance 0", m);
        if (p.thermal || p.halo_faces || p.halo_loopback)
            return lbm_fail(LBM_ERR_UNSUPPORTED, "lbm_batch_create: batched instances are isothermal cavities");
    }
    if (p0.storage < LBM_STORE_F64 || p0.storage > LBM_STORE_F16 || (p0.collision != LBM_MRT && p0.collision != LBM_BGK))
        return lbm_fail(LBM_ERR_ARG, "lbm_batch_create: storage %d / collision %d", p0.storage, p0.collision);
    if (p0.f16_scale_log2 < 0 || p0.f16_scale_log2 > 20)
        return lbm_fail(LBM_ERR_ARG, "lbm_batch_create: f16_scale_log2 %d out of [0, 20]", p0.f16_scale_log2);

    lbm_batch* b = new (std::nothrow) lbm_batch();
    if (!b) return lbm_fail(LBM_ERR_NOMEM, "lbm_batch_create: host allocation failed");
    b->ninst = count;
    b->p0 = p0;
    b->elem = p0.storage == LBM_STORE_F64 ? 8 : (p0.storage == LBM_STORE_F32 ? 4 : 2);
    const int kf = p0.f16_scale_log2 ? p0.f16_scale_log2 : 12;
    b->scale = p0.storage == LBM_STORE_F16 ? std::ldexp(1.0f, kf) : 1.0f;
    b->inv_scale = 1.0f / b->scale;
    b->host.resize(count);
    long long off = 0, moff = 0, loff = 0;
    long long tiles = 0;
    const long long page = 4096 / (long long)b->elem;
    for (int m = 0; m < count; ++m) {
        const lbm_params& p = params[m];
        BatchInst& in = b->host[m];
        in.nx = p.nx;
        in.ny = p.ny;
        in.nz = p.nz;
        in.px = (p.nx + 15) / 16 * 16;
        in.pxy = (long long)in.px * p.ny;
        const long long cells = in.pxy * p.nz;
        in.ps = (cells + page - 1) / page * page + 1024 / (long long)b->elem;
        in.off = off;
        off += 19 * in.ps;
        in.ncells = (long long)p.nx * p.ny * p.nz;
        in.macro_off = moff;
        moff += in.ncells;
        in.lid_off = loff;
        loff += (long long)p.nx * p.ny;
        in.snu = 1.0 / p.tau;
        in.sq = 8.0 * (2.0 * p.tau - 1.0) / (8.0 * p.tau - 1.0);
        in.u_lid = p.u_lid;
        in.tile_begin = (int)tiles;
        in.pad = 0;
        tiles += (in.ncells + TILE - 1) / TILE;
        if (tiles > INT_MAX) {
            delete b;
            return lbm_fail(LBM_ERR_ARG, "lbm_batch_create: more than %d tiles of 64 cells in the batch", INT_MAX);
        }
    }
    b->total_elems = off;
    b->total_cells = moff;
    b->total_lid = loff;
    b->total_tiles = (int)tiles;
    hipError_t e = hipStreamCreate(&b->stream);
    for (auto& p : b->buf)
        if (e == hipSuccess) e = hipMalloc(&p, b->elem * (size_t)off);
    if (e == hipSuccess) e = hipMalloc(&b->inst, sizeof(BatchInst) * count);
    if (e == hipSuccess) e = hipMalloc(&b->lid_rho, sizeof(double) * loff);
    if (e == hipSuccess) e = hipMalloc(&b->macro, sizeof(double) * 4 * moff);
    if (e == hipSuccess) e = hipMalloc(&b->ctr, sizeof(unsigned));
    if (e == hipSuccess)
        e = hipMemcpy(b->inst, b->host.data(), sizeof(BatchInst) * count, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        release(b);
        delete b;
        (void)hipGetLastError();
        return lbm_fail(LBM_ERR_NOMEM, "lbm_batch_create: %s", hipGetErrorString(e));
    }
    b->device_bytes = (long long)(2 * b->elem * off + sizeof(double) * (loff + 4 * moff) + sizeof(BatchInst) * count);
    b->threads = 256;
    b->blocks = default_blocks(b->threads);
    // single-domain arguments per instance, for the one-launch-per-instance baseline
    b->single.resize(count);
    for (int m = 0; m < count; ++m) {
        const BatchInst& in = b->host[m];
        StepArgs& a = b->single[m];
        a = StepArgs{};
        a.nx = in.nx;
        a.ny = in.ny;
        a.nz = in.nz;
        a.px = in.px;
        a.pxy = in.pxy;
        a.ps = in.ps;
        a.u_lid = in.u_lid;
        a.snu = in.snu;
        a.sq = in.sq;
        a.scale = b->scale;
        a.inv_scale = b->inv_scale;
        a.meq12_rho = p0.meq12_rho;
        a.lid_rho = b->lid_rho + in.lid_off;
        a.rho = b->macro + in.macro_off;
        a.u = b->macro + b->total_cells + in.macro_off;
        a.v = b->macro + 2 * b->total_cells + in.macro_off;
        a.w = b->macro + 3 * b->total_cells + in.macro_off;
    }
    *out = b;
    return LBM_OK;
}

int lbm_batch_destroy(lbm_batch* b) {
    if (!b) return LBM_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    release(b);
    delete b;
    return LBM_OK;
}

int lbm_batch_initial(lbm_batch* b) {
    if (!b) return lbm_fail(LBM_ERR_ARG, "lbm_batch_initial: null batch");
    hipLaunchKernelGGL(k_batch_fill_initial, dim3(2048), dim3(256), 0, b->stream, b->inst, b->ninst, b->p0.rho0 ? b->p0.rho0 : 1.0,
                       b->macro, b->macro + b->total_cells, b->macro + 2 * b->total_cells,
                       b->macro + 3 * b->total_cells, b->total_cells);
    BatchArgs a = batch_args(b);
    a.dst = b->buf[0];
    hipLaunchKernelGGL(batch_init_fn(b), dim3(default_blocks(256)), dim3(256), 0, b->stream, a);
    hipLaunchKernelGGL(k_counter_set, dim3(1), dim3(1), 0, b->stream, b->ctr, 1u);
    int rc = lbm_check_launch();
    if (rc) return rc;
    b->cur = 0;
    b->itc = 0;
    b->initialised = true;
    return LBM_OK;
}

int lbm_batch_step(lbm_batch* b, int nsteps) {
    if (!b) return lbm_fail(LBM_ERR_ARG, "lbm_batch_step: null batch");
    if (nsteps < 0) return lbm_fail(LBM_ERR_ARG, "lbm_batch_step: nsteps = %d", nsteps);
    if (!b->initialised) return lbm_fail(LBM_ERR_STATE, "lbm_batch_step before lbm_batch_initial");
    if (nsteps == 0) return LBM_OK;
    // graph replays for whole 100-step blocks, direct launches for the rest
    int remaining = nsteps;
    while (remaining >= GRAPH_STEPS) {
        if (!b->graph_exec || b->graph_cur != b->cur || b->graph_threads != b->threads ||
            b->graph_blocks != b->blocks || b->graph_cpt != b->cpt) {
            int rc = capture_graph(b, b->cur);
            if (rc) return rc;
        }
        HIP_TRY(hipGraphLaunch(b->graph_exec, b->stream));
        remaining -= GRAPH_STEPS;  // even, so the current buffer is unchanged
    }
    for (int s = 0; s < remaining; ++s) {
        launch_batch_step(b, b->cur, s == remaining - 1, s);
        b->cur = 1 - b->cur;
    }
    advance_counter(b, remaining);
    b->itc += nsteps;
    return lbm_check_launch();
}

int lbm_batch_step_each(lbm_batch* b, int nsteps) {
    if (!b) return lbm_fail(LBM_ERR_ARG, "lbm_batch_step_each: null batch");
    if (nsteps < 0) return lbm_fail(LBM_ERR_ARG, "lbm_batch_step_each: nsteps = %d", nsteps);
    if (!b->initialised) return lbm_fail(LBM_ERR_STATE, "lbm_batch_step_each before lbm_batch_initial");
    for (int s = 0; s < nsteps; ++s) {
        const bool macro = s == nsteps - 1;
        for (int m = 0; m < b->ninst; ++m) {
            StepArgs a = b->single[m];
            a.src = static_cast<const char*>(b->buf[b->cur]) + b->elem * b->host[m].off;
            a.dst = static_cast<char*>(b->buf[1 - b->cur]) + b->elem * b->host[m].off;
            a.seed = (unsigned)(b->itc + s + 1);
            int rc = lbm_launch_cavity_step(a, b->p0.storage, b->p0.collision, macro, b->stream);
            if (rc) return rc;
        }
        b->cur = 1 - b->cur;
    }
    advance_counter(b, nsteps);
    b->itc += nsteps;
    return lbm_check_launch();
}

int lbm_batch_autotune(lbm_batch* b, int trial_steps, int* best_threads, int* best_blocks) {
    if (!b) return lbm_fail(LBM_ERR_ARG, "lbm_batch_autotune: null batch");
    if (trial_steps < 1) trial_steps = 4;
    int rc = lbm_batch_initial(b);
    if (rc) return rc;
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
        cus = prop.multiProcessorCount;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    float best = 1e30f;
    int bt = b->threads, bb = b->blocks, bc = b->cpt;
    const int threads_c[3] = {64, 128, 256};
    const int waves_per_cu_c[5] = {4, 8, 16, 32, 64};
    for (int ci = 0; ci < 3; ++ci)
    for (int ti = 0; ti < 3; ++ti)
        for (int wi = 0; wi < 5; ++wi) {
            b->cpt = CPT_CHOICES[ci];
            b->threads = threads_c[ti];
            b->blocks = std::max(1, cus * waves_per_cu_c[wi] / (threads_c[ti] / 64));
            launch_batch_step(b, b->cur, false, 0);  // warm-up
            b->cur = 1 - b->cur;
            (void)hipEventRecord(e0, b->stream);
            for (int s = 0; s < trial_steps; ++s) {
                launch_batch_step(b, b->cur, false, s + 1);
                b->cur = 1 - b->cur;
            }
            (void)hipEventRecord(e1, b->stream);
            advance_counter(b, trial_steps + 1);
            (void)hipEventSynchronize(e1);
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (ms < best) {
                best = ms;
                bt = b->threads;
                bb = b->blocks;
                bc = b->cpt;
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->threads = bt;
    b->blocks = bb;
    b->cpt = bc;
    if (best_threads) *best_threads = bt;
    if (best_blocks) *best_blocks = bb;
    rc = lbm_check_launch();
    if (rc) return rc;
    // trial steps advanced the state: start again from the initial state
    return lbm_batch_initial(b);
}

int lbm_batch_set_launch(lbm_batch* b, int threads, int blocks) {
    if (!b) return lbm_fail(LBM_ERR_ARG, "lbm_batch_set_launch: null batch");
    if (threads % 64 != 0 || threads < 64 || threads > 256 || blocks < 1)
        return lbm_fail(LBM_ERR_ARG, "lbm_batch_set_launch: %d threads x %d blocks", threads, blocks);
    b->threads = threads;
    b->blocks = blocks;
    return LBM_OK;
}

int lbm_batch_set_cells_per_thread(lbm_batch* b, int cells_per_thread) {
    if (!b) return lbm_fail(LBM_ERR_ARG, "lbm_batch_set_cells_per_thread: null batch");
    if (cells_per_thread != 1 && cells_per_thread != 2 && cells_per_thread != 4)
        return lbm_fail(LBM_ERR_ARG, "lbm_batch_set_cells_per_thread: %d (1, 2 or 4)", cells_per_thread);
    b->cpt = cells_per_thread;
    return LBM_OK;
}

int lbm_batch_synchronize(lbm_batch* b) {
    if (!b) return lbm_fail(LBM_ERR_ARG, "lbm_batch_synchronize: null batch");
    HIP_TRY(hipStreamSynchronize(b->stream));
    return LBM_OK;
}

int lbm_batch_download_macro(lbm_batch* b, int instance, double* rho, double* u, double* v, double* w) {
    if (!b) return lbm_fail(LBM_ERR_ARG, "lbm_batch_download_macro: null batch");
    if (instance < 0 || instance >= b->ninst)
        return lbm_fail(LBM_ERR_ARG, "lbm_batch_download_macro: instance %d of %d", instance, b->ninst);
    if (!b->initialised) return lbm_fail(LBM_ERR_STATE, "lbm_batch_download_macro before lbm_batch_initial");
    const BatchInst& in = b->host[instance];
    double* dst[4] = {rho, u, v, w};
    for (int q = 0; q < 4; ++q)
        if (dst[q])
            HIP_TRY(hipMemcpyAsync(dst[q], b->macro + q * b->total_cells + in.macro_off, sizeof(double) * in.ncells,
                                   hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return LBM_OK;
}

int lbm_batch_get_info(const lbm_batch* b, lbm_batch_info* info) {
    if (!b || !info) return lbm_fail(LBM_ERR_ARG, "lbm_batch_get_info: null argument");
    info->instances = b->ninst;
    info->total_cells = b->total_cells;
    info->total_tiles = b->total_tiles;
    info->steps_done = b->itc;
    info->device_bytes = b->device_bytes;
    info->bytes_per_lup = 2.0 * Q * (double)b->elem;
    info->threads = b->threads;
    info->blocks = b->blocks;
    info->graph_steps = GRAPH_STEPS;
    info->cells_per_thread = b->cpt;
    info->stream = b->stream;
    return LBM_OK;
}

}  // extern "C"
