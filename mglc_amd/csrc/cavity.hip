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
#include <mutex>
#include <new>

#include "collide.hpp"
#include "context.hpp"
#include "halo.hpp"
#include "step.hpp"

using namespace lbm;

namespace {

// The looped-back exchange folded into the step: every stored population
// that leaves through a face is also written where the looped-back unpack
// puts it in the next step's halo (a.halo_out) -- the incoming region of the
// opposite face, same slot and plane index as k_halo_pack and k_halo_unpack
// use, and for a diagonal population leaving through an edge the same edge
// region -- for the regions in a.loop_mask (those the next step reads; all
// of them when the exchange is only rehearsed).  The step then needs no pack
// or unpack kernel.  (Across GPUs the same stores would go to the
// neighbour's halo over the peer link.)
// Only cells on the box's faces store anything and the hook is skipped (one
// branch) elsewhere.  The x faces are the one lane at each end of a row: it
// writes its five leaving populations through a base pointer computed once;
// the rows on y or z faces (whole waves) write theirs and the edges'.
template <class S>
struct LoopbackHalo {
    const StepArgs& a;
    int i, j, k;
    bool xlo, xhi, ylo, yhi, zlo, zhi;  // the cell lies on that face of the box
    __device__ __forceinline__ LoopbackHalo(const StepArgs& args, int ci, int cj, int ck)
        : a(args), i(ci), j(cj), k(ck), xlo(ci == 0), xhi(ci == args.nx - 1), ylo(cj == 0),
          yhi(cj == args.ny - 1), zlo(ck == 0), zhi(ck == args.nz - 1) {}
    __device__ __forceinline__ bool active() const { return xlo || xhi || ylo || yhi || zlo || zhi; }

    // the populations leaving through x face `leave` (sign SX), into the
    // incoming region of the opposite face
    template <int SX, class T>
    __device__ __forceinline__ void x_face(T* out, const T (&v)[Q]) const {
        constexpr int leave = SX > 0 ? LBM_FACE_XP : LBM_FACE_XM, arrive = leave ^ 1;
        if (!((a.loop_mask >> arrive) & 1u)) return;
        const long long plane = (long long)a.ny * a.nz;
        T* b = out + a.hoff[arrive] + j + (long long)a.ny * k;
        sfor<Q>([&](auto A) {
            constexpr int al = decltype(A)::value;
            if constexpr (EX[al] == SX) {
                constexpr int slot = cross_slot(arrive, al, -1);
                static_assert(slot >= 0 && slot < 5, "a population leaving through a face has one of its 5 slots");
                b[slot * plane] = v[al];
            }
        });
    }

    template <class T>
    __device__ __forceinline__ void store(const T (&v)[Q]) const {
        T* out = static_cast<T*>(a.halo_out);
        if (xhi) x_face<+1>(out, v);
        if (xlo) x_face<-1>(out, v);
        if (!(ylo || yhi || zlo || zhi)) return;  // uniform over a wave: its row
        const unsigned m = a.loop_mask;
        sfor<Q>([&](auto A) {
            constexpr int al = decltype(A)::value;
            constexpr int ex = EX[al], ey = EY[al], ez = EZ[al];
            const bool lx = ex > 0 ? xhi : (ex < 0 ? xlo : false);
            bool ly = false, lz = false;
            if constexpr (ey != 0) {
                constexpr int leave = ey > 0 ? LBM_FACE_YP : LBM_FACE_YM, arrive = leave ^ 1;
                constexpr int slot = cross_slot(arrive, al, -1);
                static_assert(slot >= 0 && slot < 5, "a population leaving through a face has one of its 5 slots");
                ly = ey > 0 ? yhi : ylo;
                if (ly && ((m >> arrive) & 1u))
                    out[a.hoff[arrive] + (long long)slot * a.nx * a.nz + i + (long long)a.nx * k] = v[al];
            }
            if constexpr (ez != 0) {
                constexpr int leave = ez > 0 ? LBM_FACE_ZP : LBM_FACE_ZM, arrive = leave ^ 1;
                constexpr int slot = cross_slot(arrive, al, -1);
                static_assert(slot >= 0 && slot < 5, "a population leaving through a face has one of its 5 slots");
                lz = ez > 0 ? zhi : zlo;
                if (lz && ((m >> arrive) & 1u))
                    out[a.hoff[arrive] + (long long)slot * a.nx * a.ny + i + (long long)a.nx * j] = v[al];
            }
            if constexpr (al >= 7) {
                // a diagonal population: on its edge when it leaves through both
                // faces (an edge always has a y or z component)
                constexpr int r = edge_region(al), fa = edge_free_axis(al);
                const bool on = (ex == 0 || lx) && (ey == 0 || ly) && (ez == 0 || lz);
                if (on && ((m >> r) & 1u)) out[a.hoff[r] + (fa == 0 ? i : (fa == 1 ? j : k))] = v[al];
            }
        });
    }
};

// One lattice update per thread.  Block = (bx, by) with bx a multiple of 64:
// one wavefront per x-row segment of 64 cells, so every population load and
// store of a wave is 64 consecutive elements.
// LOOP: the exchange looped back inside the step (LoopbackHalo).
// MX: FP16 MRT with the forward moment transform on the matrix cores.
template <class S, int COLL, bool MACRO, bool HALO, bool LOOP, bool MX = false>
__global__ __launch_bounds__(256) void k_cavity_step(const StepArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= a.nx || j >= a.ny) return;
    if constexpr (LOOP)
        cavity_cell<S, COLL, MACRO, HALO, MX>(a, i, j, k, LoopbackHalo<S>{a, i, j, k});
    else
        cavity_cell<S, COLL, MACRO, HALO, MX>(a, i, j, k);
}

// The split step (halo_overlap): the boundary shell -- every cell on a halo
// face, the cells whose populations the neighbours need next -- and the
// interior, the rest.  Halo faces are y and z faces only, so the shell is
// whole x rows: first the z planes of the halo z faces (every j), then the
// rows j of the halo y faces on the remaining k.  Each cell belongs to
// exactly one of the two kernels and is updated by the same cavity_cell as
// the fused kernel, so the split step equals the fused one bit for bit.
struct ShellRows {
    int nzf, kf[2];  // halo z faces: their planes k
    int njf, jf[2];  // halo y faces: their rows j
    int k0, nzr;     // the k range [k0, k0 + nzr) not on a halo z face
    int rows;        // nzf * ny + njf * nzr
};

template <class S, int COLL, bool MACRO, bool LOOP>
__global__ __launch_bounds__(256) void k_cavity_shell(const StepArgs a, const ShellRows r) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nx) return;
    const int row = blockIdx.y;
    int j, k;
    if (row < r.nzf * a.ny) {
        const int p = row / a.ny;
        j = row - p * a.ny;
        k = r.kf[p];
    } else {
        const int t = row - r.nzf * a.ny;
        const int q = t / r.nzr;
        j = r.jf[q];
        k = r.k0 + (t - q * r.nzr);
    }
    if constexpr (LOOP)
        cavity_cell<S, COLL, MACRO, true>(a, i, j, k, LoopbackHalo<S>{a, i, j, k});
    else
        cavity_cell<S, COLL, MACRO, true>(a, i, j, k);
}

// Interior: j in [j0, j1), k in [k0, k0 + gridDim.z); no pull from these
// cells crosses a halo face, so the halo path is compiled out.
template <class S, int COLL, bool MACRO>
__global__ __launch_bounds__(256) void k_cavity_interior(const StepArgs a, int j0, int j1, int k0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = j0 + blockIdx.y * blockDim.y + threadIdx.y;
    const int k = k0 + blockIdx.z;
    if (i >= a.nx || j >= j1) return;
    cavity_cell<S, COLL, MACRO, false>(a, i, j, k);
}

struct LeavingTable {
    int pop[NFACE][5];
};
constexpr LeavingTable make_leaving() {
    LeavingTable t{};
    for (int f = 0; f < NFACE; ++f)
        for (int q = 0; q < 5; ++q) t.pop[f][q] = leaving_pop(f, q);
    return t;
}
constexpr LeavingTable LEAVING = make_leaving();

// Pack every outgoing region of state `src` into `buf` (packed layout).
// grid.y = region; grid.x covers the largest region.
template <class T>
__global__ __launch_bounds__(256) void k_halo_pack(const T* __restrict__ src, const T* __restrict__ gsrc,
                                                   T* __restrict__ buf, const StepArgs a, unsigned mask) {
    const int r = blockIdx.y;
    if (!((mask >> r) & 1u)) return;
    const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    int i, j, k, al;
    if (r >= GREGION0) {
        // thermal: the D3Q7 population leaving through face r - GREGION0
        const int f = r - GREGION0, axis = face_axis(f), hi = f & 1;
        if (e >= a.hcount[r]) return;
        if (axis == 0) { i = hi ? a.nx - 1 : 0; j = int(e % a.ny); k = int(e / a.ny); }
        else if (axis == 1) { j = hi ? a.ny - 1 : 0; i = int(e % a.nx); k = int(e / a.nx); }
        else { k = hi ? a.nz - 1 : 0; i = int(e % a.nx); j = int(e / a.nx); }
        const long long c = i + (long long)a.px * (j + (long long)a.ny * k);
        buf[a.hoff[r] + e] = gsrc[(long long)g_leaving_pop(f) * a.ps + c];
        return;
    }
    if (r < NFACE) {
        const int axis = face_axis(r), hi = r & 1;
        const long long plane = axis == 0 ? (long long)a.ny * a.nz
                                          : (axis == 1 ? (long long)a.nx * a.nz : (long long)a.nx * a.ny);
        if (e >= 5 * plane) return;
        const int slot = int(e / plane);
        const long long p = e % plane;
        if (axis == 0) { i = hi ? a.nx - 1 : 0; j = int(p % a.ny); k = int(p / a.ny); }
        else if (axis == 1) { j = hi ? a.ny - 1 : 0; i = int(p % a.nx); k = int(p / a.nx); }
        else { k = hi ? a.nz - 1 : 0; i = int(p % a.nx); j = int(p / a.nx); }
        al = LEAVING.pop[r][slot];
    } else {
        al = r - NFACE + 7;
        const int fa = edge_free_axis(al);
        const int n = fa == 0 ? a.nx : (fa == 1 ? a.ny : a.nz);
        if (e >= n) return;
        const int t = int(e);
        i = EX[al] > 0 ? a.nx - 1 : (EX[al] < 0 ? 0 : t);
        j = EY[al] > 0 ? a.ny - 1 : (EY[al] < 0 ? 0 : t);
        k = EZ[al] > 0 ? a.nz - 1 : (EZ[al] < 0 ? 0 : t);
    }
    const long long c = i + (long long)a.px * (j + (long long)a.ny * k);
    buf[a.hoff[r] + e] = src[(long long)al * a.ps + c];
}

// Take incoming regions from `recv` into the context's halo storage.  With
// loopback, face r takes what this subdomain packed for the opposite face
// (r ^ 1) and each edge its own packed line: the exchange wrapped around on
// one GPU.
template <class T>
__global__ __launch_bounds__(256) void k_halo_unpack(const T* __restrict__ recv, T* __restrict__ halo,
                                                     const StepArgs a, int loopback, unsigned mask) {
    const int r = blockIdx.y;
    if (!((mask >> r) & 1u)) return;
    const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (e >= a.hcount[r]) return;
    const int from = !loopback ? r : (r < NFACE ? (r ^ 1) : (r >= GREGION0 ? GREGION0 + ((r - GREGION0) ^ 1) : r));
    halo[a.hoff[r] + e] = recv[a.hoff[from] + e];
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
    const unsigned h = cell_random<S>(c, a.seed);
    sfor<Q>([&](auto A) {
        constexpr int al = decltype(A)::value;
        S::st(dst + (long long)al * a.ps + c, f[al], a.scale, store_random<S, al>(h));
    });
}

// ---- FP64 residual (check.f90:9-34) -----------------------------------------------

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

// T of the thermal initial state: t_hot on the first layer along `axis`,
// t_cold on the last, 0 elsewhere (axis < 0: 0 everywhere).
__global__ void k_fill_layers(double* __restrict__ T, int nx, int ny, int nz, int axis, double t_hot,
                              double t_cold) {
    const long long n = (long long)nx * ny * nz;
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n;
         c += (long long)gridDim.x * blockDim.x) {
        const long long i = c % nx, j = (c / nx) % ny, k = c / ((long long)nx * ny);
        const long long q = axis == 0 ? i : (axis == 1 ? j : k);
        const long long nq = axis == 0 ? nx : (axis == 1 ? ny : nz);
        double t = 0.0;
        if (axis >= 0 && q == 0) t = t_hot;
        if (axis >= 0 && q == nq - 1) t = t_cold;
        T[c] = t;
    }
}

using StepFn = void (*)(const StepArgs);
using InitFn = void (*)(const StepArgs, const double*, const double*, const double*, const double*);
using ShellFn = void (*)(const StepArgs, const ShellRows);
using InteriorFn = void (*)(const StepArgs, int, int, int);

template <class S>
struct Table {
    // step[collision][macro][halo][loop]
    static constexpr StepFn step[2][2][2][2] = {
        {{{k_cavity_step<S, COLL_MRT, false, false, false>, k_cavity_step<S, COLL_MRT, false, false, true>},
          {k_cavity_step<S, COLL_MRT, false, true, false>, k_cavity_step<S, COLL_MRT, false, true, true>}},
         {{k_cavity_step<S, COLL_MRT, true, false, false>, k_cavity_step<S, COLL_MRT, true, false, true>},
          {k_cavity_step<S, COLL_MRT, true, true, false>, k_cavity_step<S, COLL_MRT, true, true, true>}}},
        {{{k_cavity_step<S, COLL_BGK, false, false, false>, k_cavity_step<S, COLL_BGK, false, false, true>},
          {k_cavity_step<S, COLL_BGK, false, true, false>, k_cavity_step<S, COLL_BGK, false, true, true>}},
         {{k_cavity_step<S, COLL_BGK, true, false, false>, k_cavity_step<S, COLL_BGK, true, false, true>},
          {k_cavity_step<S, COLL_BGK, true, true, false>, k_cavity_step<S, COLL_BGK, true, true, true>}}}};
    static constexpr InitFn init[2] = {k_cavity_init<S, COLL_MRT>, k_cavity_init<S, COLL_BGK>};
    // FP16 MRT with the matrix-core forward transform: stepx[macro][halo][loop]
    static constexpr StepFn stepx[2][2][2] = {
        {{k_cavity_step<StoreF16, COLL_MRT, false, false, false, true>,
          k_cavity_step<StoreF16, COLL_MRT, false, false, true, true>},
         {k_cavity_step<StoreF16, COLL_MRT, false, true, false, true>,
          k_cavity_step<StoreF16, COLL_MRT, false, true, true, true>}},
        {{k_cavity_step<StoreF16, COLL_MRT, true, false, false, true>,
          k_cavity_step<StoreF16, COLL_MRT, true, false, true, true>},
         {k_cavity_step<StoreF16, COLL_MRT, true, true, false, true>,
          k_cavity_step<StoreF16, COLL_MRT, true, true, true, true>}}};
    static constexpr ShellFn shell[2][2][2] = {
        {{k_cavity_shell<S, COLL_MRT, false, false>, k_cavity_shell<S, COLL_MRT, false, true>},
         {k_cavity_shell<S, COLL_MRT, true, false>, k_cavity_shell<S, COLL_MRT, true, true>}},
        {{k_cavity_shell<S, COLL_BGK, false, false>, k_cavity_shell<S, COLL_BGK, false, true>},
         {k_cavity_shell<S, COLL_BGK, true, false>, k_cavity_shell<S, COLL_BGK, true, true>}}};
    static constexpr InteriorFn interior[2][2] = {
        {k_cavity_interior<S, COLL_MRT, false>, k_cavity_interior<S, COLL_MRT, true>},
        {k_cavity_interior<S, COLL_BGK, false>, k_cavity_interior<S, COLL_BGK, true>}};
};

// FP16 MRT at the default store scale 2^12, folded into the inverse
// transform (StoreF16K12): the same bits as StoreF16, 19 multiplies fewer
// per cell.
struct TableK12 {
    using S = StoreF16K12;
    // step[macro][halo][loop]
    static constexpr StepFn step[2][2][2] = {
        {{k_cavity_step<S, COLL_MRT, false, false, false>, k_cavity_step<S, COLL_MRT, false, false, true>},
         {k_cavity_step<S, COLL_MRT, false, true, false>, k_cavity_step<S, COLL_MRT, false, true, true>}},
        {{k_cavity_step<S, COLL_MRT, true, false, false>, k_cavity_step<S, COLL_MRT, true, false, true>},
         {k_cavity_step<S, COLL_MRT, true, true, false>, k_cavity_step<S, COLL_MRT, true, true, true>}}};
    static constexpr ShellFn shell[2][2] = {
        {k_cavity_shell<S, COLL_MRT, false, false>, k_cavity_shell<S, COLL_MRT, false, true>},
        {k_cavity_shell<S, COLL_MRT, true, false>, k_cavity_shell<S, COLL_MRT, true, true>}};
    static constexpr InteriorFn interior[2] = {k_cavity_interior<S, COLL_MRT, false>,
                                               k_cavity_interior<S, COLL_MRT, true>};
};

}  // namespace


namespace {

void loopback_masks(const lbm_ctx* c, unsigned* pack, unsigned* unpack);

// FP16-stored MRT on the VALU at scale 2^12: the folded-scale kernels.
bool folds_scale(const lbm_ctx* c) {
    return c->p.storage == LBM_STORE_F16 && c->p.collision != LBM_BGK && c->p.transform != 2 &&
           c->scale == 4096.0f;
}

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
    a.halo_faces = c->p.halo_faces;
    a.halo = c->hin;
    for (int r = 0; r < NREGION_ALL; ++r) {
        a.hoff[r] = c->hoff[r];
        a.hcount[r] = c->hcount[r];
    }
    unsigned pack_mask;
    loopback_masks(c, &pack_mask, &a.loop_mask);
    return a;
}

dim3 grid_of(const lbm_ctx* c) {
    return dim3((c->p.nx + c->block.x - 1) / c->block.x, (c->p.ny + c->block.y - 1) / c->block.y,
                c->p.nz);
}

StepFn step_fn(const lbm_ctx* c, bool macro, bool loop) {
    const int coll = c->p.collision == LBM_BGK ? 1 : 0;
    const int halo = c->p.halo_faces != 0 ? 1 : 0;
    switch (c->p.storage) {
    case LBM_STORE_F32: return Table<StoreF32>::step[coll][macro][halo][loop];
    case LBM_STORE_F16:
        if (coll == COLL_MRT && c->p.transform == 2) return Table<StoreF16>::stepx[macro][halo][loop];
        if (folds_scale(c)) return TableK12::step[macro][halo][loop];
        return Table<StoreF16>::step[coll][macro][halo][loop];
    default: return Table<StoreF64>::step[coll][macro][halo][loop];
    }
}

long long halo_max_region(const lbm_ctx* c) {
    long long m = 0;
    for (int r = 0; r < NREGION_ALL; ++r) m = c->hcount[r] > m ? c->hcount[r] : m;
    return m;
}

// Regions of the context's packed layout: the D3Q19 faces and edges, and for
// a thermal context the D3Q7 faces.
int regions(const lbm_ctx* c) { return c->p.thermal ? NREGION_ALL : NREGION; }

constexpr unsigned ALL_REGIONS = (1u << NREGION_ALL) - 1u;

// Launch the pack (dst = send buffer) or unpack (src = receive buffer) kernel
// for the context's storage type, for the regions in `mask`, on `stream`.
int launch_pack(lbm_ctx* c, const void* state, void* send, unsigned mask = ALL_REGIONS,
                hipStream_t stream = nullptr, const void* gstate = nullptr) {
    const StepArgs a = make_args(c);
    const dim3 grid((unsigned)((halo_max_region(c) + 255) / 256), regions(c));
    if (!stream) stream = c->stream;
    if (!gstate) gstate = c->p.thermal ? c->gbuf[c->cur] : state;
    switch (c->p.storage) {
    case LBM_STORE_F32:
        hipLaunchKernelGGL(k_halo_pack<float>, grid, dim3(256), 0, stream, (const float*)state,
                           (const float*)gstate, (float*)send, a, mask);
        break;
    case LBM_STORE_F16:
        hipLaunchKernelGGL(k_halo_pack<_Float16>, grid, dim3(256), 0, stream, (const _Float16*)state,
                           (const _Float16*)gstate, (_Float16*)send, a, mask);
        break;
    default:
        hipLaunchKernelGGL(k_halo_pack<double>, grid, dim3(256), 0, stream, (const double*)state,
                           (const double*)gstate, (double*)send, a, mask);
    }
    return lbm_check_launch();
}

int launch_unpack(lbm_ctx* c, const void* recv, int loopback, unsigned mask = ALL_REGIONS,
                  hipStream_t stream = nullptr) {
    const StepArgs a = make_args(c);
    const dim3 grid((unsigned)((halo_max_region(c) + 255) / 256), regions(c));
    if (!stream) stream = c->stream;
    switch (c->p.storage) {
    case LBM_STORE_F32:
        hipLaunchKernelGGL(k_halo_unpack<float>, grid, dim3(256), 0, stream, (const float*)recv, (float*)c->hin,
                           a, loopback, mask);
        break;
    case LBM_STORE_F16:
        hipLaunchKernelGGL(k_halo_unpack<_Float16>, grid, dim3(256), 0, stream, (const _Float16*)recv,
                           (_Float16*)c->hin, a, loopback, mask);
        break;
    default:
        hipLaunchKernelGGL(k_halo_unpack<double>, grid, dim3(256), 0, stream, (const double*)recv, (double*)c->hin,
                           a, loopback, mask);
    }
    return lbm_check_launch();
}

// Regions a looped-back exchange must move.  A face region r is read when
// face r is a halo face and is filled from the opposite face's outgoing
// region r ^ 1; an edge region is read when both faces its population enters
// through are halo faces and is filled from the same region.  With no halo
// face the loopback rehearses the whole exchange (every region).
void loopback_masks(const lbm_ctx* c, unsigned* pack, unsigned* unpack) {
    const int m = c->p.halo_faces;
    if (m == 0) {
        *pack = *unpack = ALL_REGIONS;
        return;
    }
    unsigned pm = 0, um = 0;
    for (int f = 0; f < NFACE; ++f)
        if ((m >> f) & 1) {
            um |= 1u << f;
            pm |= 1u << (f ^ 1);
        }
    for (int al = 7; al < Q; ++al) {
        bool all = true;
        for (int axis = 0; axis < 3; ++axis)
            if (ecomp(al, axis) != 0 && !((m >> entry_face(al, axis)) & 1)) all = false;
        if (all) {
            um |= 1u << edge_region(al);
            pm |= 1u << edge_region(al);
        }
    }
    if (c->p.thermal)
        for (int f = 0; f < NFACE; ++f)
            if ((m >> f) & 1) {
                um |= 1u << (GREGION0 + f);
                pm |= 1u << (GREGION0 + (f ^ 1));
            }
    *pack = pm;
    *unpack = um;
}

ShellRows shell_rows(const lbm_ctx* c) {
    const int m = c->p.halo_faces;
    ShellRows r{};
    if ((m >> LBM_FACE_ZM) & 1) r.kf[r.nzf++] = 0;
    if ((m >> LBM_FACE_ZP) & 1) r.kf[r.nzf++] = c->p.nz - 1;
    if ((m >> LBM_FACE_YM) & 1) r.jf[r.njf++] = 0;
    if ((m >> LBM_FACE_YP) & 1) r.jf[r.njf++] = c->p.ny - 1;
    r.k0 = ((m >> LBM_FACE_ZM) & 1) ? 1 : 0;
    r.nzr = c->p.nz - r.nzf;
    r.rows = r.nzf * c->p.ny + r.njf * r.nzr;
    return r;
}

// Launch the two halves of one split step for args `a`.
int launch_shell(lbm_ctx* c, const StepArgs& a, bool macro, bool loop) {
    const ShellRows r = shell_rows(c);
    if (r.rows == 0) return LBM_OK;
    const int coll = c->p.collision == LBM_BGK ? 1 : 0;
    ShellFn fn;
    switch (c->p.storage) {
    case LBM_STORE_F32: fn = Table<StoreF32>::shell[coll][macro][loop]; break;
    case LBM_STORE_F16:
        fn = folds_scale(c) ? TableK12::shell[macro][loop] : Table<StoreF16>::shell[coll][macro][loop];
        break;
    default: fn = Table<StoreF64>::shell[coll][macro][loop];
    }
    const int bx = c->block.x;
    hipLaunchKernelGGL(fn, dim3((c->p.nx + bx - 1) / bx, r.rows), dim3(bx), 0, c->stream, a, r);
    return lbm_check_launch();
}

int launch_interior(lbm_ctx* c, const StepArgs& a, bool macro) {
    const int m = c->p.halo_faces;
    const int j0 = ((m >> LBM_FACE_YM) & 1) ? 1 : 0, j1 = c->p.ny - (((m >> LBM_FACE_YP) & 1) ? 1 : 0);
    const int k0 = ((m >> LBM_FACE_ZM) & 1) ? 1 : 0, k1 = c->p.nz - (((m >> LBM_FACE_ZP) & 1) ? 1 : 0);
    if (j1 <= j0 || k1 <= k0) return LBM_OK;
    const int coll = c->p.collision == LBM_BGK ? 1 : 0;
    InteriorFn fn;
    switch (c->p.storage) {
    case LBM_STORE_F32: fn = Table<StoreF32>::interior[coll][macro]; break;
    case LBM_STORE_F16: fn = folds_scale(c) ? TableK12::interior[macro] : Table<StoreF16>::interior[coll][macro]; break;
    default: fn = Table<StoreF64>::interior[coll][macro];
    }
    const dim3 grid((c->p.nx + c->block.x - 1) / c->block.x, (j1 - j0 + c->block.y - 1) / c->block.y, k1 - k0);
    hipLaunchKernelGGL(fn, grid, c->block, c->step_lds, c->stream, a, j0, j1, k0);
    return lbm_check_launch();
}

// With the loopback the step writes the next step's halo itself; a halo
// that does not belong to the current state (after lbm_initial,
// lbm_load_fields, lbm_load_state, a caller-driven step or unpack) is
// exchanged once first.
int fresh_loopback_halo(lbm_ctx* c) {
    if (c->halo_fresh) return LBM_OK;
    unsigned pmask, umask;
    loopback_masks(c, &pmask, &umask);
    int rc = launch_pack(c, c->buf[c->cur], c->hsend, pmask);
    if (!rc) rc = launch_unpack(c, c->hsend, 1, umask);
    if (rc) return rc;
    c->halo_fresh = true;
    return LBM_OK;
}

}  // namespace

int halo_loopback_exchange(lbm_ctx* c) {
    unsigned pmask, umask;
    loopback_masks(c, &pmask, &umask);
    int rc = launch_pack(c, c->buf[c->cur], c->hsend, pmask);
    if (!rc) rc = launch_unpack(c, c->hsend, 1, umask);
    return rc;
}

namespace {

// After a looped-back step: the halo it wrote is the one the next step reads.
void swap_halos(lbm_ctx* c) {
    void* t = c->hin;
    c->hin = c->hin_next;
    c->hin_next = t;
}

// nsteps split steps: the shell, then the interior.  With the loopback the
// shell also writes the next step's halo (LoopbackHalo) into the second halo
// buffer, and the two buffers swap after each step.
int step_split(lbm_ctx* c, int nsteps) {
    const bool loop = c->p.halo_loopback != 0;
    int rc = LBM_OK;
    if (loop && nsteps > 0) {
        rc = fresh_loopback_halo(c);
        if (rc) return rc;
    }
    for (int s = 0; s < nsteps; ++s) {
        const bool macro = s == nsteps - 1;
        StepArgs a = make_args(c);
        a.src = c->buf[c->cur];
        a.dst = c->buf[1 - c->cur];
        a.seed = (unsigned)(c->itc + s + 1);
        a.halo_out = c->hin_next;
        rc = launch_shell(c, a, macro, loop);
        if (!rc) rc = launch_interior(c, a, macro);
        if (rc) return rc;
        c->cur = 1 - c->cur;
        if (loop) swap_halos(c);
    }
    return LBM_OK;
}

InitFn init_fn(const lbm_ctx* c) {
    const int coll = c->p.collision == LBM_BGK ? 1 : 0;
    switch (c->p.storage) {
    case LBM_STORE_F32: return Table<StoreF32>::init[coll];
    case LBM_STORE_F16: return Table<StoreF16>::init[coll];
    default: return Table<StoreF64>::init[coll];
    }
}

void release(lbm_ctx* c) {
    for (void* b : c->buf)
        if (b) (void)hipFree(b);
    if (c->lid_rho) (void)hipFree(c->lid_rho);
    if (c->macro) (void)hipFree(c->macro);
    if (c->prev) (void)hipFree(c->prev);
    if (c->red) (void)hipFree(c->red);
    if (c->hin) (void)hipFree(c->hin);
    if (c->hsend) (void)hipFree(c->hsend);
    thermal_release(c);
    particles_release(c);
    if (c->hin_next) (void)hipFree(c->hin_next);
}

}  // namespace

namespace {
thread_local char g_err[512] = "";
}  // namespace

int lbm_launch_cavity_step(const StepArgs& a, int storage, int collision, bool macro, hipStream_t stream) {
    const int coll = collision == LBM_BGK ? 1 : 0;
    StepFn fn;
    switch (storage) {
    case LBM_STORE_F32: fn = Table<StoreF32>::step[coll][macro][0][0]; break;
    case LBM_STORE_F16: fn = Table<StoreF16>::step[coll][macro][0][0]; break;
    default: fn = Table<StoreF64>::step[coll][macro][0][0];
    }
    const int bx = a.nx >= 256 ? 256 : (a.nx + 63) / 64 * 64;
    const dim3 block(bx, 256 / bx, 1);
    const dim3 grid((a.nx + block.x - 1) / block.x, (a.ny + block.y - 1) / block.y, a.nz);
    hipLaunchKernelGGL(fn, grid, block, 0, stream, a);
    return lbm_check_launch();
}

// ---- scratch of the entry point that allocates nothing -------------------------------
// lbm_check_sums_noalloc keeps its per-block partial sums (the scratch of
// k_check_partial / k_reduce_pairs) in this fixed array instead of a context's
// buffer: 2 x RED_BLOCKS doubles (16 KiB), one copy per device, held until the
// process ends.  In host code `scratch` names a host-side stand-in, not a
// device address: host code never passes it to a kernel or a copy, but looks
// up each device's own address once (noalloc_scratch) and passes that.
static __device__ __align__(256) double scratch[2 * RED_BLOCKS];

namespace {
constexpr long long NOALLOC_MAX_CELLS = 1LL << 40;
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[NOALLOC_MAX_DEVICES];  // device address of `scratch` per device, once looked up

// The current device's address of `scratch`, looked up (hipGetSymbolAddress,
// with that device current) on the first call on each device.
int noalloc_scratch(double** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return lbm_fail(LBM_ERR_HIP, "lbm_check_sums_noalloc: hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return lbm_fail(LBM_ERR_UNSUPPORTED, "lbm_check_sums_noalloc: device %d beyond the first %d", dev,
                        NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return lbm_fail(LBM_ERR_HIP, "lbm_check_sums_noalloc: hipGetSymbolAddress(scratch) on device %d: %s", dev,
                            hipGetErrorString(e));
        g_scratch_dev[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_dev[dev];
    return LBM_OK;
}
}  // namespace

int lbm_fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

int lbm_check_launch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return lbm_fail(LBM_ERR_HIP, "kernel launch failed: %s", hipGetErrorString(e));
    return LBM_OK;
}

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
    if (!p || !out) return lbm_fail(LBM_ERR_ARG, "lbm_create: null argument");
    *out = nullptr;
    if (p->nx < 2 || p->ny < 2 || p->nz < 2)
        return lbm_fail(LBM_ERR_ARG, "lbm_create: grid %d x %d x %d too small", p->nx, p->ny, p->nz);
    if (!(p->tau > 0.5) || !std::isfinite(p->tau))
        return lbm_fail(LBM_ERR_TAU, "lbm_create: tau = %g must be > 0.5", p->tau);
    if (p->collision != LBM_MRT && p->collision != LBM_BGK)
        return lbm_fail(LBM_ERR_ARG, "lbm_create: collision %d unknown", p->collision);
    if (p->storage < LBM_STORE_F64 || p->storage > LBM_STORE_F16)
        return lbm_fail(LBM_ERR_ARG, "lbm_create: storage %d unknown", p->storage);
    if (p->thermal != 0 && p->thermal != 1) return lbm_fail(LBM_ERR_ARG, "lbm_create: thermal = %d", p->thermal);
    if (p->thermal) {
        if (p->collision != LBM_MRT)
            return lbm_fail(LBM_ERR_UNSUPPORTED, "lbm_create: the thermal path's flow collision is MRT "
                                                 "(evolution_f.F90)");
        if (p->u_lid != 0.0)
            return lbm_fail(LBM_ERR_ARG, "lbm_create: the buoyancy cavity has no moving lid (u_lid = %g)", p->u_lid);
        if (p->thermal_axis < -1 || p->thermal_axis > 2)
            return lbm_fail(LBM_ERR_ARG, "lbm_create: thermal_axis %d", p->thermal_axis);
        if (!(p->para_a + 6.0 > 0.0))
            return lbm_fail(LBM_ERR_TAU, "lbm_create: paraA = %g gives a non-positive diffusivity", p->para_a);
    }
    if (p->f16_scale_log2 < 0 || p->f16_scale_log2 > 20)
        return lbm_fail(LBM_ERR_ARG, "lbm_create: f16_scale_log2 %d out of [0, 20]", p->f16_scale_log2);
    if (p->transform < 0 || p->transform > 2)
        return lbm_fail(LBM_ERR_ARG, "lbm_create: transform %d (0 library's choice, 1 VALU, 2 MFMA)", p->transform);
    // default launch shape: one block per 256-cell x-row run (measured best at
    // 256^3: whole rows per workgroup share their partial edge cache lines)
    const int bx_def = p->nx >= 256 ? 256 : (p->nx + 63) / 64 * 64;
    const int bx = p->block_x ? p->block_x : bx_def;
    const int by = p->block_y ? p->block_y : 256 / bx;
    if (bx % 64 != 0 || bx * by > 256 || by < 1)
        return lbm_fail(LBM_ERR_ARG, "lbm_create: block %d x %d (x must be a multiple of 64, <= 256 threads)",
                    bx, by);

    if (p->waves_per_cu < 0 || p->waves_per_cu > 32 || p->waves_per_cu % ((bx * by + 63) / 64) != 0)
        return lbm_fail(LBM_ERR_ARG, "lbm_create: waves_per_cu %d is not a multiple of the %d waves of a block "
                                     "in [0, 32]", p->waves_per_cu, (bx * by + 63) / 64);
    if (p->halo_faces < 0 || p->halo_faces > 63)
        return lbm_fail(LBM_ERR_ARG, "lbm_create: halo_faces 0x%x is not a 6-bit face mask", p->halo_faces);
    if (p->halo_overlap != 0 && p->halo_overlap != 1)
        return lbm_fail(LBM_ERR_ARG, "lbm_create: halo_overlap = %d", p->halo_overlap);
    if (p->halo_overlap) {
        if (p->thermal)
            return lbm_fail(LBM_ERR_UNSUPPORTED, "lbm_create: the thermal path has no split step");
        if (p->halo_faces & ((1 << LBM_FACE_XM) | (1 << LBM_FACE_XP)))
            return lbm_fail(LBM_ERR_UNSUPPORTED, "lbm_create: the split step keeps x whole; decompose along y "
                                                 "and z (halo_faces 0x%x)", p->halo_faces);
        if (p->halo_faces == 0)
            return lbm_fail(LBM_ERR_ARG, "lbm_create: halo_overlap without halo faces has nothing to overlap");
        const int m = p->halo_faces;
        if (p->halo_loopback && (((m >> 2) & 1) != ((m >> 3) & 1) || ((m >> 4) & 1) != ((m >> 5) & 1)))
            return lbm_fail(LBM_ERR_ARG, "lbm_create: a looped-back split step needs both faces of each halo axis "
                                         "(halo_faces 0x%x)", m);
        const long long rows = 2LL * p->ny + 2LL * p->nz;
        if (rows > 65535)
            return lbm_fail(LBM_ERR_UNSUPPORTED, "lbm_create: %lld shell rows exceed the launch grid", rows);
    }

    lbm_ctx* c = new (std::nothrow) lbm_ctx();
    if (!c) return lbm_fail(LBM_ERR_NOMEM, "lbm_create: host allocation failed");
    c->p = *p;
    if (c->p.rho0 == 0.0) c->p.rho0 = 1.0;
    c->block = dim3(bx, by, 1);
    c->elem = p->storage == LBM_STORE_F64 ? 8 : (p->storage == LBM_STORE_F32 ? 4 : 2);
    c->px = (p->nx + 15) / 16 * 16;
    c->pxy = (long long)c->px * p->ny;
    const long long cells = c->pxy * p->nz;
    if (cells >= (1LL << 31)) {
        delete c;
        return lbm_fail(LBM_ERR_ARG, "lbm_create: %lld padded cells exceed the 32-bit cell index", cells);
    }
    // population stride: whole 4 KiB pages plus a 1 KiB stagger per stream
    const long long page = 4096 / (long long)c->elem;
    c->ps = (cells + page - 1) / page * page + 1024 / (long long)c->elem;
    c->n = (long long)p->nx * p->ny * p->nz;
    const int k = p->f16_scale_log2 ? p->f16_scale_log2 : 12;
    c->scale = p->storage == LBM_STORE_F16 ? std::ldexp(1.0f, k) : 1.0f;
    c->inv_scale = 1.0f / c->scale;

    // packed halo layout: 6 faces (5 populations x face plane), 12 edge lines
    {
        long long off = 0;
        for (int r = 0; r < NREGION; ++r) {
            long long n;
            if (r < NFACE) {
                const int axis = face_axis(r);
                n = 5LL * (axis == 0 ? (long long)p->ny * p->nz
                                     : (axis == 1 ? (long long)p->nx * p->nz : (long long)p->nx * p->ny));
            } else {
                const int fa = edge_free_axis(r - NFACE + 7);
                n = fa == 0 ? p->nx : (fa == 1 ? p->ny : p->nz);
            }
            c->hoff[r] = off;
            c->hcount[r] = n;
            off += n;
        }
        // thermal: one D3Q7 population per face plane
        for (int f = 0; p->thermal && f < NFACE; ++f) {
            const int axis = face_axis(f);
            const long long n = axis == 0 ? (long long)p->ny * p->nz
                                          : (axis == 1 ? (long long)p->nx * p->nz : (long long)p->nx * p->ny);
            c->hoff[GREGION0 + f] = off;
            c->hcount[GREGION0 + f] = n;
            off += n;
        }
        c->htotal = off;
    }

    (void)hipGetDevice(&c->device);
    // Occupancy of the memory-bound steps: with fewer waves in flight each
    // HBM stream keeps more of its locality.  Measured at 256^3 (MRT,
    // alternating runs): 16 waves per CU (four blocks of 256) instead of the
    // 32 the registers allow make the step 0.8-1.8 % faster in FP64 and
    // 2.6-3.6 % in FP32; FP16, bound by VALU issue, loses 25 % and keeps every
    // wave, and so do the kernels that read a halo (4 % slower capped: more
    // registers and work per wave).  The cap is an LDS reservation per block
    // (the CU's LDS over the blocks of the allowed waves) that the kernels
    // never touch.  An explicit waves_per_cu applies to the fused step (also
    // when it reads a halo), the split step's interior and the thermal step.
    {
        const int waves_per_block = (int)((c->block.x * c->block.y + 63) / 64);
        const int policy = (!p->thermal && p->storage != LBM_STORE_F16) ? 16 : 0;
        const int want = p->waves_per_cu ? p->waves_per_cu : policy;
        if (want > 0 && want % waves_per_block == 0) {
            int lds_cu = 0, lds_block = 0;
            unsigned reserve = 0;
            if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, c->device) ==
                    hipSuccess &&
                hipDeviceGetAttribute(&lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) ==
                    hipSuccess) {
                const int r = lds_cu / (want / waves_per_block);
                if (r > 0 && r <= lds_block) reserve = (unsigned)r;
            }
            (void)hipGetLastError();
            if (!reserve && p->waves_per_cu) {
                delete c;
                return lbm_fail(LBM_ERR_UNSUPPORTED, "lbm_create: the device cannot hold %d waves per CU by an "
                                                     "LDS reservation (%d bytes of LDS per CU, %d per block)",
                                p->waves_per_cu, lds_cu, lds_block);
            }
            c->step_lds = reserve;
            c->step_lds_halo = p->waves_per_cu ? reserve : 0u;
        }
    }
    const size_t lat = (size_t)c->ps * Q * c->elem;
    hipError_t e = hipSuccess;
    for (auto& b : c->buf)
        if (e == hipSuccess) e = hipMalloc(&b, lat);
    if (e == hipSuccess) e = hipMalloc(&c->lid_rho, sizeof(double) * p->nx * p->ny);
    if (e == hipSuccess) e = hipMalloc(&c->macro, sizeof(double) * 4 * c->n);
    if (e == hipSuccess) e = hipMalloc(&c->prev, sizeof(double) * 2 * c->n);
    if (e == hipSuccess) e = hipMalloc(&c->red, sizeof(double) * RED_MAX_SUMS * (RED_BLOCKS + 1));
    if (e == hipSuccess) e = hipMalloc(&c->hin, c->elem * c->htotal);
    if (e == hipSuccess) e = hipMalloc(&c->hsend, c->elem * c->htotal);
    if (e != hipSuccess) {
        release(c);
        delete c;
        (void)hipGetLastError();
        return lbm_fail(LBM_ERR_NOMEM, "lbm_create: hipMalloc failed: %s", hipGetErrorString(e));
    }
    c->device_bytes = (long long)(2 * lat + sizeof(double) * (p->nx * p->ny + 6 * c->n) +
                                  sizeof(double) * RED_MAX_SUMS * (RED_BLOCKS + 1) + 2 * c->elem * c->htotal);
    if (p->thermal) {
        int rc = thermal_alloc(c);
        if (rc) {
            (void)hipGetLastError();
            release(c);
            delete c;
            return rc == LBM_ERR_HIP ? LBM_ERR_NOMEM : rc;
        }
    }
    if (p->halo_loopback && !p->thermal) {
        if (hipMalloc(&c->hin_next, c->elem * c->htotal) != hipSuccess ||
            hipMemset(c->hin_next, 0, c->elem * c->htotal) != hipSuccess) {
            (void)hipGetLastError();
            release(c);
            delete c;
            return lbm_fail(LBM_ERR_NOMEM, "lbm_create: allocating the second halo buffer failed");
        }
        c->device_bytes += (long long)(c->elem * c->htotal);
    }
    // incoming halos start as zeros until the first exchange
    if (hipMemset(c->hin, 0, c->elem * c->htotal) != hipSuccess) {
        release(c);
        delete c;
        return lbm_fail(LBM_ERR_HIP, "lbm_create: clearing the halo failed");
    }
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
    if (!c) return lbm_fail(LBM_ERR_ARG, "lbm_set_stream: null context");
    c->stream = static_cast<hipStream_t>(stream);
    return LBM_OK;
}

static int load_device_fields(lbm_ctx* c, const double* d_rho, const double* d_u, const double* d_v,
                              const double* d_w) {
    int rc;
    if (c->p.thermal) {
        rc = thermal_load(c, d_rho, d_u, d_v, d_w, c->temp);
        if (rc) return rc;
        hipLaunchKernelGGL(k_fill, dim3(2048), dim3(256), 0, c->stream, c->prev_t, 2 * c->n, 0.0);
    } else {
        StepArgs a = make_args(c);
        a.dst = c->buf[0];
        hipLaunchKernelGGL(init_fn(c), grid_of(c), c->block, 0, c->stream, a, d_rho, d_u, d_v, d_w);
    }
    rc = lbm_check_launch();
    if (rc) return rc;
    // up = vp = 0 (initial.f90:51-53); a fill kernel rather than a memset so
    // the sequence stays valid under stream capture
    hipLaunchKernelGGL(k_fill, dim3(2048), dim3(256), 0, c->stream, c->prev, 2 * c->n, 0.0);
    rc = lbm_check_launch();
    if (rc) return rc;
    c->cur = 0;
    c->itc = 0;
    c->initialised = true;
    return LBM_OK;
}

int lbm_initial(lbm_ctx* c) {
    if (!c) return lbm_fail(LBM_ERR_ARG, "lbm_initial: null context");
    c->in_step = false;  // a fresh start abandons a step in progress
    c->halo_fresh = false;
    particles_release(c);  // a fresh start detaches any spheres
    // initial fields into the macro arrays, then equilibrium + first collision
    // from them.  The init kernel reads and rewrites the same macro values
    // cell by cell, so the arrays can serve as both input and output.
    const long long n = c->n, layer = (long long)c->p.nx * c->p.ny;
    const int tb = 256, nb = 2048;
    hipLaunchKernelGGL(k_fill, dim3(nb), dim3(tb), 0, c->stream, c->macro, n, c->p.rho0);
    hipLaunchKernelGGL(k_fill, dim3(nb), dim3(tb), 0, c->stream, c->macro + n, n, 0.0);
    hipLaunchKernelGGL(k_fill, dim3(nb), dim3(tb), 0, c->stream, c->macro + 2 * n, 2 * n, 0.0);
    // u = U0 on the top fluid layer k = nz of the subdomain that holds the lid
    // (initial.f90:54-61 tests coords(2) == dims(2) - 1; here: z+ is a wall)
    if (!((c->p.halo_faces >> LBM_FACE_ZP) & 1))
        hipLaunchKernelGGL(k_fill, dim3(64), dim3(tb), 0, c->stream, c->macro + n + (n - layer), layer,
                           c->p.u_lid);
    if (c->p.thermal) {
        // T = 0 except the layers next to the held walls (initial.F90:126-162)
        hipLaunchKernelGGL(k_fill_layers, dim3(2048), dim3(256), 0, c->stream, c->temp, c->p.nx, c->p.ny, c->p.nz,
                           c->p.thermal_axis, c->p.t_hot, c->p.t_cold);
    }
    int rc = lbm_check_launch();
    if (rc) return rc;
    return load_device_fields(c, c->macro, c->macro + n, c->macro + 2 * n, c->macro + 3 * n);
}

int lbm_load_fields(lbm_ctx* c, const double* rho, const double* u, const double* v, const double* w,
                    const double* T) {
    if (!c || !rho || !u || !v || !w) return lbm_fail(LBM_ERR_ARG, "lbm_load_fields: null argument");
    c->in_step = false;
    c->halo_fresh = false;
    particles_release(c);
    const size_t bytes = sizeof(double) * c->n;
    if (c->p.thermal) {
        if (T)
            HIP_TRY(hipMemcpyAsync(c->temp, T, bytes, hipMemcpyHostToDevice, c->stream));
        else
            hipLaunchKernelGGL(k_fill, dim3(2048), dim3(256), 0, c->stream, c->temp, c->n, 0.0);
    }
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
    if (!c) return lbm_fail(LBM_ERR_ARG, "lbm_step: null context");
    if (nsteps < 0) return lbm_fail(LBM_ERR_ARG, "lbm_step: nsteps = %d", nsteps);
    if (!c->initialised) return lbm_fail(LBM_ERR_STATE, "lbm_step before lbm_initial / lbm_load_fields");
    if (c->in_step) return lbm_fail(LBM_ERR_STATE, "lbm_step between lbm_step_begin and lbm_step_end");
    if (c->p.thermal) {
        int rc = thermal_step(c, nsteps);
        c->itc += nsteps;
        return rc;
    }
    if (c->part) {
        int rc = particles_step(c, nsteps);
        c->itc += nsteps;
        return rc;
    }
    if (c->p.halo_faces != 0 && !c->p.halo_loopback && nsteps > 1)
        return lbm_fail(LBM_ERR_STATE, "lbm_step: with halo faces and no loopback the halo must be exchanged "
                                   "between steps; call lbm_step(ctx, 1)");
    if (c->p.halo_overlap) {
        int rc = step_split(c, nsteps);
        if (rc) return rc;
        c->itc += nsteps;
        return lbm_check_launch();
    }
    StepArgs a = make_args(c);
    const dim3 grid = grid_of(c);
    const bool loop = c->p.halo_loopback != 0;
    if (loop && nsteps > 0) {
        int rc = fresh_loopback_halo(c);
        if (rc) return rc;
    }
    const unsigned lds = c->p.halo_faces ? c->step_lds_halo : c->step_lds;  // by default the halo-reading
                                                                            // kernel keeps every wave
    for (int s = 0; s < nsteps; ++s) {
        a.src = c->buf[c->cur];
        a.dst = c->buf[1 - c->cur];
        a.seed = (unsigned)(c->itc + s + 1);
        a.halo = c->hin;
        a.halo_out = c->hin_next;
        hipLaunchKernelGGL(step_fn(c, s == nsteps - 1, loop), grid, c->block, lds, c->stream, a);
        c->cur = 1 - c->cur;
        if (loop) swap_halos(c);
    }
    c->itc += nsteps;
    return lbm_check_launch();
}

int lbm_check_sums(lbm_ctx* c, double* e1, double* e2) {
    if (!c || !e1 || !e2) return lbm_fail(LBM_ERR_ARG, "lbm_check_sums: null argument");
    if (!c->initialised) return lbm_fail(LBM_ERR_STATE, "lbm_check before lbm_initial");
    if (c->in_step) return lbm_fail(LBM_ERR_STATE, "lbm_check between lbm_step_begin and lbm_step_end");
    if (c->p.thermal) {
        double s4[4];
        int rc = thermal_check(c, s4);
        if (rc) return rc;
        *e1 = s4[0];
        *e2 = s4[1];
        return LBM_OK;
    }
    const long long n = c->n;
    hipLaunchKernelGGL(k_check_partial, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, c->stream,
                       c->macro + n, c->macro + 2 * n, c->macro + 3 * n, c->prev, c->prev + n, n, c->red);
    hipLaunchKernelGGL(k_reduce_pairs, dim3(1), dim3(RED_THREADS), 0, c->stream, c->red, RED_BLOCKS,
                       c->red + 2 * RED_BLOCKS);
    int rc = lbm_check_launch();
    if (rc) return rc;
    double tot[2];
    HIP_TRY(hipMemcpyAsync(tot, c->red + 2 * RED_BLOCKS, sizeof tot, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    *e1 = tot[0];
    *e2 = tot[1];
    return LBM_OK;
}

int lbm_check_sums_noalloc(const double* u, const double* v, const double* w, double* up, double* vp,
                           long long n, double* sums_dev, void* hip_stream) {
    if (!u || !v || !w || !up || !vp || !sums_dev)
        return lbm_fail(LBM_ERR_ARG, "lbm_check_sums_noalloc: null argument");
    if (n < 1 || n > NOALLOC_MAX_CELLS)
        return lbm_fail(LBM_ERR_ARG, "lbm_check_sums_noalloc: n = %lld cells outside [1, 2^40]", n);
    double* partial = nullptr;
    int rc = noalloc_scratch(&partial);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(hip_stream);
    hipLaunchKernelGGL(k_check_partial, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, s, u, v, w, up, vp, n, partial);
    hipLaunchKernelGGL(k_reduce_pairs, dim3(1), dim3(RED_THREADS), 0, s, partial, RED_BLOCKS, sums_dev);
    return lbm_check_launch();
}

int lbm_check(lbm_ctx* c, double* error_u) {
    if (!error_u) return lbm_fail(LBM_ERR_ARG, "lbm_check: null argument");
    double e1 = 0.0, e2 = 0.0;
    int rc = lbm_check_sums(c, &e1, &e2);
    if (rc) return rc;
    *error_u = std::sqrt(e1) / std::sqrt(e2);
    if (std::isnan(*error_u)) return lbm_fail(LBM_ERR_NAN, "lbm_check: residual is NaN at itc = %lld", c->itc);
    return LBM_OK;
}

int lbm_check_thermal(lbm_ctx* c, double* error_u, double* error_t) {
    if (!c || !error_u || !error_t) return lbm_fail(LBM_ERR_ARG, "lbm_check_thermal: null argument");
    if (!c->p.thermal) return lbm_fail(LBM_ERR_STATE, "lbm_check_thermal on an isothermal context");
    if (!c->initialised) return lbm_fail(LBM_ERR_STATE, "lbm_check_thermal before lbm_initial");
    double s4[4];
    int rc = thermal_check(c, s4);
    if (rc) return rc;
    *error_u = std::sqrt(s4[0]) / std::sqrt(s4[1]);
    *error_t = s4[2] / s4[3];
    if (std::isnan(*error_u) || std::isnan(*error_t))
        return lbm_fail(LBM_ERR_NAN, "lbm_check_thermal: residual is NaN at itc = %lld", c->itc);
    return LBM_OK;
}

int lbm_nusselt(lbm_ctx* c, double* nu_vol, double* nu_wall) {
    if (!c || !nu_vol || !nu_wall) return lbm_fail(LBM_ERR_ARG, "lbm_nusselt: null argument");
    if (!c->p.thermal) return lbm_fail(LBM_ERR_STATE, "lbm_nusselt on an isothermal context");
    if (!c->initialised) return lbm_fail(LBM_ERR_STATE, "lbm_nusselt before lbm_initial");
    return thermal_nusselt(c, nu_vol, nu_wall);
}

int lbm_reynolds(lbm_ctx* c, double* re) {
    if (!c || !re) return lbm_fail(LBM_ERR_ARG, "lbm_reynolds: null argument");
    if (!c->initialised) return lbm_fail(LBM_ERR_STATE, "lbm_reynolds before lbm_initial");
    if (c->part) return lbm_fail(LBM_ERR_UNSUPPORTED, "lbm_reynolds: the macro fields of a particle run hold solids");
    return flow_reynolds(c, re);
}

int lbm_download_temperature(lbm_ctx* c, double* T) {
    if (!c || !T) return lbm_fail(LBM_ERR_ARG, "lbm_download_temperature: null argument");
    if (!c->p.thermal) return lbm_fail(LBM_ERR_STATE, "lbm_download_temperature on an isothermal context");
    if (!c->initialised) return lbm_fail(LBM_ERR_STATE, "lbm_download_temperature before lbm_initial");
    HIP_TRY(hipMemcpyAsync(T, c->temp, sizeof(double) * c->n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_halo_layout(const lbm_ctx* c, int region, long long* offset, long long* count) {
    if (!c || !offset || !count) return lbm_fail(LBM_ERR_ARG, "lbm_halo_layout: null argument");
    if (region < 0 || region >= regions(c))
        return lbm_fail(LBM_ERR_ARG, "lbm_halo_layout: region %d (this context has %d)", region, regions(c));
    *offset = c->hoff[region];
    *count = c->hcount[region];
    return LBM_OK;
}

int lbm_halo_buffer_bytes(const lbm_ctx* c, size_t* bytes) {
    if (!c || !bytes) return lbm_fail(LBM_ERR_ARG, "lbm_halo_buffer_bytes: null argument");
    *bytes = c->elem * (size_t)c->htotal;
    return LBM_OK;
}

int lbm_halo_pack(lbm_ctx* c, void* dev_send) {
    if (!c || !dev_send) return lbm_fail(LBM_ERR_ARG, "lbm_halo_pack: null argument");
    if (!c->initialised) return lbm_fail(LBM_ERR_STATE, "lbm_halo_pack before lbm_initial");
    // between lbm_step_begin and lbm_step_end: the shell of the step in progress
    return launch_pack(c, c->buf[c->in_step ? 1 - c->cur : c->cur], dev_send);
}

int lbm_step_begin(lbm_ctx* c) {
    if (!c) return lbm_fail(LBM_ERR_ARG, "lbm_step_begin: null context");
    if (!c->p.halo_overlap) return lbm_fail(LBM_ERR_STATE, "lbm_step_begin: the context was created without halo_overlap");
    if (!c->initialised) return lbm_fail(LBM_ERR_STATE, "lbm_step_begin before lbm_initial / lbm_load_fields");
    if (c->in_step) return lbm_fail(LBM_ERR_STATE, "lbm_step_begin: a step is already in progress");
    StepArgs a = make_args(c);
    a.src = c->buf[c->cur];
    a.dst = c->buf[1 - c->cur];
    a.seed = (unsigned)(c->itc + 1);
    int rc = launch_shell(c, a, true, false);
    if (rc) return rc;
    c->in_step = true;
    c->halo_fresh = false;  // the caller exchanges; a later looped-back step re-exchanges first
    return LBM_OK;
}

int lbm_step_end(lbm_ctx* c) {
    if (!c) return lbm_fail(LBM_ERR_ARG, "lbm_step_end: null context");
    if (!c->in_step) return lbm_fail(LBM_ERR_STATE, "lbm_step_end without lbm_step_begin");
    StepArgs a = make_args(c);
    a.src = c->buf[c->cur];
    a.dst = c->buf[1 - c->cur];
    a.seed = (unsigned)(c->itc + 1);
    int rc = launch_interior(c, a, true);
    if (rc) return rc;
    c->cur = 1 - c->cur;
    c->itc += 1;
    c->in_step = false;
    return LBM_OK;
}

int lbm_halo_unpack(lbm_ctx* c, const void* dev_recv) {
    if (!c || !dev_recv) return lbm_fail(LBM_ERR_ARG, "lbm_halo_unpack: null argument");
    c->halo_fresh = false;  // not the looped-back halo of the current state any more
    return launch_unpack(c, dev_recv, 0);
}

int lbm_download_macro(lbm_ctx* c, double* rho, double* u, double* v, double* w) {
    if (!c) return lbm_fail(LBM_ERR_ARG, "lbm_download_macro: null context");
    if (!c->initialised) return lbm_fail(LBM_ERR_STATE, "lbm_download_macro before lbm_initial");
    if (c->in_step) return lbm_fail(LBM_ERR_STATE, "lbm_download_macro between lbm_step_begin and lbm_step_end");
    double* dst[4] = {rho, u, v, w};
    const size_t bytes = sizeof(double) * c->n;
    for (int q = 0; q < 4; ++q)
        if (dst[q])
            HIP_TRY(hipMemcpyAsync(dst[q], c->macro + q * c->n, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_synchronize(lbm_ctx* c) {
    if (!c) return lbm_fail(LBM_ERR_ARG, "lbm_synchronize: null context");
    HIP_TRY(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_get_info(const lbm_ctx* c, lbm_info* info) {
    if (!c || !info) return lbm_fail(LBM_ERR_ARG, "lbm_get_info: null argument");
    info->nx = c->p.nx;
    info->ny = c->p.ny;
    info->nz = c->p.nz;
    info->pitch_x = c->px;
    info->pop_stride = c->ps;
    info->steps_done = c->itc;
    info->device_bytes = c->device_bytes;
    info->bytes_per_lup = 2.0 * Q * (double)c->elem;
    if (c->p.thermal)  // + the D3Q7 populations and the stored force (Fz; Fx, Fy too with rotation)
        info->bytes_per_lup = 2.0 * (Q + Q7) * (double)c->elem +
                              (c->p.omega_rot != 0.0 ? 6.0 : 2.0) * (double)thermal_force_elem(c);
    info->storage = c->p.storage;
    info->collision = c->p.collision;
    return LBM_OK;
}

}  // extern "C"
