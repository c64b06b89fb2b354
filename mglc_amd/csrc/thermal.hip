// thermal.hip -- buoyancy-driven cavity on one MI355X: D3Q19 MRT flow with the
// Guo body force (Boussinesq + Coriolis) coupled to a D3Q7 MRT temperature
// field, fused into one pass per step; the thermal residual and the Nusselt
// reductions.
//
// Reference step (bouyancy3d_mpi.F90:222-248, module.F90, evolution_f.F90,
// evolution_g.F90):
//   collision  (F from the previous rho, u, T; MRT + Guo source)
//   streaming, bounceback          (no-slip walls everywhere)
//   collisionT (previous u, T), streamingT, bouncebackT
//   macro (u = (sum f e + F/2) / rho), macroT (T = sum g)
// Stored between steps: f_post (19), g_post (7) and the force F of the last
// collision (macro of the next step needs F/2).  One kernel per step:
//   pull f and g (+ walls) -> macro with the stored F -> new F -> collide f
//   and g -> store.
// Device layout: the SoA layout of the isothermal path; g uses the same
// population stride; F is three FP64 arrays (only Fz is touched when the
// rotation rate is zero).
#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>

#include "collide.hpp"
#include "context.hpp"
#include "halo.hpp"
#include "step.hpp"

using namespace lbm;

namespace {

struct ThermalArgs {
    const void* fsrc;
    void* fdst;
    const void* gsrc;
    void* gdst;
    void* force;  // Fx | Fy | Fz, each ps long (FP64 with FP64 storage, else FP32)
    int nx, ny, nz, px;
    long long pxy, ps;
    double snu, sq;           // flow rates (collision.f90:94-112 assignment)
    double para_a, qd, qnu;   // D3Q7: paraA, Qd = 3 - sqrt 3, Qnu = 4 sqrt 3 - 6
    double g_beta, omega_rot, t_ref;
    double t_hot, t_cold, cw; // held walls; cw = (6 + paraA) / 21
    int taxis;                // axis of the held walls (0 x, 1 y, 2 z), -1 none
    float scale, inv_scale;   // FP16 storage: populations held as value * 2^k
    unsigned seed;            // step number: drives the stochastic rounding of FP16 stores
    double *rho, *u, *v, *w, *T;
    // subdomain halo (the members the isothermal pull from outside reads):
    // faces in halo_faces take their incoming populations from `halo`
    int halo_faces;
    const void* halo;
    long long hoff[NREGION_ALL];
    double u_lid;             // 0: no moving lid here
};

// The stored body force: FP64 beside FP64 populations, else FP32 (the force
// is ~g beta ~ 1e-8, below FP16's normal range).
template <class S>
using ForceT = std::conditional_t<std::is_same_v<S, StoreF64>, double, float>;

// A stored population read by the buoyancy step: non-temporal in FP16
// storage.  Measured on MI355X at 128^3: FP16 62.9 -> 58.9 us per step;
// FP64 and FP32 ran no faster (171.8 and 82.5 us) and fetched 4 % more
// (lines shared by neighbouring blocks evicted sooner), and the cavity, batch
// and particle kernels ran 2-17 % slower, so those keep ordinary loads.
template <class S>
__device__ __forceinline__ typename S::Real ld_stream(const typename S::T* p, float inv_scale) {
    if constexpr (S::BYTES == 2)
        return S::val(__builtin_nontemporal_load(p), inv_scale);
    else
        return S::ld(p, inv_scale);
}

// One step of the buoyancy-driven cavity.  S = StoreF64: FP64 arithmetic in
// the reference's order.  StoreF32 / StoreF16: f stored as f - w (FP16
// scaled by 2^k with stochastic rounding), g stored unshifted (scaled the
// same way in FP16), the force in FP32, all arithmetic in FP32; the macro
// fields are written as FP64 for the FP64 reductions.
template <class S, bool ROT, bool MACRO, bool HALO>
__global__ __launch_bounds__(256) void k_thermal_step(const ThermalArgs a) {
    using ST = typename S::T;
    using R = typename S::Real;
    using FT = ForceT<S>;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= a.nx || j >= a.ny) return;
    const int c = i + a.px * (j + a.ny * k);
    const ST* fsrc = static_cast<const ST*>(a.fsrc);
    const ST* gsrc = static_cast<const ST*>(a.gsrc);
    FT* force = static_cast<FT*>(a.force);

    // streaming + bounceback (evolution_f.F90:227-343): no-slip on all walls
    R f[Q];
    sfor<Q>([&](auto A) {
        constexpr int al = decltype(A)::value;
        constexpr int ex = EX[al], ey = EY[al], ez = EZ[al];
        if constexpr (al == 0) {
            f[0] = ld_stream<S>(fsrc + c, a.inv_scale);
        } else {
            bool in = true;
            if constexpr (ex != 0) in = in && (unsigned)(i - ex) < (unsigned)a.nx;
            if constexpr (ey != 0) in = in && (unsigned)(j - ey) < (unsigned)a.ny;
            if constexpr (ez != 0) in = in && (unsigned)(k - ez) < (unsigned)a.nz;
            const long long off = (long long)ex + (long long)ey * a.px + (long long)ez * a.pxy;
            if constexpr (HALO) {
                // from the halo where every crossed face borders a neighbour,
                // else bounce-back (no lid: top = false)
                f[al] = in ? ld_stream<S>(fsrc + (al * a.ps + c - off), a.inv_scale)
                           : pull_outside<S, al, true>(a, fsrc, i, j, k, c, false, 0.0);
            } else {
                f[al] = in ? ld_stream<S>(fsrc + (al * a.ps + c - off), a.inv_scale)
                           : ld_stream<S>(fsrc + (OPP[al] * a.ps + c), a.inv_scale);
            }
        }
    });
    // streamingT + bouncebackT (evolution_g.F90:60-197): adiabatic walls bounce
    // back; held walls: g(a) = -g_post(abar) + (6 + paraA)/21 Twall
    R g[Q7];
    sfor<Q7>([&](auto A) {
        constexpr int al = decltype(A)::value;
        constexpr int ex = EX[al], ey = EY[al], ez = EZ[al];
        if constexpr (al == 0) {
            g[0] = ld_stream<S>(gsrc + c, a.inv_scale);
        } else {
            constexpr int axis = ex != 0 ? 0 : (ey != 0 ? 1 : 2);
            constexpr int e = ex + ey + ez;
            const int q = axis == 0 ? i : (axis == 1 ? j : k);
            const int nq = axis == 0 ? a.nx : (axis == 1 ? a.ny : a.nz);
            const bool in = (unsigned)(q - e) < (unsigned)nq;
            const long long off = (long long)ex + (long long)ey * a.px + (long long)ez * a.pxy;
            constexpr int face = e > 0 ? 2 * axis : 2 * axis + 1;  // the face the pull enters through
            if (in) {
                g[al] = ld_stream<S>(gsrc + (al * a.ps + c - off), a.inv_scale);
            } else if (HALO && ((a.halo_faces >> face) & 1)) {
                // what arrives through `face`: the neighbour's D3Q7 population al
                const long long p = axis == 0 ? j + (long long)a.ny * k
                                              : (axis == 1 ? i + (long long)a.nx * k : i + (long long)a.nx * j);
                g[al] = ld_stream<S>(static_cast<const ST*>(a.halo) + (a.hoff[GREGION0 + face] + p), a.inv_scale);
            } else {
                const R bb = ld_stream<S>(gsrc + (OPP7[al] * a.ps + c), a.inv_scale);
                g[al] = (a.taxis == axis) ? -bb + R(a.cw) * R(e > 0 ? a.t_hot : a.t_cold) : bb;
            }
        }
    });

    // macro (evolution_f.F90:346-370) with the force of the last collision
    const R Fz0 = R(force[2 * a.ps + c]);
    const R Fx0 = ROT ? R(force[c]) : R(0.0);
    const R Fy0 = ROT ? R(force[a.ps + c]) : R(0.0);
    R s = f[0], jx = R(-0.0), jy = R(-0.0), jz = R(-0.0);
    sfor<Q - 1>([&](auto Am1) {
        constexpr int al = decltype(Am1)::value + 1;
        s = s + f[al];
        if constexpr (EX[al] != 0) jx = mac<EX[al]>(jx, f[al]);
        if constexpr (EY[al] != 0) jy = mac<EY[al]>(jy, f[al]);
        if constexpr (EZ[al] != 0) jz = mac<EZ[al]>(jz, f[al]);
    });
    const R rho = S::SHIFT ? R(1.0) + s : s;
    const R drho = S::SHIFT ? s : s - R(1.0);
    const R u = (jx + R(0.5) * Fx0) / rho;
    const R v = (jy + R(0.5) * Fy0) / rho;
    const R w = (jz + R(0.5) * Fz0) / rho;
    // macroT (evolution_g.F90:202-218)
    R T = g[0];
    sfor<Q7 - 1>([&](auto Am1) { T = T + g[decltype(Am1)::value + 1]; });
    if constexpr (MACRO) {
        const long long m = i + (long long)a.nx * (j + (long long)a.ny * k);
        a.rho[m] = S::SHIFT ? 1.0 + double(drho) : double(rho);
        a.u[m] = double(u);
        a.v[m] = double(v);
        a.w[m] = double(w);
        a.T[m] = double(T);
    }
    // next collision: force from this step's fields (evolution_f.F90:103-105)
    const R Fx = ROT ? R(-2.0) * rho * v * R(a.omega_rot) : R(0.0);
    const R Fy = ROT ? R(2.0) * rho * u * R(a.omega_rot) : R(0.0);
    const R Fz = rho * R(a.g_beta) * (T - R(a.t_ref));
    collide_mrt_forced<R, S::SHIFT>(f, rho, drho, u, v, w, Fx, Fy, Fz, R(a.snu), R(a.sq));
    collide_d3q7<R>(g, T, u, v, w, R(a.para_a), R(a.qd), R(a.qnu));
    ST* fdst = static_cast<ST*>(a.fdst);
    ST* gdst = static_cast<ST*>(a.gdst);
    const unsigned h = cell_random<S>(c, a.seed);
    sfor<Q>([&](auto A) {
        constexpr int al = decltype(A)::value;
        S::st(fdst + (al * a.ps + c), f[al], a.scale, store_random<S, al>(h));
    });
    sfor<Q7>([&](auto A) {
        constexpr int al = decltype(A)::value;
        S::st(gdst + (al * a.ps + c), g[al], a.scale, store_random<S, Q + al>(h));
    });
    force[2 * a.ps + c] = FT(Fz);
    if constexpr (ROT) {
        force[c] = FT(Fx);
        force[a.ps + c] = FT(Fy);
    }
}

// Equilibrium start (initial.F90:172-182) and the first collision from the
// given fields.  rho_in.. are compact nx*ny*nz arrays (also the macro output).
// The equilibria are formed in FP64 and the collision runs in the storage's
// arithmetic, as the step does.
template <class S, bool ROT>
__global__ __launch_bounds__(256) void k_thermal_init(const ThermalArgs a, const double* __restrict__ rho_in,
                                                      const double* __restrict__ u_in,
                                                      const double* __restrict__ v_in,
                                                      const double* __restrict__ w_in,
                                                      const double* __restrict__ T_in) {
    using R = typename S::Real;
    using FT = ForceT<S>;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= a.nx || j >= a.ny) return;
    const int c = i + a.px * (j + a.ny * k);
    const long long m = i + (long long)a.nx * (j + (long long)a.ny * k);
    const double rho = rho_in[m], u = u_in[m], v = v_in[m], w = w_in[m], T = T_in[m];
    const double us2 = u * u + v * v + w * w;
    R f[Q], g[Q7];
    sfor<Q>([&](auto A) {
        constexpr int al = decltype(A)::value;
        const double un = u * EX[al] + v * EY[al] + w * EZ[al];
        const double feq = rho * weight(al) * (1.0 + 3.0 * un + 4.5 * un * un - 1.5 * us2);
        f[al] = S::SHIFT ? R(feq - weight(al)) : R(feq);
    });
    const double wt0 = (1.0 - a.para_a) / 7.0, wt1 = (a.para_a + 6.0) / 42.0;
    sfor<Q7>([&](auto A) {
        constexpr int al = decltype(A)::value;
        const double un = u * EX[al] + v * EY[al] + w * EZ[al];
        g[al] = R((al == 0 ? wt0 : wt1) * T * (1.0 + 21.0 / (6.0 + a.para_a) * un));
    });
    const double Fx = -2.0 * rho * v * a.omega_rot;
    const double Fy = 2.0 * rho * u * a.omega_rot;
    const double Fz = rho * a.g_beta * (T - a.t_ref);
    collide_mrt_forced<R, S::SHIFT>(f, R(rho), R(rho - 1.0), R(u), R(v), R(w), R(Fx), R(Fy), R(Fz), R(a.snu),
                                    R(a.sq));
    collide_d3q7<R>(g, R(T), R(u), R(v), R(w), R(a.para_a), R(a.qd), R(a.qnu));
    using Tst = typename S::T;
    Tst* fdst = static_cast<Tst*>(a.fdst);
    Tst* gdst = static_cast<Tst*>(a.gdst);
    FT* force = static_cast<FT*>(a.force);
    const unsigned h = cell_random<S>(c, a.seed);
    sfor<Q>([&](auto A) {
        constexpr int al = decltype(A)::value;
        fdst[al * a.ps + c] = S::cvt(f[al], a.scale, store_random<S, al>(h));
    });
    sfor<Q7>([&](auto A) {
        constexpr int al = decltype(A)::value;
        gdst[al * a.ps + c] = S::cvt(g[al], a.scale, store_random<S, Q + al>(h));
    });
    force[c] = FT(Fx);
    force[a.ps + c] = FT(Fy);
    force[2 * a.ps + c] = FT(Fz);
    a.rho[m] = rho;
    a.u[m] = u;
    a.v[m] = v;
    a.w[m] = w;
    a.T[m] = T;
}

__device__ __forceinline__ double wave_sum(double x) {
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    return x;
}

// Block-reduce NS sums into partial[NS * blockIdx.x + s].
template <int NS>
__device__ __forceinline__ void block_store(double (&e)[NS], double* partial) {
    __shared__ double sh[NS][RED_THREADS / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int s = 0; s < NS; ++s) {
        const double t = wave_sum(e[s]);
        if (lane == 0) sh[s][wid] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0)
        for (int s = 0; s < NS; ++s) {
            double t = 0.0;
            for (int q = 0; q < RED_THREADS / 64; ++q) t += sh[s][q];
            partial[NS * blockIdx.x + s] = t;
        }
}

// check.F90:16-41: e1 = sum |u - up|^2 (with w), e2 = sum |u|^2,
// e5 = sum |T - Tp|, e6 = sum |T|; then up = u, ..., Tp = T.
__global__ __launch_bounds__(RED_THREADS) void k_check_thermal(const double* __restrict__ u,
                                                               const double* __restrict__ v,
                                                               const double* __restrict__ w,
                                                               const double* __restrict__ T,
                                                               double* __restrict__ up, double* __restrict__ vp,
                                                               double* __restrict__ wp, double* __restrict__ Tp,
                                                               long long n, double* __restrict__ partial) {
    double e[4] = {0.0, 0.0, 0.0, 0.0};
    for (long long c = blockIdx.x * (long long)RED_THREADS + threadIdx.x; c < n;
         c += (long long)gridDim.x * RED_THREADS) {
        const double uu = u[c], vv = v[c], ww = w[c], tt = T[c];
        const double du = uu - up[c], dv = vv - vp[c], dw = ww - wp[c];
        e[0] += du * du + dv * dv + dw * dw;
        e[1] += uu * uu + vv * vv + ww * ww;
        e[2] += fabs(tt - Tp[c]);
        e[3] += fabs(tt);
        up[c] = uu;
        vp[c] = vv;
        wp[c] = ww;
        Tp[c] = tt;
    }
    block_store<4>(e, partial);
}

// Nusselt sums: s0 = sum w T over all cells (RaNu.F90:13-21), s1 = sum
// 2 (Thot - T) over the hot-wall layer (first layer along the held axis).
__global__ __launch_bounds__(RED_THREADS) void k_nusselt(const double* __restrict__ w,
                                                         const double* __restrict__ T, int nx, int ny, int nz,
                                                         int taxis, double t_hot, double* __restrict__ partial) {
    double e[2] = {0.0, 0.0};
    const long long n = (long long)nx * ny * nz;
    for (long long c = blockIdx.x * (long long)RED_THREADS + threadIdx.x; c < n;
         c += (long long)gridDim.x * RED_THREADS) {
        const double t = T[c];
        e[0] += w[c] * t;
        const long long i = c % nx, j = (c / nx) % ny, k = c / ((long long)nx * ny);
        const long long q = taxis == 0 ? i : (taxis == 1 ? j : k);
        if (q == 0) e[1] += (t_hot - t) * 2.0;
    }
    block_store<2>(e, partial);
}

template <int NS>
__global__ __launch_bounds__(RED_THREADS) void k_reduce_sums(const double* __restrict__ partial, int nblocks,
                                                             double* __restrict__ out) {
    double e[NS];
    for (int s = 0; s < NS; ++s) e[s] = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += RED_THREADS)
        for (int s = 0; s < NS; ++s) e[s] += partial[NS * b + s];
    block_store<NS>(e, out);
}

ThermalArgs thermal_args(const lbm_ctx* c) {
    ThermalArgs a{};
    a.nx = c->p.nx;
    a.ny = c->p.ny;
    a.nz = c->p.nz;
    a.px = c->px;
    a.pxy = c->pxy;
    a.ps = c->ps;
    a.snu = 1.0 / c->p.tau;
    a.sq = 8.0 * (2.0 * c->p.tau - 1.0) / (8.0 * c->p.tau - 1.0);
    a.para_a = c->p.para_a;
    a.qd = 3.0 - std::sqrt(3.0);
    a.qnu = 4.0 * std::sqrt(3.0) - 6.0;
    a.g_beta = c->p.g_beta;
    a.omega_rot = c->p.omega_rot;
    a.t_ref = c->p.t_ref;
    a.t_hot = c->p.t_hot;
    a.t_cold = c->p.t_cold;
    a.cw = (6.0 + c->p.para_a) / 21.0;
    a.taxis = c->p.thermal_axis;
    a.scale = c->scale;
    a.inv_scale = c->inv_scale;
    a.force = c->force;
    a.rho = c->macro;
    a.u = c->macro + c->n;
    a.v = c->macro + 2 * c->n;
    a.w = c->macro + 3 * c->n;
    a.T = c->temp;
    a.halo_faces = c->p.halo_faces;
    a.halo = c->hin;
    for (int r = 0; r < NREGION_ALL; ++r) a.hoff[r] = c->hoff[r];
    a.u_lid = 0.0;
    return a;
}

dim3 thermal_grid(const lbm_ctx* c) {
    return dim3((c->p.nx + c->block.x - 1) / c->block.x, (c->p.ny + c->block.y - 1) / c->block.y,
                c->p.nz);
}

// Sum NS partial columns over RED_BLOCKS blocks into host `out`.
template <int NS>
int finish_sums(lbm_ctx* c, double* out) {
    hipLaunchKernelGGL(k_reduce_sums<NS>, dim3(1), dim3(RED_THREADS), 0, c->stream, c->red, RED_BLOCKS,
                       c->red + RED_MAX_SUMS * RED_BLOCKS);
    int rc = lbm_check_launch();
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(out, c->red + RED_MAX_SUMS * RED_BLOCKS, NS * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

}  // namespace

size_t thermal_force_elem(const lbm_ctx* c) { return c->p.storage == LBM_STORE_F64 ? sizeof(double) : sizeof(float); }

int thermal_alloc(lbm_ctx* c) {
    const size_t g_bytes = (size_t)c->ps * Q7 * c->elem;
    const size_t f_bytes = 3 * (size_t)c->ps * thermal_force_elem(c);
    HIP_TRY(hipMalloc(&c->gbuf[0], g_bytes));
    HIP_TRY(hipMalloc(&c->gbuf[1], g_bytes));
    HIP_TRY(hipMalloc(&c->force, f_bytes));
    HIP_TRY(hipMalloc(&c->temp, (size_t)c->n * sizeof(double)));
    HIP_TRY(hipMalloc(&c->prev_t, 2 * (size_t)c->n * sizeof(double)));
    c->device_bytes += (long long)(2 * g_bytes + f_bytes + 3 * c->n * sizeof(double));
    return LBM_OK;
}

void thermal_release(lbm_ctx* c) {
    for (void* b : c->gbuf)
        if (b) (void)hipFree(b);
    if (c->force) (void)hipFree(c->force);
    if (c->temp) (void)hipFree(c->temp);
    if (c->prev_t) (void)hipFree(c->prev_t);
}

namespace {

using ThermalFn = void (*)(const ThermalArgs);
using ThermalInitFn = void (*)(const ThermalArgs, const double*, const double*, const double*, const double*,
                               const double*);

template <class S>
struct ThermalTable {
    static constexpr ThermalFn step[2][2][2] = {
        {{k_thermal_step<S, false, false, false>, k_thermal_step<S, false, false, true>},
         {k_thermal_step<S, false, true, false>, k_thermal_step<S, false, true, true>}},
        {{k_thermal_step<S, true, false, false>, k_thermal_step<S, true, false, true>},
         {k_thermal_step<S, true, true, false>, k_thermal_step<S, true, true, true>}}};
    static constexpr ThermalInitFn init[2] = {k_thermal_init<S, false>, k_thermal_init<S, true>};
};

ThermalFn thermal_step_fn(const lbm_ctx* c, bool rot, bool macro, bool halo) {
    switch (c->p.storage) {
    case LBM_STORE_F32: return ThermalTable<StoreF32>::step[rot][macro][halo];
    case LBM_STORE_F16: return ThermalTable<StoreF16>::step[rot][macro][halo];
    default: return ThermalTable<StoreF64>::step[rot][macro][halo];
    }
}

ThermalInitFn thermal_init_fn(const lbm_ctx* c, bool rot) {
    switch (c->p.storage) {
    case LBM_STORE_F32: return ThermalTable<StoreF32>::init[rot];
    case LBM_STORE_F16: return ThermalTable<StoreF16>::init[rot];
    default: return ThermalTable<StoreF64>::init[rot];
    }
}

}  // namespace

int thermal_load(lbm_ctx* c, const double* d_rho, const double* d_u, const double* d_v, const double* d_w,
                 const double* d_T) {
    ThermalArgs a = thermal_args(c);
    a.fdst = c->buf[0];
    a.gdst = c->gbuf[0];
    a.seed = (unsigned)c->itc;
    hipLaunchKernelGGL(thermal_init_fn(c, c->p.omega_rot != 0.0), thermal_grid(c), c->block, 0, c->stream, a, d_rho,
                       d_u, d_v, d_w, d_T);
    return lbm_check_launch();
}

// nsteps steps.  With halo_loopback the exchange runs before every step: the
// outgoing regions of the current state are packed and unpacked into the
// opposite faces' incoming regions (the regions the step reads).
int thermal_step(lbm_ctx* c, int nsteps) {
    ThermalArgs a = thermal_args(c);
    const bool rot = c->p.omega_rot != 0.0;
    const bool halo = c->p.halo_faces != 0;
    const bool loop = c->p.halo_loopback != 0;
    const dim3 grid = thermal_grid(c);
    for (int s = 0; s < nsteps; ++s) {
        if (loop) {
            int rc = halo_loopback_exchange(c);
            if (rc) return rc;
        }
        a.fsrc = c->buf[c->cur];
        a.fdst = c->buf[1 - c->cur];
        a.gsrc = c->gbuf[c->cur];
        a.gdst = c->gbuf[1 - c->cur];
        a.seed = (unsigned)(c->itc + s + 1);
        hipLaunchKernelGGL(thermal_step_fn(c, rot, s == nsteps - 1, halo), grid, c->block, c->step_lds_halo,
                           c->stream, a);
        c->cur = 1 - c->cur;
    }
    return lbm_check_launch();
}

int thermal_check(lbm_ctx* c, double* sums4) {
    const long long n = c->n;
    hipLaunchKernelGGL(k_check_thermal, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, c->stream, c->macro + n,
                       c->macro + 2 * n, c->macro + 3 * n, c->temp, c->prev, c->prev + n, c->prev_t,
                       c->prev_t + n, n, c->red);
    int rc = lbm_check_launch();
    if (rc) return rc;
    return finish_sums<4>(c, sums4);
}

// Sum of |u|^2 over all cells (calNuRe's Reynolds sum, RaNu.F90:36-44).
__global__ __launch_bounds__(RED_THREADS) void k_kinetic(const double* __restrict__ u, const double* __restrict__ v,
                                                         const double* __restrict__ w, long long n,
                                                         double* __restrict__ partial) {
    double e[1] = {0.0};
    for (long long c = blockIdx.x * (long long)RED_THREADS + threadIdx.x; c < n;
         c += (long long)gridDim.x * RED_THREADS)
        e[0] += u[c] * u[c] + v[c] * v[c] + w[c] * w[c];
    block_store<1>(e, partial);
}

int flow_reynolds(lbm_ctx* c, double* re) {
    hipLaunchKernelGGL(k_kinetic, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, c->stream, c->macro + c->n,
                       c->macro + 2 * c->n, c->macro + 3 * c->n, c->n, c->red);
    int rc = lbm_check_launch();
    if (rc) return rc;
    double s[1];
    rc = finish_sums<1>(c, s);
    if (rc) return rc;
    const double viscosity = (c->p.tau - 0.5) / 3.0;
    *re = std::sqrt(s[0] / (double)c->n) * (double)c->p.nz / viscosity;
    return LBM_OK;
}

int thermal_nusselt(lbm_ctx* c, double* nu_vol, double* nu_wall) {
    const int ax = c->p.thermal_axis;
    if (ax < 0 || ax > 2) return lbm_fail(LBM_ERR_ARG, "lbm_nusselt: no held walls (thermal_axis = %d)", ax);
    const double* w = c->macro + 3 * c->n;  // gravity acts along z
    hipLaunchKernelGGL(k_nusselt, dim3(RED_BLOCKS), dim3(RED_THREADS), 0, c->stream, w, c->temp, c->p.nx, c->p.ny,
                       c->p.nz, ax, c->p.t_hot, c->red);
    int rc = lbm_check_launch();
    if (rc) return rc;
    double s[2];
    rc = finish_sums<2>(c, s);
    if (rc) return rc;
    const double kappa = (c->p.para_a + 6.0) / (42.0 * std::sqrt(3.0));  // inverse of module.F90:27
    const int na = ax == 0 ? c->p.nx : (ax == 1 ? c->p.ny : c->p.nz);
    const long long nwall = c->n / na;
    // calNuRe (RaNu.F90:21): <w T> nz / diffusivity + 1, whichever walls are held
    *nu_vol = s[0] / (double)c->n * (double)c->p.nz / kappa + 1.0;
    *nu_wall = s[1] / (double)nwall * (double)na / (c->p.t_hot - c->p.t_cold);
    return LBM_OK;
}
