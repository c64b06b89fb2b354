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

#include "collide.hpp"
#include "context.hpp"

using namespace lbm;

namespace {

struct ThermalArgs {
    const double* fsrc;
    double* fdst;
    const double* gsrc;
    double* gdst;
    double* force;  // Fx | Fy | Fz, each ps long
    int nx, ny, nz, px;
    long long pxy, ps;
    double snu, sq;           // flow rates (collision.f90:94-112 assignment)
    double para_a, qd, qnu;   // D3Q7: paraA, Qd = 3 - sqrt 3, Qnu = 4 sqrt 3 - 6
    double g_beta, omega_rot, t_ref;
    double t_hot, t_cold, cw; // held walls; cw = (6 + paraA) / 21
    int taxis;                // axis of the held walls (0 x, 1 y, 2 z), -1 none
    double *rho, *u, *v, *w, *T;
};

template <bool ROT, bool MACRO>
__global__ __launch_bounds__(256) void k_thermal_step(const ThermalArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= a.nx || j >= a.ny) return;
    const int c = i + a.px * (j + a.ny * k);

    // streaming + bounceback (evolution_f.F90:227-343): no-slip on all walls
    double f[Q];
    sfor<Q>([&](auto A) {
        constexpr int al = decltype(A)::value;
        constexpr int ex = EX[al], ey = EY[al], ez = EZ[al];
        if constexpr (al == 0) {
            f[0] = a.fsrc[c];
        } else {
            bool in = true;
            if constexpr (ex != 0) in = in && (unsigned)(i - ex) < (unsigned)a.nx;
            if constexpr (ey != 0) in = in && (unsigned)(j - ey) < (unsigned)a.ny;
            if constexpr (ez != 0) in = in && (unsigned)(k - ez) < (unsigned)a.nz;
            const long long off = (long long)ex + (long long)ey * a.px + (long long)ez * a.pxy;
            f[al] = in ? a.fsrc[al * a.ps + c - off] : a.fsrc[OPP[al] * a.ps + c];
        }
    });
    // streamingT + bouncebackT (evolution_g.F90:60-197): adiabatic walls bounce
    // back; held walls: g(a) = -g_post(abar) + (6 + paraA)/21 Twall
    double g[Q7];
    sfor<Q7>([&](auto A) {
        constexpr int al = decltype(A)::value;
        constexpr int ex = EX[al], ey = EY[al], ez = EZ[al];
        if constexpr (al == 0) {
            g[0] = a.gsrc[c];
        } else {
            constexpr int axis = ex != 0 ? 0 : (ey != 0 ? 1 : 2);
            constexpr int e = ex + ey + ez;
            const int q = axis == 0 ? i : (axis == 1 ? j : k);
            const int nq = axis == 0 ? a.nx : (axis == 1 ? a.ny : a.nz);
            const bool in = (unsigned)(q - e) < (unsigned)nq;
            const long long off = (long long)ex + (long long)ey * a.px + (long long)ez * a.pxy;
            if (in) {
                g[al] = a.gsrc[al * a.ps + c - off];
            } else {
                const double bb = a.gsrc[OPP7[al] * a.ps + c];
                g[al] = (a.taxis == axis) ? -bb + a.cw * (e > 0 ? a.t_hot : a.t_cold) : bb;
            }
        }
    });

    // macro (evolution_f.F90:346-370) with the force of the last collision
    const double Fz0 = a.force[2 * a.ps + c];
    const double Fx0 = ROT ? a.force[c] : 0.0;
    const double Fy0 = ROT ? a.force[a.ps + c] : 0.0;
    double rho = f[0], jx = -0.0, jy = -0.0, jz = -0.0;
    sfor<Q - 1>([&](auto Am1) {
        constexpr int al = decltype(Am1)::value + 1;
        rho = rho + f[al];
        if constexpr (EX[al] != 0) jx = mac<EX[al]>(jx, f[al]);
        if constexpr (EY[al] != 0) jy = mac<EY[al]>(jy, f[al]);
        if constexpr (EZ[al] != 0) jz = mac<EZ[al]>(jz, f[al]);
    });
    const double u = (jx + 0.5 * Fx0) / rho;
    const double v = (jy + 0.5 * Fy0) / rho;
    const double w = (jz + 0.5 * Fz0) / rho;
    // macroT (evolution_g.F90:202-218)
    double T = g[0];
    sfor<Q7 - 1>([&](auto Am1) { T = T + g[decltype(Am1)::value + 1]; });
    if constexpr (MACRO) {
        const long long m = i + (long long)a.nx * (j + (long long)a.ny * k);
        a.rho[m] = rho;
        a.u[m] = u;
        a.v[m] = v;
        a.w[m] = w;
        a.T[m] = T;
    }
    // next collision: force from this step's fields (evolution_f.F90:103-105)
    const double Fx = ROT ? -2.0 * rho * v * a.omega_rot : 0.0;
    const double Fy = ROT ? 2.0 * rho * u * a.omega_rot : 0.0;
    const double Fz = rho * a.g_beta * (T - a.t_ref);
    collide_mrt_forced(f, rho, u, v, w, Fx, Fy, Fz, a.snu, a.sq);
    collide_d3q7(g, T, u, v, w, a.para_a, a.qd, a.qnu);
    sfor<Q>([&](auto A) {
        constexpr int al = decltype(A)::value;
        __builtin_nontemporal_store(f[al], a.fdst + al * a.ps + c);
    });
    sfor<Q7>([&](auto A) {
        constexpr int al = decltype(A)::value;
        __builtin_nontemporal_store(g[al], a.gdst + al * a.ps + c);
    });
    a.force[2 * a.ps + c] = Fz;
    if constexpr (ROT) {
        a.force[c] = Fx;
        a.force[a.ps + c] = Fy;
    }
}

// Equilibrium start (initial.F90:172-182) and the first collision from the
// given fields.  rho_in.. are compact nx*ny*nz arrays (also the macro output).
template <bool ROT>
__global__ __launch_bounds__(256) void k_thermal_init(const ThermalArgs a, const double* __restrict__ rho_in,
                                                      const double* __restrict__ u_in,
                                                      const double* __restrict__ v_in,
                                                      const double* __restrict__ w_in,
                                                      const double* __restrict__ T_in) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= a.nx || j >= a.ny) return;
    const int c = i + a.px * (j + a.ny * k);
    const long long m = i + (long long)a.nx * (j + (long long)a.ny * k);
    const double rho = rho_in[m], u = u_in[m], v = v_in[m], w = w_in[m], T = T_in[m];
    const double us2 = u * u + v * v + w * w;
    double f[Q], g[Q7];
    sfor<Q>([&](auto A) {
        constexpr int al = decltype(A)::value;
        const double un = u * EX[al] + v * EY[al] + w * EZ[al];
        f[al] = rho * weight(al) * (1.0 + 3.0 * un + 4.5 * un * un - 1.5 * us2);
    });
    const double wt0 = (1.0 - a.para_a) / 7.0, wt1 = (a.para_a + 6.0) / 42.0;
    sfor<Q7>([&](auto A) {
        constexpr int al = decltype(A)::value;
        const double un = u * EX[al] + v * EY[al] + w * EZ[al];
        g[al] = (al == 0 ? wt0 : wt1) * T * (1.0 + 21.0 / (6.0 + a.para_a) * un);
    });
    const double Fx = -2.0 * rho * v * a.omega_rot;
    const double Fy = 2.0 * rho * u * a.omega_rot;
    const double Fz = rho * a.g_beta * (T - a.t_ref);
    collide_mrt_forced(f, rho, u, v, w, Fx, Fy, Fz, a.snu, a.sq);
    collide_d3q7(g, T, u, v, w, a.para_a, a.qd, a.qnu);
    sfor<Q>([&](auto A) {
        constexpr int al = decltype(A)::value;
        a.fdst[al * a.ps + c] = f[al];
    });
    sfor<Q7>([&](auto A) {
        constexpr int al = decltype(A)::value;
        a.gdst[al * a.ps + c] = g[al];
    });
    a.force[c] = Fx;
    a.force[a.ps + c] = Fy;
    a.force[2 * a.ps + c] = Fz;
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
    a.force = c->force;
    a.rho = c->macro;
    a.u = c->macro + c->n;
    a.v = c->macro + 2 * c->n;
    a.w = c->macro + 3 * c->n;
    a.T = c->temp;
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

int thermal_alloc(lbm_ctx* c) {
    const size_t g_bytes = (size_t)c->ps * Q7 * sizeof(double);
    HIP_TRY(hipMalloc(&c->gbuf[0], g_bytes));
    HIP_TRY(hipMalloc(&c->gbuf[1], g_bytes));
    HIP_TRY(hipMalloc(&c->force, 3 * (size_t)c->ps * sizeof(double)));
    HIP_TRY(hipMalloc(&c->temp, (size_t)c->n * sizeof(double)));
    HIP_TRY(hipMalloc(&c->prev_t, 2 * (size_t)c->n * sizeof(double)));
    c->device_bytes += (long long)(2 * g_bytes + (3 * c->ps + 3 * c->n) * sizeof(double));
    return LBM_OK;
}

void thermal_release(lbm_ctx* c) {
    for (void* b : c->gbuf)
        if (b) (void)hipFree(b);
    if (c->force) (void)hipFree(c->force);
    if (c->temp) (void)hipFree(c->temp);
    if (c->prev_t) (void)hipFree(c->prev_t);
}

int thermal_load(lbm_ctx* c, const double* d_rho, const double* d_u, const double* d_v, const double* d_w,
                 const double* d_T) {
    ThermalArgs a = thermal_args(c);
    a.fdst = static_cast<double*>(c->buf[0]);
    a.gdst = static_cast<double*>(c->gbuf[0]);
    if (c->p.omega_rot != 0.0)
        hipLaunchKernelGGL(k_thermal_init<true>, thermal_grid(c), c->block, 0, c->stream, a, d_rho, d_u, d_v, d_w,
                           d_T);
    else
        hipLaunchKernelGGL(k_thermal_init<false>, thermal_grid(c), c->block, 0, c->stream, a, d_rho, d_u, d_v,
                           d_w, d_T);
    return lbm_check_launch();
}

int thermal_step(lbm_ctx* c, int nsteps) {
    ThermalArgs a = thermal_args(c);
    const bool rot = c->p.omega_rot != 0.0;
    const dim3 grid = thermal_grid(c);
    for (int s = 0; s < nsteps; ++s) {
        a.fsrc = static_cast<const double*>(c->buf[c->cur]);
        a.fdst = static_cast<double*>(c->buf[1 - c->cur]);
        a.gsrc = static_cast<const double*>(c->gbuf[c->cur]);
        a.gdst = static_cast<double*>(c->gbuf[1 - c->cur]);
        const bool macro = s == nsteps - 1;
        void (*kern)(const ThermalArgs) = rot ? (macro ? k_thermal_step<true, true> : k_thermal_step<true, false>)
                                              : (macro ? k_thermal_step<false, true> : k_thermal_step<false, false>);
        hipLaunchKernelGGL(kern, grid, c->block, c->step_lds_halo, c->stream, a);
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
