// collide.hpp -- per-cell device arithmetic of the D3Q19 step: moments
// (macro), equilibria, BGK and MRT relaxation, and the population storage
// formats.  Everything here is unrolled at compile time over the lattice
// tables of lattice.hpp, so zero coefficients cost nothing and +-1
// coefficients are plain adds.
//
// Two arithmetic spaces:
//   absolute  (SHIFT = false): registers hold f itself.  FP64 storage uses
//             this, so its operation order can follow the reference.
//   shifted   (SHIFT = true): registers hold g = f - w (the deviation from the
//             rest-state weights) and rho = 1 + drho.  FP32/FP16 storage use
//             it: deviations are O(Ma) small, so the 24-bit / 11-bit
//             mantissas are spent on the part that changes.
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

#include "lattice.hpp"

namespace lbm {

// ---- compile-time loop ------------------------------------------------------
template <class F, int... Is>
__device__ __forceinline__ void sfor_impl(F&& f, std::integer_sequence<int, Is...>) {
    (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void sfor(F&& f) {
    sfor_impl(f, std::make_integer_sequence<int, N>{});
}

// acc + C*x with the multiply dropped for C = +-1.
template <int C, class R>
__device__ __forceinline__ R mac(R acc, R x) {
    if constexpr (C == 1) return acc + x;
    else if constexpr (C == -1) return acc - x;
    else return acc + R(C) * x;
}

// ---- storage formats ----------------------------------------------------------
// Stores are non-temporal: the next step reads the written lattice only after
// the whole sweep, so caching the lines on the way out buys nothing (measured
// 2-4% faster at 256^3).
// Every st() takes a 32-bit random word; only the FP16 format uses it.
struct StoreF64 {
    static constexpr int FOLD_LOG2 = 0;  // the store scale the MRT inverse folds in (0: none)
    using T = double;
    using Real = double;
    static constexpr bool SHIFT = false;
    static constexpr bool RANDOM = false;
    static constexpr int BYTES = 8;
    __device__ __forceinline__ static Real val(T v, float) { return v; }
    __device__ __forceinline__ static Real ld(const T* p, float) { return *p; }
    __device__ __forceinline__ static T cvt(Real v, float, unsigned) { return v; }
    __device__ __forceinline__ static void st(T* p, Real v, float, unsigned) { __builtin_nontemporal_store(v, p); }
};

struct StoreF32 {
    static constexpr int FOLD_LOG2 = 0;  // the store scale the MRT inverse folds in (0: none)
    using T = float;
    using Real = float;
    static constexpr bool SHIFT = true;
    static constexpr bool RANDOM = false;
    static constexpr int BYTES = 4;
    __device__ __forceinline__ static Real val(T v, float) { return v; }
    __device__ __forceinline__ static Real ld(const T* p, float) { return *p; }
    __device__ __forceinline__ static T cvt(Real v, float, unsigned) { return v; }
    __device__ __forceinline__ static void st(T* p, Real v, float, unsigned) { __builtin_nontemporal_store(v, p); }
};

// FP32 -> IEEE binary16 with stochastic rounding, on gfx950's conversion
// instruction (v_cvt_sr_f16_f32): the mantissa bits binary16 cannot hold are
// rounded up with probability equal to their value, driven by a uniform
// 32-bit word.  Rounding to nearest would round a slowly varying population
// the same way every step -- a constant spurious force that the flow
// integrates over its long viscous time; stochastic rounding makes that
// error unbiased.
__device__ __forceinline__ _Float16 to_half_stochastic(float x, unsigned rnd) {
    typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
    const half2_t h = __builtin_amdgcn_cvt_sr_f16_f32(half2_t{0, 0}, x, rnd, false);
    return h.x;
}

// Avalanche hash of a 32-bit word (the "lowbias32" integer mixer), used to
// draw the per-store random words from (cell, population, step).
__device__ __forceinline__ unsigned mix32(unsigned x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// IEEE binary16 holding (f - w) * 2^k.  The scale is an exact power of two,
// so scaling adds no rounding; the store rounds stochastically.
struct StoreF16 {
    static constexpr int FOLD_LOG2 = 0;  // the store scale the MRT inverse folds in (0: none)
    using T = _Float16;
    using Real = float;
    static constexpr bool SHIFT = true;
    static constexpr bool RANDOM = true;
    static constexpr int BYTES = 2;
    __device__ __forceinline__ static Real val(T v, float inv_scale) {
        // fma with +0 rather than a product: one mixed-precision fma
        // (v_fma_mix_f32) instead of a conversion and a multiply; the value
        // is the same (the scale is a power of two)
        return __builtin_fmaf(static_cast<float>(v), inv_scale, 0.0f);
    }
    __device__ __forceinline__ static Real ld(const T* p, float inv_scale) { return val(*p, inv_scale); }
    // the stored value: v scaled, rounded stochastically to FP16
    __device__ __forceinline__ static T cvt(Real v, float scale, unsigned rnd) {
        return to_half_stochastic(v * scale, rnd);
    }
    __device__ __forceinline__ static void st(T* p, Real v, float scale, unsigned rnd) {
        __builtin_nontemporal_store(cvt(v, scale, rnd), p);
    }
};

// FP16 storage at the default scale 2^12 with the scale folded into the MRT
// inverse transform's coefficients (exact: a power of two), so the store
// converts without a multiply; the bits equal StoreF16's at scale 2^12.
struct StoreF16K12 : StoreF16 {
    static constexpr int FOLD_LOG2 = 12;
    __device__ __forceinline__ static T cvt(Real v, float, unsigned rnd) { return to_half_stochastic(v, rnd); }
    __device__ __forceinline__ static void st(T* p, Real v, float scale, unsigned rnd) {
        __builtin_nontemporal_store(cvt(v, scale, rnd), p);
    }
};

// Random words for the stores of cell c at step `seed`: one avalanche hash
// per cell, then a Weyl step per population.  Each word is uniform over cells
// and steps, which is what makes the rounding unbiased; one hash per cell
// keeps the store ALU-cheap.
template <class S>
__device__ __forceinline__ unsigned cell_random(long long c, unsigned seed) {
    if constexpr (S::RANDOM)
        return mix32((unsigned)c ^ (seed * 0x9E3779B9u));
    else
        return 0u;
}

template <class S, int a>
__device__ __forceinline__ unsigned store_random(unsigned h) {
    if constexpr (S::RANDOM)
        return h + (unsigned)a * 0x61C88647u;
    else
        return 0u;
}

// ---- macroscopic moments (macro.f90:6-25) ---------------------------------------
// Sums run a = 0..18 in order, as the reference accumulates them.
template <class R, bool SHIFT>
struct Moments {
    R rho;   // density
    R drho;  // rho - 1 (shifted space only; absolute space leaves it unused)
    R u, v, w;
};

template <class R, bool SHIFT>
__device__ __forceinline__ Moments<R, SHIFT> moments(const R f[Q]) {
    R s = f[0], jx = R(-0.0), jy = R(-0.0), jz = R(-0.0);
    sfor<Q - 1>([&](auto am1) {
        constexpr int a = decltype(am1)::value + 1;
        s = s + f[a];
        if constexpr (EX[a] != 0) jx = mac<EX[a]>(jx, f[a]);
        if constexpr (EY[a] != 0) jy = mac<EY[a]>(jy, f[a]);
        if constexpr (EZ[a] != 0) jz = mac<EZ[a]>(jz, f[a]);
    });
    Moments<R, SHIFT> m;
    if constexpr (SHIFT) {
        m.drho = s;
        m.rho = R(1) + s;
    } else {
        m.rho = s;
        m.drho = s - R(1);
    }
    m.u = jx / m.rho;
    m.v = jy / m.rho;
    m.w = jz / m.rho;
    return m;
}

// ---- BGK (collision.f90:192-198, feq of initial.f90:63-73) ---------------------------
// f_post = f - (f - feq)/tau.  In shifted space the w-part of feq is dropped:
// geq = w (drho + rho (3 e.u + 4.5 (e.u)^2 - 1.5 u.u)).
template <class R, bool SHIFT>
__device__ __forceinline__ void collide_bgk(R f[Q], const Moments<R, SHIFT>& m, R omega) {
    const R us2 = m.u * m.u + m.v * m.v + m.w * m.w;
    sfor<Q>([&](auto A) {
        constexpr int a = decltype(A)::value;
        R un = R(-0.0);
        if constexpr (EX[a] != 0) un = mac<EX[a]>(un, m.u);
        if constexpr (EY[a] != 0) un = mac<EY[a]>(un, m.v);
        if constexpr (EZ[a] != 0) un = mac<EZ[a]>(un, m.w);
        constexpr R wa = R(weight(a));
        R feq;
        if constexpr (SHIFT)
            feq = wa * (m.drho + m.rho * (R(3) * un + R(4.5) * un * un - R(1.5) * us2));
        else
            feq = m.rho * wa * (R(1) + R(3) * un + R(4.5) * un * un - R(1.5) * us2);
        f[a] = f[a] - omega * (f[a] - feq);
    });
}

// ---- MRT (collision.f90:15-202) ------------------------------------------------------
// m = M f computed from the opposite-pair sums s_p = f_a + f_abar (even rows)
// and differences d_p = f_a - f_abar (odd rows); f_post = M^-1 m* likewise
// from even/odd partial sums E, O with f_a = E + O, f_abar = E - O.
template <class R>
__device__ __forceinline__ void mrt_forward(const R f[Q], R m[Q]) {
    R ps[NPAIR], pd[NPAIR];
    sfor<NPAIR>([&](auto P) {
        constexpr int p = decltype(P)::value;
        constexpr int a = PAIR_A[p];
        ps[p] = f[a] + f[OPP[a]];
        pd[p] = f[a] - f[OPP[a]];
    });
    sfor<Q>([&](auto K) {
        constexpr int k = decltype(K)::value;
        constexpr bool odd = row_is_odd(k);
        R acc = R(-0.0);
        if constexpr (!odd) acc = mac<basis(k, 0)>(acc, f[0]);
        sfor<NPAIR>([&](auto P) {
            constexpr int p = decltype(P)::value;
            constexpr int c = basis(k, PAIR_A[p]);
            if constexpr (c != 0) acc = mac<c>(acc, odd ? pd[p] : ps[p]);
        });
        m[k] = acc;
    });
}

// FOLD: the result times 2^FOLD, the scale folded into the coefficients
// (exact for a power of two: the same bits as scaling the result).
template <class R, int FOLD = 0>
__device__ __forceinline__ void mrt_inverse(const R m[Q], R f[Q]) {
    constexpr double SC = double(1LL << FOLD);
    {
        R acc = R(-0.0);
        sfor<Q>([&](auto K) {
            constexpr int k = decltype(K)::value;
            constexpr int c = basis(k, 0);
            if constexpr (c != 0) acc = acc + R(double(c) / double(row_norm2(k)) * SC) * m[k];
        });
        f[0] = acc;
    }
    sfor<NPAIR>([&](auto P) {
        constexpr int a = PAIR_A[decltype(P)::value];
        R e = R(-0.0), o = R(-0.0);
        sfor<Q>([&](auto K) {
            constexpr int k = decltype(K)::value;
            constexpr int c = basis(k, a);
            if constexpr (c != 0) {
                constexpr R coef = R(double(c) / double(row_norm2(k)) * SC);
                if constexpr (row_is_odd(k)) o = o + coef * m[k];
                else e = e + coef * m[k];
            }
        });
        f[a] = e + o;
        f[OPP[a]] = e - o;
    });
}

// Equilibrium moments (collision.f90:73-91).  In shifted space M w =
// (1, -11, 3, 0, ...) is subtracted, leaving drho in rows 0..2.
// meq12_rho selects the thermal form -rho/2 (v^2 - w^2) (evolution_f.F90:75);
// the cavity form has no rho (collision.f90:85).
template <class R, bool SHIFT>
__device__ __forceinline__ void mrt_meq(const Moments<R, SHIFT>& q, bool meq12_rho, R meq[Q]) {
    const R rho = q.rho, u = q.u, v = q.v, w = q.w;
    const R us2 = u * u + v * v + w * w;
    if constexpr (SHIFT) {
        meq[0] = q.drho;
        meq[1] = R(-11) * q.drho + R(19) * rho * us2;
        meq[2] = R(3) * q.drho - R(11.0 / 2.0) * rho * us2;
    } else {
        meq[0] = rho;
        meq[1] = rho * (R(-11) + R(19) * us2);
        meq[2] = rho * (R(3) - R(11.0 / 2.0) * us2);
    }
    meq[3] = rho * u;
    meq[4] = R(-2.0 / 3.0) * rho * u;
    meq[5] = rho * v;
    meq[6] = R(-2.0 / 3.0) * rho * v;
    meq[7] = rho * w;
    meq[8] = R(-2.0 / 3.0) * rho * w;
    meq[9] = rho * (R(2) * u * u - v * v - w * w);
    meq[10] = R(-0.5) * rho * (R(2) * u * u - v * v - w * w);
    meq[11] = rho * (v * v - w * w);
    meq[12] = meq12_rho ? R(-0.5) * rho * (v * v - w * w) : R(-0.5) * (v * v - w * w);
    meq[13] = rho * u * v;
    meq[14] = rho * v * w;
    meq[15] = rho * u * w;
    meq[16] = R(0);
    meq[17] = R(0);
    meq[18] = R(0);
}

// Rates (collision.f90:94-112): 0 on rho and j, Snu = 1/tau on e, eps and the
// stress rows 9..15, Sq = 8(2tau-1)/(8tau-1) on q and m16..18.
template <int k>
constexpr int rate_kind() {
    return (k == 0 || k == 3 || k == 5 || k == 7) ? 0 : ((k == 4 || k == 6 || k == 8 || k >= 16) ? 2 : 1);
}

template <class R, bool SHIFT, int FOLD = 0>
__device__ __forceinline__ void collide_mrt(R f[Q], const Moments<R, SHIFT>& q, R snu, R sq,
                                            bool meq12_rho) {
    R m[Q], meq[Q];
    mrt_forward(f, m);
    mrt_meq(q, meq12_rho, meq);
    sfor<Q>([&](auto K) {
        constexpr int k = decltype(K)::value;
        constexpr int kind = rate_kind<k>();
        if constexpr (kind == 1) m[k] = m[k] - snu * (m[k] - meq[k]);
        else if constexpr (kind == 2) m[k] = m[k] - sq * (m[k] - meq[k]);
    });
    mrt_inverse<R, FOLD>(m, f);
}

}  // namespace lbm

// ---- buoyancy-driven flow (evolution_f.F90, evolution_g.F90) -------------------------
namespace lbm {

// MRT with the Guo source in moment space (evolution_f.F90:58-129):
// m* = m - s (m - meq) + (1 - s/2) S, meq in the thermal form (meq(12) with rho).
// R = double, SHIFT = false: absolute FP64 arithmetic in the reference's
// operation order.  SHIFT = true: f holds f - w and rho = 1 + drho; M w =
// (1, -11, 3, 0, ...) leaves drho in meq rows 0..2 (the relaxation and the
// source are unchanged, being differences and forces).
template <class R, bool SHIFT>
__device__ __forceinline__ void collide_mrt_forced(R f[Q], R rho, R drho, R u, R v, R w, R Fx, R Fy, R Fz,
                                                   R snu, R sq) {
    R m[Q], meq[Q], S[Q];
    mrt_forward(f, m);
    const R us2 = u * u + v * v + w * w;
    if constexpr (SHIFT) {
        meq[0] = drho;
        meq[1] = R(-11.0) * drho + R(19.0) * rho * us2;
        meq[2] = R(3.0) * drho - R(11.0 / 2.0) * rho * us2;
    } else {
        meq[0] = rho;
        meq[1] = R(-11.0) * rho + R(19.0) * rho * us2;
        meq[2] = R(3.0) * rho - R(11.0 / 2.0) * rho * us2;
    }
    meq[3] = rho * u;
    meq[4] = R(-2.0 / 3.0) * meq[3];
    meq[5] = rho * v;
    meq[6] = R(-2.0 / 3.0) * meq[5];
    meq[7] = rho * w;
    meq[8] = R(-2.0 / 3.0) * meq[7];
    meq[9] = rho * (R(2.0) * u * u - v * v - w * w);
    meq[10] = R(-0.5) * meq[9];
    meq[11] = rho * (v * v - w * w);
    meq[12] = R(-0.5) * meq[11];
    meq[13] = rho * (u * v);
    meq[14] = rho * (v * w);
    meq[15] = rho * (w * u);
    meq[16] = meq[17] = meq[18] = R(0.0);
    const R uF = u * Fx + v * Fy + w * Fz;
    S[0] = R(0.0);
    S[1] = R(38.0) * uF;
    S[2] = R(-11.0) * uF;
    S[3] = Fx;
    S[4] = R(-2.0 / 3.0) * Fx;
    S[5] = Fy;
    S[6] = R(-2.0 / 3.0) * Fy;
    S[7] = Fz;
    S[8] = R(-2.0 / 3.0) * Fz;
    S[9] = R(4.0) * u * Fx - R(2.0) * v * Fy - R(2.0) * w * Fz;
    S[10] = R(-2.0) * u * Fx + v * Fy + w * Fz;
    S[11] = R(2.0) * v * Fy - R(2.0) * w * Fz;
    S[12] = -v * Fy + w * Fz;
    S[13] = u * Fy + v * Fx;
    S[14] = v * Fz + w * Fy;
    S[15] = u * Fz + w * Fx;
    S[16] = S[17] = S[18] = R(0.0);
    sfor<Q>([&](auto K) {
        constexpr int k = decltype(K)::value;
        constexpr int kind = rate_kind<k>();
        if constexpr (kind == 0) m[k] = m[k] + S[k];
        else if constexpr (kind == 1) m[k] = m[k] - snu * (m[k] - meq[k]) + (R(1.0) - R(0.5) * snu) * S[k];
        else m[k] = m[k] - sq * (m[k] - meq[k]) + (R(1.0) - R(0.5) * sq) * S[k];
    });
    mrt_inverse(m, f);
}

// D3Q7 MRT for the temperature (evolution_g.F90:1-57): n = N g,
// neq = (T, T u, T v, T w, T paraA, 0, 0), rates (0, Qd x3, Qnu x3),
// g_post = N^-1 n*, N^-1 = N^T diag(1/|row|^2).  g is stored unshifted.
template <class R>
__device__ __forceinline__ void collide_d3q7(R g[Q7], R T, R u, R v, R w, R para_a, R qd, R qnu) {
    R n[Q7];
    sfor<Q7>([&](auto K) {
        constexpr int k = decltype(K)::value;
        R acc = R(-0.0);
        sfor<Q7>([&](auto A) {
            constexpr int a = decltype(A)::value;
            constexpr int c = basis7(k, a);
            if constexpr (c != 0) acc = mac<c>(acc, g[a]);
        });
        n[k] = acc;
    });
    const R neq[Q7] = {T, T * u, T * v, T * w, T * para_a, R(0.0), R(0.0)};
    sfor<Q7>([&](auto K) {
        constexpr int k = decltype(K)::value;
        if constexpr (k >= 1 && k <= 3) n[k] = n[k] - qd * (n[k] - neq[k]);
        else if constexpr (k >= 4) n[k] = n[k] - qnu * (n[k] - neq[k]);
    });
    sfor<Q7>([&](auto A) {
        constexpr int a = decltype(A)::value;
        R acc = R(-0.0);
        sfor<Q7>([&](auto K) {
            constexpr int k = decltype(K)::value;
            constexpr int c = basis7(k, a);
            if constexpr (c != 0) acc = acc + R(double(c) / double(row_norm2_7(k))) * n[k];
        });
        g[a] = acc;
    });
}

}  // namespace lbm
