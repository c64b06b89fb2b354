// step.hpp -- the per-cell update of the isothermal D3Q19 step, shared by the
// single-domain kernel (cavity.hip) and the batched persistent kernel
// (batch.hip).  Internal to libmglc_lbm.so.
// `Args` below is any argument type with the members of StepArgs the path
// uses (the batched kernel passes a per-instance view).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

#include "collide.hpp"
#include "context.hpp"
#include "halo.hpp"
#include "mfma.hpp"

namespace lbm {

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
    int halo_faces;     // bit f: face f reads the halo instead of bouncing back
    const void* halo;   // incoming regions, packed layout (HALO kernels only)
    long long hoff[NREGION_ALL];    // element offset of each region in the packed layout
    long long hcount[NREGION_ALL];  // elements in each region (D3Q7 regions: thermal contexts only)
    unsigned seed;              // step number: drives the stochastic rounding of FP16 stores
    void* halo_out;             // loopback: the halo of the next step, written by the step itself
    unsigned loop_mask;         // loopback: the regions of halo_out the next step reads (bit = region)
};

// No extra use of a cell's stored populations.  A hook's active() says
// whether its store() has anything to do for this cell at all; store() sees
// the 19 values as stored.
struct NoPost {
    __device__ __forceinline__ static constexpr bool active() { return false; }
    template <class T>
    __device__ __forceinline__ void store(const T (&)[Q]) const {}
};

enum { COLL_MRT = 0, COLL_BGK = 1 };

template <class S, int COLL, class Args>
__device__ __forceinline__ void collide(typename S::Real f[Q],
                                        const Moments<typename S::Real, S::SHIFT>& q,
                                        const Args& a) {
    using R = typename S::Real;
    if constexpr (COLL == COLL_BGK) {
        static_assert(S::FOLD_LOG2 == 0, "a folded store scale is an MRT form");
        collide_bgk<R, S::SHIFT>(f, q, R(a.snu));
    } else {
        // f returned times 2^FOLD_LOG2 when the store folds its scale in
        collide_mrt<R, S::SHIFT, S::FOLD_LOG2>(f, q, R(a.snu), R(a.sq), a.meq12_rho != 0);
    }
}

// Value pulled from outside the domain for population al at cell (i, j, k):
// from the halo when every face the pull crosses borders a neighbour
// subdomain, else halfway bounce-back from this cell's opposite population
// (f(a, x) = f_post(abar, x), bounce_back.f90:6-83), with the moving-lid term
// when the pull crosses the top wall.  A wall crossed anywhere wins, as the
// reference's bounceback overwrites what streaming pulled from its halo.
template <class S, int al, bool HALO, class Args>
__device__ __forceinline__ typename S::Real pull_outside(const Args& a, const typename S::T* src,
                                                         int i, int j, int k, int c, bool top,
                                                         double rho_lid) {
    using T = typename S::T;
    using R = typename S::Real;
    constexpr int ex = EX[al], ey = EY[al], ez = EZ[al];
    bool lid = top;  // the z+ face is the moving lid when it is a wall
    if constexpr (HALO) {
        bool cx = false, cy = false, cz = false;
        if constexpr (ex != 0) cx = (unsigned)(i - ex) >= (unsigned)a.nx;
        if constexpr (ey != 0) cy = (unsigned)(j - ey) >= (unsigned)a.ny;
        if constexpr (ez != 0) cz = (unsigned)(k - ez) >= (unsigned)a.nz;
        constexpr int fx = ex != 0 ? entry_face(al, 0) : 0;
        constexpr int fy = ey != 0 ? entry_face(al, 1) : 0;
        constexpr int fz = ez != 0 ? entry_face(al, 2) : 0;
        const bool wall = (cx && !((a.halo_faces >> fx) & 1)) || (cy && !((a.halo_faces >> fy) & 1)) ||
                          (cz && !((a.halo_faces >> fz) & 1));
        if (!wall) {
            const T* h = static_cast<const T*>(a.halo);
            long long idx;
            if (int(cx) + int(cy) + int(cz) == 2) {
                constexpr int free_axis = al >= 7 ? edge_free_axis(al) : 0;
                const int t = free_axis == 0 ? i : (free_axis == 1 ? j : k);
                idx = a.hoff[al >= 7 ? edge_region(al) : 0] + t;
            } else if (cx) {
                idx = a.hoff[fx] + (long long)cross_slot(fx, al, -1) * a.ny * a.nz + (j - ey) +
                      (long long)a.ny * (k - ez);
            } else if (cy) {
                idx = a.hoff[fy] + (long long)cross_slot(fy, al, -1) * a.nx * a.nz + (i - ex) +
                      (long long)a.nx * (k - ez);
            } else {
                idx = a.hoff[fz] + (long long)cross_slot(fz, al, -1) * a.nx * a.ny + (i - ex) +
                      (long long)a.nx * (j - ey);
            }
            return S::ld(h + idx, a.inv_scale);
        }
        lid = top && !((a.halo_faces >> LBM_FACE_ZP) & 1);
    }
    R val = S::ld(src + (long long)OPP[al] * a.ps + c, a.inv_scale);
    if constexpr (ez == -1 && ex != 0) {
        // moving lid (bounce_back.f90:77-78): f(14) = f_post(11) - rho/6 U0,
        // f(13) = f_post(12) - rho/6 (-U0); the lid is applied last, so it
        // wins over the side walls on these two populations.
        if (lid) {
            const double t = rho_lid / 6.0 * a.u_lid;
            val = (ex > 0) ? R(double(val) + t) : R(double(val) - t);
        }
    }
    return val;
}

// One FP16-stored MRT cell of a whole wavefront away from the y and z faces,
// with the forward moment transform on the matrix cores: the populations stay
// in their stored FP16 form (the transform takes them exactly), the moments
// and the relaxation are FP32 in the stored scale 2^k, and the inverse
// transform (VALU) gives the scaled post-collision values that the
// stochastic rounding stores directly.  Same pulls, walls, stores and hook as
// the VALU path; the forward sums run in the matrix cores' order.
template <bool MACRO, bool HALO, class Args, class Post>
__device__ __forceinline__ void mfma_cell(const Args& a, const _Float16* src, _Float16* dst, int i, int j, int k,
                                          int c, const Post& post) {
    using S = StoreF16;
    _Float16 h[Q];
    sfor<Q>([&](auto A) {
        constexpr int al = decltype(A)::value;
        const long long off = (long long)EX[al] + (long long)EY[al] * a.px + (long long)EZ[al] * a.pxy;
        h[al] = src[(long long)al * a.ps + c - off];
    });
    if (i == 0 || i == a.nx - 1) {
        // an x-face lane: halfway bounce-back or the halo (no lid here: not
        // the top layer), back in the stored form (exact: a power-of-two scale)
        sfor<Q>([&](auto A) {
            constexpr int al = decltype(A)::value;
            if constexpr (EX[al] != 0) {
                if ((unsigned)(i - EX[al]) >= (unsigned)a.nx)
                    h[al] = (_Float16)(pull_outside<S, al, HALO>(a, src, i, j, k, c, false, 0.0) * a.scale);
            }
        });
    }
    const int lane = __lane_id();
    float m[Q];
    mrt_forward_mfma(h, m, lane);
    Moments<float, true> q;
    q.drho = m[0] * a.inv_scale;
    q.rho = 1.0f + q.drho;
    q.u = m[3] * a.inv_scale / q.rho;
    q.v = m[5] * a.inv_scale / q.rho;
    q.w = m[7] * a.inv_scale / q.rho;
    if constexpr (MACRO) {
        const long long mm = i + (long long)a.nx * (j + (long long)a.ny * k);
        a.rho[mm] = 1.0 + double(q.drho);
        a.u[mm] = double(q.u);
        a.v[mm] = double(q.v);
        a.w[mm] = double(q.w);
    }
    float meq[Q];
    mrt_meq<float, true>(q, a.meq12_rho != 0, meq);
    const float snu = float(a.snu), sq = float(a.sq);
    sfor<Q>([&](auto K) {
        constexpr int kk = decltype(K)::value;
        constexpr int kind = rate_kind<kk>();
        if constexpr (kind == 1) m[kk] = m[kk] - snu * (m[kk] - a.scale * meq[kk]);
        else if constexpr (kind == 2) m[kk] = m[kk] - sq * (m[kk] - a.scale * meq[kk]);
    });
    float f[Q];
    mrt_inverse(m, f);
    const unsigned hr = cell_random<S>(c, a.seed);
    _Float16 stored[Q];
    sfor<Q>([&](auto A) {
        constexpr int al = decltype(A)::value;
        stored[al] = to_half_stochastic(f[al], store_random<S, al>(hr));
        __builtin_nontemporal_store(stored[al], dst + (long long)al * a.ps + c);
    });
    if (post.active()) post.store(stored);
}

// The update of cell (i, j, k): pull the 19 populations (bounce-back or halo
// where the pull leaves the domain), macro, collision, store.  `post` sees
// the stored values (post.store(values)), e.g. to write them to a halo too,
// after the last store: the collision and its stores stay one straight-line
// block in every kernel, so the compiler contracts the arithmetic the same
// way and every kernel built on this function produces the same bits.
// MX: the forward MRT moment transform of whole FP16 wavefronts away from the
// y and z faces on the matrix cores (mfma.hpp); other waves use the VALU path.
// XSEL: a lane on an x face points its unconditional x-shifted pulls at its
// own column of the pulled row (the value is replaced by the wall's), so the
// wave reads no line of a neighbouring row; for the short rows of the batched
// instances (32-96 cells) that saves a line per row and plane.
template <class S, int COLL, bool MACRO, bool HALO, bool MX = false, bool XSEL = false, class Args, class Post = NoPost>
__device__ __forceinline__ void cavity_cell(const Args& a, int i, int j, int k, const Post& post = Post{}) {
    using T = typename S::T;
    using R = typename S::Real;
    const int c = i + a.px * (j + a.ny * k);
    const T* __restrict__ src = static_cast<const T*>(a.src);
    T* __restrict__ dst = static_cast<T*>(a.dst);
    const bool top = (k == a.nz - 1);
    const int lid_idx = i + a.nx * j;
    double rho_lid = 0.0;
    if (top) rho_lid = a.lid_rho[lid_idx];

    R f[Q];
    if ((unsigned)(j - 1) < (unsigned)(a.ny - 2) && (unsigned)(k - 1) < (unsigned)(a.nz - 2)) {
        // Not on a y or z face: every pull from a neighbour row and plane that
        // exists, so every pull's address lies inside its population's plane
        // -- also for a lane on an x face, whose pull then reads a
        // neighbouring row's element and is replaced below.  No per-pull test.
        if constexpr (MX) {
            static_assert(std::is_same_v<S, StoreF16> && COLL == COLL_MRT,
                          "matrix-core transforms: FP16-stored MRT");
            if (__builtin_amdgcn_read_exec() == ~0ull) {
                mfma_cell<MACRO, HALO>(a, src, dst, i, j, k, c, post);
                return;
            }
        }
        sfor<Q>([&](auto A) {
            constexpr int al = decltype(A)::value;
            long long off = (long long)EX[al] + (long long)EY[al] * a.px + (long long)EZ[al] * a.pxy;
            if constexpr (XSEL && EX[al] != 0) {
                if ((unsigned)(i - EX[al]) >= (unsigned)a.nx) off -= EX[al];
            }
            f[al] = S::ld(src + ((long long)al * a.ps + c - off), a.inv_scale);
        });
        if (i == 0 || i == a.nx - 1) {
            sfor<Q>([&](auto A) {
                constexpr int al = decltype(A)::value;
                if constexpr (EX[al] != 0) {
                    if ((unsigned)(i - EX[al]) >= (unsigned)a.nx)
                        f[al] = pull_outside<S, al, HALO>(a, src, i, j, k, c, top, rho_lid);
                }
            });
        }
    } else {
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
                    f[al] = pull_outside<S, al, HALO>(a, src, i, j, k, c, top, rho_lid);
                }
            }
        });
    }

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
    const unsigned h = cell_random<S>(c, a.seed);
    T stored[Q];
    sfor<Q>([&](auto A) {
        constexpr int al = decltype(A)::value;
        stored[al] = S::cvt(f[al], a.scale, store_random<S, al>(h));
        __builtin_nontemporal_store(stored[al], dst + (long long)al * a.ps + c);
    });
    if (post.active()) post.store(stored);
}


}  // namespace lbm

// Launch the single-domain step kernel (walls only, no halo) for `a` on
// `stream` with the default launch shape for its size (cavity.hip).
int lbm_launch_cavity_step(const lbm::StepArgs& a, int storage, int collision, bool macro, hipStream_t stream);
