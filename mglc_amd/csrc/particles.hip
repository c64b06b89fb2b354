// particles.hip -- rigid spheres coupled to the D3Q19 fluid on one MI355X, as
// the reference couples discs to D2Q9 (MPI/Micro_particles/fortran/case4/mpi).
//
// Per step, all on the device (no host round trip):
//   k_links           one block per sphere: the boundary links (fluid node x,
//                     direction a with x + e_a inside the sphere) of the
//                     sphere's shell, compacted in a fixed order into the
//                     sphere's slab (or the last list, while no node it can
//                     see changed owner); then one lane per link: wall
//                     distance q by bisection, quadratic interpolated
//                     bounce-back with the moving-wall term
//                     (particle_bounceback.F90:37-141), the value stored in
//                     the solid node's own slot of the source buffer (the
//                     slot the fluid node pulls from), and the link's
//                     momentum exchange and torque (particle_force.F90:
//                     31-92), summed in FP64 in a fixed order: force and
//                     torque are deterministic
//   k_particle_step   every node: plain pull streaming with halfway bounce-back
//                     at the box walls, collision, per block the sums of rho
//                     and count over fluid nodes; the macro fields on the last
//                     step of a call
//   k_particle_update one wave per sphere: spring and buoyancy forces and the
//                     kinematics (particle_force.F90:96-191,
//                     particle_update.F90:26-42); stop codes for
//                     interpenetration and wall penetration; the neighbours
//                     within LINK_GAP
//   k_particle_mark   new solid map, rho_s and u = 0 in solids
//                     (particle_update.F90:69-91); flags the link lists that
//                     see a node changing owner
//   k_refill_scan     one block per sphere: its uncovered nodes (old solid,
//                     new fluid) appended to a list
//   k_fluid_chain     rhoAvg over the new fluid nodes (particle_bounceback.F90:
//                     15-35) from the fluid kernel's block sums, corrected for
//                     the nodes the map update covered and uncovered
//   k_refill_proc     one lane per listed node: refilled by 3-point
//                     extrapolation along the most outward lattice direction
//                     and moment replacement with the body velocity
//                     (particle_update.F90:110-186), then collided; the change
//                     of rho goes into the next step's rhoAvg
// The first step after the spheres are attached takes rhoAvg from a full pass
// (k_fluid_sum, k_fluid_first).
// Solid nodes are streamed and collided like fluid ones (their values are
// never read), which keeps the fluid kernel's waves uniform.
// The 3-D choices (volume 4/3 pi r^3, I = 2/5 m r^2, springs on all six
// walls, rest equilibrium for stencil points that are not fluid nodes) are
// those of the CPU restatement the GPU path is tested against.
#include <hip/hip_runtime.h>

#include <cmath>
#include <algorithm>
#include <new>
#include <vector>

#include "collide.hpp"
#include "context.hpp"
#include "sphere_geom.hpp"
#include "step.hpp"

using namespace lbm;

// rhoAvg carried from step to step (particle_bounceback.F90:15-35 takes it
// over the fluid nodes of the last macro).  After the map update of a step:
// the fluid kernel's sum over the old map's fluid nodes, less the newly
// covered nodes (their rho before rho_s), plus rho_s for each uncovered node
// -- the sum over the new map; the refill then adds (rho - rho_s) for each
// refilled node -- the sum the next step's streaming uses.
struct MeanChain {
    double sum, count;       // fluid sum and count after the map update
    unsigned long long cov;  // fixed point: rho of the newly covered nodes
    unsigned long long ref;  // fixed point: rho - rho_s of the refilled nodes
    int ncov;                // newly covered nodes
    int pad;
};

struct ParticleState {
    int np = 0;
    lbm_particle_params pp{};
    double *xc[2] = {nullptr, nullptr}, *U[2] = {nullptr, nullptr}, *Om[2] = {nullptr, nullptr};
    double* r = nullptr;
    double* F = nullptr;     // hydrodynamic force (3 np) of the last step
    double* T = nullptr;     // hydrodynamic torque (3 np)
    double* Ftot = nullptr;  // total force (3 np)
    int* owner[2] = {nullptr, nullptr};
    int cur = 0;             // current map / sphere state
    // boundary-link list: a slab per sphere, sized for its largest shell
    unsigned* items = nullptr;
    unsigned char* item_fluid1 = nullptr;  // per entry: the node x - e_a is in the box and fluid
    long long* slab_off = nullptr;  // np: first entry of each slab
    int* slab_cap = nullptr;        // np: entries of each slab
    // link-list reuse: a list stays valid while the map near its sphere does
    int* link_count = nullptr;  // np: entries of the last list built
    int* link_lo = nullptr;     // 3 np: the bounding-box origin its entries are relative to
    int* changed = nullptr;     // 2 np + 2, by step parity: the map update changed a node the sphere's list
                                //  can see; the last two words: every list is rebuilt
    int* nbrs = nullptr;        // LINK_NBRS np: the spheres within LINK_GAP of each (k_particle_update)
    int* nnbr = nullptr;        // np: their number (above LINK_NBRS: too many, always rebuild)
    int* refill = nullptr;        // refill_cap pairs (node, sphere) of uncovered nodes
    int* refill_count = nullptr;  // entries this step
    int refill_cap = 0;
    double* scal = nullptr;  // rhoAvg of the refill (k_fluid_chain)
    double* red = nullptr;   // reduction partials (k_fluid_sum)
    int* err = nullptr;      // device {stop code (0 = running), step of the stop, number of the step being made,
                             // a sphere moved 1/4 node this step}
    int* alone = nullptr;    // np flags of k_particle_update for k_particle_mark
    double* cpart = nullptr;      // k_particle_step's per-block fluid sums
    double* cpart2 = nullptr;     // their slice sums (k_refill_scan)
    MeanChain* chain = nullptr;   // rhoAvg carried from step to step
    bool chain_started = false;   // the first step starts it with a full pass
    bool stopped = false;    // a stop was reported: no further steps
    // graphs of GRAPH_STEPS steps without macro output, one per pair of
    // (fluid buffer, map) parities at their start, captured on cap_stream
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t graph[4] = {nullptr, nullptr, nullptr, nullptr};
};

namespace {

constexpr double PI = 3.14159265358979323846;
constexpr int REFILL_BLOCKS = 128;  // k_refill_proc: lanes walk the list with a grid stride
constexpr int GRAPH_STEPS = 10;     // steps per graph replay; even, so the buffers end where they started
constexpr int SUM_BLOCKS = 256;      // k_refill_scan: at least this many blocks (slices of the fluid sums)
constexpr int UPDATE_CHUNK = 1024;  // k_particle_update: spheres staged in LDS at a time (32 KiB)
constexpr int LINK_NBRS = 16;     // neighbours a sphere's link list watches (more: rebuilt every step)
// surface gap within which another sphere can change the map near a list: a
// list reads owners up to r + 2.92 from its centre (x - e_a of a shell node),
// plus 1/4 node of motion for each of the two spheres, with room to spare
constexpr double LINK_GAP = 4.5;
constexpr int UPDATE_WAVES = 4;     // k_particle_update: waves (spheres) per block, which share the staging
                                    // (16, staging a quarter as often on a quarter of the CUs: 18 -> 23 us)
constexpr int FLUID_BLOCKS = 4096;  // k_fluid_sum: enough waves in flight to stream the map and rho

inline int scan_blocks(int np) { return np > SUM_BLOCKS ? np : SUM_BLOCKS; }

struct PArgs {
    const void* src;
    void* dst;
    void* src_w;  // the source buffer, written by k_links at solid nodes only
    int nx, ny, nz, px;
    long long pxy, ps;
    double snu, sq;
    float scale, inv_scale;
    int meq12_rho;
    double rho0, rho_solid;
    const int* owner;      // map of this step: sphere index, -1 for fluid
    const int* owner_new;  // map after the sphere update (refill)
    const double *xc, *U, *Om;           // sphere state of this step
    const double *xc_new, *U_new, *Om_new;  // after the update (refill)
    const double* r;
    const double* rho_avg;  // rhoAvg of the refill (k_fluid_chain)
    double *rho, *u, *v, *w;
    int* err;  // {code, step, steps completed, some sphere moved 1/4 node}: first stop wins
    int step_off;  // the step a stop reports: steps completed + step_off
    double* cpart;     // k_particle_step's per-block fluid sums (rho, count)
    double* cpart2;    // their slice sums, one pair per k_refill_scan block
    long long ncb;     // k_particle_step's blocks
    MeanChain* chain;  // rhoAvg from one step to the next
    int* alone;  // per sphere: no other sphere within one node (k_particle_update)
    int np;
    // boundary links
    unsigned* items;
    unsigned char* item_fluid1;
    const long long* slab_off;
    const int* slab_cap;
    int* link_count;
    int* link_lo;
    int* changed;  // [parity][np], then [parity] rebuild every list
    int* nbrs;
    int* nnbr;
    int par;       // the state's parity this step (which half of changed the map update sets)
    // sphere update
    double gravity, threshold_wall, stiff_wall, threshold_particle, stiff_particle;
    double *F, *T, *Ftot;
    double *xc_out, *U_out, *Om_out;
    // uncovered nodes to refill
    int* refill;
    int* refill_count;
    int refill_cap;
};

using lbm::calq;

// Fixed point for the fluid-density corrections: values near 1 summed as
// integers (2^40 per unit), so that sums taken in any order by atomics are
// the same on every run; a node's rounding is below 1e-12.
constexpr double FIX_ONE = 1099511627776.0;  // 2^40
__device__ __forceinline__ unsigned long long to_fix(double x) {
    return (unsigned long long)__double2ll_rn(x * FIX_ONE);
}

__device__ __forceinline__ double dist2(double dx, double dy, double dz) { return sphere_dist2(dx, dy, dz); }

// The spin part of the body velocity at a point, W = Omega x (x - c); the
// body velocity is U + W (CaseIV_DKT.F90:509-510 forms temp1, temp2 the same
// way for a disc spinning about z).
__device__ __forceinline__ void body_spin(const double* xc, const double* Om, int id, double x, double y, double z,
                                          double& wx, double& wy, double& wz) {
    const double rx = x - xc[3 * id], ry = y - xc[3 * id + 1], rz = z - xc[3 * id + 2];
    wx = Om[3 * id + 1] * rz - Om[3 * id + 2] * ry;
    wy = Om[3 * id + 2] * rx - Om[3 * id] * rz;
    wz = Om[3 * id] * ry - Om[3 * id + 1] * rx;
}

__device__ __forceinline__ void body_velocity(const double* xc, const double* U, const double* Om, int id, double x,
                                              double y, double z, double& ux, double& uy, double& uz) {
    double wx, wy, wz;
    body_spin(xc, Om, id, x, y, z, wx, wy, wz);
    ux = U[3 * id] + wx;
    uy = U[3 * id + 1] + wy;
    uz = U[3 * id + 2] + wz;
}

// Record a stop: the first code raised in a step call wins, with its step:
// the device count of completed steps (advanced by k_fluid_chain, between
// the map update and the refill) plus the launch's step_off.
__device__ __forceinline__ void stop(const PArgs& a, int code) {
    if (atomicCAS(a.err, 0, code) == 0) atomicExch(a.err + 1, *(volatile const int*)(a.err + 2) + a.step_off);
}

// Every particle kernel returns at once after a stop, so the state stays that
// of the stopping step's launches.
__device__ __forceinline__ bool stopped(const PArgs& a) { return *(volatile const int*)a.err != 0; }

__device__ __forceinline__ bool in_box(const PArgs& a, int i, int j, int k) {
    return (unsigned)i < (unsigned)a.nx && (unsigned)j < (unsigned)a.ny && (unsigned)k < (unsigned)a.nz;
}

// rhoAvg of a step's streaming: the chain after the last step's refill.
__device__ __forceinline__ double stream_rho_avg(const PArgs& a) {
    const MeanChain* ch = a.chain;
    return (ch->sum + (double)(long long)ch->ref / FIX_ONE) / ch->count;
}

// Absolute population a of a stored state at storage index c.
template <class S>
__device__ __forceinline__ double abs_pop(const typename S::T* buf, long long ps, long long c, int a, float inv) {
    const double g = double(S::ld(buf + (long long)a * ps + c, inv));
    return S::SHIFT ? g + weight(a) : g;
}

// Bounding box (0-based cell indices, clamped) of sphere `id` with `margin`.
__device__ __forceinline__ void sphere_box(const PArgs& a, const double* xc, int id, int margin, int lo[3], int hi[3]) {
    const int n[3] = {a.nx, a.ny, a.nz};
    for (int d = 0; d < 3; ++d) {
        lo[d] = (int)floor(xc[3 * id + d] - a.r[id]) - 1 - margin;
        hi[d] = (int)ceil(xc[3 * id + d] + a.r[id]) - 1 + margin;
        lo[d] = lo[d] < 0 ? 0 : lo[d];
        hi[d] = hi[d] > n[d] - 1 ? n[d] - 1 : hi[d];
    }
}

// Rows k of column (i, j) (0-based, clamped to [k_lo, k_hi]) that can hold
// nodes within distance R of the centre (cx, cy, cz) (1-based coordinates):
// the chord |z - cz| <= sqrt(R^2 - dx^2 - dy^2), widened by a node on each
// side so rounding cannot lose one.  Callers apply the exact test per node.
__device__ __forceinline__ void column_rows(double cx, double cy, double cz, double R, int i, int j, int k_lo, int k_hi,
                                            int& k0, int& k1) {
    const double dx = i + 1.0 - cx, dy = j + 1.0 - cy;
    const double rho2 = dx * dx + dy * dy;
    if (rho2 > (R + 1.0) * (R + 1.0)) {
        k0 = 1;
        k1 = 0;
        return;
    }
    const double h = sqrt(fmax(R * R - rho2, 0.0));
    k0 = (int)floor(cz - h) - 2;
    k1 = (int)ceil(cz + h);
    k0 = k0 < k_lo ? k_lo : k0;
    k1 = k1 > k_hi ? k_hi : k1;
}

// A link list entry: node offsets in the sphere's bounding box (9 bits each;
// boxes are at most 2 r + 5 <= 505 nodes wide) and the direction (5 bits).
__device__ __forceinline__ unsigned link_entry(int di, int dj, int dk, int al) {
    return ((unsigned)di << 23) | ((unsigned)dj << 14) | ((unsigned)dk << 5) | (unsigned)al;
}

__device__ __forceinline__ double wave_sum(double x) {
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    return x;
}

// Block-wide sums of x and y (RED_THREADS threads), in a fixed order; every
// thread gets the totals.
__device__ __forceinline__ void block_pair_sum(double& x, double& y) {
    __shared__ double sh[2][RED_THREADS / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    x = wave_sum(x);
    y = wave_sum(y);
    if (lane == 0) {
        sh[0][wid] = x;
        sh[1][wid] = y;
    }
    __syncthreads();
    x = y = 0.0;
    for (int w = 0; w < RED_THREADS / 64; ++w) {
        x += sh[0][w];
        y += sh[1][w];
    }
}

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_CHUNK = 2048;  // bounding-box nodes tested per block round

// A sphere's state of this step, read once per block by the link pass.
struct SphereNow {
    double c[3], u[3], om[3], r;
    double rho_avg;  // of the fluid, for the moving-wall term
};

// One boundary link, entry `it` of the sphere `sp` whose bounding box starts at
// lo (fluid node x, direction a with x + e_a inside the sphere), in two
// parts, so that one link's gathers are in flight while the next link's wall
// distance is computed:
//   link_begin: wall distance q and boundary point x0 (calQ,
//     particle_bounceback.F90:101-141), then the loads the interpolation
//     needs;
//   link_end: f(abar, x) by quadratic interpolated bounce-back with the
//     moving-wall term (particle_bounceback.F90:37-96), stored in the source
//     buffer at the solid node x + e_a, direction abar (exactly where the
//     fluid node's pull along abar reads), and the momentum exchange
//     (e_a - u_w) f_post(a, x) - (e_abar - u_w) f(abar, x) with its torque
//     (particle_force.F90:31-92), added to acc.
struct LinkWork {
    long long c;
    int al;
    bool near, fl1, in2;
    int o2;
    double x0, y0, z0, qq, fa, s1, s2;
};

template <class S>
__device__ __forceinline__ LinkWork link_begin(const PArgs& a, const SphereNow& sp, const int (&lo)[3], unsigned it,
                                               bool fl1) {
    using T = typename S::T;
    const T* src = static_cast<const T*>(a.src);
    LinkWork w;
    const int al = (int)(it & 31u), bd = OPP[al];
    w.al = al;
    const int i = lo[0] + (int)(it >> 23), j = lo[1] + (int)((it >> 14) & 511u),
              k = lo[2] + (int)((it >> 5) & 511u);
    const long long c = i + (long long)a.px * (j + (long long)a.ny * k);
    w.c = c;
    if (!calq(sp.c[0], sp.c[1], sp.c[2], sp.r, i + 1.0, j + 1.0, k + 1.0, EX[al], EY[al], EZ[al], w.x0, w.y0, w.z0,
              w.qq))
        stop(a, 3);
    const bool near = w.qq < 0.5;
    w.near = near;
    const int i1 = i - EX[al], j1 = j - EY[al], k1 = k - EZ[al];
    const int i2 = i1 - EX[al], j2 = j1 - EY[al], k2 = k1 - EZ[al];
    // x - e_a: in the box and fluid is the scan's flag; x - 2 e_a is read
    // only by the q < 1/2 branch
    w.fl1 = fl1;
    w.in2 = in_box(a, i2, j2, k2);
    const int ci1 = min(max(i1, 0), a.nx - 1), cj1 = min(max(j1, 0), a.ny - 1), ck1 = min(max(k1, 0), a.nz - 1);
    const int ci2 = min(max(i2, 0), a.nx - 1), cj2 = min(max(j2, 0), a.ny - 1), ck2 = min(max(k2, 0), a.nz - 1);
    const long long c1 = ci1 + (long long)a.px * (cj1 + (long long)a.ny * ck1);
    const long long c2 = ci2 + (long long)a.px * (cj2 + (long long)a.ny * ck2);
    w.o2 = near ? a.owner[ci2 + (long long)a.nx * (cj2 + (long long)a.ny * ck2)] : 0;
    w.fa = abs_pop<S>(src, a.ps, c, al, a.inv_scale);
    w.s1 = abs_pop<S>(src, a.ps, c1, near ? al : bd, a.inv_scale);
    w.s2 = abs_pop<S>(src, a.ps, near ? c2 : c, near ? al : bd, a.inv_scale);
    return w;
}

template <class S>
__device__ __forceinline__ void link_end(const PArgs& a, const SphereNow& sp, const LinkWork& w, double (&acc)[6]) {
    using T = typename S::T;
    using R = typename S::Real;
    T* src_w = static_cast<T*>(a.src_w);
    const int al = w.al, bd = OPP[al];
    const double cx = sp.c[0], cy = sp.c[1], cz = sp.c[2];
    const double x0 = w.x0, y0 = w.y0, z0 = w.z0, qq = w.qq, fa = w.fa, s1 = w.s1, s2 = w.s2;
    const bool fl1 = w.fl1, fl2 = w.in2 && w.o2 < 0;
    const double rx0 = x0 - cx, ry0 = y0 - cy, rz0 = z0 - cz;
    const double wx = sp.om[1] * rz0 - sp.om[2] * ry0, wy = sp.om[2] * rx0 - sp.om[0] * rz0,
                 wz = sp.om[0] * ry0 - sp.om[1] * rx0;  // body_spin
    const double ux = sp.u[0], uy = sp.u[1], uz = sp.u[2];
    const double eu = EX[bd] * (ux + wx) + EY[bd] * (uy + wy) + EZ[bd] * (uz + wz);
    const double rho_avg = sp.rho_avg;
    double val;
    if (w.near) {
        const double f1 = fl1 ? s1 : weight(al) * a.rho0;
        const double f2 = fl2 ? s2 : weight(al) * a.rho0;
        val = qq * (1.0 + 2.0 * qq) * fa + (1.0 - 4.0 * qq * qq) * f1 - qq * (1.0 - 2.0 * qq) * f2 +
              6.0 * weight(al) * rho_avg * eu;
    } else {
        const double fb = s2, fb1 = fl1 ? s1 : weight(bd) * a.rho0;
        const double iq = 1.0 / qq, i12 = 1.0 / (1.0 + 2.0 * qq);
        val = fa * iq * i12 + fb * (2.0 * qq - 1.0) * iq - fb1 * (2.0 * qq - 1.0) * i12 +
              6.0 * weight(al) * rho_avg * iq * i12 * eu;
    }
    const long long cs = w.c + EX[al] + (long long)EY[al] * a.px + (long long)EZ[al] * a.pxy;
    S::st(src_w + (long long)bd * a.ps + cs, S::SHIFT ? R(val - weight(bd)) : R(val), a.scale, 0u);
    const double dFx = (EX[al] - ux - wx) * fa - (EX[bd] - ux - wx) * val;
    const double dFy = (EY[al] - uy - wy) * fa - (EY[bd] - uy - wy) * val;
    const double dFz = (EZ[al] - uz - wz) * fa - (EZ[bd] - uz - wz) * val;
    const double rx = x0 - cx, ry = y0 - cy, rz = z0 - cz;
    acc[0] += dFx;
    acc[1] += dFy;
    acc[2] += dFz;
    acc[3] += ry * dFz - rz * dFy;
    acc[4] += rz * dFx - rx * dFz;
    acc[5] += rx * dFy - ry * dFx;
}

// The boundary links of one sphere and what they exchange, one block per
// sphere (all of a sphere's gathers go through one CU's L1):
//  1. the links (fluid node x, direction a with x + e_a owned by the sphere)
//     of the shell r < |x - c| <= r + 1.5 (the map's own distance test), in a
//     fixed order into the sphere's slab, each with a flag: x - e_a is in the
//     box and fluid (the scan has loaded that neighbour's owner word).  The bounding box is walked in
//     chunks of SCAN_CHUNK nodes (x fastest): the distance test alone picks
//     the chunk's shell nodes, which are compacted in LDS in box order; then
//     every lane takes one shell node and loads the owner words of the node
//     and its 18 neighbours in one go, so the gathers run with full waves
//     (about one box node in seven is in the shell).  Entries are written in
//     one pass at a block-wide running offset: per round of 256 shell nodes
//     the waves in order, inside a wave direction by direction, lane by lane;
//     A link is (x fluid, x + e_a owned by the sphere), a function of the map
//     alone, so step 1 runs only when the last map update changed the owner
//     of a node the list can see (k_particle_mark and k_refill_scan flag the
//     spheres within reach of each node they change, flag_near).  Otherwise
//     the last list is used as it stands, with the box origin it was built in
//     -- the same links, in the order of that build;
//  2. one lane per link, two at a time (entries e and e + 256, e = 64 w +
//     512 t + lane for wave w): link_begin of both, then link_end of both;
//  3. force and torque: per lane in entry order, per wave by a fixed shuffle
//     tree, then the four waves in order -- the same sums on every run.
// Measured on MI355X at 192^3 with 2000 spheres: ordering the entries
// direction by direction over the whole sphere, or giving each XCD the
// spheres of one region of the box, fetched less or the same and ran no
// faster, and neither did more or fewer registers (occupancy); the link pass
// is bound by its scattered gathers, so what counts is fewer of them and
// more of them in flight.
template <class S>
__global__ __launch_bounds__(SCAN_THREADS) void k_links(const PArgs a) {
    __shared__ int halt;  // read once, so the whole block leaves together (barriers follow)
    __shared__ unsigned shell[SCAN_CHUNK];
    __shared__ int part[SCAN_CHUNK / SCAN_THREADS][SCAN_THREADS / 64];
    __shared__ int wave_tot[SCAN_THREADS / 64];
    __shared__ int rebuild;
    const int id = blockIdx.x;
    if (threadIdx.x == 0) {
        // the list from the last step holds unless the last map update
        // changed the owner of a node the list can see (flag_near), or asked
        // for every list to be rebuilt
        rebuild = (a.changed[(long long)(1 - a.par) * a.np + id] != 0 ||
                   a.changed[2LL * a.np + (1 - a.par)] != 0) ? 1 : 0;
        halt = stopped(a);
    }
    __syncthreads();
    if (halt) return;
    int lo[3], hi[3];
    sphere_box(a, a.xc, id, 1, lo, hi);
    const unsigned bx = hi[0] - lo[0] + 1, by = hi[1] - lo[1] + 1, bz = hi[2] - lo[2] + 1;
    const unsigned total = (hi[0] >= lo[0] && hi[1] >= lo[1] && hi[2] >= lo[2]) ? bx * by * bz : 0u;
    const double cx = a.xc[3 * id], cy = a.xc[3 * id + 1], cz = a.xc[3 * id + 2], rad = a.r[id];
    const double rin2 = rad * rad, rout2 = (rad + 1.5) * (rad + 1.5);
    unsigned* list = a.items + a.slab_off[id];
    unsigned char* fluid1 = a.item_fluid1 + a.slab_off[id];
    const int cap = a.slab_cap[id];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    constexpr int PER = SCAN_CHUNK / SCAN_THREADS;
    auto coords = [&](unsigned t, int& i, int& j, int& k) {
        i = lo[0] + (int)(t % bx);
        j = lo[1] + (int)((t / bx) % by);
        k = lo[2] + (int)(t / (bx * by));
    };
    int pos = 0;  // entries written so far (block-uniform)
    if (!rebuild) {  // the last list, relative to the box it was built in
        pos = a.link_count[id];
        for (int d = 0; d < 3; ++d) lo[d] = a.link_lo[3 * id + d];
    }
    for (unsigned base = 0; rebuild && base < total; base += SCAN_CHUNK) {
        // 1. shell nodes of the chunk, compacted in box order: node
        //    base + u * 256 + tid goes after every (u' < u) and every earlier
        //    wave of the same u
        bool in[PER];
        unsigned long long bal[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const unsigned t = base + u * SCAN_THREADS + threadIdx.x;
            in[u] = false;
            if (t < total) {
                int i, j, k;
                coords(t, i, j, k);
                const double d2 = dist2(i + 1.0 - cx, j + 1.0 - cy, k + 1.0 - cz);
                in[u] = d2 > rin2 && d2 <= rout2;
            }
            bal[u] = __ballot(in[u]);
            if (lane == 0) part[u][wid] = (int)__popcll(bal[u]);
        }
        __syncthreads();
        int nshell = 0;
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            int before = nshell;
            for (int w = 0; w < SCAN_THREADS / 64; ++w) {
                if (w < wid) before += part[u][w];
                nshell += part[u][w];
            }
            if (in[u]) {
                const int within = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(bal[u] >> 32),
                                                                  __builtin_amdgcn_mbcnt_lo((unsigned)bal[u], 0u));
                shell[before + within] = base + u * SCAN_THREADS + threadIdx.x;
            }
        }
        __syncthreads();
        // 2. one lane per shell node: its links, then the entries
        for (int r0 = 0; r0 < nshell; r0 += SCAN_THREADS) {
            const int q = r0 + (int)threadIdx.x;
            unsigned mk = 0u, e0 = 0u, fl1mask = 0u;
            if (q < nshell) {
                int i, j, k;
                coords(shell[q], i, j, k);
                e0 = link_entry(i - lo[0], j - lo[1], k - lo[2], 0);
                const int own = a.owner[i + (long long)a.nx * (j + (long long)a.ny * k)];
                int nb[Q];
                sfor<Q - 1>([&](auto B) {
                    constexpr int al = decltype(B)::value + 1;
                    const int ni = min(max(i + EX[al], 0), a.nx - 1), nj = min(max(j + EY[al], 0), a.ny - 1),
                              nk = min(max(k + EZ[al], 0), a.nz - 1);
                    nb[al] = a.owner[ni + (long long)a.nx * (nj + (long long)a.ny * nk)];
                });
                if (own < 0) {
                    sfor<Q - 1>([&](auto B) {
                        constexpr int al = decltype(B)::value + 1;
                        if (in_box(a, i + EX[al], j + EY[al], k + EZ[al]) && nb[al] == id) mk |= 1u << al;
                        // the link's first stencil point x - e_a (the neighbour
                        // along the opposite direction): in the box and fluid
                        if (in_box(a, i - EX[al], j - EY[al], k - EZ[al]) && nb[OPP[al]] < 0) fl1mask |= 1u << al;
                    });
                }
            }
            int wcnt = __popc(mk);
            for (int off = 32; off > 0; off >>= 1) wcnt += __shfl_xor(wcnt, off, 64);
            if (lane == 0) wave_tot[wid] = wcnt;
            __syncthreads();
            int at = pos, all = 0;
            for (int w = 0; w < SCAN_THREADS / 64; ++w) {
                if (w < wid) at += wave_tot[w];
                all += wave_tot[w];
            }
            sfor<Q - 1>([&](auto B) {
                constexpr int al = decltype(B)::value + 1;
                const bool has = (mk >> al) & 1u;
                const unsigned long long b = __ballot(has);
                const int before = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(b >> 32),
                                                                  __builtin_amdgcn_mbcnt_lo((unsigned)b, 0u));
                if (has && at + before < cap) {
                    list[at + before] = e0 | (unsigned)al;
                    fluid1[at + before] = (unsigned char)((fl1mask >> al) & 1u);
                }
                at += (int)__popcll(b);
            });
            pos += all;
            __syncthreads();  // wave_tot is rewritten by the next round
        }
    }
    // pos: the sphere's entries (block-uniform)
    if (pos > cap) {  // cannot happen for a slab sized by the shell bound
        if (threadIdx.x == 0) stop(a, 5);
        return;
    }
    if (rebuild && threadIdx.x == 0) {
        a.link_count[id] = pos;
        for (int d = 0; d < 3; ++d) a.link_lo[3 * id + d] = lo[d];
    }
    // the entries written above are visible to every wave of the block (the
    // last barrier of step 1)
    double acc[6] = {0, 0, 0, 0, 0, 0};
    SphereNow sp;
    for (int d = 0; d < 3; ++d) {
        sp.c[d] = a.xc[3 * id + d];
        sp.u[d] = a.U[3 * id + d];
        sp.om[d] = a.Om[3 * id + d];
    }
    sp.r = rad;
    sp.rho_avg = stream_rho_avg(a);
    for (int e = threadIdx.x; e < pos; e += 2 * SCAN_THREADS) {
        const int e1 = e + SCAN_THREADS;
        const int ee = min(e1, pos - 1);  // past the end: a copy of the last link, computed and dropped
        const unsigned it0 = list[e], it1 = list[ee];
        const bool f0 = fluid1[e] != 0, f1 = fluid1[ee] != 0;
        const LinkWork w0 = link_begin<S>(a, sp, lo, it0, f0);
        const LinkWork w1 = link_begin<S>(a, sp, lo, it1, f1);
        link_end<S>(a, sp, w0, acc);
        if (e1 < pos) link_end<S>(a, sp, w1, acc);
    }
    __shared__ double sums[6][SCAN_THREADS / 64];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        const double v = wave_sum(acc[q]);
        if (lane == 0) sums[q][wid] = v;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        double v = sums[threadIdx.x][0];
        for (int w = 1; w < SCAN_THREADS / 64; ++w) v += sums[threadIdx.x][w];
        if (threadIdx.x < 3) a.F[3 * id + threadIdx.x] = v;
        else a.T[3 * id + threadIdx.x - 3] = v;
    }
}

// The 19 populations node (i, j, k) pulls from the source buffer: pull
// streaming (fluid.F90:75-161) with halfway bounce-back at the box walls -- a
// pull out of a solid node reads the interpolated bounce-back value k_links
// stored there.  The fluid kernel and the map update (for the rho of a node
// it covers) take the same values through this one function.
template <class S>
__device__ __forceinline__ void pull_node(const PArgs& a, int i, int j, int k, typename S::Real (&f)[Q]) {
    using T = typename S::T;
    const int c = i + a.px * (j + a.ny * k);
    const T* __restrict__ src = static_cast<const T*>(a.src);
    sfor<Q>([&](auto B) {
        constexpr int bd = decltype(B)::value;
        constexpr int ex = EX[bd], ey = EY[bd], ez = EZ[bd];
        if constexpr (bd == 0) {
            f[0] = S::ld(src + c, a.inv_scale);
        } else {
            bool in = true;
            if constexpr (ex != 0) in = in && (unsigned)(i - ex) < (unsigned)a.nx;
            if constexpr (ey != 0) in = in && (unsigned)(j - ey) < (unsigned)a.ny;
            if constexpr (ez != 0) in = in && (unsigned)(k - ez) < (unsigned)a.nz;
            const long long off = (long long)ex + (long long)ey * a.px + (long long)ez * a.pxy;
            f[bd] = in ? S::ld(src + (long long)bd * a.ps + (c - off), a.inv_scale)
                       : S::ld(src + (long long)OPP[bd] * a.ps + c, a.inv_scale);
        }
    });
}

// rho of fluid node (i, j, k) in this step's streaming, as the fluid kernel
// computes it (the same pulls and moments, so the same bits).
template <class S>
__device__ __forceinline__ double pulled_rho(const PArgs& a, int i, int j, int k) {
    using R = typename S::Real;
    R f[Q];
    pull_node<S>(a, i, j, k, f);
    const Moments<R, S::SHIFT> mo = moments<R, S::SHIFT>(f);
    return S::SHIFT ? 1.0 + double(mo.drho) : double(mo.rho);
}

// Every node: the pulls (pull_node), then macro (fluid.F90:164-184) on fluid
// nodes -- rho into the block's fluid sums every step (rhoAvg), rho, u, v, w
// into the macro fields on the last step of a call -- and the next collision
// (fluid.F90:1-72).  Solid nodes run the same arithmetic on values nobody
// reads.
template <class S, int COLL, bool MACRO>
__global__ __launch_bounds__(256) void k_particle_step(const PArgs a) {
    using T = typename S::T;
    using R = typename S::Real;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (stopped(a)) return;  // uniform: no kernel of this launch raises a stop
    double rho_fluid = 0.0, n_fluid = 0.0;  // this node's part of the fluid sums
    if (i < a.nx && j < a.ny) {
        const long long m = i + (long long)a.nx * (j + (long long)a.ny * k);
        const int c = i + a.px * (j + a.ny * k);
        R f[Q];
        pull_node<S>(a, i, j, k, f);
        const Moments<R, S::SHIFT> mo = moments<R, S::SHIFT>(f);
        if (a.owner[m] < 0) {
            const double rho = S::SHIFT ? 1.0 + double(mo.drho) : double(mo.rho);
            rho_fluid = rho;
            n_fluid = 1.0;
            if constexpr (MACRO) {  // other steps: rho goes into the block sums only
                a.rho[m] = rho;
                a.u[m] = double(mo.u);
                a.v[m] = double(mo.v);
                a.w[m] = double(mo.w);
            }
        }
        collide<S, COLL>(f, mo, a);
        T* __restrict__ dst = static_cast<T*>(a.dst);
        sfor<Q>([&](auto B) {
            constexpr int bd = decltype(B)::value;
            S::st(dst + (long long)bd * a.ps + c, f[bd], a.scale, 0u);
        });
    }
    // the block's sums of rho and count over its fluid nodes (fixed order)
    __shared__ double sh[2][4];
    const int t = threadIdx.x + blockDim.x * threadIdx.y, lane = t & 63, wid = t >> 6;
    rho_fluid = wave_sum(rho_fluid);
    n_fluid = wave_sum(n_fluid);
    if (lane == 0) {
        sh[0][wid] = rho_fluid;
        sh[1][wid] = n_fluid;
    }
    __syncthreads();
    if (t == 0) {
        const int nw = (int)((blockDim.x * blockDim.y + 63) / 64);
        double s0 = 0.0, s1 = 0.0;
        for (int w = 0; w < nw; ++w) {
            s0 += sh[0][w];
            s1 += sh[1][w];
        }
        const long long b = blockIdx.x + (long long)gridDim.x * (blockIdx.y + (long long)gridDim.y * blockIdx.z);
        a.cpart[2 * b] = s0;
        a.cpart[2 * b + 1] = s1;
    }
}

// Streamed population al of node (i, j, k) (in the box), as k_particle_step
// pulls it: from the neighbour upstream, or bounced back at the box wall.
// One load at a selected address (the weights of al and its opposite agree).
template <class S, int al>
__device__ __forceinline__ double streamed_pop_at(const PArgs& a, const typename S::T* src, int i, int j, int k) {
    const long long c = i + (long long)a.px * (j + (long long)a.ny * k);
    long long idx = c;
    if constexpr (al != 0) {
        const int si = i - EX[al], sj = j - EY[al], sk = k - EZ[al];
        idx = in_box(a, si, sj, sk) ? (long long)al * a.ps + si + (long long)a.px * (sj + (long long)a.ny * sk)
                                    : (long long)OPP[al] * a.ps + c;
    }
    const double g = double(S::ld(src + idx, a.inv_scale));
    return S::SHIFT ? g + weight(al) : g;
}

// One wave per sphere: the lanes split the pair loop of the particle springs,
// the wave sums their parts (fixed shuffle tree), lane 0 adds the wall
// springs and buoyancy and advances the sphere.
__global__ __launch_bounds__(64 * UPDATE_WAVES) void k_particle_update(const PArgs a) {
    // the other spheres' centres and radii, UPDATE_CHUNK at a time, staged in
    // LDS by the whole block (coalesced) and read by each wave's lanes
    __shared__ double4 others[UPDATE_CHUNK];
    const int lane = threadIdx.x & 63;
    const int p = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (blockIdx.x == 0 && threadIdx.x == 0) *a.refill_count = 0;  // the refill list of this step starts empty
    // p and the stop flag are uniform over the wave (the flag read by lane 0),
    // so the whole wave takes part in the shuffles below; every wave takes
    // part in the staging barriers
    const bool live = p < a.np && !__builtin_amdgcn_readfirstlane(stopped(a) ? 1 : 0);
    const double rho_avg = stream_rho_avg(a);
    const int pp = live ? p : 0;
    const double rp = a.r[pp];
    const double V = 4.0 / 3.0 * PI * rp * rp * rp;
    double fs[3] = {0.0, 0.0, 0.0};
    const double scale_p = V * (a.rho_solid - rho_avg) * a.gravity / a.stiff_particle;
    const double px = a.xc[3 * pp], py = a.xc[3 * pp + 1], pz = a.xc[3 * pp + 2];
    // the sphere's own state for lane 0's update after the pair loop, loaded
    // now so that its latency runs under the loop
    double Fl[3], Tl[3], Ul[3], Oml[3];
    for (int d = 0; d < 3; ++d) {
        Fl[d] = a.F[3 * pp + d];
        Tl[d] = a.T[3 * pp + d];
        Ul[d] = a.U[3 * pp + d];
        Oml[d] = a.Om[3 * pp + d];
    }
    const double pc[3] = {px, py, pz};
    bool close = false;  // another sphere within one node of this one
    int nnb = 0;         // spheres within LINK_GAP of this one (wave-uniform)
    for (int q0 = 0; q0 < a.np; q0 += UPDATE_CHUNK) {
        const int len = min(UPDATE_CHUNK, a.np - q0);
        __syncthreads();  // the previous chunk is read
        for (int t = threadIdx.x; t < len; t += blockDim.x) {
            const int q = q0 + t;
            others[t] = double4{a.xc[3 * q], a.xc[3 * q + 1], a.xc[3 * q + 2], a.r[q]};
        }
        __syncthreads();
        if (!live) continue;
        // lane by lane over q = lane, lane + 64, ... as one pass over all
        // spheres would (chunks are whole multiples of 64)
        for (int t = lane; t < len; t += 64) {
            const int q = q0 + t;
            const double4 o = others[t];
            const double dx = px - o.x, dy = py - o.y, dz = pz - o.z;
            const double rs = rp + o.w;
            // the neighbours whose map changes can touch this sphere's link
            // list, in q order (every lane still active in this pass takes
            // part, so nnb stays the same on them)
            const double rl = rs + LINK_GAP;
            const bool nb = q != p && fabs(dx) < rl && fabs(dy) < rl && fabs(dz) < rl && dist2(dx, dy, dz) < rl * rl;
            const unsigned long long bb = __ballot(nb);
            if (nb) {
                const int at = nnb + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(bb >> 32),
                                                                    __builtin_amdgcn_mbcnt_lo((unsigned)bb, 0u));
                if (at < LINK_NBRS) a.nbrs[(long long)p * LINK_NBRS + at] = q;
            }
            nnb += (int)__popcll(bb);
            if (q == p) continue;
            const double r1 = rs + 1.0;
            if (fabs(dx) < r1 && fabs(dy) < r1 && fabs(dz) < r1 && dist2(dx, dy, dz) < r1 * r1) close = true;
            const double reach = rs + a.threshold_particle;
            if (fabs(dx) >= reach || fabs(dy) >= reach || fabs(dz) >= reach) continue;  // dij >= reach too
            const double dij = sqrt(dist2(dx, dy, dz));
            if (dij >= reach) continue;
            if (dij < rs) {
                stop(a, 1);
                continue;
            }
            const double g = (dij - rp - o.w - a.threshold_particle) / a.threshold_particle;
            const double mag = scale_p * (g * g);
            fs[0] += mag * dx / dij;
            fs[1] += mag * dy / dij;
            fs[2] += mag * dz / dij;
        }
    }
    if (!live) return;
    for (int d = 0; d < 3; ++d) fs[d] = wave_sum(fs[d]);
    const bool alone = __ballot(close) == 0ull;
    if (lane != 0) return;
    a.nnbr[p] = nnb;
    a.changed[(long long)a.par * a.np + p] = 0;  // this step's map update sets it again
    if (p == 0) a.changed[2LL * a.np + a.par] = 0;
    // wall springs summed apart from the particle springs and added last, as
    // the reference keeps Fwxij apart from Fxij (particle_force.F90:192-193)
    const double scale_w = V * (a.rho_solid - a.rho0) * a.gravity / a.stiff_wall;
    const double N[3] = {(double)a.nx, (double)a.ny, (double)a.nz};
    double fw[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < 3; ++d) {
        const double lo = pc[d] - rp - 1.0;
        const double hi = N[d] - pc[d] - rp;
        if (lo < 0.0 || hi < 0.0) stop(a, 2);
        if (lo < a.threshold_wall) {
            const double g = (lo - a.threshold_wall) / a.threshold_wall;
            fw[d] += scale_w * (g * g);
        }
        if (hi < a.threshold_wall) {
            const double g = (hi - a.threshold_wall) / a.threshold_wall;
            fw[d] -= scale_w * (g * g);
        }
    }
    double Fn[3] = {Fl[0] + fs[0] + fw[0], Fl[1] + fs[1] + fw[1],
                    Fl[2] - (a.rho_solid - rho_avg) * V * a.gravity + fs[2] + fw[2]};
    const double mass = a.rho_solid * V, inertia = 0.4 * mass * rp * rp;
    double moved = 0.0;
    for (int d = 0; d < 3; ++d) {
        const double acc = Fn[d] / mass;
        const double Uold = Ul[d];
        a.U_out[3 * p + d] = Uold + acc;
        const double x = pc[d] + Uold + 0.5 * acc;
        a.xc_out[3 * p + d] = x;
        a.Om_out[3 * p + d] = Oml[d] + Tl[d] / inertia;
        a.Ftot[3 * p + d] = Fn[d];
        moved += (x - pc[d]) * (x - pc[d]);
    }
    // k_particle_mark may claim this sphere's new nodes with plain stores if
    // no other sphere was within one node of it and no sphere moves 1/4 node
    // or more (or is not finite): the new balls are then at least half a node
    // apart and share no node
    a.alone[p] = alone ? 1 : 0;
    if (!(moved < 0.0625)) atomicOr(a.err + 3, 1);
}

// Node (i, j, k) (0-based) is in the ball of sphere id with centres xc: the
// walk's own test (the same expression, so the same answer for any node).
__device__ __forceinline__ bool in_ball(const PArgs& a, const double* xc, int id, int i, int j, int k) {
    const double cx = xc[3 * id], cy = xc[3 * id + 1], cz = xc[3 * id + 2], rr = a.r[id] * a.r[id];
    return dist2(i + 1.0 - cx, j + 1.0 - cy, k + 1.0 - cz) <= rr;
}

// The map update changes the owner of node (i, j, k) (0-based), which sphere t
// claims or loses: flag the spheres whose link lists can see the node for a
// rebuild next step -- t and those of t's neighbours (k_particle_update's
// list) whose new ball comes within LIST_REACH of it.  A list reads owners up
// to 2.92 nodes beyond its sphere's nodes (x - e_a of a shell node), so no
// other sphere can see the node: one that did would come within
// r_t + r + 3.25 of t, inside the neighbours' margin.  If t's neighbour list
// overflowed, every list is rebuilt.
constexpr double LIST_REACH = 3.0;
__device__ __forceinline__ void flag_near(const PArgs& a, int t, int i, int j, int k) {
    int* chg = a.changed + (long long)a.par * a.np;
    chg[t] = 1;
    const int nn = a.nnbr[t];
    if (nn > LINK_NBRS) {
        a.changed[2LL * a.np + a.par] = 1;
        return;
    }
    for (int q = 0; q < nn; ++q) {
        const int u = a.nbrs[(long long)t * LINK_NBRS + q];
        const double R = a.r[u] + LIST_REACH;
        if (dist2(i + 1.0 - a.xc_new[3 * u], j + 1.0 - a.xc_new[3 * u + 1], k + 1.0 - a.xc_new[3 * u + 2]) <= R * R)
            chg[u] = 1;
    }
}

// Nodes of sphere id (d^2 <= r^2, initial.F90:78-89), walked column by
// column of its bounding box, four rows of a column at a time: load(m) for
// the four nodes first, so their loads are in flight together, then
// f(i, j, k, m, loaded value) for each.
template <class L, class F>
__device__ __forceinline__ void for_sphere_nodes(const PArgs& a, const double* xc, int id, L&& load, F&& f) {
    using V = decltype(load(0, 0, 0, 0LL));
    int lo[3], hi[3];
    sphere_box(a, xc, id, 1, lo, hi);
    const int bx = hi[0] - lo[0] + 1, by = hi[1] - lo[1] + 1;
    const int ncol = (bx > 0 && by > 0 && hi[2] >= lo[2]) ? bx * by : 0;
    const double cx = xc[3 * id], cy = xc[3 * id + 1], cz = xc[3 * id + 2], rr = a.r[id] * a.r[id];
    for (int col = threadIdx.x; col < ncol; col += blockDim.x) {
        const int i = lo[0] + col % bx, j = lo[1] + col / bx;
        int k0, k1;
        column_rows(cx, cy, cz, a.r[id], i, j, lo[2], hi[2], k0, k1);
        for (int kb = k0; kb <= k1; kb += 4) {
            bool in[4];
            long long m[4];
            V v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int k = kb + u;
                in[u] = k <= k1 && dist2(i + 1.0 - cx, j + 1.0 - cy, k + 1.0 - cz) <= rr;
                m[u] = i + (long long)a.nx * (j + (long long)a.ny * k);
                if (in[u]) v[u] = load(i, j, k, m[u]);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (in[u]) f(i, j, kb + u, m[u], v[u]);
        }
    }
}

// New map from the updated spheres; solids at rest with rho_s.  A newly
// covered node (fluid in the old map) is handled by its first claimer only:
// its rho of this step's streaming (pulled_rho: the fluid kernel's value,
// recomputed from the same pulls, since the fluid kernel writes the rho field
// on the last step of a call only) goes into the chain's correction, then
// rho_s and u = 0 are written.
template <class S>
__global__ __launch_bounds__(256) void k_particle_mark(const PArgs a, int* owner_out) {
    if (stopped(a)) return;
    const int id = blockIdx.x;
    // a node claimed twice goes to the higher index (atomicMax); a sphere that
    // shares no node with another (k_particle_update's test) stores plainly
    const bool plain = a.alone[id] != 0 && *(volatile const int*)(a.err + 3) == 0;
    unsigned long long cov = 0ull;  // this lane's newly covered nodes, then the wave's
    int ncov = 0;
    int* chg = a.changed + (long long)a.par * a.np;  // spheres whose lists see a node this update changes
    for_sphere_nodes(
        a, a.xc_new, id,
        // the old owner; a sphere with no other within a node held every
        // node of its old ball alone, so those need no load
        [&](int i, int j, int k, long long m) { return plain && in_ball(a, a.xc, id, i, j, k) ? id : a.owner[m]; },
        [&](int i, int j, int k, long long m, int own) {
            bool first = true;
            if (plain) owner_out[m] = id;
            else first = atomicMax(owner_out + m, id) < 0;
            if (own != id) {  // the node changes owner
                flag_near(a, id, i, j, k);
                if (own >= 0) chg[own] = 1;  // a node another sphere held (rare: overlapping claims)
            }
            if (own < 0 && first) {  // nodes solid before already hold rho_s, u = 0
                cov += to_fix(pulled_rho<S>(a, i, j, k));
                ++ncov;
                a.rho[m] = a.rho_solid;
                a.u[m] = 0.0;
                a.v[m] = 0.0;
                a.w[m] = 0.0;
            }
        });
    // the neighbour lists' margin holds for spheres that move less than 1/4
    // node per step; if one moved more (k_particle_update's flag), every list
    // is built afresh
    if (threadIdx.x == 0 && *(volatile const int*)(a.err + 3) != 0) a.changed[2LL * a.np + a.par] = 1;
    for (int off = 32; off > 0; off >>= 1) {
        cov += __shfl_down(cov, off, 64);
        ncov += __shfl_down(ncov, off, 64);
    }
    if ((threadIdx.x & 63) == 0 && ncov) {
        atomicAdd(&a.chain->cov, cov);
        atomicAdd(&a.chain->ncov, ncov);
    }
}

// Initial map (from the current state).
__global__ __launch_bounds__(256) void k_particle_init_mark(const PArgs a, int* owner_out) {
    const int id = blockIdx.x;
    for_sphere_nodes(
        a, a.xc, id, [](int, int, int, long long) { return 0; },
        [&](int, int, int, long long m, int) { atomicMax(owner_out + m, id); });
}

// rho = rho_s in solids and rho0 in the fluid, u = 0, and the first
// collision of f = w rho0 on fluid nodes written to dst.
template <class S, int COLL>
__global__ __launch_bounds__(256) void k_particle_init_fields(const PArgs a) {
    using T = typename S::T;
    using R = typename S::Real;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= a.nx || j >= a.ny) return;
    const long long m = i + (long long)a.nx * (j + (long long)a.ny * k);
    const bool solid = a.owner[m] >= 0;
    a.rho[m] = solid ? a.rho_solid : a.rho0;
    a.u[m] = 0.0;
    a.v[m] = 0.0;
    a.w[m] = 0.0;
    if (solid) return;
    R f[Q];
    sfor<Q>([&](auto B) {
        constexpr int bd = decltype(B)::value;
        const double fe = weight(bd) * a.rho0;
        f[bd] = S::SHIFT ? R(fe - weight(bd)) : R(fe);
    });
    Moments<R, S::SHIFT> mo;
    mo.rho = R(a.rho0);
    mo.drho = R(a.rho0 - 1.0);
    mo.u = mo.v = mo.w = R(0);
    collide<S, COLL>(f, mo, a);
    const int c = i + a.px * (j + a.ny * k);
    T* __restrict__ dst = static_cast<T*>(a.dst);
    sfor<Q>([&](auto B) {
        constexpr int bd = decltype(B)::value;
        S::st(dst + (long long)bd * a.ps + c, f[bd], a.scale, 0u);
    });
}

// Refill of one uncovered node (i, j, k) = m of sphere id: inside the
// sphere's old position, fluid in the new map.
template <class S, int COLL>
__device__ __forceinline__ void refill_node(const PArgs& a, int i, int j, int k, long long m, int id) {
    using T = typename S::T;
    using R = typename S::Real;
    const T* src = static_cast<const T*>(a.src);
    const double rho_avg = *a.rho_avg;
    // extrapolation direction: most outward lattice direction w.r.t. the new centre
    const double rx = i + 1.0 - a.xc_new[3 * id], ry = j + 1.0 - a.xc_new[3 * id + 1],
                 rz = k + 1.0 - a.xc_new[3 * id + 2];
    const double rn = sqrt(dist2(rx, ry, rz));
    double best = 0.0;
    int ec = 0;
    for (int al = 1; al < Q; ++al) {
        const double tn = __dadd_rn(__dadd_rn(rx * EX[al], ry * EY[al]), rz * EZ[al]) / rn;
        if (tn > best) {
            best = tn;
            ec = al;
        }
    }
    if (ec == 0) {
        stop(a, 4);
        return;
    }
    // the stencil points x + s e_c, s = 1, 2, 3: the streamed populations of
    // those fluid in both maps, the rest equilibrium w rho0 for the others.
    // A refill is one lane's chain of gathers, so every load is issued (at a
    // point clamped into the box) before any is used.
    int pi[3], pj[3], pk[3];
    bool in[3];
    int own[3][2];
    for (int s = 0; s < 3; ++s) {
        pi[s] = i + (s + 1) * EX[ec];
        pj[s] = j + (s + 1) * EY[ec];
        pk[s] = k + (s + 1) * EZ[ec];
        in[s] = in_box(a, pi[s], pj[s], pk[s]);
        pi[s] = min(max(pi[s], 0), a.nx - 1);
        pj[s] = min(max(pj[s], 0), a.ny - 1);
        pk[s] = min(max(pk[s], 0), a.nz - 1);
        const long long pm = pi[s] + (long long)a.nx * (pj[s] + (long long)a.ny * pk[s]);
        own[s][0] = a.owner[pm];
        own[s][1] = a.owner_new[pm];
    }
    double v[3][Q];
    for (int s = 0; s < 3; ++s)
        sfor<Q>([&](auto B) {
            constexpr int al = decltype(B)::value;
            v[s][al] = streamed_pop_at<S, al>(a, src, pi[s], pj[s], pk[s]);
        });
    double fr[Q];
    for (int s = 0; s < 3; ++s) {
        const bool ok = in[s] && own[s][0] < 0 && own[s][1] < 0;
        const double coef = s == 0 ? 3.0 : (s == 1 ? -3.0 : 1.0);
        for (int al = 0; al < Q; ++al) {
            const double x = ok ? v[s][al] : weight(al) * a.rho0;
            fr[al] = s == 0 ? coef * x : fr[al] + coef * x;
        }
    }
    // momentum moments replaced by rhoAvg times the body velocity
    double mm[Q];
    mrt_forward(fr, mm);
    double ux, uy, uz;
    body_velocity(a.xc_new, a.U_new, a.Om_new, id, i + 1.0, j + 1.0, k + 1.0, ux, uy, uz);
    mm[3] = rho_avg * ux;
    mm[5] = rho_avg * uy;
    mm[7] = rho_avg * uz;
    mrt_inverse(mm, fr);
    R f[Q];
    for (int al = 0; al < Q; ++al) f[al] = S::SHIFT ? R(fr[al] - weight(al)) : R(fr[al]);
    const Moments<R, S::SHIFT> mo = moments<R, S::SHIFT>(f);
    const double rho_new = S::SHIFT ? 1.0 + double(mo.drho) : double(mo.rho);
    atomicAdd(&a.chain->ref, to_fix(rho_new - a.rho_solid));
    a.rho[m] = rho_new;
    a.u[m] = double(mo.u);
    a.v[m] = double(mo.v);
    a.w[m] = double(mo.w);
    collide<S, COLL>(f, mo, a);
    const long long c = i + (long long)a.px * (j + (long long)a.ny * k);
    T* __restrict__ dst = static_cast<T*>(a.dst);
    for (int al = 0; al < Q; ++al) S::st(dst + (long long)al * a.ps + c, f[al], a.scale, 0u);
}

// The uncovered nodes of each sphere, appended to the refill list.  The
// list's order does not matter: a node's refill reads only the previous step's
// populations, the maps and the sphere states, and writes only its own cell.
// Every block also sums its slice of the fluid kernel's per-block sums
// (block order within the slice) into cpart2, for k_fluid_chain; the grid has
// at least SUM_BLOCKS blocks so that the slices stay short.
__global__ __launch_bounds__(RED_THREADS) void k_refill_scan(const PArgs a) {
    if (stopped(a)) return;
    const int id = blockIdx.x;
    if (id < a.np)
        for_sphere_nodes(
            a, a.xc, id,
            // a node of the new ball was claimed by the map update: not
            // uncovered, no load (only the trailing crescent is checked)
            [&](int i, int j, int k, long long m) {
                return in_ball(a, a.xc_new, id, i, j, k) ? int2{id, id} : int2{a.owner[m], a.owner_new[m]};
            },
            [&](int i, int j, int k, long long m, int2 own) {
                if (!(own.x == id && own.y < 0)) return;
                flag_near(a, id, i, j, k);  // the sphere lost the node
                const int slot = atomicAdd(a.refill_count, 1);
                if (slot < a.refill_cap) {
                    a.refill[2 * slot] = (int)m;
                    a.refill[2 * slot + 1] = id;
                } else {
                    stop(a, 6);
                }
            });
    const long long lo = a.ncb * blockIdx.x / gridDim.x, hi = a.ncb * (blockIdx.x + 1) / gridDim.x;
    const double2* cp = reinterpret_cast<const double2*>(a.cpart);
    double s = 0.0, cnt = 0.0;
    for (long long b = lo + threadIdx.x; b < hi; b += RED_THREADS) {
        const double2 v = cp[b];
        s += v.x;
        cnt += v.y;
    }
    block_pair_sum(s, cnt);
    if (threadIdx.x == 0) {
        a.cpart2[2 * blockIdx.x] = s;
        a.cpart2[2 * blockIdx.x + 1] = cnt;
    }
}

template <class S, int COLL>
__global__ __launch_bounds__(256) void k_refill_proc(const PArgs a) {
    if (stopped(a)) return;
    const int count = min(*a.refill_count, a.refill_cap);
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < count; t += gridDim.x * blockDim.x) {
        const long long m = a.refill[2 * t];
        const int id = a.refill[2 * t + 1];
        const int i = (int)(m % a.nx), j = (int)((m / a.nx) % a.ny), k = (int)(m / ((long long)a.nx * a.ny));
        refill_node<S, COLL>(a, i, j, k, m, id);
    }
}

// Sum of rho and count over fluid nodes, per block.
__global__ __launch_bounds__(RED_THREADS) void k_fluid_sum(const double* __restrict__ rho, const int* __restrict__ owner,
                                                           long long n, double* __restrict__ partial) {
    // four nodes per thread and pass: one 16-byte owner load, two 16-byte
    // rho loads (both arrays start on hipMalloc's alignment); the n % 4 last
    // nodes go to the first threads of block 0.  Fixed grid and order, so the
    // sums are the same on every run.
    double s = 0.0, cnt = 0.0;
    const long long groups = n / 4;
    const int4* o4 = reinterpret_cast<const int4*>(owner);
    const double2* r2 = reinterpret_cast<const double2*>(rho);
    for (long long g = blockIdx.x * (long long)RED_THREADS + threadIdx.x; g < groups;
         g += (long long)gridDim.x * RED_THREADS) {
        const int4 o = o4[g];
        const double2 ra = r2[2 * g], rb = r2[2 * g + 1];
        if (o.x < 0) { s += ra.x; cnt += 1.0; }
        if (o.y < 0) { s += ra.y; cnt += 1.0; }
        if (o.z < 0) { s += rb.x; cnt += 1.0; }
        if (o.w < 0) { s += rb.y; cnt += 1.0; }
    }
    if (blockIdx.x == 0 && threadIdx.x < (unsigned)(n - 4 * groups)) {
        const long long c = 4 * groups + threadIdx.x;
        if (owner[c] < 0) { s += rho[c]; cnt += 1.0; }
    }
    block_pair_sum(s, cnt);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = s;
        partial[2 * blockIdx.x + 1] = cnt;
    }
}

// The first step's rhoAvg from a full pass (k_fluid_sum's partials, summed
// in block order), which starts the chain.  One block.
__global__ __launch_bounds__(RED_THREADS) void k_fluid_first(const double* __restrict__ partial, int nblocks,
                                                             MeanChain* chain) {
    double s = 0.0, cnt = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += RED_THREADS) {
        s += partial[2 * b];
        cnt += partial[2 * b + 1];
    }
    block_pair_sum(s, cnt);
    if (threadIdx.x == 0) {
        chain->sum = s;
        chain->count = cnt;
        chain->cov = chain->ref = 0ull;
        chain->ncov = 0;
    }
}

// After the map update and the uncovered-node list: the chain's sum and
// count over the new map's fluid nodes from the fluid kernel's per-block sums
// (k_refill_scan's slice sums, in order) and the corrections; rhoAvg of the refill into *out; the
// corrections cleared for the next step; the count of completed steps moves
// on and the step's "a sphere moved 1/4 node" flag is cleared.  One block.
__global__ __launch_bounds__(RED_THREADS) void k_fluid_chain(const PArgs a, int nslices, double* out) {
    if (stopped(a)) return;
    double s = 0.0, cnt = 0.0;
    for (int b = threadIdx.x; b < nslices; b += RED_THREADS) {
        s += a.cpart2[2 * b];
        cnt += a.cpart2[2 * b + 1];
    }
    block_pair_sum(s, cnt);
    if (threadIdx.x == 0) {
        MeanChain* ch = a.chain;
        const int unc = min(*a.refill_count, a.refill_cap);
        const double sum = s - (double)(long long)ch->cov / FIX_ONE + unc * a.rho_solid;
        const double count = cnt - ch->ncov + unc;
        ch->sum = sum;
        ch->count = count;
        ch->cov = ch->ref = 0ull;
        ch->ncov = 0;
        *out = sum / count;
        ++a.err[2];
        a.err[3] = 0;
    }
}


__global__ void k_fill_int(int* __restrict__ p, long long n, int v) {
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x)
        p[c] = v;
}

dim3 cell_grid(const lbm_ctx* c) {
    return dim3((c->p.nx + c->block.x - 1) / c->block.x, (c->p.ny + c->block.y - 1) / c->block.y, c->p.nz);
}

long long cell_blocks(const lbm_ctx* c) {
    const dim3 g = cell_grid(c);
    return (long long)g.x * g.y * g.z;
}

PArgs particle_args(const lbm_ctx* c) {
    const ParticleState* s = c->part;
    PArgs a{};
    a.nx = c->p.nx;
    a.ny = c->p.ny;
    a.nz = c->p.nz;
    a.px = c->px;
    a.pxy = c->pxy;
    a.ps = c->ps;
    a.snu = 1.0 / c->p.tau;
    a.sq = 8.0 * (2.0 * c->p.tau - 1.0) / (8.0 * c->p.tau - 1.0);
    a.scale = c->scale;
    a.inv_scale = c->inv_scale;
    a.meq12_rho = c->p.meq12_rho;
    a.rho0 = c->p.rho0;
    a.rho_solid = s->pp.rho_solid;
    a.owner = s->owner[s->cur];
    a.owner_new = s->owner[1 - s->cur];
    a.xc = s->xc[s->cur];
    a.U = s->U[s->cur];
    a.Om = s->Om[s->cur];
    a.xc_new = a.xc_out = s->xc[1 - s->cur];
    a.U_new = a.U_out = s->U[1 - s->cur];
    a.Om_new = a.Om_out = s->Om[1 - s->cur];
    a.r = s->r;
    a.rho_avg = s->scal;
    a.rho = c->macro;
    a.u = c->macro + c->n;
    a.v = c->macro + 2 * c->n;
    a.w = c->macro + 3 * c->n;
    a.err = s->err;
    a.alone = s->alone;
    a.cpart = s->cpart;
    a.cpart2 = s->cpart2;
    a.ncb = cell_blocks(c);
    a.chain = s->chain;
    a.np = s->np;
    a.gravity = s->pp.gravity;
    a.threshold_wall = s->pp.threshold_wall;
    a.stiff_wall = s->pp.stiff_wall;
    a.threshold_particle = s->pp.threshold_particle;
    a.stiff_particle = s->pp.stiff_particle;
    a.F = s->F;
    a.T = s->T;
    a.Ftot = s->Ftot;
    a.items = s->items;
    a.item_fluid1 = s->item_fluid1;
    a.slab_off = s->slab_off;
    a.slab_cap = s->slab_cap;
    a.link_count = s->link_count;
    a.link_lo = s->link_lo;
    a.changed = s->changed;
    a.nbrs = s->nbrs;
    a.nnbr = s->nnbr;
    a.par = s->cur;
    a.refill = s->refill;
    a.refill_count = s->refill_count;
    a.refill_cap = s->refill_cap;
    return a;
}


__global__ void k_set_int(int* p, int v) { *p = v; }


// The launches of one step on stream st; both parities flip.  The first step
// after the spheres are attached starts the rhoAvg chain with a full pass
// over the map and the fluid density.
template <class S, int COLL>
void particle_launches(lbm_ctx* c, hipStream_t st, bool macro) {
    ParticleState* s = c->part;
    const int np = s->np;
    if (!s->chain_started) {
        hipLaunchKernelGGL(k_fluid_sum, dim3(FLUID_BLOCKS), dim3(RED_THREADS), 0, st, c->macro, s->owner[s->cur], c->n,
                           s->red);
        hipLaunchKernelGGL(k_fluid_first, dim3(1), dim3(RED_THREADS), 0, st, s->red, FLUID_BLOCKS, s->chain);
        s->chain_started = true;
    }
    PArgs a = particle_args(c);
    a.src = c->buf[c->cur];
    a.src_w = c->buf[c->cur];
    a.dst = c->buf[1 - c->cur];
    a.step_off = 1;  // before k_fluid_chain: the step under way is completed + 1
    hipLaunchKernelGGL(k_links<S>, dim3(np), dim3(SCAN_THREADS), 0, st, a);
    if (macro)
        hipLaunchKernelGGL((k_particle_step<S, COLL, true>), cell_grid(c), c->block, 0, st, a);
    else
        hipLaunchKernelGGL((k_particle_step<S, COLL, false>), cell_grid(c), c->block, 0, st, a);
    hipLaunchKernelGGL(k_particle_update, dim3((np + UPDATE_WAVES - 1) / UPDATE_WAVES), dim3(64 * UPDATE_WAVES), 0, st,
                       a);
    hipLaunchKernelGGL(k_fill_int, dim3(2048), dim3(256), 0, st, s->owner[1 - s->cur], c->n, -1);
    hipLaunchKernelGGL(k_particle_mark<S>, dim3(np), dim3(256), 0, st, a, s->owner[1 - s->cur]);
    hipLaunchKernelGGL(k_refill_scan, dim3(scan_blocks(np)), dim3(RED_THREADS), 0, st, a);
    hipLaunchKernelGGL(k_fluid_chain, dim3(1), dim3(RED_THREADS), 0, st, a, scan_blocks(np), s->scal);
    a.step_off = 0;
    hipLaunchKernelGGL((k_refill_proc<S, COLL>), dim3(REFILL_BLOCKS), dim3(256), 0, st, a);
    s->cur = 1 - s->cur;
    c->cur = 1 - c->cur;
}

void launch_step(lbm_ctx* c, hipStream_t st, bool macro) {
    const bool bgk = c->p.collision == LBM_BGK;
    if (c->p.storage == LBM_STORE_F32) {
        if (bgk) particle_launches<StoreF32, COLL_BGK>(c, st, macro);
        else particle_launches<StoreF32, COLL_MRT>(c, st, macro);
    } else {
        if (bgk) particle_launches<StoreF64, COLL_BGK>(c, st, macro);
        else particle_launches<StoreF64, COLL_MRT>(c, st, macro);
    }
}

// GRAPH_STEPS steps without macro output as one graph (captured on first use
// for the current parities), launched on the context's stream.  The kernels
// take only pointers fixed for the life of the particle state, so a graph
// stays valid until the spheres are attached again.
int replay_graph(lbm_ctx* c) {
    ParticleState* s = c->part;
    const int key = 2 * c->cur + s->cur;
    if (!s->graph[key]) {
        if (!s->cap_stream) HIP_TRY(hipStreamCreateWithFlags(&s->cap_stream, hipStreamNonBlocking));
        HIP_TRY(hipStreamBeginCapture(s->cap_stream, hipStreamCaptureModeThreadLocal));
        for (int q = 0; q < GRAPH_STEPS; ++q) launch_step(c, s->cap_stream, false);  // even: parities restored
        hipGraph_t g = nullptr;
        HIP_TRY(hipStreamEndCapture(s->cap_stream, &g));
        hipError_t e = hipGraphInstantiate(&s->graph[key], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            s->graph[key] = nullptr;
            return lbm_fail(LBM_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
        }
    }
    HIP_TRY(hipGraphLaunch(s->graph[key], c->stream));
    return LBM_OK;
}

void free_state(ParticleState* s) {
    for (hipGraphExec_t& g : s->graph)
        if (g) (void)hipGraphExecDestroy(g);
    if (s->cap_stream) (void)hipStreamDestroy(s->cap_stream);
    for (int q = 0; q < 2; ++q) {
        if (s->xc[q]) (void)hipFree(s->xc[q]);
        if (s->U[q]) (void)hipFree(s->U[q]);
        if (s->Om[q]) (void)hipFree(s->Om[q]);
        if (s->owner[q]) (void)hipFree(s->owner[q]);
    }
    for (void* p : {(void*)s->r, (void*)s->F, (void*)s->T, (void*)s->Ftot, (void*)s->items, (void*)s->item_fluid1, (void*)s->slab_off,
                    (void*)s->slab_cap, (void*)s->link_count, (void*)s->link_lo, (void*)s->changed, (void*)s->nbrs,
                    (void*)s->nnbr, (void*)s->scal,
                    (void*)s->red, (void*)s->err, (void*)s->alone, (void*)s->cpart, (void*)s->cpart2, (void*)s->chain, (void*)s->refill, (void*)s->refill_count})
        if (p) (void)hipFree(p);
}

}  // namespace

void particles_release(lbm_ctx* c) {
    if (!c->part) return;
    free_state(c->part);
    delete c->part;
    c->part = nullptr;
}

int particles_step(lbm_ctx* c, int nsteps) {
    ParticleState* s = c->part;
    if (s->stopped) return lbm_fail(LBM_ERR_STATE, "lbm_step: the particle run has stopped; attach the spheres again");
    // the count of completed steps (stop records) restarts from the context's own
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, c->stream, s->err + 2, (int)c->itc);
    // graph replays for whole blocks of the steps before the last, direct
    // launches for the rest; the last step writes the macro fields
    int q = 0;
    if (!s->chain_started && nsteps > 0) {  // the full pass that starts the rhoAvg chain stays out of the graphs
        launch_step(c, c->stream, nsteps == 1);
        q = 1;
    }
    for (; nsteps - 1 - q >= GRAPH_STEPS; q += GRAPH_STEPS) {
        int rc = replay_graph(c);
        if (rc) return rc;
    }
    for (; q < nsteps; ++q) launch_step(c, c->stream, q == nsteps - 1);
    int rc = lbm_check_launch();
    if (rc) return rc;
    int err[2] = {0, 0};
    HIP_TRY(hipMemcpyAsync(err, s->err, sizeof err, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (err[0] == 0) return LBM_OK;
    s->stopped = true;
    static const char* const why[] = {
        "", "particle-particle interpenetration (particle_force.F90:117-124)",
        "a particle penetrates a wall (particle_force.F90:136-186)",
        "wall distance q outside [0, 1] (particle_bounceback.F90:128-138)",
        "no extrapolation direction for an uncovered node (non-finite sphere state)",
        "boundary-link list of a sphere overflowed", "list of uncovered nodes overflowed"};
    return lbm_fail(LBM_ERR_PARTICLE, "particle run stopped at step %d: %s", err[1],
                    err[0] >= 1 && err[0] <= 6 ? why[err[0]] : "unknown stop code");
}

extern "C" {

int lbm_particles_init(lbm_ctx* c, const lbm_particle_params* pp, const double* centers, const double* radii,
                       const double* U, const double* omega) {
    if (!c || !pp || !centers || !radii) return lbm_fail(LBM_ERR_ARG, "lbm_particles_init: null argument");
    if (pp->np < 1) return lbm_fail(LBM_ERR_ARG, "lbm_particles_init: np = %d", pp->np);
    if (c->p.thermal || c->p.halo_faces || c->p.halo_loopback || c->p.u_lid != 0.0 ||
        (c->p.storage != LBM_STORE_F64 && c->p.storage != LBM_STORE_F32))
        return lbm_fail(LBM_ERR_UNSUPPORTED, "lbm_particles_init: needs an isothermal FP64/FP32 context with walls "
                                             "at rest and no halo");
    if (!(pp->stiff_wall > 0.0) || !(pp->stiff_particle > 0.0) || !(pp->threshold_wall > 0.0) ||
        !(pp->threshold_particle > 0.0))
        return lbm_fail(LBM_ERR_ARG, "lbm_particles_init: spring thresholds and stiffnesses must be positive");
    // a bounding box must stay below 2^27 nodes (link list entries pack the
    // node offset into 27 bits)
    for (int q = 0; q < pp->np; ++q) {
        if (!(radii[q] > 0.0 && radii[q] <= 250.0))
            return lbm_fail(LBM_ERR_ARG, "lbm_particles_init: radius %d = %g outside (0, 250]", q, radii[q]);
        for (int d = 0; d < 3; ++d)
            if (!std::isfinite(centers[3 * q + d]))
                return lbm_fail(LBM_ERR_ARG, "lbm_particles_init: centre %d is not finite", q);
    }
    particles_release(c);
    ParticleState* s = new (std::nothrow) ParticleState();
    if (!s) return lbm_fail(LBM_ERR_NOMEM, "lbm_particles_init: host allocation failed");
    s->np = pp->np;
    s->pp = *pp;
    const size_t v3 = sizeof(double) * 3 * pp->np;
    hipError_t e = hipSuccess;
    for (int q = 0; q < 2; ++q) {
        if (e == hipSuccess) e = hipMalloc(&s->xc[q], v3);
        if (e == hipSuccess) e = hipMalloc(&s->U[q], v3);
        if (e == hipSuccess) e = hipMalloc(&s->Om[q], v3);
        if (e == hipSuccess) e = hipMalloc(&s->owner[q], sizeof(int) * c->n);
    }
    if (e == hipSuccess) e = hipMalloc(&s->r, sizeof(double) * pp->np);
    if (e == hipSuccess) e = hipMalloc(&s->F, v3);
    if (e == hipSuccess) e = hipMalloc(&s->T, v3);
    if (e == hipSuccess) e = hipMalloc(&s->Ftot, v3);
    // link slabs: a sphere of radius r has at most 18 links per node of its
    // shell r < d <= r + 1.5, and those nodes' unit cells lie inside the
    // shell r - 0.87 < d <= r + 2.37, so its volume bounds their number
    std::vector<long long> off(pp->np);
    std::vector<int> cap(pp->np);
    long long total_cap = 0;
    for (int q = 0; q < pp->np; ++q) {
        const double ro = radii[q] + 2.4, ri = radii[q] > 0.9 ? radii[q] - 0.9 : 0.0;
        const double nodes = 4.0 / 3.0 * 3.14159265358979323846 * (ro * ro * ro - ri * ri * ri);
        const long long c18 = 18 * ((long long)nodes + 1);
        cap[q] = (int)((c18 + 63) / 64 * 64);
        off[q] = total_cap;
        total_cap += cap[q];
    }
    if (e == hipSuccess) e = hipMalloc(&s->items, sizeof(unsigned) * (size_t)total_cap);
    if (e == hipSuccess) e = hipMalloc(&s->item_fluid1, (size_t)total_cap);
    if (e == hipSuccess) e = hipMalloc(&s->slab_off, sizeof(long long) * pp->np);
    if (e == hipSuccess) e = hipMalloc(&s->slab_cap, sizeof(int) * pp->np);
    if (e == hipSuccess) e = hipMemcpy(s->slab_off, off.data(), sizeof(long long) * pp->np, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->slab_cap, cap.data(), sizeof(int) * pp->np, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&s->link_count, sizeof(int) * pp->np);
    if (e == hipSuccess) e = hipMalloc(&s->link_lo, sizeof(int) * 3 * pp->np);
    if (e == hipSuccess) e = hipMalloc(&s->changed, sizeof(int) * (2 * (size_t)pp->np + 2));
    if (e == hipSuccess) e = hipMalloc(&s->nbrs, sizeof(int) * LINK_NBRS * (size_t)pp->np);
    if (e == hipSuccess) e = hipMalloc(&s->nnbr, sizeof(int) * pp->np);
    // every list is built on the first step: all spheres changed, too many neighbours
    if (e == hipSuccess) e = hipMemset(s->changed, 0x01, sizeof(int) * (2 * (size_t)pp->np + 2));
    if (e == hipSuccess) e = hipMemset(s->nnbr, 0x7f, sizeof(int) * pp->np);
    if (e == hipSuccess) e = hipMalloc(&s->scal, sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&s->red, sizeof(double) * 2 * FLUID_BLOCKS);
    if (e == hipSuccess) e = hipMalloc(&s->err, 4 * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&s->alone, sizeof(int) * pp->np);
    if (e == hipSuccess) e = hipMalloc(&s->cpart, sizeof(double) * 2 * (size_t)cell_blocks(c));
    if (e == hipSuccess) e = hipMalloc(&s->cpart2, sizeof(double) * 2 * (size_t)scan_blocks(pp->np));
    if (e == hipSuccess) e = hipMalloc(&s->chain, sizeof(MeanChain));
    if (e == hipSuccess) e = hipMemset(s->chain, 0, sizeof(MeanChain));

    // uncovered nodes lie inside the old spheres: at most the nodes of all balls
    {
        double cap = 64.0;
        for (int q = 0; q < pp->np; ++q) {
            const double rb = radii[q] + 1.0;
            cap += 4.0 / 3.0 * PI * rb * rb * rb;
        }
        s->refill_cap = (int)std::min(cap, 1.0e9);
    }
    if (e == hipSuccess) e = hipMalloc(&s->refill, sizeof(int) * 2 * (size_t)s->refill_cap);
    if (e == hipSuccess) e = hipMalloc(&s->refill_count, sizeof(int));
    if (e == hipSuccess) e = hipMemcpy(s->xc[0], centers, v3, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->r, radii, sizeof(double) * pp->np, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = U ? hipMemcpy(s->U[0], U, v3, hipMemcpyHostToDevice) : hipMemset(s->U[0], 0, v3);
    if (e == hipSuccess)
        e = omega ? hipMemcpy(s->Om[0], omega, v3, hipMemcpyHostToDevice) : hipMemset(s->Om[0], 0, v3);
    if (e == hipSuccess) e = hipMemset(s->F, 0, v3);
    if (e == hipSuccess) e = hipMemset(s->T, 0, v3);
    if (e == hipSuccess) e = hipMemset(s->Ftot, 0, v3);
    if (e == hipSuccess) e = hipMemset(s->err, 0, 4 * sizeof(int));
    // the copies and fills above complete before any launch on the context's
    // stream (which need not synchronise with the null stream)
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        free_state(s);
        delete s;
        (void)hipGetLastError();
        return lbm_fail(LBM_ERR_NOMEM, "lbm_particles_init: %s", hipGetErrorString(e));
    }
    c->part = s;
    s->cur = 0;
    c->cur = 0;
    PArgs a = particle_args(c);
    hipLaunchKernelGGL(k_fill_int, dim3(2048), dim3(256), 0, c->stream, s->owner[0], c->n, -1);
    hipLaunchKernelGGL(k_particle_init_mark, dim3(pp->np), dim3(256), 0, c->stream, a, s->owner[0]);
    a.dst = c->buf[0];
    if (c->p.storage == LBM_STORE_F32) {
        if (c->p.collision == LBM_BGK)
            hipLaunchKernelGGL((k_particle_init_fields<StoreF32, COLL_BGK>), cell_grid(c), c->block, 0, c->stream, a);
        else
            hipLaunchKernelGGL((k_particle_init_fields<StoreF32, COLL_MRT>), cell_grid(c), c->block, 0, c->stream, a);
    } else {
        if (c->p.collision == LBM_BGK)
            hipLaunchKernelGGL((k_particle_init_fields<StoreF64, COLL_BGK>), cell_grid(c), c->block, 0, c->stream, a);
        else
            hipLaunchKernelGGL((k_particle_init_fields<StoreF64, COLL_MRT>), cell_grid(c), c->block, 0, c->stream, a);
    }
    // residual history up = vp = 0: 2 n doubles of zero bits
    hipLaunchKernelGGL(k_fill_int, dim3(2048), dim3(256), 0, c->stream, reinterpret_cast<int*>(c->prev), 4 * c->n, 0);
    int rc = lbm_check_launch();
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->itc = 0;
    c->initialised = true;
    return LBM_OK;
}

int lbm_particles_get(lbm_ctx* c, double* centers, double* U, double* omega, double* F, double* T) {
    if (!c || !c->part) return lbm_fail(LBM_ERR_STATE, "lbm_particles_get: no particles attached");
    ParticleState* s = c->part;
    const size_t v3 = sizeof(double) * 3 * s->np;
    if (centers) HIP_TRY(hipMemcpyAsync(centers, s->xc[s->cur], v3, hipMemcpyDeviceToHost, c->stream));
    if (U) HIP_TRY(hipMemcpyAsync(U, s->U[s->cur], v3, hipMemcpyDeviceToHost, c->stream));
    if (omega) HIP_TRY(hipMemcpyAsync(omega, s->Om[s->cur], v3, hipMemcpyDeviceToHost, c->stream));
    if (F) HIP_TRY(hipMemcpyAsync(F, s->Ftot, v3, hipMemcpyDeviceToHost, c->stream));
    if (T) HIP_TRY(hipMemcpyAsync(T, s->T, v3, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

int lbm_particles_map(lbm_ctx* c, int* owner) {
    if (!c || !c->part || !owner) return lbm_fail(LBM_ERR_STATE, "lbm_particles_map: no particles attached");
    HIP_TRY(hipMemcpyAsync(owner, c->part->owner[c->part->cur], sizeof(int) * c->n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return LBM_OK;
}

}  // extern "C"
