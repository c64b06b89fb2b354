// probe.hip -- the attainable HBM bandwidth of the step's memory access
// pattern, measured on the device the library runs on, so that a step's
// bytes per second can be read against what the memory system delivers for
// the same traffic rather than only against the data-sheet peak.
//
//   pattern 0: one-stream copy of the bytes of 19 * n^3 elements (one step
//              of an n^3 lattice), one 8-byte word per thread
//   pattern 1: the D3Q19 pull of the step kernel with no arithmetic: each
//              thread reads the 19 populations of its cell from their
//              upstream neighbours (population planes laid out as in
//              lbm_create: x padded to 16, whole 4 KiB pages plus a 1 KiB
//              stagger per plane) and writes them to its own cell with
//              nontemporal stores, as the step kernel does; the best of four
//              occupancies (all waves, 16, 12, 8 waves per CU)
//   pattern 2: the buoyancy step's traffic (FP64 only): the D3Q19 pull, the
//              D3Q7 pull of the temperature populations and one force plane
//              read and rewritten in place (27 planes in, 27 out)
#include <hip/hip_runtime.h>

#include <cstdint>

#include "collide.hpp"
#include "context.hpp"

using namespace lbm;

namespace {

__global__ __launch_bounds__(256) void k_probe_copy(const uint64_t* __restrict__ src, uint64_t* __restrict__ dst,
                                                    long long n) {
    const long long t = blockIdx.x * 256LL + threadIdx.x;
    if (t < n) dst[t] = src[t];
}

// src and dst point `pad` elements into their allocations; pad exceeds the
// largest pull offset, so every read stays inside the allocation.  Planes
// [0, 19) are pulled as D3Q19, the next Q7P planes as D3Q7 (the first seven
// velocities), the last PLAIN planes are read and written at the cell itself.
template <class T, int Q7P, int PLAIN>
__global__ __launch_bounds__(256) void k_probe_pull(const T* __restrict__ src, T* __restrict__ dst, int nx, int ny,
                                                    int px, long long pxy, long long ps) {
    constexpr int NP = Q + Q7P + PLAIN;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= nx || j >= ny) return;
    const long long c = i + (long long)px * j + pxy * k;
    T v[NP];
    sfor<NP>([&](auto A) {
        constexpr int p = decltype(A)::value;
        constexpr int al = p < Q ? p : (p < Q + Q7P ? p - Q : 0);
        const long long off = (long long)EX[al] + (long long)EY[al] * px + (long long)EZ[al] * pxy;
        v[p] = src[p * ps + c - off];
    });
    // nontemporal stores, as the step kernels store their populations
    sfor<NP>([&](auto A) {
        constexpr int p = decltype(A)::value;
        __builtin_nontemporal_store(v[p], dst + p * ps + c);
    });
}

template <class T>
int probe(int pattern, int n, int reps, double* gbps) {
    const int planes = pattern == 2 ? Q + 7 + 1 : Q;
    const int px = (n + 15) / 16 * 16;
    const long long pxy = (long long)px * n, cells = pxy * n;
    const long long page = 4096 / (long long)sizeof(T);
    const long long ps = (cells + page - 1) / page * page + 1024 / (long long)sizeof(T);
    const long long pad = (pxy + px + 1 + page - 1) / page * page;
    const long long elems = planes * ps + 2 * pad;
    const long long n3 = (long long)n * n * n;
    T* a = nullptr;
    T* b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = LBM_OK;
    hipError_t e = hipMalloc(&a, elems * sizeof(T));
    if (e == hipSuccess) e = hipMalloc(&b, elems * sizeof(T));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        rc = lbm_fail(LBM_ERR_NOMEM, "lbm_bandwidth_probe: hipMalloc of 2 x %lld bytes failed",
                      elems * (long long)sizeof(T));
    }
    if (!rc && (hipMemset(a, 0, elems * sizeof(T)) != hipSuccess || hipMemset(b, 0, elems * sizeof(T)) != hipSuccess ||
                hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = lbm_fail(LBM_ERR_HIP, "lbm_bandwidth_probe: set-up failed");
    if (!rc) {
        const int bx = n >= 256 ? 256 : (n + 63) / 64 * 64;
        const dim3 block(bx, 256 / bx, 1);
        const dim3 grid((n + block.x - 1) / block.x, (n + block.y - 1) / block.y, n);
        // pattern 0: the step's bytes in whole 8-byte words (the allocation
        // holds Q * ps >= Q * n3 elements plus 2 * pad, so rounding up fits)
        const long long words = ((long long)Q * n3 * (long long)sizeof(T) + 7) / 8;
        auto launch = [&](int r, unsigned lds) {
            const T* src = (r & 1) ? b : a;
            T* dst = (r & 1) ? a : b;
            if (pattern == 0)
                hipLaunchKernelGGL(k_probe_copy, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, s,
                                   (const uint64_t*)src, (uint64_t*)dst, words);
            else if (pattern == 1)
                hipLaunchKernelGGL((k_probe_pull<T, 0, 0>), grid, block, lds, s, src + pad, dst + pad, n, n, px,
                                   pxy, ps);
            else
                hipLaunchKernelGGL((k_probe_pull<T, 7, 1>), grid, block, lds, s, src + pad, dst + pad, n, n, px,
                                   pxy, ps);
        };
        // the pulls are timed with all the waves the registers allow and with
        // 16, 12 and 8 waves per CU (an LDS reservation per block of 4
        // waves): the best of them is what the pattern can deliver
        unsigned caps[4] = {0u, 0u, 0u, 0u};
        int ncap = 1;
        int lds_cu = 0, lds_block = 0, dev = 0;
        (void)hipGetDevice(&dev);
        if (pattern != 0 &&
            hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) == hipSuccess &&
            hipDeviceGetAttribute(&lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess) {
            for (int blocks : {4, 3, 2}) {
                const int reserve = lds_cu / blocks;
                if (reserve > 0 && reserve <= lds_block) caps[ncap++] = (unsigned)reserve;
            }
        }
        (void)hipGetLastError();
        double best = 0.0;
        for (int q = 0; q < ncap && !rc; ++q) {
            for (int r = 0; r < 3; ++r) launch(r, caps[q]);
            (void)hipEventRecord(e0, s);
            for (int r = 0; r < reps; ++r) launch(r, caps[q]);
            (void)hipEventRecord(e1, s);
            float ms = 0.0f;
            if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess || !(ms > 0.0f)) {
                rc = lbm_fail(LBM_ERR_HIP, "lbm_bandwidth_probe: timing failed");
            } else {
                const double g = (pattern == 0 ? 16.0 * (double)words
                                               : 2.0 * planes * (double)sizeof(T) * (double)n3) /
                                 (ms * 1e-3 / reps) / 1e9;
                best = g > best ? g : best;
            }
        }
        if (!rc) *gbps = best;
    }
    if (e1) (void)hipEventDestroy(e1);
    if (e0) (void)hipEventDestroy(e0);
    if (s) (void)hipStreamDestroy(s);
    if (b) (void)hipFree(b);
    if (a) (void)hipFree(a);
    return rc;
}

}  // namespace

extern "C" int lbm_bandwidth_probe(int pattern, int storage, int n, int reps, double* gbps) {
    if (!gbps) return lbm_fail(LBM_ERR_ARG, "lbm_bandwidth_probe: null result pointer");
    if (pattern < 0 || pattern > 2) return lbm_fail(LBM_ERR_ARG, "lbm_bandwidth_probe: pattern %d", pattern);
    if (pattern == 2 && storage != LBM_STORE_F64)
        return lbm_fail(LBM_ERR_ARG, "lbm_bandwidth_probe: the buoyancy pattern is FP64");
    if (n < 16 || n > 512) return lbm_fail(LBM_ERR_ARG, "lbm_bandwidth_probe: n = %d outside [16, 512]", n);
    if (reps < 1 || reps > 1000) return lbm_fail(LBM_ERR_ARG, "lbm_bandwidth_probe: reps = %d", reps);
    switch (storage) {
    case LBM_STORE_F64: return probe<double>(pattern, n, reps, gbps);
    case LBM_STORE_F32: return probe<float>(pattern, n, reps, gbps);
    case LBM_STORE_F16: return probe<uint16_t>(pattern, n, reps, gbps);
    default: return lbm_fail(LBM_ERR_ARG, "lbm_bandwidth_probe: storage %d", storage);
    }
}
