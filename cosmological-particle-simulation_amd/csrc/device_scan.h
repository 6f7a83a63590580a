// Device-wide scans (prefix sums, minima, maxima) of libtpm, hand-written for wave64.
//
// A scan over n values runs as three kernels with a fixed grid of kScanParts workgroups, so it
// needs only kScanParts partial values of temporary storage and never a host synchronisation or a
// memset (it can be captured into a HIP graph):
//   1. each workgroup reduces its contiguous chunk of the input to one partial;
//   2. one workgroup scans the partials (exclusive);
//   3. each workgroup scans its chunk again, starting from its partial, and writes the output.
// Input and output go through functors (in(i) -> T, out(i, T)), so transformed, counted or
// reversed sequences are scanned without materialising them. The operator must be associative
// and commutative (sum, min, max).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tpm {

constexpr int kScanParts = 1024;  // workgroups of phases 1 and 3
constexpr int kScanItems = 8;     // consecutive values per thread in phase 3

struct ScanSum {
    __device__ __forceinline__ uint32_t operator()(uint32_t a, uint32_t b) const { return a + b; }
};
struct ScanMin {
    __device__ __forceinline__ uint32_t operator()(uint32_t a, uint32_t b) const { return a < b ? a : b; }
};
struct ScanMax {
    __device__ __forceinline__ uint32_t operator()(uint32_t a, uint32_t b) const { return a > b ? a : b; }
};

// Reads in[i].
struct ScanLoad {
    const uint32_t *p;
    __device__ __forceinline__ uint32_t operator()(size_t i) const { return p[i]; }
};
// Writes out[i] = v.
struct ScanStore {
    uint32_t *p;
    __device__ __forceinline__ void operator()(size_t i, uint32_t v) const { p[i] = v; }
};

// Exclusive scan of one value per thread over a 256-thread workgroup (four wavefronts): wavefront
// scans by lane shifts, then the wavefront totals through LDS. `total` receives the reduction.
template <typename Op>
__device__ __forceinline__ uint32_t block_exclusive_256(uint32_t v, uint32_t id, Op op, uint32_t &total,
                                                         uint32_t *sh4) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x = op(y, x);
    }
    if (lane == 63) sh4[w] = x;
    __syncthreads();
    uint32_t pre = id;
    for (int k = 0; k < w; ++k) pre = op(pre, sh4[k]);
    total = op(op(sh4[0], sh4[1]), op(sh4[2], sh4[3]));
    __syncthreads();
    uint32_t ex = __shfl_up(x, 1, 64);
    if (lane == 0) ex = id;
    return op(pre, ex);
}

__device__ __forceinline__ void scan_part_range(size_t n, uint32_t b, size_t &c0, size_t &c1) {
    const size_t per = (n + kScanParts - 1) / kScanParts;
    c0 = b * per < n ? b * per : n;
    c1 = c0 + per < n ? c0 + per : n;
}

template <typename In, typename Op>
__global__ void __launch_bounds__(256) k_dscan_reduce(In in, size_t n, uint32_t id, Op op, uint32_t *__restrict__ part) {
    __shared__ uint32_t sh4[4];
    size_t c0, c1;
    scan_part_range(n, blockIdx.x, c0, c1);
    uint32_t s = id;
    for (size_t i = c0 + threadIdx.x; i < c1; i += 256) s = op(s, in(i));
    uint32_t total;
    (void)block_exclusive_256(s, id, op, total, sh4);
    if (threadIdx.x == 0) part[blockIdx.x] = total;
}

template <typename Op>
__global__ void __launch_bounds__(256) k_dscan_spine(uint32_t *__restrict__ part, uint32_t id, Op op) {
    // kScanParts partials, four per thread (blocked).
    __shared__ uint32_t sh4[4];
    static_assert(kScanParts == 4 * 256, "spine layout");
    uint32_t v[4];
    uint32_t s = id;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        v[q] = part[4 * threadIdx.x + q];
        s = op(s, v[q]);
    }
    uint32_t total;
    uint32_t run = block_exclusive_256(s, id, op, total, sh4);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        part[4 * threadIdx.x + q] = run;
        run = op(run, v[q]);
    }
}

template <typename In, typename Out, typename Op>
__global__ void __launch_bounds__(256) k_dscan_down(In in, Out out, size_t n, uint32_t id, Op op,
                                                    const uint32_t *__restrict__ part, int inclusive) {
    __shared__ uint32_t sh4[4];
    size_t c0, c1;
    scan_part_range(n, blockIdx.x, c0, c1);
    uint32_t carry = part[blockIdx.x];
    for (size_t t0 = c0; t0 < c1; t0 += 256 * kScanItems) {
        const size_t base = t0 + (size_t)threadIdx.x * kScanItems;
        uint32_t v[kScanItems];
        uint32_t s = id;
#pragma unroll
        for (int q = 0; q < kScanItems; ++q) {
            v[q] = base + q < c1 ? in(base + q) : id;
            s = op(s, v[q]);
        }
        uint32_t total;
        uint32_t run = op(carry, block_exclusive_256(s, id, op, total, sh4));
#pragma unroll
        for (int q = 0; q < kScanItems; ++q) {
            if (inclusive) run = op(run, v[q]);
            if (base + q < c1) out(base + q, run);
            if (!inclusive) run = op(run, v[q]);
        }
        carry = op(carry, total);
    }
}

// Scan of n values (exclusive, or inclusive when `inclusive`), enqueued on `s`; `part` holds
// kScanParts words of scratch. Returns the launch status.
template <typename In, typename Out, typename Op>
hipError_t device_scan(In in, Out out, size_t n, uint32_t id, Op op, bool inclusive, uint32_t *part, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL((k_dscan_reduce<In, Op>), dim3(kScanParts), dim3(256), 0, s, in, n, id, op, part);
    hipLaunchKernelGGL((k_dscan_spine<Op>), dim3(1), dim3(256), 0, s, part, id, op);
    hipLaunchKernelGGL((k_dscan_down<In, Out, Op>), dim3(kScanParts), dim3(256), 0, s, in, out, n, id, op,
                       (const uint32_t *)part, inclusive ? 1 : 0);
    return hipGetLastError();
}

}  // namespace tpm
