// Device-wide scans (prefix sums, minima, maxima) of libtpm, hand-written for wave64.
//
// A scan over n values is ONE kernel (decoupled look-back): each workgroup reduces its contiguous
// chunk of the input, publishes that reduction in a status word, finds the reduction of every
// earlier chunk by reading its predecessors' status words (a predecessor that has finished its own
// look-back publishes its inclusive prefix, which ends the search), publishes its own inclusive
// prefix and scans its chunk again from there. It never synchronises with the host and clears no
// memory (it can be captured into a HIP graph), so a scan costs one launch where a reduce, a scan of
// the partials and a down-sweep cost three: the scans and octree levels of a small box's evaluation
// are bound by launches, not by bytes.
//
// The state of a scan lives in a slot of kLbSlotWords 64-bit words: a ticket counter, then one
// status word per workgroup. Workgroups take their chunk in the order they start (a ticket from the
// counter, so every earlier chunk belongs to a workgroup that is running or done, and a look-back
// always ends). The launch's last ticket advances the counter to the next multiple of
// kLbMaxBlocks, so ticket >> kLbShift numbers the launches of a slot; status words carry that
// number (the epoch) and words of earlier launches read as not yet published. Launches that share
// a slot must be ordered (one stream); slots start zeroed.
//
// Input and output go through functors (in(i) -> T, out(i, T)), so transformed, counted or
// reversed sequences are scanned without materialising them. The operator must be associative
// and commutative (sum, min, max).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tpm {

constexpr int kScanParts = 1024;  // most workgroups of one scan
constexpr int kScanItems = 8;     // consecutive values per thread in the down-sweep
constexpr int kLbShift = 13;
constexpr uint32_t kLbMaxBlocks = 1u << kLbShift;  // most workgroups of one look-back launch
constexpr size_t kLbSlotWords = 1 + (size_t)kLbMaxBlocks;

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

// ------------------------------------------------------------------------- look-back protocol

// Status word: value (bits 0-31), flag (bits 32-33: 0 not published, 1 the chunk's reduction, 2 the
// inclusive prefix up to the chunk), epoch (bits 34-63).
__device__ __forceinline__ unsigned long long lb_pack(uint32_t value, uint32_t flag, uint32_t ep) {
    return (unsigned long long)value | ((unsigned long long)flag << 32) | ((unsigned long long)ep << 34);
}
__device__ __forceinline__ uint32_t lb_flag(unsigned long long s, uint32_t ep) {
    return (uint32_t)(s >> 34) == ep ? (uint32_t)(s >> 32) & 3u : 0u;
}
__device__ __forceinline__ void lb_store(unsigned long long *p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long lb_load(const unsigned long long *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The workgroup's chunk number b (in start order, < G) and the launch's epoch. Exactly G workgroups
// of the launch call it, once each, at their start.
__device__ __forceinline__ void lb_ticket(unsigned long long *slot, uint32_t G, uint32_t &b, uint32_t &ep) {
    __shared__ unsigned long long s_ticket;
    if (threadIdx.x == 0) {
        const unsigned long long t = atomicAdd(slot, 1ull);
        // The last ticket of the launch: every ticket of it is taken, so no workgroup of this launch
        // reads the counter again.
        if ((uint32_t)(t & (kLbMaxBlocks - 1)) == G - 1) atomicAdd(slot, (unsigned long long)(kLbMaxBlocks - G));
        s_ticket = t;
    }
    __syncthreads();
    const unsigned long long t = s_ticket;
    b = (uint32_t)(t & (kLbMaxBlocks - 1));
    ep = (uint32_t)(t >> kLbShift) & 0x3FFFFFFFu;
}

// Publishes chunk b's reduction `total`, looks back for the reduction of chunks 0..b-1 (returned,
// `id` for b = 0) and publishes the inclusive prefix. Called by the whole workgroup (the first
// wavefront does the work); ends with a barrier.
template <typename Op>
__device__ __forceinline__ uint32_t lb_prefix(unsigned long long *status, uint32_t b, uint32_t ep, uint32_t total,
                                              uint32_t id, Op op) {
    __shared__ uint32_t s_carry;
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        uint32_t carry = id;
        if (b == 0) {
            if (lane == 0) lb_store(status, lb_pack(total, 2u, ep));
        } else {
            if (lane == 0) lb_store(status + b, lb_pack(total, 1u, ep));
            // 64 predecessors at a time, nearest in lane 0; stop at the first inclusive prefix.
            long long base = (long long)b - 1;
            for (;;) {
                const long long j = base - lane;
                unsigned long long s = 0;
                uint32_t fl = 2u;
                for (;;) {
                    if (j >= 0) {
                        s = lb_load(status + j);
                        fl = lb_flag(s, ep);
                    }
                    if (__all(fl != 0u)) break;
                    __builtin_amdgcn_s_sleep(1);
                }
                const unsigned long long incl = __ballot(fl == 2u);
                const int stop = incl ? __ffsll(incl) - 1 : 63;
                uint32_t v = (j >= 0 && lane <= stop) ? (uint32_t)s : id;
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) v = op(v, (uint32_t)__shfl_xor(v, off, 64));
                carry = op(carry, v);
                if (incl) break;
                base -= 64;
            }
            if (lane == 0) lb_store(status + b, lb_pack(op(carry, total), 2u, ep));
        }
        if (lane == 0) s_carry = carry;
    }
    __syncthreads();
    return s_carry;
}

// ------------------------------------------------------------------------------------ scans

__device__ __forceinline__ void scan_chunk_range(size_t n, uint32_t G, uint32_t b, size_t &c0, size_t &c1) {
    const size_t per = (n + G - 1) / G;
    c0 = b * per < n ? b * per : n;
    c1 = c0 + per < n ? c0 + per : n;
}

template <typename In, typename Out, typename Op>
__global__ void __launch_bounds__(256) k_dscan(In in, Out out, size_t n, uint32_t id, Op op,
                                               unsigned long long *__restrict__ slot, uint32_t G, int inclusive) {
    __shared__ uint32_t sh4[4];
    // Workgroups past the last non-empty chunk end at once and take no ticket.
    const size_t per = (n + G - 1) / G;
    const uint32_t chunks = (uint32_t)((n + per - 1) / per);
    if (blockIdx.x >= chunks) return;
    uint32_t b, ep;
    lb_ticket(slot, chunks, b, ep);
    size_t c0, c1;
    scan_chunk_range(n, G, b, c0, c1);
    uint32_t s = id;
    for (size_t i = c0 + threadIdx.x; i < c1; i += 256) s = op(s, in(i));
    uint32_t total;
    (void)block_exclusive_256(s, id, op, total, sh4);
    uint32_t carry = lb_prefix(slot + 1, b, ep, total, id, op);
    for (size_t t0 = c0; t0 < c1; t0 += 256 * kScanItems) {
        const size_t base = t0 + (size_t)threadIdx.x * kScanItems;
        uint32_t v[kScanItems];
        uint32_t r = id;
#pragma unroll
        for (int q = 0; q < kScanItems; ++q) {
            v[q] = base + q < c1 ? in(base + q) : id;
            r = op(r, v[q]);
        }
        uint32_t tile;
        uint32_t run = op(carry, block_exclusive_256(r, id, op, tile, sh4));
#pragma unroll
        for (int q = 0; q < kScanItems; ++q) {
            if (inclusive) run = op(run, v[q]);
            if (base + q < c1) out(base + q, run);
            if (!inclusive) run = op(run, v[q]);
        }
        carry = op(carry, tile);
    }
}

// Scan of n values (exclusive, or inclusive when `inclusive`), enqueued on `s` with the look-back
// slot `slot` (kLbSlotWords words; not used by another stream meanwhile). In place (in and out the
// same array) is allowed. Returns the launch status.
template <typename In, typename Out, typename Op>
hipError_t device_scan(In in, Out out, size_t n, uint32_t id, Op op, bool inclusive, unsigned long long *slot,
                       hipStream_t s) {
    if (n == 0) return hipSuccess;
    // About 2048 values per workgroup, at most kScanParts workgroups.
    const size_t want = (n + 256 * kScanItems - 1) / (256 * kScanItems);
    const uint32_t G = (uint32_t)(want < (size_t)kScanParts ? want : (size_t)kScanParts);
    hipLaunchKernelGGL((k_dscan<In, Out, Op>), dim3(G), dim3(256), 0, s, in, out, n, id, op, slot, G,
                       inclusive ? 1 : 0);
    return hipGetLastError();
}

}  // namespace tpm
