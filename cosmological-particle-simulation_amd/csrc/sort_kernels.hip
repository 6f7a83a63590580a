// Stable LSD radix sort of (key, index) pairs: the Morton ordering behind the octree.
//
// Replaces the octant counting sort of Tree::build (src/tree.rs:215-236), which partitions each
// node's particles into its eight octants in input order. Sorting every particle once by its Morton
// key, stably, gives every octree node a contiguous run in the same octant order (bit 0 = x,
// bit 1 = y, bit 2 = z), with input order kept below the sorted depth.
//
// Each pass orders by one 8-bit digit (reduce-then-scan, no look-back waits, no memsets):
//   k_radix_up    per tile of 4096 keys, the count of each digit, written digit-major
//                 (table[d * tiles + tile]);
//   device_scan   exclusive sum over that table: where digit d of tile t starts in the output;
//   k_radix_down  per tile, the stable rank of every key among the tile's keys of its digit, found
//                 wavefront by wavefront: eight ballots give each lane the mask of lanes holding the
//                 same digit, whose lower-lane count is its rank and whose first lane advances the
//                 wavefront's counter of that digit. Keys are staged through LDS in digit order so
//                 the global writes run along the digit's output range.
// The tile's keys are taken in wavefront-round order (round r of wavefront w holds elements
// (w * kItems + r) * 64 + lane), which is their input order, so equal digits keep input order.
// The Morton sort of positions computes the keys from the positions in its first counting pass; a
// gather then puts the positions (and full keys) in sorted order.
#include "device_scan.h"
#include "morton_keys.h"
#include "tpm_internal.h"

namespace tpm {

constexpr int kRadixThreads = 256;
constexpr int kRadixWaves = kRadixThreads / 64;
constexpr int kRadixItems = 16;
constexpr int kRadixTile = kRadixThreads * kRadixItems;  // 4096 keys
constexpr int kRadixDigits = 256;
static_assert(kRadixThreads == kRadixDigits, "one thread per digit in the tile prefix step");

size_t radix_tiles(size_t n) { return (n + kRadixTile - 1) / kRadixTile; }

// Lanes of the wavefront holding the same 8-bit digit as this lane (within `active`).
__device__ __forceinline__ uint64_t same_digit_lanes(uint32_t d, uint64_t active) {
    uint64_t m = active;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const uint32_t bit = (d >> b) & 1u;
        const uint64_t bb = __ballot(bit);
        m &= bit ? bb : ~bb;
    }
    return m;
}

// Number of set bits of m below this lane.
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

template <typename K>
__device__ __forceinline__ uint32_t radix_digit(K key, int shift) {
    return (uint32_t)(key >> shift) & 0xFFu;
}

// Counts the digits of one wavefront's keys in its LDS counters cnt[d]; digit[r] and rank[r] receive
// round r's digit and its stable rank among the wavefront's keys of that digit so far.
template <typename K>
__device__ __forceinline__ void wave_rank(const K (&key)[kRadixItems], const bool (&valid)[kRadixItems], int shift,
                                          uint32_t *cnt, uint32_t (&digit)[kRadixItems],
                                          uint32_t (&rank)[kRadixItems]) {
#pragma unroll
    for (int r = 0; r < kRadixItems; ++r) {
        const uint64_t active = __ballot(valid[r]);
        const uint32_t d = radix_digit<K>(key[r], shift);
        const uint64_t m = same_digit_lanes(d, active);
        const uint32_t below = lanes_below(m);
        uint32_t base = 0;
        if (valid[r]) base = cnt[d];
        if (valid[r] && below == 0) cnt[d] = base + (uint32_t)__popcll(m);
        digit[r] = d;
        rank[r] = base + below;
    }
}

// Morton key of a position as a sort key: the top 30 bits (32-bit keys) or all 63.
template <typename K>
__device__ __forceinline__ K sort_key(const double *p, double box);
template <>
__device__ __forceinline__ uint32_t sort_key<uint32_t>(const double *p, double box) {
    return (uint32_t)(morton_key(p, box) >> kHiShift);
}
template <>
__device__ __forceinline__ uint64_t sort_key<uint64_t>(const double *p, double box) {
    return morton_key(p, box);
}

// The full 63-bit key of a sort key (32-bit keys: the levels below 10 cleared).
__device__ __forceinline__ uint64_t full_key(uint32_t k) { return (uint64_t)k << kHiShift; }
__device__ __forceinline__ uint64_t full_key(uint64_t k) { return k; }

// Digit counts of one tile, written digit-major. FROM_POS: the keys are first computed from the
// positions and written to `keys` (the first pass of a Morton sort). Counts commute, so LDS atomics
// count them.
template <typename K, bool FROM_POS>
__global__ void __launch_bounds__(kRadixThreads) k_radix_up(K *__restrict__ keys, const double *__restrict__ pos,
                                                            double box, size_t n, int shift,
                                                            uint32_t *__restrict__ table, uint32_t tiles) {
    __shared__ uint32_t cnt[kRadixDigits];
    __shared__ double stage[FROM_POS ? 3 * kRadixThreads : 1];
    cnt[threadIdx.x] = 0u;
    __syncthreads();
    const size_t tile0 = (size_t)blockIdx.x * kRadixTile;
    for (int r = 0; r < kRadixItems; ++r) {
        const size_t r0 = tile0 + (size_t)r * kRadixThreads;
        if (r0 >= n) break;  // uniform across the workgroup
        const size_t i = r0 + threadIdx.x;
        K k = 0;
        if (FROM_POS) {
            // The 256 positions of the round, 6 KB contiguous, read as 16-byte vectors into LDS
            // (r0 is a multiple of 256, so the run starts 16-byte aligned).
            const size_t words = 3 * min((size_t)kRadixThreads, n - r0);
            const double *src = pos + 3 * r0;
            for (size_t w = 2 * (size_t)threadIdx.x; w < words; w += 2 * kRadixThreads) {
                if (w + 1 < words) {
                    const double2 v = *reinterpret_cast<const double2 *>(src + w);
                    stage[w] = v.x;
                    stage[w + 1] = v.y;
                } else {
                    stage[w] = src[w];
                }
            }
            __syncthreads();
            if (i < n) {
                k = sort_key<K>(stage + 3 * threadIdx.x, box);
                keys[i] = k;
            }
            __syncthreads();
        } else if (i < n) {
            k = keys[i];
        }
        if (i < n) atomicAdd(&cnt[radix_digit<K>(k, shift)], 1u);
    }
    __syncthreads();
    table[(size_t)threadIdx.x * tiles + blockIdx.x] = cnt[threadIdx.x];
}

// One pass: keys_out/vals_out = keys/vals in stable order of the digit at `shift`; vals == null
// means the identity (vals[i] = i). offsets = the scanned table. WIDE (the last pass over 32-bit
// keys): keys_out receives the full 64-bit keys (the levels below 10 cleared).
template <typename K, bool WIDE>
__global__ void __launch_bounds__(kRadixThreads) k_radix_down(const K *__restrict__ keys, const uint32_t *__restrict__ vals,
                                                              size_t n, int shift, const uint32_t *__restrict__ offsets,
                                                              uint32_t tiles, void *__restrict__ keys_out,
                                                              uint32_t *__restrict__ vals_out) {
    __shared__ uint32_t cnt[kRadixWaves][kRadixDigits];
    __shared__ uint32_t tile_off[kRadixDigits];
    __shared__ uint32_t gbase[kRadixDigits];
    __shared__ uint32_t sh4[4];
    __shared__ K skey[kRadixTile];
    __shared__ uint32_t sval[kRadixTile];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int d = threadIdx.x; d < kRadixWaves * kRadixDigits; d += kRadixThreads) (&cnt[0][0])[d] = 0u;
    __syncthreads();
    // Consecutive tiles write consecutive runs of each digit's output range: giving each XCD a
    // contiguous range of tiles lets its L2 merge the runs into whole lines.
    const uint32_t tile = xcd_block(blockIdx.x, gridDim.x);
    const size_t tile0 = (size_t)tile * kRadixTile;
    K key[kRadixItems];
    uint32_t val[kRadixItems];
    bool valid[kRadixItems];
#pragma unroll
    for (int r = 0; r < kRadixItems; ++r) {
        const size_t i = tile0 + (size_t)(w * kRadixItems + r) * 64 + lane;
        valid[r] = i < n;
        key[r] = valid[r] ? keys[i] : (K)0;
        val[r] = valid[r] ? (vals ? vals[i] : (uint32_t)i) : 0u;
    }
    uint32_t digit[kRadixItems], rank[kRadixItems];
    wave_rank<K>(key, valid, shift, cnt[w], digit, rank);
    __syncthreads();
    // Per digit: the wavefronts' exclusive prefixes, the tile's total and its start in the tile.
    {
        const int d = threadIdx.x;
        uint32_t s = 0;
#pragma unroll
        for (int q = 0; q < kRadixWaves; ++q) {
            const uint32_t c = cnt[q][d];
            cnt[q][d] = s;
            s += c;
        }
        uint32_t total;
        const uint32_t off = block_exclusive_256(s, 0u, ScanSum{}, total, sh4);
        tile_off[d] = off;
        gbase[d] = offsets[(size_t)d * tiles + tile] - off;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRadixItems; ++r) {
        if (!valid[r]) continue;
        const uint32_t lr = tile_off[digit[r]] + cnt[w][digit[r]] + rank[r];
        skey[lr] = key[r];
        sval[lr] = val[r];
    }
    __syncthreads();
    const uint32_t count = (uint32_t)min((size_t)kRadixTile, n - tile0);
    for (uint32_t j = threadIdx.x; j < count; j += kRadixThreads) {
        const K k = skey[j];
        const size_t g = (size_t)gbase[radix_digit<K>(k, shift)] + j;
        if (WIDE)
            static_cast<uint64_t *>(keys_out)[g] = full_key(k);
        else
            static_cast<K *>(keys_out)[g] = k;
        vals_out[g] = sval[j];
    }
}

// Positions in sorted order, spos[k] = pos[idx[k]], and, from 32-bit sort keys, the full keys
// (keys64[k], levels below 10 cleared). A workgroup takes 512 consecutive slots, two per thread
// (256 apart, their loads in flight together), and stages the positions through LDS so that the
// stores are whole 16-byte vectors along spos. Each XCD takes a contiguous run of the sorted order (a
// compact region of the box), so the input lines that neighbouring cubes share stay in its L2.
template <typename K>
__global__ void __launch_bounds__(256) k_gather_positions(const double *__restrict__ pos, const uint32_t *__restrict__ idx,
                                                          const K *__restrict__ skeys, size_t n,
                                                          double *__restrict__ spos, uint64_t *__restrict__ keys64) {
    __shared__ double stage[3 * 512];
    const size_t s0 = (size_t)xcd_block(blockIdx.x, gridDim.x) * 512;
    const int t = threadIdx.x;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const size_t k = s0 + t + 256 * q;
        if (k < n) {
            const size_t j = idx[k];
            stage[3 * (t + 256 * q)] = pos[3 * j];
            stage[3 * (t + 256 * q) + 1] = pos[3 * j + 1];
            stage[3 * (t + 256 * q) + 2] = pos[3 * j + 2];
            if (keys64) keys64[k] = full_key(skeys[k]);
        }
    }
    __syncthreads();
    const size_t count = min((size_t)512, n - s0);
    const size_t words = 3 * count;
    double *out = spos + 3 * s0;  // 16-byte aligned: s0 is a multiple of 512
    for (size_t w = 2 * (size_t)t; w < words; w += 512) {
        if (w + 1 < words) {
            *reinterpret_cast<double2 *>(out + w) = make_double2(stage[w], stage[w + 1]);
        } else {
            out[w] = stage[w];
        }
    }
}

static inline unsigned tile_grid(size_t n) { return (unsigned)radix_tiles(n); }

// Morton sort of n positions by key bits [lo, hi) of K-bit sort keys (see the top of the file):
// afterwards c.idx[k] is the caller index of sorted slot k, c.spos its position and c.keys its
// full key. The key and index buffers ping-pong; the last pass writes the full keys into the buffer
// it does not read, which becomes c.keys.
template <typename K>
static void morton_sort(Ctx &c, const double *d_pos, size_t n, int lo, int hi) {
    TPM_REQUIRE(lo < hi, "empty radix bit range");
    TPM_REQUIRE(radix_tiles(n) * kRadixDigits <= c.radix_table_words, "radix table too small");
    const uint32_t tiles = (uint32_t)radix_tiles(n);
    const size_t entries = (size_t)tiles * kRadixDigits;
    K *ka = reinterpret_cast<K *>(c.keys), *kb = reinterpret_cast<K *>(c.keys_alt);
    uint32_t *va = c.idx, *vb = c.idx_alt;
    const double box = c.p.box_size;
    for (int shift = lo; shift < hi; shift += 8) {
        const bool first = shift == lo;
        if (first)
            hipLaunchKernelGGL((k_radix_up<K, true>), dim3(tile_grid(n)), dim3(kRadixThreads), 0, c.stream, ka, d_pos, box,
                               n, shift, c.radix_table, tiles);
        else
            hipLaunchKernelGGL((k_radix_up<K, false>), dim3(tile_grid(n)), dim3(kRadixThreads), 0, c.stream, ka,
                               (const double *)nullptr, box, n, shift, c.radix_table, tiles);
        TPM_HIP(hipGetLastError());
        TPM_HIP(device_scan(ScanLoad{c.radix_table}, ScanStore{c.radix_table}, entries, 0u, ScanSum{}, false,
                            c.scan_part, c.stream));
        const uint32_t *vin = first ? nullptr : va;
        // The last pass over 32-bit keys writes the full 64-bit keys (both buffers hold 64-bit keys).
        if (sizeof(K) == 4 && shift + 8 >= hi)
            hipLaunchKernelGGL((k_radix_down<K, true>), dim3(tile_grid(n)), dim3(kRadixThreads), 0, c.stream, ka, vin,
                               n, shift, c.radix_table, tiles, (void *)kb, vb);
        else
            hipLaunchKernelGGL((k_radix_down<K, false>), dim3(tile_grid(n)), dim3(kRadixThreads), 0, c.stream, ka, vin,
                               n, shift, c.radix_table, tiles, (void *)kb, vb);
        TPM_HIP(hipGetLastError());
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    // ka: the sorted keys, full 64-bit keys in either case.
    uint64_t *full = reinterpret_cast<uint64_t *>(ka);
    hipLaunchKernelGGL(k_gather_positions<K>, dim3((unsigned)((n + 511) / 512)), dim3(256), 0, c.stream, d_pos, va,
                       (const K *)nullptr, n, c.spos, (uint64_t *)nullptr);
    TPM_HIP(hipGetLastError());
    c.keys = full;
    c.keys_alt = reinterpret_cast<uint64_t *>(kb);
    c.idx = va;
    c.idx_alt = vb;
}

void exclusive_sum_u32(Ctx &c, const uint32_t *in, uint32_t *out, size_t n) {
    TPM_HIP(device_scan(ScanLoad{in}, ScanStore{out}, n, 0u, ScanSum{}, false, c.scan_part, c.stream));
}

void morton_sort_positions(Ctx &c, const double *d_pos, size_t n, int depth) {
    if (n == 0) return;
    if (depth <= 10)
        morton_sort<uint32_t>(c, d_pos, n, 30 - 3 * depth, 30);
    else
        morton_sort<uint64_t>(c, d_pos, n, 3 * (kKeyBits - depth), 3 * kKeyBits);
}

}  // namespace tpm
