// Morton ordering and the short-range octree.
//
// Replaces Tree::build / build_node (src/tree.rs:171-250). The reference builds its octree
// recursively on one thread, counting-sorting each node's particles into the eight octants
// (bit 0 = x >= centre, bit 1 = y, bit 2 = z; src/tree.rs:216-218) and stopping at 8 particles
// (LEAF_SIZE, src/tree.rs:12). Sorting all particles once by a 63-bit Morton key whose 3-bit
// digits are exactly those octant numbers puts every octree node's particles in one contiguous
// run, in the same octant order. The tree is then built level by level in parallel: a node
// with more than 8 particles finds its children's runs by binary search on the key digit of
// the next level. Node cubes use the reference's FP32 centre/half arithmetic (centre +- half/2),
// and centres of mass are accumulated bottom-up in FP64.
#include <algorithm>
#include <cmath>
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "tpm_internal.h"

namespace tpm {

__device__ __forceinline__ uint64_t spread3(uint32_t v) {
    uint64_t x = v & 0x1fffffu;
    x = (x | x << 32) & 0x1f00000000ffffULL;
    x = (x | x << 16) & 0x1f0000ff0000ffULL;
    x = (x | x << 8) & 0x100f00f00f00f00fULL;
    x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
    x = (x | x << 2) & 0x1249249249249249ULL;
    return x;
}

__global__ void k_morton(const double *__restrict__ pos, size_t n, double box, uint64_t *__restrict__ keys,
                         uint32_t *__restrict__ idx) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // floor(x / L * 2^21): the division comes first, so a position exactly on a dyadic plane
    // of the box maps to the upper cell, like the reference's `p >= centre` octant test.
    const double q0 = floor((pos[3 * i] / box) * (double)(1u << kKeyBits));
    const double q1 = floor((pos[3 * i + 1] / box) * (double)(1u << kKeyBits));
    const double q2 = floor((pos[3 * i + 2] / box) * (double)(1u << kKeyBits));
    const double qmax = (double)((1u << kKeyBits) - 1);
    const uint32_t a = (uint32_t)fmin(fmax(q0, 0.0), qmax);
    const uint32_t b = (uint32_t)fmin(fmax(q1, 0.0), qmax);
    const uint32_t c = (uint32_t)fmin(fmax(q2, 0.0), qmax);
    keys[i] = spread3(a) | (spread3(b) << 1) | (spread3(c) << 2);
    idx[i] = (uint32_t)i;
}

__global__ void k_gather(const double *__restrict__ pos, const uint32_t *__restrict__ idx, size_t n,
                         double *__restrict__ spos) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t j = idx[i];
    spos[3 * i] = pos[3 * j];
    spos[3 * i + 1] = pos[3 * j + 1];
    spos[3 * i + 2] = pos[3 * j + 2];
}

// ---------------------------------------------------------------------------- tree build

__device__ __forceinline__ int digit(uint64_t key, int level) {
    // Octant digit of `key` at tree level `level` (1 = children of the root).
    return (int)((key >> (3 * (kKeyBits - level))) & 7u);
}

// First index in [lo, hi) whose digit at `level` is >= o (keys sorted, common prefix above).
__device__ __forceinline__ uint32_t lower_digit(const uint64_t *keys, uint32_t lo, uint32_t hi, int level, int o) {
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (digit(keys[mid], level) < o)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Pass 1 over one level: decide which nodes split and count their non-empty octants.
__global__ void k_count_children(const uint64_t *__restrict__ keys, TreeView t, size_t lb, size_t le, int level,
                                 uint32_t ref_leaf_rule, uint32_t *__restrict__ nchild) {
    const size_t q = lb + (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= le) return;
    const uint2 r = t.range[q];
    const uint32_t cnt = r.y - r.x;
    uint32_t mask = 0, nc = 0;
    if (cnt > (uint32_t)kLeafSize && level < kMaxDepth) {
        uint32_t b[9];
        b[0] = r.x;
        b[8] = r.y;
        for (int o = 1; o < 8; ++o) b[o] = lower_digit(keys, b[o - 1], r.y, level + 1, o);
        for (int o = 0; o < 8; ++o)
            if (b[o + 1] > b[o]) {
                mask |= 1u << o;
                ++nc;
            }
    }
    // Walk-leaf: no children, or (reference rule) no child in octant 0 (src/tree.rs:280).
    const uint32_t walk_leaf = (nc == 0 || (ref_leaf_rule && !(mask & 1u))) ? 1u : 0u;
    uint4 l = t.link[q];
    l.y = nc;
    l.z = mask;
    l.w = (uint32_t)level | (walk_leaf << 8);
    t.link[q] = l;
    nchild[q - lb] = nc;
}

// Pass 2: create the children of each split node at first = le + offset[q].
__global__ void k_make_children(const uint64_t *__restrict__ keys, TreeView t, size_t lb, size_t le, int level,
                                const uint32_t *__restrict__ offset) {
    const size_t q = lb + (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= le) return;
    uint4 l = t.link[q];
    if (l.y == 0) {
        l.x = 0xFFFFFFFFu;
        t.link[q] = l;
        return;
    }
    const uint2 r = t.range[q];
    uint32_t b[9];
    b[0] = r.x;
    b[8] = r.y;
    for (int o = 1; o < 8; ++o) b[o] = lower_digit(keys, b[o - 1], r.y, level + 1, o);
    const uint32_t first = (uint32_t)(le + offset[q - lb]);
    l.x = first;
    t.link[q] = l;
    const float4 pb = t.box[q];
    const float quarter = pb.w / 2.0f;
    uint32_t ci = first;
    for (int o = 0; o < 8; ++o) {
        if (b[o + 1] == b[o]) continue;
        t.range[ci] = make_uint2(b[o], b[o + 1]);
        // src/tree.rs:243-245: centre[k] +- quarter in FP32.
        t.box[ci] = make_float4(pb.x + ((o & 1) ? quarter : -quarter), pb.y + ((o & 2) ? quarter : -quarter),
                                pb.z + ((o & 4) ? quarter : -quarter), quarter);
        t.link[ci] = make_uint4(0xFFFFFFFFu, 0, 0, (uint32_t)(level + 1));
        ++ci;
    }
}

// Centres of mass of one level, deepest level first: leaves from their particles, split nodes
// from their children (count-weighted), all in FP64.
__global__ void k_com_level(const double *__restrict__ spos, TreeView t, size_t lb, size_t le) {
    const size_t q = lb + (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= le) return;
    const uint4 l = t.link[q];
    const uint2 r = t.range[q];
    double sx = 0, sy = 0, sz = 0;
    if (l.y == 0) {
        for (uint32_t j = r.x; j < r.y; ++j) {
            sx += spos[3 * (size_t)j];
            sy += spos[3 * (size_t)j + 1];
            sz += spos[3 * (size_t)j + 2];
        }
    } else {
        for (uint32_t cix = 0; cix < l.y; ++cix) {
            const double4 cc = t.com[l.x + cix];
            sx += cc.x * cc.w;
            sy += cc.y * cc.w;
            sz += cc.z * cc.w;
        }
    }
    const double cnt = (double)(r.y - r.x);
    t.com[q] = make_double4(sx / cnt, sy / cnt, sz / cnt, cnt);
}

static inline unsigned blocks(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

void sort_particles(Ctx &c, const double *d_pos, size_t n) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_morton, dim3(blocks(n, 256)), dim3(256), 0, c.stream, d_pos, n, c.p.box_size, c.keys, c.idx);
    TPM_HIP(hipGetLastError());
    rocprim::double_buffer<uint64_t> kb(c.keys, c.keys_alt);
    rocprim::double_buffer<uint32_t> vb(c.idx, c.idx_alt);
    size_t bytes = c.sort_tmp_bytes;
    TPM_HIP(rocprim::radix_sort_pairs(c.sort_tmp, bytes, kb, vb, n, 0, 3 * kKeyBits, c.stream));
    if (kb.current() != c.keys) std::swap(c.keys, c.keys_alt);
    if (vb.current() != c.idx) std::swap(c.idx, c.idx_alt);
    hipLaunchKernelGGL(k_gather, dim3(blocks(n, 256)), dim3(256), 0, c.stream, d_pos, c.idx, n, c.spos);
    TPM_HIP(hipGetLastError());
}

static void grow_tree(Ctx &c, size_t need) {
    Tree &t = c.tree;
    if (need <= t.cap) return;
    size_t cap = t.cap ? t.cap : 1024;
    while (cap < need) cap *= 2;
    uint2 *range;
    uint4 *link;
    float4 *box;
    double4 *com;
    TPM_HIP(hipMalloc(&range, cap * sizeof(uint2)));
    TPM_HIP(hipMalloc(&link, cap * sizeof(uint4)));
    TPM_HIP(hipMalloc(&box, cap * sizeof(float4)));
    TPM_HIP(hipMalloc(&com, cap * sizeof(double4)));
    if (t.count) {
        TPM_HIP(hipMemcpyAsync(range, t.range, t.count * sizeof(uint2), hipMemcpyDeviceToDevice, c.stream));
        TPM_HIP(hipMemcpyAsync(link, t.link, t.count * sizeof(uint4), hipMemcpyDeviceToDevice, c.stream));
        TPM_HIP(hipMemcpyAsync(box, t.box, t.count * sizeof(float4), hipMemcpyDeviceToDevice, c.stream));
        TPM_HIP(hipStreamSynchronize(c.stream));
    }
    (void)hipFree(t.range);
    (void)hipFree(t.link);
    (void)hipFree(t.box);
    (void)hipFree(t.com);
    t.range = range;
    t.link = link;
    t.box = box;
    t.com = com;
    t.cap = cap;
}

static void grow_scan(Ctx &c, size_t need) {
    if (need <= c.scan_cap) return;
    size_t cap = c.scan_cap ? c.scan_cap : 4096;
    while (cap < need) cap *= 2;
    (void)hipFree(c.scan_tmp);
    TPM_HIP(hipMalloc(&c.scan_tmp, 2 * cap * sizeof(uint32_t) + 64));
    c.scan_cap = cap;
    size_t bytes = 0;
    TPM_HIP(rocprim::exclusive_scan(nullptr, bytes, c.scan_tmp, c.scan_tmp + cap, 0u, cap, rocprim::plus<uint32_t>(),
                                    c.stream));
    if (bytes > c.scan_tmp_bytes) {
        (void)hipFree(c.scan_tmp_storage);
        TPM_HIP(hipMalloc(&c.scan_tmp_storage, bytes));
        c.scan_tmp_bytes = bytes;
    }
}

void build_tree(Ctx &c, size_t n) {
    Tree &t = c.tree;
    t.level_begin.clear();
    t.count = 0;
    grow_tree(c, std::max<size_t>(1024, n / 2 + 64));
    // Root: the whole box, centre L/2 and half L/2 in FP32 (src/tree.rs:180-181).
    const float half = (float)c.p.box_size / 2.0f;
    const uint2 rr = make_uint2(0, (uint32_t)n);
    const float4 rb = make_float4(half, half, half, half);
    const uint4 rl = make_uint4(0xFFFFFFFFu, 0, 0, 0);
    TPM_HIP(hipMemcpyAsync(t.range, &rr, sizeof(rr), hipMemcpyHostToDevice, c.stream));
    TPM_HIP(hipMemcpyAsync(t.box, &rb, sizeof(rb), hipMemcpyHostToDevice, c.stream));
    TPM_HIP(hipMemcpyAsync(t.link, &rl, sizeof(rl), hipMemcpyHostToDevice, c.stream));
    t.count = 1;
    t.level_begin.push_back(0);
    const uint32_t ref_rule = (c.p.flags & TPM_FLAG_REF_LEAF_RULE) ? 1u : 0u;
    size_t lb = 0, le = 1;
    for (int level = 0; level <= kMaxDepth && le > lb; ++level) {
        const size_t cnt = le - lb;
        grow_scan(c, cnt);
        uint32_t *nch = c.scan_tmp;
        uint32_t *off = c.scan_tmp + c.scan_cap;
        hipLaunchKernelGGL(k_count_children, dim3(blocks(cnt, 256)), dim3(256), 0, c.stream, c.keys, t.view(), lb, le, level,
                           ref_rule, nch);
        TPM_HIP(hipGetLastError());
        size_t bytes = c.scan_tmp_bytes;
        TPM_HIP(rocprim::exclusive_scan(c.scan_tmp_storage, bytes, nch, off, 0u, cnt, rocprim::plus<uint32_t>(),
                                        c.stream));
        uint32_t tail[2];
        TPM_HIP(hipMemcpyAsync(&tail[0], off + cnt - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
        TPM_HIP(hipMemcpyAsync(&tail[1], nch + cnt - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
        TPM_HIP(hipStreamSynchronize(c.stream));
        const size_t added = (size_t)tail[0] + tail[1];
        grow_tree(c, le + added);
        hipLaunchKernelGGL(k_make_children, dim3(blocks(cnt, 256)), dim3(256), 0, c.stream, c.keys, t.view(), lb, le, level,
                           off);
        TPM_HIP(hipGetLastError());
        t.count = le + added;
        lb = le;
        le = le + added;
        t.level_begin.push_back(lb);
    }
    if (t.level_begin.back() != t.count) t.level_begin.push_back(t.count);
    // Centres of mass, deepest level first.
    for (int lv = (int)t.level_begin.size() - 2; lv >= 0; --lv) {
        const size_t b = t.level_begin[lv], e = t.level_begin[lv + 1];
        if (e <= b) continue;
        hipLaunchKernelGGL(k_com_level, dim3(blocks(e - b, 256)), dim3(256), 0, c.stream, c.spos, t.view(), b, e);
        TPM_HIP(hipGetLastError());
    }
}

// ------------------------------------------------------------------------- walk groups

__global__ void k_group_seg(const uint64_t *__restrict__ keys, size_t n, int shift, uint32_t *__restrict__ segval) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool start = i == 0 || (keys[i] >> shift) != (keys[i - 1] >> shift);
    segval[i] = start ? (uint32_t)i : 0u;
}

__global__ void k_group_chunk(const uint32_t *__restrict__ segstart, size_t n, uint32_t *__restrict__ flag) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flag[i] = ((uint32_t)i - segstart[i]) % (uint32_t)kGroup == 0 ? 1u : 0u;
}

__global__ void k_group_starts(const uint32_t *__restrict__ flag, const uint32_t *__restrict__ gid, size_t n,
                               uint32_t *__restrict__ gstart) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (flag[i]) gstart[gid[i]] = (uint32_t)i;
    if (i == n - 1) gstart[gid[i] + flag[i]] = (uint32_t)n;
}

size_t group_scan_bytes(size_t max_n) {
    size_t a = 0, b = 0;
    uint32_t *p = nullptr;
    if (rocprim::inclusive_scan(nullptr, a, p, p, max_n, rocprim::maximum<uint32_t>()) != hipSuccess ||
        rocprim::exclusive_scan(nullptr, b, p, p, 0u, max_n, rocprim::plus<uint32_t>()) != hipSuccess)
        throw Error{TPM_E_DEVICE, "scan size query failed"};
    return std::max(a, b);
}

void build_groups(Ctx &c, size_t n, const uint64_t *keys) {
    if (!keys) keys = c.keys;
    if (n == 0) {
        c.ngroups = 0;
        return;
    }
    // Group cubes hold about 512 particles at the mean density (8 particle spacings on a side
    // when mesh = particle grid), and are never larger than L/4.
    int level = 2;
    while (level < kMaxDepth && ((double)n / std::pow(8.0, level + 1)) >= 512.0) ++level;
    c.group_level = level;
    const int shift = 3 * (kKeyBits - level);
    uint32_t *segval = c.gscratch;
    uint32_t *tmp = c.gscratch + c.max_n;
    hipLaunchKernelGGL(k_group_seg, dim3(blocks(n, 256)), dim3(256), 0, c.stream, keys, n, shift, segval);
    TPM_HIP(hipGetLastError());
    size_t bytes = c.gscan_bytes;
    TPM_HIP(rocprim::inclusive_scan(c.gscan_tmp, bytes, segval, tmp, n, rocprim::maximum<uint32_t>(), c.stream));
    hipLaunchKernelGGL(k_group_chunk, dim3(blocks(n, 256)), dim3(256), 0, c.stream, tmp, n, segval);
    TPM_HIP(hipGetLastError());
    bytes = c.gscan_bytes;
    TPM_HIP(rocprim::exclusive_scan(c.gscan_tmp, bytes, segval, tmp, 0u, n, rocprim::plus<uint32_t>(), c.stream));
    hipLaunchKernelGGL(k_group_starts, dim3(blocks(n, 256)), dim3(256), 0, c.stream, segval, tmp, n, c.gstart);
    TPM_HIP(hipGetLastError());
    uint32_t last[2];
    TPM_HIP(hipMemcpyAsync(&last[0], tmp + n - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
    TPM_HIP(hipMemcpyAsync(&last[1], segval + n - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
    TPM_HIP(hipStreamSynchronize(c.stream));
    c.ngroups = (size_t)last[0] + last[1];
}

hipError_t rocprim_exclusive_scan_u32(void *tmp, size_t &bytes, const uint32_t *in, uint32_t *out, size_t n,
                                      hipStream_t s) {
    return rocprim::exclusive_scan(tmp, bytes, in, out, 0u, n, rocprim::plus<uint32_t>(), s);
}

size_t sort_temp_bytes(size_t max_n) {
    size_t bytes = 0;
    rocprim::double_buffer<uint64_t> kb(nullptr, nullptr);
    rocprim::double_buffer<uint32_t> vb(nullptr, nullptr);
    hipError_t e = rocprim::radix_sort_pairs(nullptr, bytes, kb, vb, max_n, 0, 3 * kKeyBits);
    if (e != hipSuccess) throw Error{TPM_E_DEVICE, "radix_sort_pairs size query failed"};
    return bytes;
}

}  // namespace tpm
