// Morton ordering and the short-range octree.
//
// Replaces Tree::build / build_node (src/tree.rs:171-250). The reference builds its octree
// recursively on one thread, counting-sorting each node's particles into the eight octants
// (bit 0 = x >= centre, bit 1 = y, bit 2 = z; src/tree.rs:216-218) and stopping at 8 particles
// (LEAF_SIZE, src/tree.rs:12). Sorting all particles once by a 63-bit Morton key whose 3-bit
// digits are exactly those octant numbers puts every octree node's particles in one contiguous
// run, in the same octant order. The tree is then built level by level in parallel: a node
// with more than 8 particles finds its children's runs from the key digit of the next level
// (binary searches, or for at most 64 particles a count of the digits). Node cubes use the
// reference's FP32 centre/half arithmetic (centre +- half/2), and centres of mass are
// accumulated bottom-up in FP64.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "device_scan.h"
#include "morton_keys.h"
#include "tpm_internal.h"

namespace tpm {

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

// Octant boundaries of the sorted keys [lo, hi) at `level`: b[o] = first index whose digit is >= o.
// A range of at most 64 keys counts its digits in one pass of independent loads (a 7-bit count
// per octant in one 64-bit word); a larger one uses binary searches.
__device__ __forceinline__ void digit_bounds(const uint64_t *keys, uint32_t lo, uint32_t hi, int level,
                                             uint32_t b[9]) {
    b[0] = lo;
    b[8] = hi;
    if (hi - lo <= 64) {
        const int sh = 3 * (kKeyBits - level);
        uint64_t f = 0;
        uint32_t j = lo;
        for (; j + 4 <= hi; j += 4) {
            const uint64_t k0 = keys[j], k1 = keys[j + 1], k2 = keys[j + 2], k3 = keys[j + 3];
            f += (1ull << (7 * ((k0 >> sh) & 7u))) + (1ull << (7 * ((k1 >> sh) & 7u))) +
                 (1ull << (7 * ((k2 >> sh) & 7u))) + (1ull << (7 * ((k3 >> sh) & 7u)));
        }
        for (; j < hi; ++j) f += 1ull << (7 * ((keys[j] >> sh) & 7u));
        uint32_t at = lo;
        for (int o = 1; o < 8; ++o) {
            at += (uint32_t)((f >> (7 * (o - 1))) & 127u);
            b[o] = at;
        }
    } else {
        for (int o = 1; o < 8; ++o) b[o] = lower_digit(keys, b[o - 1], hi, level, o);
    }
}

// The octree is built level by level entirely on the device (no host synchronisation, so the
// build can be captured in a HIP graph): lv[l] is the first node of level l, lv[l + 1] its end.
// One kernel per level counts each node's children, places them by a single-pass (look-back) scan
// of the counts and creates them; its last workgroup writes the next level's end to lv[l + 2].

// Octant boundaries of a node of `level` (cube prefix of its first key) from the table of the first
// sorted index of every level-`tab_level` cube, for level < tab_level: no search at all.
__device__ __forceinline__ void table_bounds(const uint64_t *keys, const uint32_t *tab, int tab_level, uint32_t first,
                                             int level, uint32_t b[9]) {
    const uint64_t cube = keys[first] >> (3 * (kKeyBits - level));
    const int sh = 3 * (tab_level - level - 1);
#pragma unroll
    for (int o = 0; o <= 8; ++o) b[o] = tab[(size_t)((cube << 3) + (uint64_t)o) << sh];
}

// Pass 1 for node q of `level`: decide whether it splits, count its non-empty octants and record
// count, octant mask, level and walk-leaf status in its link (returns the count). Nodes of at most
// `keep_whole` particles that would split get no children (the group walk never opens them) but
// keep their walk-leaf status from the split they would have.
__device__ __forceinline__ uint32_t count_node(const uint64_t *__restrict__ keys, TreeView t, uint32_t q, int level,
                                               int max_depth, uint32_t ref_leaf_rule, uint32_t keep_whole,
                                               const uint32_t *__restrict__ tab, int tab_level) {
    const uint2 r = t.range[q];
    const uint32_t cnt = r.y - r.x;
    uint32_t mask = 0, nc = 0;
    if (cnt > (uint32_t)kLeafSize && level < max_depth) {
        uint32_t b[9];
        if (level < tab_level)
            table_bounds(keys, tab, tab_level, r.x, level, b);
        else
            digit_bounds(keys, r.x, r.y, level + 1, b);
        for (int o = 0; o < 8; ++o)
            if (b[o + 1] > b[o]) {
                mask |= 1u << o;
                ++nc;
            }
    }
    // Walk-leaf: no children, or (reference rule) no child in octant 0 (src/tree.rs:280).
    const uint32_t walk_leaf = (nc == 0 || (ref_leaf_rule && !(mask & 1u))) ? 1u : 0u;
    if (cnt <= keep_whole) nc = 0;
    uint4 l = t.link[q];
    l.y = nc;
    l.z = mask;
    l.w = (uint32_t)level | (walk_leaf << 8);
    t.link[q] = l;
    return nc;
}

// Pass 2 for node q: create its children at [first, first + count) (count from pass 1).
__device__ __forceinline__ void make_children(const uint64_t *__restrict__ keys, TreeView t, uint32_t q, int level,
                                              uint64_t first, uint32_t cap, const uint32_t *__restrict__ tab,
                                              int tab_level) {
    uint4 l = t.link[q];
    if (l.y == 0) {
        l.x = 0xFFFFFFFFu;
        t.link[q] = l;
        return;
    }
    if (first + l.y > cap) {
        // Over capacity (reported through info[1]): the node stays a leaf so the walk stays exact.
        l.x = 0xFFFFFFFFu;
        l.y = 0;
        l.w |= 1u << 8;
        t.link[q] = l;
        return;
    }
    const uint2 r = t.range[q];
    uint32_t b[9];
    if (level < tab_level)
        table_bounds(keys, tab, tab_level, r.x, level, b);
    else
        digit_bounds(keys, r.x, r.y, level + 1, b);
    l.x = (uint32_t)first;
    t.link[q] = l;
    const float4 pb = t.box[q];
    const float quarter = pb.w / 2.0f;
    uint32_t ci = (uint32_t)first;
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

// One level of the build: each of the G workgroups takes a contiguous chunk of the level's nodes
// (in 256-node tiles), counts their children (pass 1), obtains the number of children of all
// earlier chunks by look-back (`slot`, see device_scan.h), then creates its nodes' children in node
// order from there (pass 2). The grid is sized by the level's largest possible node count, so
// most workgroups of a deep level usually have no chunk: only the first `chunks` workgroups take a
// ticket (at least one, when the level is empty), the others end at once. The workgroup of the
// last chunk writes the next level's end (clamped to the capacity, with info[1] set when children
// had to be dropped) to lv[level + 2] and info[0].
__global__ void __launch_bounds__(256) k_build_level(const uint64_t *__restrict__ keys, TreeView t,
                                                     uint32_t *__restrict__ lv, int level, int max_depth,
                                                     uint32_t ref_leaf_rule, uint32_t keep_whole,
                                                     const uint32_t *__restrict__ tab, int tab_level, uint32_t cap,
                                                     uint32_t *__restrict__ info, unsigned long long *__restrict__ slot,
                                                     uint32_t G) {
    __shared__ uint32_t sh4[4];
    const uint32_t lb = lv[level], le = lv[level + 1];
    const uint32_t tiles = (le - lb + 255) / 256;
    const uint32_t per = tiles ? (tiles + G - 1) / G : 1u;
    const uint32_t chunks = tiles ? (tiles + per - 1) / per : 1u;  // non-empty chunks (one if none)
    if (blockIdx.x >= chunks) return;  // uniform across the workgroup; takes no ticket
    uint32_t b, ep;
    lb_ticket(slot, chunks, b, ep);
    const uint32_t t0 = min(tiles, b * per), t1 = min(tiles, t0 + per);
    const uint32_t last = chunks - 1;
    uint32_t mine = 0;
    for (uint32_t tile = t0; tile < t1; ++tile) {
        const uint32_t q = lb + tile * 256 + threadIdx.x;
        if (q < le) mine += count_node(keys, t, q, level, max_depth, ref_leaf_rule, keep_whole, tab, tab_level);
    }
    uint32_t total;
    (void)block_exclusive_256(mine, 0u, ScanSum{}, total, sh4);
    uint32_t carry = lb_prefix(slot + 1, b, ep, total, 0u, ScanSum{});
    for (uint32_t tile = t0; tile < t1; ++tile) {
        const uint32_t q = lb + tile * 256 + threadIdx.x;
        // This thread wrote node q's count in pass 1.
        const uint32_t nc = q < le ? t.link[q].y : 0u;
        uint32_t tile_total;
        const uint32_t ex = block_exclusive_256(nc, 0u, ScanSum{}, tile_total, sh4);
        if (q < le) make_children(keys, t, q, level, (uint64_t)le + carry + ex, cap, tab, tab_level);
        carry += tile_total;
    }
    if (b == last && threadIdx.x == 0) {
        uint64_t end = (uint64_t)le + carry;
        if (end > cap) {
            info[1] = 1u;  // capacity exceeded: children past the capacity are not created
            end = cap;
        }
        lv[level + 2] = (uint32_t)end;
        info[0] = (uint32_t)end;
    }
}

// Centres of mass of one level, deepest level first: leaves from their particles, split nodes
// from their children (count-weighted), all in FP64.
__global__ void __launch_bounds__(256) k_com_level(const double *__restrict__ spos, TreeView t,
                                                   const uint32_t *__restrict__ lv, int level) {
    const uint32_t lb = lv[level], le = lv[level + 1];
    for (uint32_t q = lb + blockIdx.x * blockDim.x + threadIdx.x; q < le; q += gridDim.x * blockDim.x) {
        const uint4 l = t.link[q];
        const uint2 r = t.range[q];
        double sx = 0, sy = 0, sz = 0;
        if (l.y == 0) {
            // Two independent partial sums per axis, so consecutive loads need not wait on each other.
            double tx = 0, ty = 0, tz = 0;
            uint32_t j = r.x;
            for (; j + 2 <= r.y; j += 2) {
                sx += spos[3 * (size_t)j];
                sy += spos[3 * (size_t)j + 1];
                sz += spos[3 * (size_t)j + 2];
                tx += spos[3 * (size_t)j + 3];
                ty += spos[3 * (size_t)j + 4];
                tz += spos[3 * (size_t)j + 5];
            }
            if (j < r.y) {
                sx += spos[3 * (size_t)j];
                sy += spos[3 * (size_t)j + 1];
                sz += spos[3 * (size_t)j + 2];
            }
            sx += tx;
            sy += ty;
            sz += tz;
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
}

__global__ void k_tree_root(TreeView t, uint32_t n, float half, uint32_t *__restrict__ lv,
                            uint32_t *__restrict__ info) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    t.range[0] = make_uint2(0, n);
    t.box[0] = make_float4(half, half, half, half);
    t.link[0] = make_uint4(0xFFFFFFFFu, 0, 0, 0);
    lv[0] = 0;
    lv[1] = 1;
    for (int l = 2; l < kMaxDepth + 3; ++l) lv[l] = 1;
    info[0] = 1;
    info[1] = 0;
}

static inline unsigned blocks(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

int walk_group_level(const Ctx &c, size_t n) {
    // Group cubes hold about 512 particles at the mean density (8 particle spacings on a side
    // when mesh = particle grid), and are never larger than L/4.
    int level = 2;
    while (level < kMaxDepth && ((double)n / std::pow(8.0, level + 1)) >= 512.0) ++level;
    return std::max(2, std::min(kMaxDepth, level + c.group_level_shift));
}

int sort_depth(const Ctx &c, size_t n) {
    return std::min(kMaxDepth, std::max(c.tree_depth, std::max(c.dep_level, walk_group_level(c, n))));
}

void sort_particles(Ctx &c, const double *d_pos, size_t n) {
    // Only the key bits of the octree's levels are sorted (a shallower tree sorts fewer bits; the
    // sort is stable, so the order below that depth is the input order).
    morton_sort_positions(c, d_pos, n, sort_depth(c, n));
}

__global__ void k_cube_starts(const uint64_t *__restrict__ keys, uint32_t n, int shift, uint32_t cubes,
                              uint32_t *__restrict__ starts);

__global__ void k_fill_u32(uint32_t *__restrict__ p, size_t count, uint32_t v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}

// tab[c] = i for the first sorted particle i of every non-empty cube c (one writer per cube).
__global__ void k_cube_mark(const uint64_t *__restrict__ keys, uint32_t n, int shift, uint32_t *__restrict__ tab) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t c = keys[i] >> shift;
    if (i == 0 || (keys[i - 1] >> shift) != c) tab[c] = i;
}

// The table read back to front (suffix minima are prefix minima of the reversed table).
struct ReversedLoad {
    const uint32_t *p;
    size_t count;
    __device__ __forceinline__ uint32_t operator()(size_t i) const { return p[count - 1 - i]; }
};
struct ReversedStore {
    uint32_t *p;
    size_t count;
    __device__ __forceinline__ void operator()(size_t i, uint32_t v) const { p[count - 1 - i] = v; }
};

// Tree arrays for up to `cap` nodes (allocated once per capacity; never resized during a build).
static void ensure_tree(Ctx &c, size_t cap) {
    Tree &t = c.tree;
    if (cap <= t.cap) return;
    (void)hipFree(t.range);
    (void)hipFree(t.link);
    (void)hipFree(t.box);
    (void)hipFree(t.com);
    (void)hipFree(c.tree_scan);
    TPM_HIP(hipMalloc(&t.range, cap * sizeof(uint2)));
    TPM_HIP(hipMalloc(&t.link, cap * sizeof(uint4)));
    TPM_HIP(hipMalloc(&t.box, cap * sizeof(float4)));
    TPM_HIP(hipMalloc(&t.com, cap * sizeof(double4)));
    // scratch of the cube table's scan (cap), level table, info words, cube table (< cap)
    TPM_HIP(hipMalloc(&c.tree_scan, (2 * cap + kMaxDepth + 11) * sizeof(uint32_t)));
    t.cap = cap;
}

size_t tree_capacity(size_t n) { return std::max<size_t>(1024, 2 * n + 1024); }

void build_tree(Ctx &c, size_t n, bool group_walk_only) {
    Tree &t = c.tree;
    const int depth = sort_depth(c, n);  // the levels the sort ordered
    ensure_tree(c, tree_capacity(n));
    uint32_t *off = c.tree_scan;
    uint32_t *lv = off + t.cap;
    uint32_t *info = lv + kMaxDepth + 3;
    c.tree_info = info;
    // Root: the whole box, centre L/2 and half L/2 in FP32 (src/tree.rs:180-181).
    const float half = (float)c.p.box_size / 2.0f;
    hipLaunchKernelGGL(k_tree_root, dim3(1), dim3(64), 0, c.stream, t.view(), (uint32_t)n, half, lv, info);
    TPM_HIP(hipGetLastError());
    const uint32_t ref_rule = (c.p.flags & TPM_FLAG_REF_LEAF_RULE) ? 1u : 0u;
    const uint32_t keep_whole = group_walk_only ? kWalkLeaf : 0u;
    // The levels above T take their octant boundaries from a table of the first sorted index of
    // every level-T cube, instead of searching per node. T is the deepest level with at most n / 8
    // cubes; the table is built particle-parallel: every cube's first particle marks its entry
    // (entries preset to n), then a suffix minimum gives empty cubes the next cube's start.
    int tab_level = 0;
    while (tab_level < depth && std::pow(8.0, (double)(tab_level + 1)) * 8.0 <= (double)n) ++tab_level;
    uint32_t *tab = info + 8;
    c.tree_tab = tab;
    c.tree_tab_level = tab_level;
    c.tree_tab_n = n;
    if (tab_level > 0) {
        const size_t entries = ((size_t)1 << (3 * tab_level)) + 1;
        hipLaunchKernelGGL(k_fill_u32, dim3(blocks(entries, 256)), dim3(256), 0, c.stream, off, entries, (uint32_t)n);
        hipLaunchKernelGGL(k_cube_mark, dim3(blocks(n, 256)), dim3(256), 0, c.stream, c.keys, (uint32_t)n,
                           3 * (kKeyBits - tab_level), off);
        TPM_HIP(hipGetLastError());
        TPM_HIP(device_scan(ReversedLoad{off, entries}, ReversedStore{tab, entries}, entries, 0xFFFFFFFFu, ScanMin{},
                            true, c.scan_part, c.stream));
    }
    // A level holds at most min(8^level, n) nodes (each holds a particle); the node-parallel kernels
    // get about one thread per possible node, so the deep levels' latency-bound loops (key loads,
    // record loads) have their iterations in flight at once, and the shallow levels launch little.
    auto level_grid = [&](int level) {
        const double bound = std::min(std::pow(8.0, (double)level), (double)n);
        return (unsigned)std::max(1.0, std::min(std::ceil(bound / 256.0), 65536.0));
    };
    for (int level = 0; level <= depth; ++level) {
        // One workgroup per 256 possible nodes, at most kLbMaxBlocks (larger levels: several tiles
        // per workgroup).
        const unsigned grid = std::min<unsigned>(level_grid(level), kLbMaxBlocks);
        hipLaunchKernelGGL(k_build_level, dim3(grid), dim3(256), 0, c.stream, c.keys, t.view(), lv, level, depth,
                           ref_rule, keep_whole, tab, tab_level, (uint32_t)t.cap, info, c.scan_part, grid);
        TPM_HIP(hipGetLastError());
    }
    // Centres of mass, deepest level first (levels past the deepest are empty ranges).
    for (int level = depth + 1; level >= 0; --level) {
        hipLaunchKernelGGL(k_com_level, dim3(level_grid(level)), dim3(256), 0, c.stream, c.spos, t.view(), lv, level);
        TPM_HIP(hipGetLastError());
    }
}

// ------------------------------------------------------------------------- walk groups

// Segment starts (the first particle of each level-l cube) for the max-scan: i where the cube
// changes, else 0; the scan then gives every particle the first index of its cube.
struct GroupSegStart {
    const uint64_t *keys;
    int shift;
    __device__ uint32_t operator()(size_t i) const {
        return (i == 0 || (keys[i] >> shift) != (keys[i - 1] >> shift)) ? (uint32_t)i : 0u;
    }
};

// Group starts: every kGroup-th particle of a cube, counted from the cube's first one.
struct GroupChunkStart {
    const uint32_t *segstart;
    __device__ uint32_t operator()(size_t i) const {
        return ((uint32_t)i - segstart[i]) % (uint32_t)kGroup == 0 ? 1u : 0u;
    }
};

__global__ void k_group_starts(const uint32_t *__restrict__ segstart, const uint32_t *__restrict__ gid, size_t n,
                               uint32_t *__restrict__ gstart, uint32_t *__restrict__ ngroups) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t flag = ((uint32_t)i - segstart[i]) % (uint32_t)kGroup == 0 ? 1u : 0u;
    if (flag) gstart[gid[i]] = (uint32_t)i;
    if (i == n - 1) {
        gstart[gid[i] + flag] = (uint32_t)n;
        *ngroups = gid[i] + flag;
    }
}

// First sorted index of every level-l cube (starts[cubes] = n): binary search on the sorted keys.
__global__ void k_cube_starts(const uint64_t *__restrict__ keys, uint32_t n, int shift, uint32_t cubes,
                              uint32_t *__restrict__ starts) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > cubes) return;
    if (k == cubes) {
        starts[k] = n;
        return;
    }
    const uint64_t key = (uint64_t)k << shift;
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (keys[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    starts[k] = lo;
}

// Groups per cube: ceil(count / kGroup) (entry `cubes` is 0, so its exclusive scan is the total).
struct CubeGroups {
    const uint32_t *starts;  // entry k << sh is cube k's first particle
    uint32_t cubes;
    int sh;
    __device__ uint32_t operator()(size_t k) const {
        return k < cubes ? (starts[(k + 1) << sh] - starts[k << sh] + (uint32_t)kGroup - 1) / (uint32_t)kGroup : 0u;
    }
};

// Cube k's groups start every kGroup particles from its first; the last thread writes the end.
__global__ void k_cube_gstart(const uint32_t *__restrict__ starts, int sh, const uint32_t *__restrict__ gbase,
                              uint32_t cubes, uint32_t *__restrict__ gstart, uint32_t *__restrict__ ngroups) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= cubes) return;
    const uint32_t s0 = starts[(size_t)k << sh], s1 = starts[(size_t)(k + 1) << sh];
    uint32_t g = gbase[k];
    for (uint32_t p = s0; p < s1; p += (uint32_t)kGroup) gstart[g++] = p;
    if (k == cubes - 1) {
        gstart[gbase[cubes]] = starts[(size_t)cubes << sh];
        *ngroups = gbase[cubes];
    }
}

void build_groups(Ctx &c, size_t n, const uint64_t *keys) {
    if (!keys) keys = c.keys;
    if (n == 0) {
        zero_fill(c, c.ngroups_dev, sizeof(uint32_t));
        return;
    }
    const int level = walk_group_level(c, n);
    c.group_level = level;
    const int shift = 3 * (kKeyBits - level);
    const double cubes_d = std::pow(8.0, (double)level);
    if (cubes_d + 1.0 <= (double)c.max_n) {
        // Few cubes (about one per 64 or 512 particles): the first particle of every cube by binary
        // search, the groups per cube, an exclusive scan over the cubes and the group starts.
        const uint32_t cubes = (uint32_t)cubes_d;
        const uint32_t *starts = c.gscratch;
        int sh = 0;
        uint32_t *gbase = c.gscratch + c.max_n;
        if (keys == c.keys && n == c.tree_tab_n && level <= c.tree_tab_level && c.tree_tab) {
            // The octree build's table of the same or finer cubes: cube k starts at entry k << sh.
            starts = c.tree_tab;
            sh = 3 * (c.tree_tab_level - level);
        } else {
            hipLaunchKernelGGL(k_cube_starts, dim3(blocks((size_t)cubes + 1, 256)), dim3(256), 0, c.stream, keys,
                               (uint32_t)n, shift, cubes, c.gscratch);
        }
        TPM_HIP(device_scan(CubeGroups{starts, cubes, sh}, ScanStore{gbase}, (size_t)cubes + 1, 0u, ScanSum{}, false,
                            c.scan_part, c.stream));
        hipLaunchKernelGGL(k_cube_gstart, dim3(blocks(cubes, 256)), dim3(256), 0, c.stream, starts, sh, gbase, cubes,
                           c.gstart, c.ngroups_dev);
        TPM_HIP(hipGetLastError());
        return;
    }
    uint32_t *gid = c.gscratch;
    uint32_t *segstart = c.gscratch + c.max_n;
    // First index of each particle's cube (max-scan of the cube starts), then the group number of
    // each particle (sum-scan of the group starts); both scans read their input through iterators.
    TPM_HIP(device_scan(GroupSegStart{keys, shift}, ScanStore{segstart}, n, 0u, ScanMax{}, true, c.scan_part,
                        c.stream));
    TPM_HIP(device_scan(GroupChunkStart{segstart}, ScanStore{gid}, n, 0u, ScanSum{}, false, c.scan_part, c.stream));
    hipLaunchKernelGGL(k_group_starts, dim3(blocks(n, 256)), dim3(256), 0, c.stream, segstart, gid, n, c.gstart,
                       c.ngroups_dev);
    TPM_HIP(hipGetLastError());
}

}  // namespace tpm
