// Cloud-in-cell deposit and interpolation on the periodic mesh.
//
// Replaces Mesh::deposit / Mesh::interpolate (src/mesh.rs:29-77) and the density contrast of
// Simulation::density_contrast (src/integrator.rs:117-126).
//
// Deposit: the reference adds FP32 weights serially, which fixes its summation order. Here
// every particle's eight weights are quantised to 2^-32 units of mass, the last corner taking
// the remainder so each particle adds exactly round(mass * 2^32), and the quanta are summed
// with 64-bit integer atomics. Integer addition commutes, so the grid is bit-exact from run to
// run whatever the order of the atomics, and the total mass is exact.
#include <algorithm>
#include <climits>
#include <cmath>
#include <mutex>

#include "tpm_internal.h"

namespace tpm {

// src/mesh.rs:29-59 stencil: cell-centred CIC with periodic wrap. T = float reproduces the
// reference's FP32 arithmetic on positions that were FP32; T = double is the FP64 form.
template <typename T>
struct Stencil {
    int64_t idx[8];
    T w[8];
};

// Periodic wrap of an unwrapped cell index (rem_euclid). For positions in [0, L) the stencil
// corners lie in [-1, n], so one compare-and-add covers them; anything else takes the modulo.
__device__ __forceinline__ int64_t wrap_cell(int64_t v, int n) {
    if (v >= 0 && v < n) return v;
    if (v == -1) return n - 1;
    if (v == n) return 0;
    const int64_t r = v % n;
    return r < 0 ? r + n : r;
}

// Stencil base (unwrapped) and fraction on one axis: s = x / h - 0.5, base = floor(s).
template <typename T>
__device__ __forceinline__ void cic_axis(T x, T h, int64_t &base, T &frac) {
    const T s = x / h - (T)0.5;
    const T i0 = floor(s);
    base = (int64_t)i0;
    frac = s - i0;
}

template <typename T>
__device__ __forceinline__ void cic_corners(const int64_t base[3], const T frac[3], int n, Stencil<T> &st) {
    int64_t ix[2], iy[2], iz[2];
#pragma unroll
    for (int d = 0; d < 2; ++d) {
        ix[d] = wrap_cell(base[0] + d, n);
        iy[d] = wrap_cell(base[1] + d, n);
        iz[d] = wrap_cell(base[2] + d, n);
    }
    int c = 0;
#pragma unroll
    for (int dx = 0; dx < 2; ++dx)
#pragma unroll
        for (int dy = 0; dy < 2; ++dy)
#pragma unroll
            for (int dz = 0; dz < 2; ++dz) {
                const T wx = dx == 0 ? (T)1 - frac[0] : frac[0];
                const T wy = dy == 0 ? (T)1 - frac[1] : frac[1];
                const T wz = dz == 0 ? (T)1 - frac[2] : frac[2];
                st.idx[c] = (ix[dx] * n + iy[dy]) * n + iz[dz];
                st.w[c] = (wx * wy) * wz;
                ++c;
            }
}

template <typename T>
__device__ __forceinline__ void cic_stencil(T x0, T x1, T x2, int n, T box, Stencil<T> &st) {
    const T h = box / (T)n;
    int64_t base[3];
    T frac[3];
    cic_axis<T>(x0, h, base[0], frac[0]);
    cic_axis<T>(x1, h, base[1], frac[1]);
    cic_axis<T>(x2, h, base[2], frac[2]);
    cic_corners<T>(base, frac, n, st);
}

// Quantum of one corner: the nearest integer (ties to even, as llrint) of x = w * mass * 2^32.
// When 0 <= mass * 2^32 < 2^52 (`small`: uniform across the launch), x + 2^52 is that integer
// plus 2^52, exactly, in the low mantissa bits of the FP64 sum: one addition and one 64-bit
// subtraction instead of llrint's conversion sequence, with the same result.
__device__ __forceinline__ unsigned long long quanta(double x, bool small) {
    if (small) return (unsigned long long)__double_as_longlong(x + 0x1p52) - 0x4330000000000000ull;
    return (unsigned long long)llrint(x);
}

template <typename T>
__global__ void __launch_bounds__(256) k_deposit(const double *__restrict__ pos, size_t np, int n, double box,
                                                 double mass, unsigned long long *__restrict__ grid) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    Stencil<T> st;
    cic_stencil<T>((T)pos[3 * i], (T)pos[3 * i + 1], (T)pos[3 * i + 2], n, (T)box, st);
    const double scale = mass * 0x1p32;
    const bool small = scale >= 0.0 && scale < 0x1p52;
    const unsigned long long total = (unsigned long long)llrint(scale);
    unsigned long long used = 0;
#pragma unroll
    for (int c = 0; c < 7; ++c) {
        const unsigned long long q = quanta((double)st.w[c] * scale, small);
        used += q;
        atomicAdd(grid + st.idx[c], q);
    }
    // Remainder to the last corner; modular arithmetic keeps the cell sums exact even if a
    // rounding makes this quantum "negative".
    atomicAdd(grid + st.idx[7], total - used);
}

// Same quanta as k_deposit, accumulated per workgroup in an LDS tile first. A workgroup of
// kDepThreads takes up to kDepItem particles of one aligned Morton cube of the sorted order
// (cubes of at most kDepSide cells on a side, so the CIC footprint, with a margin of two cells for
// the rounding of the keys, always fits the tile; dense cubes are split into several items), adds
// their 8 quanta with LDS atomics and then adds each non-zero tile cell to the grid with one
// global atomic. A particle whose stencil lies outside the tile (only possible for a position
// outside [0, L)) adds directly to the grid. Integer sums commute, so the grid is bit-identical to
// k_deposit's for any order and any split. A cube of 16 cells a side holds about 4,096 particles
// at the mean density: its tile of at most 21^3 cells (74 KB) is cleared
// and flushed once per 4,096 particles, about 1.7 tile cells per particle against 3.4 for cubes of
// 8 cells, and the workgroup's fixed costs (the cube's bounds, the flush's reciprocals) are spread
// over eight times as many particles.
constexpr int kDepTileSide = kDepSide + 5;  // 2 cells below, 1 above, 1 for a ceil, 1 stencil corner
constexpr int kDepTile = kDepTileSide * kDepTileSide * kDepTileSide;
constexpr int kDepThreads = 1024;
constexpr int kDepBatch = 2;
constexpr uint32_t kDepItem = 2u * kDepSide * kDepSide * kDepSide;  // 8 particles per thread
constexpr size_t kDepExtra = 1024;  // workgroups for the dense cubes' extra items

// First sorted index of every level-`level` Morton cube (starts[nodes] = np).
__global__ void k_node_starts(const uint64_t *__restrict__ keys, size_t np, int shift, uint32_t nodes,
                              uint32_t *__restrict__ starts) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nodes) return;
    if (k == nodes) {
        starts[k] = (uint32_t)np;
        return;
    }
    const uint64_t key = (uint64_t)k << shift;
    size_t lo = 0, hi = np;
    while (lo < hi) {
        const size_t mid = (lo + hi) >> 1;
        if (keys[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    starts[k] = (uint32_t)lo;
}

// Extra work items per cube beyond its first: ceil(count / kDepItem) - 1, 0 for an empty cube
// (entry `nodes` is 0, so its exclusive scan is the total).
__global__ void k_node_items(const uint32_t *__restrict__ starts, uint32_t nodes, uint32_t *__restrict__ items) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nodes) return;
    const uint32_t cnt = k == nodes ? 0u : starts[k + 1] - starts[k];
    items[k] = cnt > kDepItem ? (cnt - 1) / kDepItem : 0u;
}

__device__ __forceinline__ uint32_t morton_axis(uint32_t k, int axis, int level) {
    uint32_t v = 0;
    for (int b = 0; b < level; ++b) v |= ((k >> (3 * b + axis)) & 1u) << b;
    return v;
}

// One work item of the deposit, by the whole workgroup (uniform control flow; the tile is free
// again when it returns). Item b < nodes is the first item of cube b; item nodes + e is extra item
// e of the dense cubes: extra item e belongs to the cube k with item_start[k] <= e <
// item_start[k + 1] (item_start: exclusive scan of the extra items per cube), as its item
// 1 + e - item_start[k].
template <typename T>
__device__ __forceinline__ void deposit_item(uint32_t b, unsigned long long *tile, const double *__restrict__ pos,
                                             const uint32_t *__restrict__ starts,
                                             const uint32_t *__restrict__ item_start, uint32_t nodes, int level,
                                             int n, double box, double mass, unsigned long long *__restrict__ grid) {
    uint32_t k, sub;
    if (b < nodes) {
        k = b;
        sub = 0;
        if (starts[k] == starts[k + 1]) return;
    } else {
        const uint32_t e = b - nodes;
        if (e >= item_start[nodes]) return;
        uint32_t lo_k = 0, hi_k = nodes;  // last k in [0, nodes) with item_start[k] <= e
        while (lo_k < hi_k) {
            const uint32_t mid = (lo_k + hi_k + 1) >> 1;
            if (item_start[mid] <= e)
                lo_k = mid;
            else
                hi_k = mid - 1;
        }
        k = lo_k;
        sub = 1 + e - item_start[k];
    }
    const uint32_t p0 = starts[k] + sub * kDepItem;
    const uint32_t p1 = min(starts[k + 1], p0 + kDepItem);
    const int tid = threadIdx.x;
    const T h = (T)box / (T)n;
    // Cells of the cube [k_a n / 2^level, (k_a + 1) n / 2^level), widened by two cells below and
    // one above (plus the +1 stencil corner).
    int lo[3], dim[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int64_t ka = morton_axis(k, a, level);
        lo[a] = (int)((ka * n) >> level) - 2;
        const int hi = (int)(((ka + 1) * n + (1 << level) - 1) >> level) + 1;
        dim[a] = hi - lo[a] + 1;
    }
    const int cells = dim[0] * dim[1] * dim[2];
    for (int c = tid; c < cells; c += kDepThreads) tile[c] = 0ull;
    __syncthreads();
    const double scale = mass * 0x1p32;
    const bool small = scale >= 0.0 && scale < 0x1p52;
    const unsigned long long total = (unsigned long long)llrint(scale);
    if (starts[k + 1] - starts[k] > kDepItem) {
        // A dense cube (more than one item: at least two particles per cell on average). Each
        // thread takes a contiguous run of the sorted order, whose consecutive particles in dense
        // regions share their CIC base cell, and sums their quanta per corner in registers, adding
        // them to the tile only when the base cell changes: the LDS atomics of particles crowded
        // into few cells, which serialise on equal addresses, become one per run. Integer sums:
        // the grid is bit-identical to adding every quantum on its own.
        const uint32_t run = (p1 - p0 + kDepThreads - 1) / kDepThreads;
        const uint32_t q0 = p0 + tid * run, q1 = min(p1, q0 + run);
        const int s1 = dim[2], s0 = dim[1] * dim[2];
        const int corner[8] = {0, 1, s1, s1 + 1, s0, s0 + 1, s0 + s1, s0 + s1 + 1};
        int cur = -1;
        unsigned long long sum[8];
        for (uint32_t i = q0; i < q1; ++i) {
            int64_t base[3];
            T frac[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) cic_axis<T>((T)pos[3 * (size_t)i + a], h, base[a], frac[a]);
            bool inside = true;
            int off[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int64_t o = base[a] - lo[a];
                inside = inside && o >= 0 && o <= dim[a] - 2;
                off[a] = (int)o;
            }
            Stencil<T> st;
            if (!inside) cic_corners<T>(base, frac, n, st);
            const int t0 = inside ? (off[0] * dim[1] + off[1]) * dim[2] + off[2] : -1;
            if (inside && t0 != cur) {
                if (cur >= 0)
#pragma unroll
                    for (int c = 0; c < 8; ++c) atomicAdd(&tile[cur + corner[c]], sum[c]);
                cur = t0;
#pragma unroll
                for (int c = 0; c < 8; ++c) sum[c] = 0ull;
            }
            unsigned long long used = 0;
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const int dx = (c >> 2) & 1, dy = (c >> 1) & 1, dz = c & 1;
                const T wx = dx == 0 ? (T)1 - frac[0] : frac[0];
                const T wy = dy == 0 ? (T)1 - frac[1] : frac[1];
                const T wz = dz == 0 ? (T)1 - frac[2] : frac[2];
                const T w = (wx * wy) * wz;
                const unsigned long long q = c < 7 ? quanta((double)w * scale, small) : total - used;
                used += q;
                if (inside)
                    sum[c] += q;
                else
                    atomicAdd(grid + st.idx[c], q);
            }
        }
        if (cur >= 0)
#pragma unroll
            for (int c = 0; c < 8; ++c) atomicAdd(&tile[cur + corner[c]], sum[c]);
    } else {
        const int s1 = dim[2], s0 = dim[1] * dim[2];
        auto deposit_one = [&](double x0, double x1, double x2) {
            int64_t base[3];
            T frac[3];
            cic_axis<T>((T)x0, h, base[0], frac[0]);
            cic_axis<T>((T)x1, h, base[1], frac[1]);
            cic_axis<T>((T)x2, h, base[2], frac[2]);
            bool inside = true;
            int off[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const int64_t o = base[a] - lo[a];
                inside = inside && o >= 0 && o <= dim[a] - 2;
                off[a] = (int)o;
            }
            Stencil<T> st;
            if (!inside) cic_corners<T>(base, frac, n, st);
            const int t0 = (off[0] * dim[1] + off[1]) * dim[2] + off[2];
            unsigned long long used = 0;
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const int dx = (c >> 2) & 1, dy = (c >> 1) & 1, dz = c & 1;
                const T wx = dx == 0 ? (T)1 - frac[0] : frac[0];
                const T wy = dy == 0 ? (T)1 - frac[1] : frac[1];
                const T wz = dz == 0 ? (T)1 - frac[2] : frac[2];
                const T w = (wx * wy) * wz;
                const unsigned long long q = c < 7 ? quanta((double)w * scale, small) : total - used;
                used += q;
                if (inside)
                    atomicAdd(&tile[t0 + dx * s0 + dy * s1 + dz], q);
                else
                    atomicAdd(grid + st.idx[c], q);
            }
        };
        // kDepBatch particles per thread and round (a cube at the mean density is one round), all
        // loaded before any is deposited, so that their global loads are in flight together (the
        // indices past the item clamped into it).
        for (uint32_t i = p0 + tid; i < p1; i += kDepBatch * kDepThreads) {
            double x[kDepBatch][3];
#pragma unroll
            for (int u = 0; u < kDepBatch; ++u) {
                const size_t j = min(i + u * kDepThreads, p1 - 1);
                x[u][0] = pos[3 * j];
                x[u][1] = pos[3 * j + 1];
                x[u][2] = pos[3 * j + 2];
            }
#pragma unroll
            for (int u = 0; u < kDepBatch; ++u)
                if (i + u * kDepThreads < p1) deposit_one(x[u][0], x[u][1], x[u][2]);
        }
    }
    __syncthreads();
    // Tile cell c -> (ix, iy, iz) by multiply-high with per-workgroup reciprocals m = ceil(2^32 / d),
    // exact here (c * (m d - 2^32) < c d < 2^32 for c < 21^3, d <= 21^2); the tile spans at most 21
    // cells per axis, from 2 below the mesh, so for meshes of 32 cells or more one conditional add
    // or subtract wraps it.
    const uint32_t d12 = (uint32_t)(dim[1] * dim[2]), d2 = (uint32_t)dim[2];
    const uint32_t m12 = (uint32_t)((0x100000000ull + d12 - 1) / d12), m2 = (uint32_t)((0x100000000ull + d2 - 1) / d2);
    const bool small_mesh = n < 32;
    for (int c = tid; c < cells; c += kDepThreads) {
        const unsigned long long v = tile[c];
        if (!v) continue;
        const uint32_t ix = __umulhi((uint32_t)c, m12);
        const uint32_t r = (uint32_t)c - ix * d12;
        const uint32_t iy = __umulhi(r, m2);
        const uint32_t iz = r - iy * d2;
        int gx = lo[0] + (int)ix, gy = lo[1] + (int)iy, gz = lo[2] + (int)iz;
        if (small_mesh) {
            gx = (int)wrap_cell(gx, n);
            gy = (int)wrap_cell(gy, n);
            gz = (int)wrap_cell(gz, n);
        } else {
            gx += gx < 0 ? n : (gx >= n ? -n : 0);
            gy += gy < 0 ? n : (gy >= n ? -n : 0);
            gz += gz < 0 ? n : (gz >= n ? -n : 0);
        }
        atomicAdd(grid + (size_t)(gx * n + gy) * (size_t)n + (size_t)gz, v);
    }
    __syncthreads();  // the tile's reads are done before the next item clears it
}

// Workgroup b < nodes takes the first item of cube xcd_block(b, nodes) (each XCD a contiguous
// range of cubes); the remaining workgroups take the dense cubes' extra items in turn, as many as
// the device-side count says (none for an input without dense cubes), so the launch needs no
// workgroup per possible extra item.
template <typename T>
__global__ void __launch_bounds__(kDepThreads) k_deposit_cube(const double *__restrict__ pos,
                                                              const uint32_t *__restrict__ starts,
                                                              const uint32_t *__restrict__ item_start, uint32_t nodes,
                                                              int level, int n, double box, double mass,
                                                              unsigned long long *__restrict__ grid) {
    __shared__ unsigned long long tile[kDepTile];
    const uint32_t b = blockIdx.x;
    if (b < nodes) {
        deposit_item<T>(xcd_block(b, nodes), tile, pos, starts, item_start, nodes, level, n, box, mass, grid);
        return;
    }
    const uint32_t extra = item_start[nodes];
    for (uint32_t e = b - nodes; e < extra; e += gridDim.x - nodes)
        deposit_item<T>(nodes + e, tile, pos, starts, item_start, nodes, level, n, box, mass, grid);
}

__global__ void k_grid_to_float(const unsigned long long *__restrict__ grid, size_t m, float mean32,
                                float *__restrict__ delta, float *__restrict__ mass) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const float c = (float)((double)grid[i] * 0x1p-32);
    if (mass) mass[i] = c;
    if (delta) delta[i] = c / mean32 - 1.0f;
}

// src/mesh.rs:72-77 interpolate (sum over the 8 corners in stencil order) for the three force
// grids, at sorted positions; results scattered to the caller's order.
template <typename T>
__global__ void __launch_bounds__(256) k_interp3(const double *__restrict__ pos, const uint32_t *__restrict__ order,
                                                 size_t np, int n, double box, const float *__restrict__ g, size_t m,
                                                 float *__restrict__ acc) {
    // Two particles per thread (256 apart) with their gathers in flight together. g holds
    // (x, y, z) per mesh point: one record per corner.
    const size_t base = (size_t)xcd_block(blockIdx.x, gridDim.x) * 512 + threadIdx.x;
    if (base >= np) return;
    size_t ii[2] = {base, min(base + 256, np - 1)};
    Stencil<T> st[2];
#pragma unroll
    for (int u = 0; u < 2; ++u)
        cic_stencil<T>((T)pos[3 * ii[u]], (T)pos[3 * ii[u] + 1], (T)pos[3 * ii[u] + 2], n, (T)box, st[u]);
    float v[2][8][3];
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const float *gc = g + 3 * (size_t)st[u].idx[c];
            v[u][c][0] = gc[0];
            v[u][c][1] = gc[1];
            v[u][c][2] = gc[2];
        }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        if (u == 1 && base + 256 >= np) break;
        float sx = 0.0f, sy = 0.0f, sz = 0.0f;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const float w = (float)st[u].w[c];
            sx = sx + v[u][c][0] * w;
            sy = sy + v[u][c][1] * w;
            sz = sz + v[u][c][2] * w;
        }
        const size_t o = order ? order[ii[u]] : ii[u];
        acc[3 * o] = sx;
        acc[3 * o + 1] = sy;
        acc[3 * o + 2] = sz;
    }
}

template <typename T>
__global__ void k_interp1(const double *__restrict__ pos, size_t np, int n, double box,
                          const float *__restrict__ cells, float *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    Stencil<T> st;
    cic_stencil<T>((T)pos[3 * i], (T)pos[3 * i + 1], (T)pos[3 * i + 2], n, (T)box, st);
    float s = 0.0f;
#pragma unroll
    for (int c = 0; c < 8; ++c) s = s + cells[st.idx[c]] * (float)st.w[c];
    out[i] = s;
}

// FP64 interpolation of the judge's force grids at listed sorted indices.
__global__ void k_interp_judge(const double *__restrict__ pos, const uint32_t *__restrict__ targets, size_t mt, int n,
                               double box, const double *__restrict__ g, size_t m, double *__restrict__ out) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= mt) return;
    const size_t i = targets[t];
    Stencil<double> st;
    cic_stencil<double>(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], n, box, st);
    for (int comp = 0; comp < 3; ++comp) {
        double s = 0.0;
        for (int c = 0; c < 8; ++c) s += g[comp * m + st.idx[c]] * st.w[c];
        out[3 * t + comp] = s;
    }
}

// Per-step outputs (src/runner.rs:262-284 snapshot, src/frame.rs:56-64 rsd_shift) for FP64 state:
// positions rounded to float, speed |p| / a, density max(1 + delta(x), 0) by the FP64 stencil with
// weights rounded to float and summed in float (the FP64 interpolation convention), redshift-space
// shift p / (a^2 E). Outputs may be null.
__global__ void k_snapshot(const double *__restrict__ pos, const double *__restrict__ mom, size_t np, double a,
                           double inv_a2e, int n, double box, const float *__restrict__ delta,
                           float *__restrict__ opos, float *__restrict__ ospeed, float *__restrict__ odens,
                           float *__restrict__ orsd) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const double x0 = pos[3 * i], x1 = pos[3 * i + 1], x2 = pos[3 * i + 2];
    if (opos) {
        opos[3 * i] = (float)x0;
        opos[3 * i + 1] = (float)x1;
        opos[3 * i + 2] = (float)x2;
    }
    if (ospeed || orsd) {
        const double p0 = mom[3 * i], p1 = mom[3 * i + 1], p2 = mom[3 * i + 2];
        if (ospeed) ospeed[i] = (float)(sqrt(p0 * p0 + p1 * p1 + p2 * p2) / a);
        if (orsd) {
            orsd[3 * i] = (float)(p0 * inv_a2e);
            orsd[3 * i + 1] = (float)(p1 * inv_a2e);
            orsd[3 * i + 2] = (float)(p2 * inv_a2e);
        }
    }
    if (odens) {
        Stencil<double> st;
        cic_stencil<double>(x0, x1, x2, n, box, st);
        float s = 0.0f;
#pragma unroll
        for (int c = 0; c < 8; ++c) s = s + delta[st.idx[c]] * (float)st.w[c];
        odens[i] = fmaxf(1.0f + s, 0.0f);
    }
}

// src/frame.rs:33-52 downsample_grid: texel (x, y, z) of the m^3 output (x fastest) is 1 + the mean
// of the f^3 mesh block at (x f, y f, z f), summed in the reference's loop order in float.
__global__ void k_downsample(const float *__restrict__ delta, int n, int f, int m, float norm,
                             float *__restrict__ out) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t mm = (size_t)m * m * m;
    if (t >= mm) return;
    const int x = (int)(t % m), y = (int)((t / m) % m), z = (int)(t / ((size_t)m * m));
    float sum = 0.0f;
    for (int dx = 0; dx < f; ++dx)
        for (int dy = 0; dy < f; ++dy)
            for (int dz = 0; dz < f; ++dz) {
                const size_t ix = (size_t)x * f + dx, iy = (size_t)y * f + dy, iz = (size_t)z * f + dz;
                sum += delta[(ix * n + iy) * n + iz];
            }
    out[t] = 1.0f + sum * norm;
}

__global__ void k_zero32(uint32_t *__restrict__ p, size_t count) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

__global__ void k_f32_to_f64(const float *__restrict__ src, double *__restrict__ dst, size_t count) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) dst[i] = (double)src[i];
}

__global__ void k_check_finite(const float *__restrict__ v, size_t count, int *flag) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count && !isfinite(v[i])) *flag = 1;
}

static inline unsigned blocks(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

void deposit_fixed(Ctx &c, size_t n, double mass, bool f32_weights, const double *pos_src, bool sorted) {
    zero_fill(c, c.grid, c.m * sizeof(unsigned long long));
    if (n == 0) return;
    if (sorted) {
        // Morton cubes of at most 8 cells on a side; the sorted keys are those of sort_particles.
        const uint32_t nodes = 1u << (3 * c.dep_level);
        const int shift = 3 * (kKeyBits - c.dep_level);
        uint32_t *items = c.dep_start + nodes + 1;
        uint32_t *item_start = items + nodes + 1;
        hipLaunchKernelGGL(k_node_starts, dim3(blocks(nodes + 1, 256)), dim3(256), 0, c.stream, c.keys, n, shift,
                           nodes, c.dep_start);
        hipLaunchKernelGGL(k_node_items, dim3(blocks(nodes + 1, 256)), dim3(256), 0, c.stream, c.dep_start, nodes,
                           items);
        exclusive_sum_u32(c, items, item_start, nodes + 1);
        // One workgroup per cube for its first item, plus up to kDepExtra for the at most
        // n / kDepItem extra items.
        const size_t grid_items = (size_t)nodes + std::min<size_t>(n / kDepItem + 1, kDepExtra);
        if (f32_weights)
            hipLaunchKernelGGL(k_deposit_cube<float>, dim3((unsigned)grid_items), dim3(kDepThreads), 0, c.stream,
                               pos_src, c.dep_start, item_start, nodes, c.dep_level, c.n, c.p.box_size, mass, c.grid);
        else
            hipLaunchKernelGGL(k_deposit_cube<double>, dim3((unsigned)grid_items), dim3(kDepThreads), 0, c.stream,
                               pos_src, c.dep_start, item_start, nodes, c.dep_level, c.n, c.p.box_size, mass, c.grid);
        TPM_HIP(hipGetLastError());
        return;
    }
    if (f32_weights)
        hipLaunchKernelGGL(k_deposit<float>, dim3(blocks(n, 256)), dim3(256), 0, c.stream, pos_src, n, c.n,
                           c.p.box_size, mass, c.grid);
    else
        hipLaunchKernelGGL(k_deposit<double>, dim3(blocks(n, 256)), dim3(256), 0, c.stream, pos_src, n, c.n,
                           c.p.box_size, mass, c.grid);
    TPM_HIP(hipGetLastError());
}

void grid_to_float(Ctx &c, size_t n_particles, float *delta_out, float *mass_out) {
    const float mean32 = (float)n_particles / (float)c.m;
    hipLaunchKernelGGL(k_grid_to_float, dim3(blocks(c.m, 256)), dim3(256), 0, c.stream, c.grid, c.m, mean32,
                       delta_out, mass_out);
    TPM_HIP(hipGetLastError());
}

void interp_force(Ctx &c, size_t n, bool f32_weights, float *d_acc) {
    if (n == 0) return;
    if (f32_weights)
        hipLaunchKernelGGL(k_interp3<float>, dim3(blocks(n, 512)), dim3(256), 0, c.stream, c.spos, c.idx, n, c.n,
                           c.p.box_size, c.force, c.m, d_acc);
    else
        hipLaunchKernelGGL(k_interp3<double>, dim3(blocks(n, 512)), dim3(256), 0, c.stream, c.spos, c.idx, n, c.n,
                           c.p.box_size, c.force, c.m, d_acc);
    TPM_HIP(hipGetLastError());
}

void interp_one(Ctx &c, const float *d_cells, const double *d_pos, size_t n, bool f32_weights, float *d_out) {
    if (n == 0) return;
    if (f32_weights)
        hipLaunchKernelGGL(k_interp1<float>, dim3(blocks(n, 256)), dim3(256), 0, c.stream, d_pos, n, c.n,
                           c.p.box_size, d_cells, d_out);
    else
        hipLaunchKernelGGL(k_interp1<double>, dim3(blocks(n, 256)), dim3(256), 0, c.stream, d_pos, n, c.n,
                           c.p.box_size, d_cells, d_out);
    TPM_HIP(hipGetLastError());
}

void interp_judge_f64(Ctx &c, const uint32_t *d_targets_sorted, size_t m, double *d_out) {
    if (m == 0) return;
    hipLaunchKernelGGL(k_interp_judge, dim3(blocks(m, 256)), dim3(256), 0, c.stream, c.spos, d_targets_sorted, m, c.n,
                       c.p.box_size, c.dforce, c.m, d_out);
    TPM_HIP(hipGetLastError());
}

void snapshot(Ctx &c, const double *pos, const double *mom, size_t n, double a, double e, const float *delta,
              float *opos, float *ospeed, float *odens, float *orsd) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_snapshot, dim3(blocks(n, 256)), dim3(256), 0, c.stream, pos, mom, n, a, 1.0 / (a * a * e),
                       c.n, c.p.box_size, delta, opos, ospeed, odens, orsd);
    TPM_HIP(hipGetLastError());
}

int downsample_side(int n, int *f_out) {
    int f = (n + 127) / 128;
    if (f < 1) f = 1;
    if (f_out) *f_out = f;
    return n / f;
}

void downsample_grid(Ctx &c, const float *delta, float *out) {
    int f = 1;
    const int m = downsample_side(c.n, &f);
    const float norm = 1.0f / (float)(f * f * f);
    const size_t mm = (size_t)m * m * m;
    hipLaunchKernelGGL(k_downsample, dim3(blocks(mm, 256)), dim3(256), 0, c.stream, delta, c.n, f, m, norm, out);
    TPM_HIP(hipGetLastError());
}

void zero_fill(Ctx &c, void *ptr, size_t bytes) {
    const size_t words = bytes / 4;
    if (words == 0) return;
    const unsigned grid = (unsigned)std::min<size_t>(blocks(words, 256), 8192);
    hipLaunchKernelGGL(k_zero32, dim3(grid), dim3(256), 0, c.stream, (uint32_t *)ptr, words);
    TPM_HIP(hipGetLastError());
}

void convert_f32_to_f64(Ctx &c, const float *src, double *dst, size_t count) {
    if (count == 0) return;
    hipLaunchKernelGGL(k_f32_to_f64, dim3(blocks(count, 256)), dim3(256), 0, c.stream, src, dst, count);
    TPM_HIP(hipGetLastError());
}

void check_finite(Ctx &c, const float *d_acc, size_t count) {
    zero_fill(c, c.flag, sizeof(int));
    if (count == 0) return;
    hipLaunchKernelGGL(k_check_finite, dim3(blocks(count, 256)), dim3(256), 0, c.stream, d_acc, count, c.flag);
    TPM_HIP(hipGetLastError());
}

// ------------------------------------------------------------- deposit without allocation

// The fixed-point mass grid of tpm_deposit_dev_f64_noalloc: a file-scope device array of 16 MiB
// (meshes of up to 128 cells a side) instead of memory the library allocates. Each GPU holds its
// own copy from the first use until the process ends. On the host the array's name is not a
// device address: each device's address comes from hipGetSymbolAddress, looked up once per
// device (with that device current) on the first call there, and only that address reaches
// kernels. Every call on one GPU shares the array, so calls that overlap on one GPU race.
constexpr int kNoAllocMaxMesh = 128;
static __device__ __align__(256) unsigned long long scratch[(size_t)kNoAllocMaxMesh * kNoAllocMaxMesh *
                                                             kNoAllocMaxMesh];

// Address of `scratch` on the current device (null and a message on failure).
static unsigned long long *scratch_on_current_device(const char **why) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static unsigned long long *addr[kMaxDevices] = {};
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        *why = "no current HIP device";
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            *why = "hipGetSymbolAddress of the deposit scratch grid failed";
            return nullptr;
        }
        addr[dev] = static_cast<unsigned long long *>(p);
    }
    return addr[dev];
}

static int noalloc_fail(int code, const char *msg) {
    set_global_error(msg);
    return code;
}

}  // namespace tpm

extern "C" int tpm_deposit_dev_f64_noalloc(const double *d_pos, size_t n, int32_t mesh_n, double box_size,
                                           double mass, float *d_cells, void *stream) {
    using namespace tpm;
    if (mesh_n < 2 || mesh_n > kNoAllocMaxMesh)
        return noalloc_fail(TPM_E_INVALID, "mesh_n must be in [2, 128] for the allocation-free deposit");
    if (!(box_size > 0.0) || !std::isfinite(box_size)) return noalloc_fail(TPM_E_INVALID, "box_size must be positive");
    if (!std::isfinite(mass)) return noalloc_fail(TPM_E_INVALID, "mass must be finite");
    if (!d_cells || (n > 0 && !d_pos)) return noalloc_fail(TPM_E_INVALID, "null device pointer");
    if (n > (size_t)UINT32_MAX) return noalloc_fail(TPM_E_INVALID, "at most 2^32 - 1 particles");
    const char *why = "";
    unsigned long long *grid = scratch_on_current_device(&why);
    if (!grid) return noalloc_fail(TPM_E_DEVICE, why);
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const size_t m = (size_t)mesh_n * mesh_n * mesh_n;
    // The grid keeps the previous call's sums: clear the cells this call uses (a fill kernel, not
    // a memset, so the call can be captured into a HIP graph).
    const size_t words = 2 * m;
    hipLaunchKernelGGL(k_zero32, dim3((unsigned)std::min<size_t>(blocks(words, 256), 8192)), dim3(256), 0, s,
                       reinterpret_cast<uint32_t *>(grid), words);
    if (n > 0)
        hipLaunchKernelGGL(k_deposit<double>, dim3(blocks(n, 256)), dim3(256), 0, s, d_pos, n, (int)mesh_n, box_size,
                           mass, grid);
    hipLaunchKernelGGL(k_grid_to_float, dim3(blocks(m, 256)), dim3(256), 0, s, grid, m, 1.0f, (float *)nullptr,
                       d_cells);
    if (hipGetLastError() != hipSuccess) return noalloc_fail(TPM_E_DEVICE, "kernel launch failed");
    return TPM_OK;
}
