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

#include "tpm_internal.h"

namespace tpm {

// src/mesh.rs:29-59 stencil: cell-centred CIC with periodic wrap. T = float reproduces the
// reference's FP32 arithmetic on positions that were FP32; T = double is the FP64 form.
template <typename T>
struct Stencil {
    int64_t idx[8];
    T w[8];
};

template <typename T>
__device__ __forceinline__ void cic_stencil(T x0, T x1, T x2, int n, T box, Stencil<T> &st) {
    const T h = box / (T)n;
    const T xs[3] = {x0, x1, x2};
    int64_t base[3];
    T frac[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const T s = xs[k] / h - (T)0.5;
        const T i0 = floor(s);
        base[k] = (int64_t)i0;
        frac[k] = s - i0;
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
                int64_t ix = (base[0] + dx) % n;
                int64_t iy = (base[1] + dy) % n;
                int64_t iz = (base[2] + dz) % n;
                ix += ix < 0 ? n : 0;
                iy += iy < 0 ? n : 0;
                iz += iz < 0 ? n : 0;
                st.idx[c] = (ix * n + iy) * n + iz;
                st.w[c] = (wx * wy) * wz;
                ++c;
            }
}

template <typename T>
__global__ void __launch_bounds__(256) k_deposit(const double *__restrict__ pos, size_t np, int n, double box,
                                                 double mass, unsigned long long *__restrict__ grid) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    Stencil<T> st;
    cic_stencil<T>((T)pos[3 * i], (T)pos[3 * i + 1], (T)pos[3 * i + 2], n, (T)box, st);
    const double scale = mass * 0x1p32;
    const unsigned long long total = (unsigned long long)llrint(scale);
    unsigned long long used = 0;
#pragma unroll
    for (int c = 0; c < 7; ++c) {
        const unsigned long long q = (unsigned long long)llrint((double)st.w[c] * scale);
        used += q;
        atomicAdd(grid + st.idx[c], q);
    }
    // Remainder to the last corner; modular arithmetic keeps the cell sums exact even if a
    // rounding makes this quantum "negative".
    atomicAdd(grid + st.idx[7], total - used);
}

// Same quanta as k_deposit, accumulated per workgroup in an LDS tile first. A workgroup takes
// kDepChunk consecutive particles of the Morton order, which cover a compact block of cells; if
// that block (plus the +1 CIC corner) fits the tile, the 8 quanta of each particle are added
// with LDS atomics and each non-zero tile cell is then added to the grid with one global atomic.
// Chunks whose block does not fit (they straddle a large Morton boundary) add directly. Integer
// sums commute, so the grid is bit-identical to k_deposit's.
constexpr int kDepChunk = 1024;
constexpr int kDepTile = 4096;

template <typename T>
__global__ void __launch_bounds__(256) k_deposit_tiled(const double *__restrict__ pos, size_t np, int n, double box,
                                                       double mass, unsigned long long *__restrict__ grid) {
    __shared__ unsigned long long tile[kDepTile];
    __shared__ int red[6][8];
    const size_t c0 = (size_t)blockIdx.x * kDepChunk;
    const int tid = threadIdx.x;
    const T h = (T)box / (T)n;
    // Pass 1: bounding block of the stencil bases (unwrapped cell indices, base in [-1, n-1]).
    int mn[3] = {INT_MAX, INT_MAX, INT_MAX}, mx[3] = {INT_MIN, INT_MIN, INT_MIN};
    for (int k = tid; k < kDepChunk; k += 256) {
        const size_t i = c0 + k;
        if (i >= np) break;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const T sv = (T)pos[3 * i + a] / h - (T)0.5;
            const int b = (int)floor(sv);
            mn[a] = min(mn[a], b);
            mx[a] = max(mx[a], b);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        for (int d = 32; d >= 1; d >>= 1) {
            mn[a] = min(mn[a], __shfl_xor(mn[a], d, 64));
            mx[a] = max(mx[a], __shfl_xor(mx[a], d, 64));
        }
    }
    const int wave = tid >> 6, lane = tid & 63;
    if (lane == 0)
        for (int a = 0; a < 3; ++a) {
            red[a][wave] = mn[a];
            red[3 + a][wave] = mx[a];
        }
    __syncthreads();
    int lo[3], dim[3];
    for (int a = 0; a < 3; ++a) {
        int m0 = red[a][0], m1 = red[3 + a][0];
        for (int wv = 1; wv < 4; ++wv) {
            m0 = min(m0, red[a][wv]);
            m1 = max(m1, red[3 + a][wv]);
        }
        lo[a] = m0;
        dim[a] = m1 - m0 + 2;
    }
    const long long cells = (long long)dim[0] * dim[1] * dim[2];
    const double scale = mass * 0x1p32;
    const unsigned long long total = (unsigned long long)llrint(scale);
    const bool use_tile = cells <= kDepTile && dim[0] > 0;
    if (use_tile) {
        for (int c = tid; c < (int)cells; c += 256) tile[c] = 0ull;
        __syncthreads();
    }
    for (int k = tid; k < kDepChunk; k += 256) {
        const size_t i = c0 + k;
        if (i >= np) break;
        Stencil<T> st;
        cic_stencil<T>((T)pos[3 * i], (T)pos[3 * i + 1], (T)pos[3 * i + 2], n, (T)box, st);
        int b[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) b[a] = (int)floor((T)pos[3 * i + a] / h - (T)0.5) - lo[a];
        unsigned long long used = 0;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const unsigned long long q =
                c < 7 ? (unsigned long long)llrint((double)st.w[c] * scale) : total - used;
            used += q;
            if (use_tile) {
                const int dx = (c >> 2) & 1, dy = (c >> 1) & 1, dz = c & 1;
                const int li = ((b[0] + dx) * dim[1] + (b[1] + dy)) * dim[2] + (b[2] + dz);
                atomicAdd(&tile[li], q);
            } else {
                atomicAdd(grid + st.idx[c], q);
            }
        }
    }
    if (!use_tile) return;
    __syncthreads();
    for (int c = tid; c < (int)cells; c += 256) {
        const unsigned long long v = tile[c];
        if (!v) continue;
        const int iz = c % dim[2];
        const int iy = (c / dim[2]) % dim[1];
        const int ix = c / (dim[2] * dim[1]);
        int gx = (lo[0] + ix) % n, gy = (lo[1] + iy) % n, gz = (lo[2] + iz) % n;
        gx += gx < 0 ? n : 0;
        gy += gy < 0 ? n : 0;
        gz += gz < 0 ? n : 0;
        atomicAdd(grid + ((size_t)gx * n + gy) * n + gz, v);
    }
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
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    Stencil<T> st;
    cic_stencil<T>((T)pos[3 * i], (T)pos[3 * i + 1], (T)pos[3 * i + 2], n, (T)box, st);
    float w[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) w[c] = (float)st.w[c];
    const size_t o = order ? order[i] : i;
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) {
        const float *gc = g + comp * m;
        float s = 0.0f;
#pragma unroll
        for (int c = 0; c < 8; ++c) s = s + gc[st.idx[c]] * w[c];
        acc[3 * o + comp] = s;
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
        const unsigned nb = (unsigned)((n + kDepChunk - 1) / kDepChunk);
        if (f32_weights)
            hipLaunchKernelGGL(k_deposit_tiled<float>, dim3(nb), dim3(256), 0, c.stream, pos_src, n, c.n,
                               c.p.box_size, mass, c.grid);
        else
            hipLaunchKernelGGL(k_deposit_tiled<double>, dim3(nb), dim3(256), 0, c.stream, pos_src, n, c.n,
                               c.p.box_size, mass, c.grid);
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
        hipLaunchKernelGGL(k_interp3<float>, dim3(blocks(n, 256)), dim3(256), 0, c.stream, c.spos, c.idx, n, c.n,
                           c.p.box_size, c.force, c.m, d_acc);
    else
        hipLaunchKernelGGL(k_interp3<double>, dim3(blocks(n, 256)), dim3(256), 0, c.stream, c.spos, c.idx, n, c.n,
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

}  // namespace tpm
