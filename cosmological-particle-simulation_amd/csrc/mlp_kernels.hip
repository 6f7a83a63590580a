// Learned short-range correction: a 4-layer MLP (64 -> 128 -> 128 -> 128 -> 4, SiLU) with BF16
// weights and activations on MFMA and FP32 accumulation, trained on the GPU by its own forward
// and backward kernels. No reference counterpart (the reference has only the exact tree walk);
// the walk supplies its training targets, its fallback and the judge of its error.
//
// Per particle the input is 33 features, zero-padded to 64: log(1 + delta) of the particle's cell and,
// relative to it, of the 26 cells around it (the 3x3x3 stencil; differences keep the gradients that
// BF16 would round away on top of a common level), the PM force at the particle and the particle's
// offset from its cell centre (cell units). The inference runs all of it as one kernel
// (k_mlp_fused, below the training kernels). Forces enter and leave in the compressed units slog(F / (K / h^2)),
// slog(x) = sign(x) log(1 + |x|), since short-range forces in dense regions span many decades.
// Outputs: the short-range force (3 components, compressed units) and s = log(predicted squared
// error in those units). The force is trained by mean squared error and s by the
// Gaussian negative log-likelihood with the error held constant,
//   L = mean_b [ |e_b|^2 / 3 + sg(|e_b|^2) exp(-s_b) / 3 + s_b ],  e = y - t,
// so exp(s / 2) estimates the RMS error of a component (its optimum is |e|^2 / 3) without damping
// the force's gradient; particles whose estimate exceeds the tolerance fall back to the walk.
//
// Matrix products use one MFMA kernel: C[M x N] = A[M x K] * Bt[N x K]^T with both operands
// K-contiguous in BF16 (M, N multiples of 32; K a multiple of 16), one wavefront per 32 x 32 tile,
// v_mfma_f32_32x32x16_bf16 over K. Lane l supplies row (l mod 32) of A and row (l mod 32) of Bt,
// elements k0 + 8 (l / 32) .. + 7; it receives C rows 8 g + 4 (l / 32) + r (g, r < 4) of column
// (l mod 32).
#include <algorithm>
#include <cmath>
#include <map>
#include <vector>

#include "tpm_internal.h"

namespace tpm {

typedef __attribute__((__vector_size__(8 * sizeof(short)))) short bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

// Round-to-nearest-even FP32 -> BF16 (finite inputs): the hardware conversion on the device
// (v_cvt_pk_bf16_f32), the same rounding in integer arithmetic on the host.
__host__ __device__ inline uint16_t to_bf16(float x) {
#if __HIP_DEVICE_COMPILE__
    return __builtin_bit_cast(uint16_t, (__bf16)x);
#else
    uint32_t u;
    memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
#endif
}

__host__ __device__ inline float from_bf16(uint16_t b) {
    const uint32_t u = (uint32_t)b << 16;
    float x;
    memcpy(&x, &u, 4);
    return x;
}

__global__ void __launch_bounds__(64) k_gemm_bf16(const uint16_t *__restrict__ A, int lda, const uint16_t *__restrict__ Bt,
                                                  int ldb, float *__restrict__ C, int ldc, int K, float beta) {
    const int lane = threadIdx.x;
    const int m0 = blockIdx.x * 32, n0 = blockIdx.y * 32;
    const int half = lane >> 5, r32 = lane & 31;
    const uint16_t *ap = A + (size_t)(m0 + r32) * lda + 8 * half;
    const uint16_t *bp = Bt + (size_t)(n0 + r32) * ldb + 8 * half;
    f32x16 acc;
    for (int i = 0; i < 16; ++i) acc[i] = 0.f;
    for (int k0 = 0; k0 < K; k0 += 16) {
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(ap + k0);
        const bf16x8 b = *reinterpret_cast<const bf16x8 *>(bp + k0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    for (int g = 0; g < 4; ++g)
        for (int r = 0; r < 4; ++r) {
            const size_t at = (size_t)(m0 + 8 * g + 4 * half + r) * ldc + n0 + r32;
            C[at] = beta == 0.f ? acc[4 * g + r] : acc[4 * g + r] + beta * C[at];
        }
}

__device__ inline float slog(float x) { return copysignf(log1pf(fabsf(x)), x); }
__device__ inline float sexp(float y) { return copysignf(expm1f(fabsf(y)), y); }

// SiLU z sigmoid(z), the sigmoid by the hardware exponential and reciprocal (about 1 ulp each; the
// result is rounded to BF16 next).
__device__ inline float sigmoid(float z) { return __builtin_amdgcn_rcpf(1.f + __expf(-z)); }
__device__ inline float silu(float z) { return z * sigmoid(z); }
__device__ inline float silu_grad(float z) {
    const float s = sigmoid(z);
    return s * (1.f + z * (1.f - s));
}

// Z = C + bias (stored FP32); Y = act(Z) stored as BF16 row-major [rows x n] and transposed [n x rows].
__global__ void k_bias_act(const float *__restrict__ C, const float *__restrict__ bias, int rows, int n, int act,
                           float *__restrict__ Z, uint16_t *__restrict__ Y, uint16_t *__restrict__ Yt) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)rows * n) return;
    const int col = (int)(i % n);
    const size_t row = i / n;
    const float z = C[i] + bias[col];
    if (Z) Z[i] = z;
    const float y = act ? silu(z) : z;
    const uint16_t b = to_bf16(y);
    if (Y) Y[i] = b;
    if (Yt) Yt[(size_t)col * rows + row] = b;
}

// Loss and its gradient at the output layer (n_out = 32 padded columns, 0..2 force, 3 = log var).
// target: 3 floats per row; weight: per-row weight (0 for padding rows). dZ written BF16 both layouts.
__global__ void k_loss_grad(const float *__restrict__ Zout, const float *__restrict__ target,
                            const float *__restrict__ weight, int rows, int nout, float inv_count,
                            uint16_t *__restrict__ dZ, uint16_t *__restrict__ dZt, float *__restrict__ loss_part) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (row < rows) {
        const float *z = Zout + (size_t)row * nout;
        const float w = weight[row] * inv_count;
        // s enters exp() clamped to [-16, 16] so that neither term overflows.
        const float s = z[3];
        const float sc = fminf(fmaxf(s, -16.f), 16.f);
        const float es = __expf(-sc);
        float e2 = 0.f;
        float g[4];
        for (int c = 0; c < 3; ++c) {
            const float e = z[c] - target[3 * row + c];
            e2 += e * e;
            g[c] = w * 2.f * e / 3.f;
        }
        g[3] = w * (1.f - (s == sc ? e2 * es / 3.f : 0.f));
        l = w * (e2 / 3.f + e2 * es / 3.f + s);
        for (int c = 0; c < nout; ++c) {
            const uint16_t b = to_bf16(c < 4 ? g[c] : 0.f);
            dZ[(size_t)row * nout + c] = b;
            dZt[(size_t)c * rows + row] = b;
        }
    }
    // Block partial sums of the loss (deterministic order).
    __shared__ float sh[256];
    sh[threadIdx.x] = l;
    __syncthreads();
    for (int w2 = 128; w2 > 0; w2 >>= 1) {
        if (threadIdx.x < w2) sh[threadIdx.x] += sh[threadIdx.x + w2];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = sh[0];
}

// dZ = dY * silu'(Z) for a hidden layer; dY given FP32 (from the GEMM), output BF16 both layouts.
__global__ void k_act_grad(const float *__restrict__ dY, const float *__restrict__ Z, int rows, int n,
                           uint16_t *__restrict__ dZ, uint16_t *__restrict__ dZt) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)rows * n) return;
    const int col = (int)(i % n);
    const size_t row = i / n;
    const uint16_t b = to_bf16(dY[i] * silu_grad(Z[i]));
    dZ[i] = b;
    dZt[(size_t)col * rows + row] = b;
}

// db[c] = sum over rows of dZ[row][c] (BF16 input, FP32 sum in row order).
__global__ void k_bias_grad(const uint16_t *__restrict__ dZt, int rows, int n, float *__restrict__ db) {
    const int c = blockIdx.x;
    __shared__ float sh[256];
    float s = 0.f;
    for (int r = threadIdx.x; r < rows; r += 256) s += from_bf16(dZt[(size_t)c * rows + r]);
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w2 = 128; w2 > 0; w2 >>= 1) {
        if (threadIdx.x < w2) sh[threadIdx.x] += sh[threadIdx.x + w2];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[c] = sh[0];
}

// Xt[k][r] = X[r][k] (BF16).
__global__ void k_transpose16(const uint16_t *__restrict__ X, int rows, int k, uint16_t *__restrict__ Xt) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)rows * k) return;
    Xt[(size_t)(i % k) * rows + i / k] = X[i];
}

// Copies sample rows perm[b0 + r] of the training set into the batch buffers; rows past `count`
// are zero with weight 0.
__global__ void k_gather_batch(const uint16_t *__restrict__ Xall, const float *__restrict__ Tall,
                               const uint32_t *__restrict__ perm, size_t b0, size_t count, int rows, uint16_t *__restrict__ X,
                               float *__restrict__ T, float *__restrict__ W) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)rows * 64) return;
    const size_t r = i / 64, k = i % 64;
    const bool live = b0 + r < count;
    const size_t src = live ? perm[b0 + r] : 0;
    X[i] = live ? Xall[src * 64 + k] : (uint16_t)0;
    if (k < 3) T[3 * r + k] = live ? Tall[3 * src + k] : 0.f;
    if (k == 0) W[r] = live ? 1.f : 0.f;
}

// Adam on the FP32 master parameters, then refresh the BF16 copies (Wt row-major [n x k] and W
// row-major [k x n]) of every weight matrix.
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m1,
                       float *__restrict__ m2, size_t count, float lr, float b1, float b2, float c1, float c2,
                       float eps) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    p[i] -= lr * (a * c1) / (sqrtf(b * c2) + eps);
}

__global__ void k_to_bf16_both(const float *__restrict__ Wt, int n, int k, uint16_t *__restrict__ Wt16,
                               uint16_t *__restrict__ W16) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * k) return;
    const int col = (int)(i % k);
    const size_t row = i / k;
    const uint16_t b = to_bf16(Wt[i]);
    Wt16[i] = b;
    W16[(size_t)col * n + row] = b;
}

// Periodic wrap of a cell index; one compare-and-add for the indices of positions in [0, L]
// (at most one mesh side out), the modulo otherwise.
__device__ __forceinline__ int wrap_mesh(int v, int n) {
    if (v >= 0 && v < n) return v;
    if (v < 0 && v >= -n) return v + n;
    if (v >= n && v < 2 * n) return v - n;
    const int r = v % n;
    return r < 0 ? r + n : r;
}

// The 33 features of sorted particle i: log(1 + delta) on the 3x3x3 cells around its cell (x, y, z
// loops outer to inner; the 26 neighbours minus the particle's own cell; stencil_features), then the PM force interpolated at it (CIC, FP64 stencil) in
// compressed units and its offset from the cell centre in cells (pm_offset_features).
__device__ inline void stencil_features(const double *__restrict__ spos, size_t i, const float *__restrict__ delta,
                                        int n, double box, float f[27]) {
    const double h = box / n;
    int c[3];
    for (int a = 0; a < 3; ++a) c[a] = (int)floor(spos[3 * i + a] / h);
    int q = 0;
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz) {
                const int ix = wrap_mesh(c[0] + dx, n), iy = wrap_mesh(c[1] + dy, n), iz = wrap_mesh(c[2] + dz, n);
                const float d = delta[((size_t)ix * n + iy) * n + iz];
                f[q++] = __logf(fmaxf(1.f + d, 1e-3f));
            }
    // The neighbours relative to the particle's own cell (f[13]): the differences carry the gradients,
    // which BF16 would lose on top of a common level.
    for (int k = 0; k < 27; ++k)
        if (k != 13) f[k] -= f[13];
}

__device__ inline void pm_offset_features(const double *__restrict__ spos, size_t i, const float *__restrict__ force,
                                          int n, double box, float inv_fscale, float f[6]) {
    const double h = box / n;
    const double s0 = spos[3 * i] / h - 0.5, s1 = spos[3 * i + 1] / h - 0.5, s2 = spos[3 * i + 2] / h - 0.5;
    const double b0 = floor(s0), b1 = floor(s1), b2 = floor(s2);
    const double f0 = s0 - b0, f1 = s1 - b1, f2 = s2 - b2;
    float g[3] = {0.f, 0.f, 0.f};
    for (int dx = 0; dx < 2; ++dx)
        for (int dy = 0; dy < 2; ++dy)
            for (int dz = 0; dz < 2; ++dz) {
                const double w = (dx ? f0 : 1 - f0) * (dy ? f1 : 1 - f1) * (dz ? f2 : 1 - f2);
                const int ix = wrap_mesh((int)b0 + dx, n), iy = wrap_mesh((int)b1 + dy, n),
                          iz = wrap_mesh((int)b2 + dz, n);
                const size_t cell = ((size_t)ix * n + iy) * n + iz;
                for (int a = 0; a < 3; ++a) g[a] += (float)w * force[3 * cell + a];
            }
    for (int a = 0; a < 3; ++a) f[a] = slog(g[a] * inv_fscale);
    for (int a = 0; a < 3; ++a) {
        const double sc = spos[3 * i + a] / h;
        f[3 + a] = (float)(sc - floor(sc) - 0.5);
    }
}

__device__ inline void particle_features(const double *__restrict__ spos, size_t i, const float *__restrict__ delta,
                                         const float *__restrict__ force, int n, double box, float inv_fscale,
                                         float f[33]) {
    stencil_features(spos, i, delta, n, box, f);
    pm_offset_features(spos, i, force, n, box, inv_fscale, f + 27);
}

// Features of sorted particles (rows of X, BF16, zero padded to ldx).
__global__ void k_features(const double *__restrict__ spos, const uint32_t *__restrict__ sel, size_t first,
                           size_t count, const float *__restrict__ delta, const float *__restrict__ force, int n,
                           double box, size_t m, float inv_fscale, uint16_t *__restrict__ X, int ldx) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const size_t i = sel ? sel[t] : first + t;
    float f[33];
    particle_features(spos, i, delta, force, n, box, inv_fscale, f);
    uint16_t *x = X + t * (size_t)ldx;
    int q = 0;
    for (; q < 33; ++q) x[q] = to_bf16(f[q]);
    for (; q < ldx; ++q) x[q] = 0;
}

// Short-range force targets (sorted particle i -> acc_sr[idx[i]]) / fscale for the selection.
__global__ void k_targets(const uint32_t *__restrict__ sel, size_t count, const uint32_t *__restrict__ order,
                          const float *__restrict__ acc_sr, float inv_fscale, float *__restrict__ T) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const size_t o = order[sel ? sel[t] : t];
    for (int a = 0; a < 3; ++a) T[3 * t + a] = slog(acc_sr[3 * o + a] * inv_fscale);
}

// Apply the prediction: where the estimated component RMS error exp(s/2) is within tol * |a|,
// add the predicted short-range force (x fscale) to acc[idx[i]]; otherwise flag i for the walk.
__global__ void k_apply(const float *__restrict__ Zout, int nout, size_t count, const uint32_t *__restrict__ order,
                        float fscale, float tol, float *__restrict__ acc, uint32_t *__restrict__ reject) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float *z = Zout + i * (size_t)nout;
    const size_t o = order[i];
    const float sx = sexp(z[0]) * fscale, sy = sexp(z[1]) * fscale, sz = sexp(z[2]) * fscale;
    const float ax = acc[3 * o] + sx, ay = acc[3 * o + 1] + sy, az = acc[3 * o + 2] + sz;
    const float amag = sqrtf(ax * ax + ay * ay + az * az);
    // First-order propagation of the estimated RMS error through F = fscale sign(y) expm1(|y|),
    // taken at three standard deviations so the accepted estimates meet the tolerance.
    const float ymax = fmaxf(fabsf(z[0]), fmaxf(fabsf(z[1]), fabsf(z[2])));
    const float err = 3.f * sqrtf(3.f) * __expf(0.5f * fminf(z[3], 80.f) + ymax) * fscale;
    const bool ok = err <= tol * amag;
    if (ok) {
        acc[3 * o] = ax;
        acc[3 * o + 1] = ay;
        acc[3 * o + 2] = az;
    }
    reject[i] = ok ? 0u : 1u;
}

// ------------------------------------------------------------- fused inference (one kernel)
//
// Features, the four layers and the gate of every particle in one kernel, the weights resident in
// LDS. Layers are computed transposed, C[out x particle] = Wt[out x in] * X[particle x in]^T: the
// weights are the MFMA's A operand and the activations its B operand. A wavefront takes 32 particles
// at a time; lane l works on particle l mod 32, supplies elements k0 + 8 (l / 32) .. + 7 of its row
// and receives C rows 8 g + 4 (l / 32) + r of the same particle. After bias and SiLU, lanes l and
// l + 32 exchange half their rows (v_permlane32_swap), after which each holds 8 consecutive
// outputs of its particle and stores them as one 16-byte row segment of the wavefront's activation
// block in LDS, where the next layer reads its operand. The arithmetic is that of mlp_forward (FP32
// accumulation over k in the same order, bias and SiLU in FP32, BF16 between layers).
constexpr int kFusedWaves = 8;
// LDS rows of BF16 are stored in 16-byte chunks, chunk c of row r at position c ^ (r mod chunks), so
// the 32 rows a fragment read touches spread over the banks without padding.
constexpr int kLdsW0 = 0, kLdsW1 = kLdsW0 + 128 * 64, kLdsW2 = kLdsW1 + 128 * 128, kLdsW3 = kLdsW2 + 128 * 128,
              kLdsWEnd = kLdsW3 + 32 * 128;
constexpr int kLdsAct = kLdsWEnd;                              // per wavefront: 32 rows of 128
constexpr int kLdsActEnd = kLdsAct + kFusedWaves * 32 * 128;

// Element offset of (row, chunk) in a swizzled block with `chunks` 16-byte chunks per row.
template <int CHUNKS>
__device__ __forceinline__ int swz(int row, int chunk) {
    return (row * CHUNKS + (chunk ^ (row & (CHUNKS - 1)))) * 8;
}

struct FusedArgs {
    const double *spos;
    const uint32_t *idx;     // sorted -> caller order
    const float *delta, *force;
    int n;
    double box;
    const uint16_t *wt[4];   // BF16 Wt [out x in] per layer
    const float *bias[4];    // FP32 biases
    const uint32_t *sel;     // sorted indices of the particles, or null: [first, first + count)
    size_t first, count;
    float fscale, inv_fscale, tol, kappa;
    float *acc;              // caller order: += the prediction where accepted
    uint32_t *reject;        // sorted order: 1 = walk it
    float *pred;             // calibration mode (non-null): 4 floats per particle (force, sigma), no gate
    const uint16_t *rows;    // rows mode (non-null): BF16 feature rows [count x 64] instead of particles;
    float *zout;             // the four raw outputs per row go to zout
};

__device__ __forceinline__ uint32_t pack_bf16(float a, float b) {
    return (uint32_t)to_bf16(a) | ((uint32_t)to_bf16(b) << 16);
}

// One layer of the wavefront's 32 particles: the operand comes from the activation block `act` (NK
// k-steps); the NM output blocks' accumulation chains run interleaved. With ACT the outputs go back
// into `act` (bias, SiLU, BF16, lane exchange); without, block 0's first four rows plus bias go to z.
template <int NM, int NK, bool ACT>
__device__ __forceinline__ void fused_layer(const uint16_t *__restrict__ w, const float *__restrict__ bias,
                                            uint16_t *act, float *z, int r32, int half) {
    bf16x8 xin[NK];
#pragma unroll
    for (int ks = 0; ks < NK; ++ks) xin[ks] = *reinterpret_cast<const bf16x8 *>(act + swz<16>(r32, 2 * ks + half));
    f32x16 acc[NM];
#pragma unroll
    for (int mt = 0; mt < NM; ++mt)
#pragma unroll
        for (int q = 0; q < 16; ++q) acc[mt][q] = 0.f;
#pragma unroll
    for (int ks = 0; ks < NK; ++ks)
#pragma unroll
        for (int mt = 0; mt < NM; ++mt) {
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(w + swz<NK * 2>(mt * 32 + r32, 2 * ks + half));
            acc[mt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, xin[ks], acc[mt], 0, 0, 0);
        }
    if (ACT) {
#pragma unroll
        for (int mt = 0; mt < NM; ++mt)
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                // rows 16 p + 0..15 of block mt: groups g = 2p and 2p + 1
                float v[2][4];
#pragma unroll
                for (int gg = 0; gg < 2; ++gg)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int g = 2 * p + gg;
                        v[gg][r] = silu(acc[mt][4 * g + r] + bias[mt * 32 + 8 * g + 4 * half + r]);
                    }
                const uint32_t a0 = pack_bf16(v[0][0], v[0][1]), a1 = pack_bf16(v[0][2], v[0][3]);
                const uint32_t b0 = pack_bf16(v[1][0], v[1][1]), b1 = pack_bf16(v[1][2], v[1][3]);
                const auto s0 = __builtin_amdgcn_permlane32_swap(a0, b0, false, false);
                const auto s1 = __builtin_amdgcn_permlane32_swap(a1, b1, false, false);
                // lower lane: outputs 32 mt + 16 p + 0..7 (chunk 4 mt + 2 p); upper lane: + 8..15
                uint4 seg;
                seg.x = s0[0];
                seg.y = s1[0];
                seg.z = s0[1];
                seg.w = s1[1];
                *reinterpret_cast<uint4 *>(act + swz<16>(r32, 4 * mt + 2 * p + half)) = seg;
            }
    } else {
#pragma unroll
        for (int r = 0; r < 4; ++r) z[r] = acc[0][r] + bias[4 * half + r];
    }
}

__global__ void __launch_bounds__(64 * kFusedWaves) k_mlp_fused(FusedArgs A) {
    extern __shared__ __align__(16) uint16_t lds[];
    float *lbias = reinterpret_cast<float *>(lds + kLdsActEnd);  // 128 + 128 + 128 + 32
    // Weights (swizzled rows) and biases into LDS.
    {
        const int nin[4] = {64, 128, 128, 128}, nout[4] = {128, 128, 128, 32};
        const int base[4] = {kLdsW0, kLdsW1, kLdsW2, kLdsW3};
        for (int l = 0; l < 4; ++l) {
            const int chunks = nin[l] / 8, total = nout[l] * chunks;
            const uint4 *src = reinterpret_cast<const uint4 *>(A.wt[l]);
            for (int t = threadIdx.x; t < total; t += blockDim.x) {
                const int row = t / chunks, ch = t % chunks;
                *reinterpret_cast<uint4 *>(lds + base[l] + (row * chunks + (ch ^ (row & (chunks - 1)))) * 8) = src[t];
            }
        }
        for (int t = threadIdx.x; t < 416; t += blockDim.x) {
            const int l = t < 128 ? 0 : t < 256 ? 1 : t < 384 ? 2 : 3;
            lbias[t] = A.bias[l][t - 128 * l];
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r32 = lane & 31, half = lane >> 5;
    uint16_t *act = lds + kLdsAct + wave * 32 * 128;
    const size_t tiles = (A.count + 31) / 32;
    for (size_t tile = (size_t)blockIdx.x * kFusedWaves + wave; tile < tiles; tile += (size_t)gridDim.x * kFusedWaves) {
        const size_t t = tile * 32 + r32;
        const bool live = t < A.count;
        const size_t tt = live ? t : A.count - 1;
        const size_t i = A.sel ? A.sel[tt] : A.first + tt;
        if (A.rows) {
            // Given feature rows: the lower lane copies chunks 0-3, the upper 4-7.
            const uint4 *src = reinterpret_cast<const uint4 *>(A.rows + tt * 64);
#pragma unroll
            for (int q = 0; q < 4; ++q)
                *reinterpret_cast<uint4 *>(act + swz<16>(r32, 4 * half + q)) = src[4 * half + q];
        } else if (!half) {
            // Features 0-26 (the density stencil) by the lower lane ...
            float f[27];
            stencil_features(A.spos, i, A.delta, A.n, A.box, f);
#pragma unroll
            for (int q = 0; q < 3; ++q)
                *reinterpret_cast<uint4 *>(act + swz<16>(r32, q)) =
                    make_uint4(pack_bf16(f[8 * q], f[8 * q + 1]), pack_bf16(f[8 * q + 2], f[8 * q + 3]),
                               pack_bf16(f[8 * q + 4], f[8 * q + 5]), pack_bf16(f[8 * q + 6], f[8 * q + 7]));
            uint16_t *c3 = act + swz<16>(r32, 3);
            c3[0] = to_bf16(f[24]);
            c3[1] = to_bf16(f[25]);
            c3[2] = to_bf16(f[26]);
        } else {
            // ... 27-32 (PM force, offset) and the zero padding by the upper lane.
            float f[6];
            pm_offset_features(A.spos, i, A.force, A.n, A.box, A.inv_fscale, f);
            uint16_t *c3 = act + swz<16>(r32, 3);
#pragma unroll
            for (int q = 0; q < 5; ++q) c3[3 + q] = to_bf16(f[q]);
            *reinterpret_cast<uint4 *>(act + swz<16>(r32, 4)) = make_uint4(pack_bf16(f[5], 0.f), 0u, 0u, 0u);
#pragma unroll
            for (int q = 5; q < 8; ++q) *reinterpret_cast<uint4 *>(act + swz<16>(r32, q)) = make_uint4(0u, 0u, 0u, 0u);
        }
        float z[4];
        fused_layer<4, 4, true>(lds + kLdsW0, lbias, act, z, r32, half);
        fused_layer<4, 8, true>(lds + kLdsW1, lbias + 128, act, z, r32, half);
        fused_layer<4, 8, true>(lds + kLdsW2, lbias + 256, act, z, r32, half);
        fused_layer<1, 8, false>(lds + kLdsW3, lbias + 384, act, z, r32, half);
        if (half || !live) continue;
        if (A.zout) {
#pragma unroll
            for (int r = 0; r < 4; ++r) A.zout[4 * t + r] = z[r];
            continue;
        }
        // Lower lanes hold outputs 0-3 of their particle: the gate of k_apply with the calibrated
        // multiplier kappa of the error estimate.
        const float sx = sexp(z[0]) * A.fscale, sy = sexp(z[1]) * A.fscale, sz = sexp(z[2]) * A.fscale;
        const float ymax = fmaxf(fabsf(z[0]), fmaxf(fabsf(z[1]), fabsf(z[2])));
        const float sigma = sqrtf(3.f) * __expf(0.5f * fminf(z[3], 80.f) + ymax) * A.fscale;
        if (A.pred) {
            A.pred[4 * t] = sx;
            A.pred[4 * t + 1] = sy;
            A.pred[4 * t + 2] = sz;
            A.pred[4 * t + 3] = sigma;
            continue;
        }
        const size_t o = A.idx[i];
        const float ax = A.acc[3 * o] + sx, ay = A.acc[3 * o + 1] + sy, az = A.acc[3 * o + 2] + sz;
        const float amag = sqrtf(ax * ax + ay * ay + az * az);
        const bool ok = isfinite(amag) && isfinite(sigma) && A.kappa * sigma <= A.tol * amag;
        if (ok) {
            A.acc[3 * o] = ax;
            A.acc[3 * o + 1] = ay;
            A.acc[3 * o + 2] = az;
        }
        A.reject[i] = ok ? 0u : 1u;
    }
}

constexpr size_t kFusedLds = (size_t)kLdsActEnd * sizeof(uint16_t) + 416 * sizeof(float);
static_assert(kFusedLds <= 160 * 1024, "fused inference LDS");

// Calibration pairs of the listed sorted particles, relative to the predicted total force
// |a| = |PM + F_pred| (the gate's normalisation): est[t] = sigma / |a|, err[t] = |F_pred - F_walk| / |a|.
__global__ void k_calib_pairs(const float *__restrict__ pred, const uint32_t *__restrict__ sel, size_t count,
                              const uint32_t *__restrict__ order, const float *__restrict__ acc_sr,
                              const float *__restrict__ acc_pm, float *__restrict__ est, float *__restrict__ err) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const size_t o = order[sel[t]];
    float d2 = 0.f, a2 = 0.f;
    for (int c = 0; c < 3; ++c) {
        const float d = pred[4 * t + c] - acc_sr[3 * o + c];
        const float a = acc_pm[3 * o + c] + pred[4 * t + c];
        d2 += d * d;
        a2 += a * a;
    }
    const float amag = fmaxf(sqrtf(a2), 1e-30f);
    est[t] = pred[4 * t + 3] / amag;
    err[t] = sqrtf(d2) / amag;
}

// ------------------------------------------------------------------------------------ host

static inline unsigned blocks(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

// Parameter offsets in the FP32 master vector (Wt_l row-major [n_out x n_in], then b_l).
static size_t w_off(int l) {
    size_t o = 0;
    for (int j = 0; j < l; ++j) o += (size_t)kMlpOut[j] * kMlpIn[j] + kMlpOut[j];
    return o;
}
size_t mlp_param_count() { return w_off(4); }

void mlp_init(Ctx &c, uint64_t seed) {
    Mlp &L = c.mlp;
    if (!L.params) {
        const size_t np = mlp_param_count();
        TPM_HIP(hipMalloc(&L.params, np * sizeof(float)));
        TPM_HIP(hipMalloc(&L.grads, np * sizeof(float)));
        TPM_HIP(hipMalloc(&L.m1, np * sizeof(float)));
        TPM_HIP(hipMalloc(&L.m2, np * sizeof(float)));
        for (int l = 0; l < 4; ++l) {
            const size_t e = (size_t)kMlpOut[l] * kMlpIn[l];
            TPM_HIP(hipMalloc(&L.wt16[l], e * sizeof(uint16_t)));
            TPM_HIP(hipMalloc(&L.w16[l], e * sizeof(uint16_t)));
        }
    }
    // Deterministic He-style init from a counter hash; output rows beyond 4 stay zero.
    std::vector<float> h(mlp_param_count(), 0.f);
    uint64_t st = seed * 0x9E3779B97F4A7C15ULL + 1;
    auto next = [&]() {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        return ((double)(st >> 11) + 0.5) * 0x1p-53;
    };
    for (int l = 0; l < 4; ++l) {
        const int rows = l == 3 ? 4 : kMlpOut[l];
        const int fan_in = l == 0 ? 33 : kMlpIn[l];
        const double sc = std::sqrt(6.0 / fan_in) * (l == 3 ? 0.1 : 1.0);
        for (int r = 0; r < rows; ++r)
            for (int k = 0; k < (l == 0 ? 33 : kMlpIn[l]); ++k)
                h[w_off(l) + (size_t)r * kMlpIn[l] + k] = (float)((2.0 * next() - 1.0) * sc);
    }
    TPM_HIP(hipMemcpy(L.params, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    TPM_HIP(hipMemset(L.m1, 0, h.size() * sizeof(float)));
    TPM_HIP(hipMemset(L.m2, 0, h.size() * sizeof(float)));
    L.step = 0;
    mlp_refresh_bf16(c);
}

size_t mlp_weight_offset(int l) { return w_off(l); }

void mlp_refresh_bf16(Ctx &c) {
    Mlp &L = c.mlp;
    for (int l = 0; l < 4; ++l) {
        const size_t e = (size_t)kMlpOut[l] * kMlpIn[l];
        hipLaunchKernelGGL(k_to_bf16_both, dim3(blocks(e, 256)), dim3(256), 0, c.stream, L.params + w_off(l),
                           kMlpOut[l], kMlpIn[l], L.wt16[l], L.w16[l]);
        TPM_HIP(hipGetLastError());
    }
}

static void ensure_work(Ctx &c, size_t rows) {
    Mlp &L = c.mlp;
    if (rows <= L.cap) return;
    for (void *p : L.work)
        if (p) (void)hipFree(p);
    L.cap = rows;
    // X / Xt (64), per hidden layer Y / Yt (128), Z (fp32, 128) x 3, C (fp32, 128), out Z (32),
    // dZ / dZt (128), dY (fp32, 128), targets (3), weights (1), loss partials.
    TPM_HIP(hipMalloc(&L.X, rows * 64 * 2));
    TPM_HIP(hipMalloc(&L.Xt, rows * 64 * 2));
    for (int l = 0; l < 3; ++l) {
        TPM_HIP(hipMalloc(&L.Y[l], rows * 128 * 2));
        TPM_HIP(hipMalloc(&L.Yt[l], rows * 128 * 2));
        TPM_HIP(hipMalloc(&L.Z[l], rows * 128 * 4));
    }
    TPM_HIP(hipMalloc(&L.C, rows * 128 * 4));
    TPM_HIP(hipMalloc(&L.Zout, rows * 32 * 4));
    TPM_HIP(hipMalloc(&L.dZ, rows * 128 * 2));
    TPM_HIP(hipMalloc(&L.dZt, rows * 128 * 2));
    TPM_HIP(hipMalloc(&L.dY, rows * 128 * 4));
    TPM_HIP(hipMalloc(&L.T, rows * 3 * 4));
    TPM_HIP(hipMalloc(&L.Wrow, rows * 4));
    TPM_HIP(hipMalloc(&L.loss_part, (rows / 256 + 2) * 4));
    void *w[] = {L.X, L.Xt, L.Y[0], L.Yt[0], L.Z[0], L.Y[1], L.Yt[1], L.Z[1], L.Y[2], L.Yt[2], L.Z[2], L.C,
                 L.Zout, L.dZ, L.dZt, L.dY, L.T, L.Wrow, L.loss_part};
    L.work.assign(w, w + sizeof(w) / sizeof(w[0]));
}

static void gemm(Ctx &c, const uint16_t *A, int lda, const uint16_t *Bt, int ldb, float *C, int ldc, int M, int N,
                 int K) {
    hipLaunchKernelGGL(k_gemm_bf16, dim3(M / 32, N / 32), dim3(64), 0, c.stream, A, lda, Bt, ldb, C, ldc, K, 0.f);
    TPM_HIP(hipGetLastError());
}

// Forward over `rows` (a multiple of 32) rows of L.X; writes L.Zout (rows x 32, FP32). Training
// also keeps each hidden layer's pre-activation (FP32) and transposed activation for backward.
void mlp_forward(Ctx &c, size_t rows, bool train) {
    Mlp &L = c.mlp;
    const uint16_t *in = L.X;
    for (int l = 0; l < 4; ++l) {
        gemm(c, in, kMlpIn[l], L.wt16[l], kMlpIn[l], l == 3 ? L.Zout : L.C, kMlpOut[l], (int)rows, kMlpOut[l],
             kMlpIn[l]);
        const float *bias = L.params + w_off(l) + (size_t)kMlpOut[l] * kMlpIn[l];
        const size_t e = rows * kMlpOut[l];
        if (l < 3) {
            hipLaunchKernelGGL(k_bias_act, dim3(blocks(e, 256)), dim3(256), 0, c.stream, L.C, bias, (int)rows,
                               kMlpOut[l], 1, train ? L.Z[l] : (float *)nullptr, L.Y[l],
                               train ? L.Yt[l] : (uint16_t *)nullptr);
            in = L.Y[l];
        } else {
            hipLaunchKernelGGL(k_bias_act, dim3(blocks(e, 256)), dim3(256), 0, c.stream, L.Zout, bias, (int)rows,
                               kMlpOut[l], 0, L.Zout, (uint16_t *)nullptr, (uint16_t *)nullptr);
        }
        TPM_HIP(hipGetLastError());
    }
}

// Forward + backward for rows of (L.X, L.T, L.Wrow); gradients into L.grads; returns the loss.
double mlp_forward_backward(Ctx &c, size_t rows, size_t count, bool want_loss) {
    Mlp &L = c.mlp;
    mlp_forward(c, rows, true);
    // Xt for the first layer's weight gradient.
    hipLaunchKernelGGL(k_transpose16, dim3(blocks(rows * kMlpIn[0], 256)), dim3(256), 0, c.stream, L.X, (int)rows,
                       kMlpIn[0], L.Xt);
    TPM_HIP(hipGetLastError());
    const unsigned nlb = blocks(rows, 256);
    hipLaunchKernelGGL(k_loss_grad, dim3(nlb), dim3(256), 0, c.stream, L.Zout, L.T, L.Wrow, (int)rows, 32,
                       1.f / (float)count, L.dZ, L.dZt, L.loss_part);
    TPM_HIP(hipGetLastError());
    for (int l = 3; l >= 0; --l) {
        const int nin = kMlpIn[l], nout = kMlpOut[l];
        float *gW = L.grads + w_off(l);
        float *gb = gW + (size_t)nout * nin;
        const uint16_t *xin_t = l == 0 ? L.Xt : L.Yt[l - 1];
        // dWt[n][k] = sum_b dZ[b][n] X[b][k]: A = dZt [nout x rows], Bt = Xt [nin x rows].
        gemm(c, L.dZt, (int)rows, xin_t, (int)rows, gW, nin, nout, nin, (int)rows);
        hipLaunchKernelGGL(k_bias_grad, dim3(nout), dim3(256), 0, c.stream, L.dZt, (int)rows, nout, gb);
        TPM_HIP(hipGetLastError());
        if (l == 0) break;
        // dY[b][k] = sum_n dZ[b][n] W[k][n]: A = dZ [rows x nout], Bt = W (row-major [nin x nout]).
        gemm(c, L.dZ, nout, L.w16[l], nout, L.dY, nin, (int)rows, nin, nout);
        const size_t e = rows * nin;
        hipLaunchKernelGGL(k_act_grad, dim3(blocks(e, 256)), dim3(256), 0, c.stream, L.dY, L.Z[l - 1], (int)rows, nin,
                           L.dZ, L.dZt);
        TPM_HIP(hipGetLastError());
    }
    if (!want_loss) return 0.0;
    std::vector<float> lp(nlb);
    TPM_HIP(hipMemcpyAsync(lp.data(), L.loss_part, nlb * sizeof(float), hipMemcpyDeviceToHost, c.stream));
    TPM_HIP(hipStreamSynchronize(c.stream));
    double loss = 0.0;
    for (float v : lp) loss += v;
    return loss;
}

void mlp_adam_step(Ctx &c, float lr) {
    Mlp &L = c.mlp;
    L.step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f / (1.f - std::pow(b1, (float)L.step));
    const float c2 = 1.f / (1.f - std::pow(b2, (float)L.step));
    const size_t np = mlp_param_count();
    hipLaunchKernelGGL(k_adam, dim3(blocks(np, 256)), dim3(256), 0, c.stream, L.params, L.grads, L.m1, L.m2, np, lr,
                       b1, b2, c1, c2, 1e-8f);
    TPM_HIP(hipGetLastError());
    mlp_refresh_bf16(c);
}

void mlp_prepare(Ctx &c, size_t rows) { ensure_work(c, rows); }

void mlp_features(Ctx &c, const uint32_t *d_sel, size_t first, size_t count, uint16_t *X, float inv_fscale) {
    if (count == 0) return;
    hipLaunchKernelGGL(k_features, dim3(blocks(count, 256)), dim3(256), 0, c.stream, c.spos, d_sel, first, count,
                       c.delta, c.force, c.n, c.p.box_size, c.m, inv_fscale, X, 64);
    TPM_HIP(hipGetLastError());
}

__global__ void k_pack_rows(const float *__restrict__ x, int nfeat, size_t count, int rows, uint16_t *__restrict__ X,
                            const float *__restrict__ t, float *__restrict__ T, float *__restrict__ W) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)rows * 64) return;
    const size_t r = i / 64, k = i % 64;
    const bool live = r < count;
    X[i] = (live && (int)k < nfeat) ? to_bf16(x[r * nfeat + k]) : (uint16_t)0;
    if (T && k < 3) T[3 * r + k] = live ? t[3 * r + k] : 0.f;
    if (W && k == 0) W[r] = live ? 1.f : 0.f;
}

void mlp_pack_rows(Ctx &c, const float *d_x, const float *d_t, size_t count, size_t rows) {
    Mlp &L = c.mlp;
    hipLaunchKernelGGL(k_pack_rows, dim3(blocks(rows * 64, 256)), dim3(256), 0, c.stream, d_x, 33, count, (int)rows,
                       L.X, d_t, d_t ? L.T : (float *)nullptr, d_t ? L.Wrow : (float *)nullptr);
    TPM_HIP(hipGetLastError());
}

void mlp_targets(Ctx &c, const uint32_t *d_sel, size_t count, const float *acc_sr, float inv_fscale, float *T) {
    if (count == 0) return;
    hipLaunchKernelGGL(k_targets, dim3(blocks(count, 256)), dim3(256), 0, c.stream, d_sel, count, c.idx, acc_sr,
                       inv_fscale, T);
    TPM_HIP(hipGetLastError());
}

void mlp_gather_batch(Ctx &c, const uint16_t *Xall, const float *Tall, const uint32_t *perm, size_t b0,
                      size_t count, size_t rows) {
    Mlp &L = c.mlp;
    hipLaunchKernelGGL(k_gather_batch, dim3(blocks(rows * 64, 256)), dim3(256), 0, c.stream, Xall, Tall, perm, b0,
                       count, (int)rows, L.X, L.T, L.Wrow);
    TPM_HIP(hipGetLastError());
}

static void launch_fused(Ctx &c, FusedArgs &a) {
    if (a.count == 0) return;
    Mlp &L = c.mlp;
    static bool lds_set = false;
    if (!lds_set) {
        TPM_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_mlp_fused),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)kFusedLds));
        lds_set = true;
    }
    for (int l = 0; l < 4; ++l) {
        a.wt[l] = L.wt16[l];
        a.bias[l] = L.params + w_off(l) + (size_t)kMlpOut[l] * kMlpIn[l];
    }
    const size_t tiles = (a.count + 31) / 32;
    const unsigned grid = (unsigned)std::min<size_t>((tiles + kFusedWaves - 1) / kFusedWaves,
                                                     (size_t)std::max(1, c.num_cu));
    hipLaunchKernelGGL(k_mlp_fused, dim3(grid), dim3(64 * kFusedWaves), kFusedLds, c.stream, a);
    TPM_HIP(hipGetLastError());
}

// Fused inference over sorted particles sel[0..count) (or [first, first + count) when sel is null):
// with pred == null the gate (acc and reject), else the calibration outputs (4 floats per particle
// into pred: the predicted short-range force and the estimate sigma of its error).
void mlp_infer_fused(Ctx &c, const uint32_t *sel, size_t first, size_t count, float fscale, float tol, float kappa,
                     float *d_acc, uint32_t *reject, float *pred) {
    FusedArgs a{};
    a.spos = c.spos;
    a.idx = c.idx;
    a.delta = c.delta;
    a.force = c.force;
    a.n = c.n;
    a.box = c.p.box_size;
    a.sel = sel;
    a.first = first;
    a.count = count;
    a.fscale = fscale;
    a.inv_fscale = 1.f / fscale;
    a.tol = tol;
    a.kappa = kappa;
    a.acc = d_acc;
    a.reject = reject;
    a.pred = pred;
    launch_fused(c, a);
}

// The network on given BF16 feature rows (rows x 64): the four raw outputs per row into zout.
void mlp_predict_rows(Ctx &c, const uint16_t *rows, size_t count, float *zout) {
    FusedArgs a{};
    a.rows = rows;
    a.zout = zout;
    a.count = count;
    a.fscale = 1.f;
    a.inv_fscale = 1.f;
    launch_fused(c, a);
}

// Calibration of the gate on the listed (held-out) sorted particles: keeps their (estimate, error)
// pairs, relative to |a|, sorted by estimate, for mlp_kappa.
void mlp_calibrate(Ctx &c, const uint32_t *d_sel, size_t count, float fscale, const float *acc_sr,
                   const float *acc_pm) {
    Mlp &L = c.mlp;
    L.cal_est.clear();
    L.cal_err.clear();
    L.kappa_cache.clear();
    if (count == 0) return;
    float *buf = nullptr;
    TPM_HIP(hipMalloc(&buf, count * 6 * sizeof(float)));
    std::vector<float> est(count), err(count);
    try {
        mlp_infer_fused(c, d_sel, 0, count, fscale, 0.f, 0.f, nullptr, nullptr, buf);
        hipLaunchKernelGGL(k_calib_pairs, dim3(blocks(count, 256)), dim3(256), 0, c.stream, buf, d_sel, count, c.idx,
                           acc_sr, acc_pm, buf + 4 * count, buf + 5 * count);
        TPM_HIP(hipGetLastError());
        TPM_HIP(hipMemcpyAsync(est.data(), buf + 4 * count, count * sizeof(float), hipMemcpyDeviceToHost, c.stream));
        TPM_HIP(hipMemcpyAsync(err.data(), buf + 5 * count, count * sizeof(float), hipMemcpyDeviceToHost, c.stream));
        TPM_HIP(hipStreamSynchronize(c.stream));
    } catch (...) {
        (void)hipFree(buf);
        throw;
    }
    (void)hipFree(buf);
    std::vector<size_t> ord(count);
    for (size_t t = 0; t < count; ++t) ord[t] = t;
    std::sort(ord.begin(), ord.end(), [&](size_t x, size_t y) { return est[x] < est[y]; });
    L.cal_est.resize(count);
    L.cal_err.resize(count);
    for (size_t t = 0; t < count; ++t) {
        L.cal_est[t] = std::isfinite(est[ord[t]]) ? est[ord[t]] : INFINITY;
        L.cal_err[t] = std::isfinite(err[ord[t]]) ? err[ord[t]] : INFINITY;
    }
}

// The gate's multiplier for tolerance tol: the gate accepts a particle when kappa sigma <= tol |a|,
// i.e. when its relative estimate is at most tau = tol / kappa. tau is the largest calibration
// estimate such that, among the calibration particles with estimates up to tau (at least 100 of
// them), 99% have an actual error within tol; kappa = +inf (nothing accepted) when there is none.
// Before any calibration the fixed multiplier 3 sqrt(3) applies.
double mlp_kappa(Ctx &c, double tol) {
    Mlp &L = c.mlp;
    if (L.cal_est.empty()) return L.kappa;
    if (!(tol > 0)) return INFINITY;
    auto hit = L.kappa_cache.find(tol);
    if (hit != L.kappa_cache.end()) return hit->second;
    const size_t m = L.cal_est.size();
    // Candidate prefix sizes on a geometric grid; the largest prefix that passes wins.
    double best_tau = -1.0;
    std::vector<float> e;
    for (double f = 1.0; f >= 1e-4; f *= 0.9) {
        const size_t k = (size_t)std::ceil(f * (double)m);
        if (k < 100) break;
        e.assign(L.cal_err.begin(), L.cal_err.begin() + k);
        const size_t q = std::min(k - 1, (size_t)std::ceil(0.99 * (double)k) - 1);
        std::nth_element(e.begin(), e.begin() + q, e.end());
        if (e[q] <= tol && std::isfinite(L.cal_est[k - 1])) {
            best_tau = k == m ? INFINITY : L.cal_est[k - 1];  // every calibration particle qualifies
            break;
        }
    }
    const double kappa = best_tau > 0 ? tol / best_tau : INFINITY;  // 0 when all qualify: accept all
    L.kappa_cache[tol] = kappa;
    return kappa;
}

void mlp_apply(Ctx &c, size_t count, size_t first, float fscale, float tol, float *d_acc, uint32_t *reject) {
    Mlp &L = c.mlp;
    if (count == 0) return;
    hipLaunchKernelGGL(k_apply, dim3(blocks(count, 256)), dim3(256), 0, c.stream, L.Zout, 32, count, c.idx + first,
                       fscale, tol, d_acc, reject + first);
    TPM_HIP(hipGetLastError());
}

}  // namespace tpm
