// Learned short-range correction: a 4-layer MLP (64 -> 128 -> 128 -> 128 -> 4, SiLU) with BF16
// weights and activations on MFMA and FP32 accumulation, trained on the GPU by its own forward
// and backward kernels. No reference counterpart (the reference has only the exact tree walk);
// the walk supplies its training targets, its fallback and the judge of its error.
//
// Per particle the input is 33 features, zero-padded to 64: log(1 + delta) on the 3x3x3 cells
// around the particle's cell, the PM force at the particle and the particle's offset from its
// cell centre (cell units). Forces enter and leave in the compressed units slog(F / (K / h^2)),
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
#include <cmath>
#include <vector>

#include "tpm_internal.h"

namespace tpm {

typedef __attribute__((__vector_size__(8 * sizeof(short)))) short bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;

// Round-to-nearest-even FP32 -> BF16 (finite inputs).
__host__ __device__ inline uint16_t to_bf16(float x) {
    uint32_t u;
    memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
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

__device__ inline float silu(float z) { return z / (1.f + __expf(-z)); }
__device__ inline float silu_grad(float z) {
    const float s = 1.f / (1.f + __expf(-z));
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

// Features of sorted particles: 27 log(1 + delta) cells, PM force / fscale, offset in the cell.
__global__ void k_features(const double *__restrict__ spos, const uint32_t *__restrict__ sel, size_t first,
                           size_t count, const float *__restrict__ delta, const float *__restrict__ force, int n,
                           double box, size_t m, float inv_fscale, uint16_t *__restrict__ X, int ldx) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const size_t i = sel ? sel[t] : first + t;
    const double h = box / n;
    int c[3];
    float off[3];
    for (int a = 0; a < 3; ++a) {
        const double s = spos[3 * i + a] / h;
        const double f = floor(s);
        c[a] = (int)f;
        off[a] = (float)(s - f - 0.5);
    }
    uint16_t *x = X + t * (size_t)ldx;
    int q = 0;
    for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dz = -1; dz <= 1; ++dz) {
                int ix = (c[0] + dx) % n, iy = (c[1] + dy) % n, iz = (c[2] + dz) % n;
                ix += ix < 0 ? n : 0;
                iy += iy < 0 ? n : 0;
                iz += iz < 0 ? n : 0;
                const float d = delta[((size_t)ix * n + iy) * n + iz];
                x[q++] = to_bf16(logf(fmaxf(1.f + d, 1e-3f)));
            }
    // PM force interpolated at the particle (CIC, FP64 stencil), scaled.
    {
        const double s0 = spos[3 * i] / h - 0.5, s1 = spos[3 * i + 1] / h - 0.5, s2 = spos[3 * i + 2] / h - 0.5;
        const double b0 = floor(s0), b1 = floor(s1), b2 = floor(s2);
        const double f0 = s0 - b0, f1 = s1 - b1, f2 = s2 - b2;
        float g[3] = {0.f, 0.f, 0.f};
        for (int dx = 0; dx < 2; ++dx)
            for (int dy = 0; dy < 2; ++dy)
                for (int dz = 0; dz < 2; ++dz) {
                    const double w = (dx ? f0 : 1 - f0) * (dy ? f1 : 1 - f1) * (dz ? f2 : 1 - f2);
                    int ix = ((int)b0 + dx) % n, iy = ((int)b1 + dy) % n, iz = ((int)b2 + dz) % n;
                    ix += ix < 0 ? n : 0;
                    iy += iy < 0 ? n : 0;
                    iz += iz < 0 ? n : 0;
                    const size_t cell = ((size_t)ix * n + iy) * n + iz;
                    for (int a = 0; a < 3; ++a) g[a] += (float)w * force[3 * cell + a];
                }
        for (int a = 0; a < 3; ++a) x[q++] = to_bf16(slog(g[a] * inv_fscale));
    }
    for (int a = 0; a < 3; ++a) x[q++] = to_bf16(off[a]);
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

void mlp_apply(Ctx &c, size_t count, size_t first, float fscale, float tol, float *d_acc, uint32_t *reject) {
    Mlp &L = c.mlp;
    if (count == 0) return;
    hipLaunchKernelGGL(k_apply, dim3(blocks(count, 256)), dim3(256), 0, c.stream, L.Zout, 32, count, c.idx + first,
                       fscale, tol, d_acc, reject + first);
    TPM_HIP(hipGetLastError());
}

}  // namespace tpm
