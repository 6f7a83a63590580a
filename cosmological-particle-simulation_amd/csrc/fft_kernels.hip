// PM solve on the mesh: r2c forward transform of the density contrast, Gaussian-split
// Green's function and spectral gradient, inverse transforms to three force grids.
//
// Replaces PoissonSolver::acceleration (src/poisson.rs:33-88) and Fft3 (src/fft.rs:11-82).
// The reference transforms n^3 complex values with rustfft line by line; here the grid is
// real, so only the half spectrum kz = 0..n/2 is kept (hp = n/2 + 1 values per z-line) and
// every 1D transform runs inside LDS on a batch of lines:
//   Z   r2c along z      (fixed-point mass grid -> delta -> half spectrum S)
//   Y   forward along y  (S in place)
//   X   forward along x, Green's function, inverse along x. Because k_y and k_z are
//       constant on an x-line, only two inverse x-transforms are needed:
//         A = iFFT_x(phi * (-i k_x))   (the x component, written over S)
//         B = iFFT_x(phi)              (shared by the y and z components)
//   Y'  inverse along y: A in place; B -> C = iFFT_y(B * (-i k_y)) and B = iFFT_y(B)
//   Z'  c2r along z: gx = c2r(A), gy = c2r(C), gz = c2r(B * (-i k_z))
// The Green's function factor -(1.5 Omega_m / a) exp(-k^2 r_s^2) / (k^2 n^3) is evaluated
// in FP64 per mode, with the reference's wavenumber convention (Nyquist -> 0, src/fft.rs:86-95)
// and its k^2 == 0 skip (src/poisson.rs:63-65).
#include <cmath>

#include "tpm_internal.h"

namespace tpm {

template <typename R>
struct Cx;
template <>
struct Cx<float> {
    using T = float2;
};
template <>
struct Cx<double> {
    using T = double2;
};

struct PlanArg {
    int n, nst;
    int radix[24];
};

template <typename C>
__device__ __forceinline__ C cmul(C a, C b) {
    C r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}

template <typename C>
__device__ __forceinline__ C cadd(C a, C b) {
    C r;
    r.x = a.x + b.x;
    r.y = a.y + b.y;
    return r;
}

template <typename C>
__device__ __forceinline__ C csub(C a, C b) {
    C r;
    r.x = a.x - b.x;
    r.y = a.y - b.y;
    return r;
}

// Twiddle exp(sign 2 pi i t / n) from the forward table (sign -1) or its conjugate.
template <typename C>
__device__ __forceinline__ C twid(const C *tw, int t, bool inv) {
    C w = tw[t];
    if (inv) w.y = -w.y;
    return w;
}

// Unnormalised transform of `nlines` lines of length pl.n held in LDS buffer `a` (row pitch
// ld), ping-ponging with `b`. Stockham autosort, one stage per radix: in[j + r*n/R] times
// exp(sign 2 pi i (j mod Ns) r / (Ns R)), an R-point DFT, out[(j/Ns) Ns R + (j mod Ns) + q Ns].
// Returns the buffer holding the result. Ends with a barrier.
template <typename R, typename C>
__device__ C *fft_lds(C *a, C *b, int nlines, int ld, const PlanArg &pl, const C *tw, bool inv) {
    const int n = pl.n;
    const R s = inv ? R(1) : R(-1);
    int ns = 1;
    for (int st = 0; st < pl.nst; ++st) {
        const int rad = pl.radix[st];
        const int stride = n / rad;
        const int twstep = n / (ns * rad);
        const int total = nlines * stride;
        for (int w = threadIdx.x; w < total; w += blockDim.x) {
            const int line = w / stride;
            const int j = w - line * stride;
            const int k = j % ns;
            const C *src = a + line * ld;
            C *dst = b + line * ld;
            const int base = (j - k) * rad + k;
            if (rad == 4) {
                C v0 = src[j];
                C v1 = src[j + stride];
                C v2 = src[j + 2 * stride];
                C v3 = src[j + 3 * stride];
                if (k) {
                    v1 = cmul(v1, twid(tw, k * twstep, inv));
                    v2 = cmul(v2, twid(tw, 2 * k * twstep, inv));
                    v3 = cmul(v3, twid(tw, 3 * k * twstep, inv));
                }
                C t0 = cadd(v0, v2), t1 = csub(v0, v2), t2 = cadd(v1, v3), t3 = csub(v1, v3);
                C it3;
                it3.x = -s * t3.y;
                it3.y = s * t3.x;
                dst[base] = cadd(t0, t2);
                dst[base + ns] = cadd(t1, it3);
                dst[base + 2 * ns] = csub(t0, t2);
                dst[base + 3 * ns] = csub(t1, it3);
            } else if (rad == 2) {
                C v0 = src[j];
                C v1 = src[j + stride];
                if (k) v1 = cmul(v1, twid(tw, k * twstep, inv));
                dst[base] = cadd(v0, v1);
                dst[base + ns] = csub(v0, v1);
            } else if (rad == 3) {
                C v0 = src[j];
                C v1 = src[j + stride];
                C v2 = src[j + 2 * stride];
                if (k) {
                    v1 = cmul(v1, twid(tw, k * twstep, inv));
                    v2 = cmul(v2, twid(tw, 2 * k * twstep, inv));
                }
                const R h = R(0.86602540378443864676) * s;
                C t1 = cadd(v1, v2), t2 = csub(v1, v2);
                C m;
                m.x = v0.x - R(0.5) * t1.x;
                m.y = v0.y - R(0.5) * t1.y;
                C it;
                it.x = -h * t2.y;
                it.y = h * t2.x;
                dst[base] = cadd(v0, t1);
                dst[base + ns] = cadd(m, it);
                dst[base + 2 * ns] = csub(m, it);
            } else {
                // Any other prime radix: direct R-point DFT with table twiddles.
                const int wr = n / rad;
                for (int q = 0; q < rad; ++q) {
                    C acc;
                    acc.x = R(0);
                    acc.y = R(0);
                    for (int r = 0; r < rad; ++r) {
                        C v = src[j + r * stride];
                        if (k && r) v = cmul(v, twid(tw, k * r * twstep, inv));
                        const int e = (r * q) % rad;
                        if (e) v = cmul(v, twid(tw, e * wr, inv));
                        acc = cadd(acc, v);
                    }
                    dst[base + q * ns] = acc;
                }
            }
        }
        __syncthreads();
        C *t = a;
        a = b;
        b = t;
        ns *= rad;
    }
    return a;
}

// ------------------------------------------------------------------------------ Z forward

// Source of the z pass: the fixed-point mass grid (converted to delta = c/mean - 1 as the
// reference does in FP32, src/integrator.rs:120-124) or a density-contrast grid.
template <typename R>
struct ZSrc {
    const unsigned long long *grid;  // fixed-point counts, or null
    const float *delta;              // density contrast, or null
    float mean32;                    // N / n^3 in float (reference arithmetic)
    double mean64;
    float *delta_out;                // optional copy of delta (float)
};

template <typename R>
__device__ __forceinline__ R zsrc_value(const ZSrc<R> &s, size_t i);

template <>
__device__ __forceinline__ float zsrc_value<float>(const ZSrc<float> &s, size_t i) {
    if (s.grid) {
        const float c = (float)((double)s.grid[i] * 0x1p-32);
        return c / s.mean32 - 1.0f;
    }
    return s.delta[i];
}

template <>
__device__ __forceinline__ double zsrc_value<double>(const ZSrc<double> &s, size_t i) {
    if (s.grid) return ((double)s.grid[i] * 0x1p-32) / s.mean64 - 1.0;
    return (double)s.delta[i];
}

template <typename R>
__global__ void __launch_bounds__(256) k_fft_z_fwd(ZSrc<R> src, typename Cx<R>::T *spec, PlanArg pl,
                                                   const typename Cx<R>::T *tw, int nb, int hp) {
    using C = typename Cx<R>::T;
    extern __shared__ __align__(16) unsigned char smem[];
    const int n = pl.n;
    const int ld = n + 1;
    C *a = reinterpret_cast<C *>(smem);
    C *b = a + nb * ld;
    const size_t nlines = (size_t)n * n;
    const size_t l0 = (size_t)blockIdx.x * nb;
    for (int e = threadIdx.x; e < nb * n; e += blockDim.x) {
        const int li = e / n, t = e - li * n;
        const size_t l = l0 + li;
        C v;
        v.x = R(0);
        v.y = R(0);
        if (l < nlines) {
            const size_t gi = l * n + t;
            v.x = zsrc_value<R>(src, gi);
            if (src.delta_out) src.delta_out[gi] = (float)v.x;
        }
        a[li * ld + t] = v;
    }
    __syncthreads();
    C *res = fft_lds<R>(a, b, nb, ld, pl, tw, false);
    for (int e = threadIdx.x; e < nb * hp; e += blockDim.x) {
        const int li = e / hp, kz = e - li * hp;
        const size_t l = l0 + li;
        if (l < nlines) spec[l * hp + kz] = res[li * ld + kz];
    }
}

// ------------------------------------------------------------------------------ Y passes

// mode 0: forward in place. mode 1: inverse in place. mode 2: inverse of B in place plus
// C = iFFT_y(B * (-i k_y)).
template <typename R>
__global__ void __launch_bounds__(256) k_fft_y(typename Cx<R>::T *spec, typename Cx<R>::T *out2, PlanArg pl,
                                               const typename Cx<R>::T *tw, const double *kvec, int nb,
                                               int hp, int mode) {
    using C = typename Cx<R>::T;
    extern __shared__ __align__(16) unsigned char smem[];
    const int n = pl.n;
    const int ld = n + 1;
    const int mult = mode == 2 ? 2 : 1;
    C *a = reinterpret_cast<C *>(smem);
    C *b = a + mult * nb * ld;
    const int x = blockIdx.y;
    const int kz0 = blockIdx.x * nb;
    for (int e = threadIdx.x; e < nb * n; e += blockDim.x) {
        const int t = e / nb, li = e - t * nb;
        const int kz = kz0 + li;
        C v;
        v.x = R(0);
        v.y = R(0);
        if (kz < hp) v = spec[((size_t)x * n + t) * hp + kz];
        a[li * ld + t] = v;
        if (mode == 2) {
            // (v) * (0 - i k_y): re = v.y k, im = -v.x k
            const R ky = (R)kvec[t];
            C m;
            m.x = v.y * ky;
            m.y = -v.x * ky;
            a[(nb + li) * ld + t] = m;
        }
    }
    __syncthreads();
    C *res = fft_lds<R>(a, b, mult * nb, ld, pl, tw, mode != 0);
    for (int e = threadIdx.x; e < nb * n; e += blockDim.x) {
        const int t = e / nb, li = e - t * nb;
        const int kz = kz0 + li;
        if (kz < hp) {
            const size_t gi = ((size_t)x * n + t) * hp + kz;
            spec[gi] = res[li * ld + t];
            if (mode == 2) out2[gi] = res[(nb + li) * ld + t];
        }
    }
}

// --------------------------------------------------------------------- X pass with Green

// Scalar factor G(k) applied to every mode before the gradient (-i k_a).
//   mode 0, gravity (src/poisson.rs:61-67): G = -c0 exp(-k^2 c1) / k^2, c0 = 1.5 Omega_m / (a n^3), c1 = r_s^2.
//   mode 1, Zel'dovich displacement: G = -c0 sqrt(k^-2 exp(-|k| / c1)) / k^2, the power law
//           P(k) = k^-2 exp(-k / k_c) with c1 = k_c and c0 = sqrt(n^3 / V) / n^3 (the amplitude is
//           rescaled afterwards to a target displacement RMS).
//   mode 2, Zel'dovich displacement with the BBKS spectrum (src/ics.rs:20-33, src/cosmology.rs:93-103):
//           G = -c0 sqrt(k^n_s) T(k) / k^2, c1 = n_s, c2 = Gamma = Omega_m h, c0 including sqrt(amplitude).
struct GreenArgs {
    int mode;
    double c0, c1, c2;
};

__device__ __forceinline__ double bbks_transfer(double k, double gamma) {
    const double q = k / gamma;
    if (q == 0.0) return 1.0;
    const double a = 16.1 * q, b = 5.46 * q, c = 6.71 * q;
    const double poly = 1.0 + 3.89 * q + a * a + b * b * b + (c * c) * (c * c);
    return log1p(2.34 * q) / (2.34 * q) * pow(poly, -0.25);
}

__device__ __forceinline__ double green_factor(const GreenArgs &g, double k2) {
    if (g.mode == 0) return -g.c0 / k2 * exp(-k2 * g.c1);
    const double k = sqrt(k2);
    if (g.mode == 2) return -g.c0 * pow(k, 0.5 * g.c1) * bbks_transfer(k, g.c2) / k2;
    return -g.c0 * sqrt(exp(-k / g.c1) / k2) / k2;
}

template <typename R>
__global__ void __launch_bounds__(256) k_fft_x_green(typename Cx<R>::T *spec, typename Cx<R>::T *outB, PlanArg pl,
                                                     const typename Cx<R>::T *tw, const double *kvec, int nb,
                                                     int hp, GreenArgs green) {
    using C = typename Cx<R>::T;
    extern __shared__ __align__(16) unsigned char smem[];
    const int n = pl.n;
    const int ld = n + 1;
    C *a = reinterpret_cast<C *>(smem);
    C *b = a + 2 * nb * ld;
    const int y = blockIdx.y;
    const int kz0 = blockIdx.x * nb;
    const size_t xstride = (size_t)n * hp;
    for (int e = threadIdx.x; e < nb * n; e += blockDim.x) {
        const int t = e / nb, li = e - t * nb;
        const int kz = kz0 + li;
        C v;
        v.x = R(0);
        v.y = R(0);
        if (kz < hp) v = spec[(size_t)t * xstride + (size_t)y * hp + kz];
        a[li * ld + t] = v;
    }
    __syncthreads();
    C *res = fft_lds<R>(a, b, nb, ld, pl, tw, false);
    // res holds delta_k(kx = t, ky = y, kz); build phi (lines 0..nb) and phi * (-i k_x) (lines nb..2nb)
    // in the other buffer.
    C *dst = (res == a) ? b : a;
    const double ky = kvec[y];
    for (int e = threadIdx.x; e < nb * n; e += blockDim.x) {
        const int li = e / n, t = e - li * n;
        const int kz = kz0 + li;
        const double kx = kvec[t];
        const double kzv = kz < hp ? kvec[kz] : 0.0;
        const double k2 = kx * kx + ky * ky + kzv * kzv;
        C v = res[li * ld + t];
        C phi;
        phi.x = R(0);
        phi.y = R(0);
        if (k2 != 0.0 && kz < hp) {
            const R g = (R)green_factor(green, k2);
            phi.x = v.x * g;
            phi.y = v.y * g;
        }
        dst[li * ld + t] = phi;
        const R kxr = (R)kx;
        C m;
        m.x = phi.y * kxr;
        m.y = -phi.x * kxr;
        dst[(nb + li) * ld + t] = m;
    }
    __syncthreads();
    C *other = (dst == a) ? b : a;
    C *res2 = fft_lds<R>(dst, other, 2 * nb, ld, pl, tw, true);
    for (int e = threadIdx.x; e < nb * n; e += blockDim.x) {
        const int t = e / nb, li = e - t * nb;
        const int kz = kz0 + li;
        if (kz < hp) {
            const size_t gi = (size_t)t * xstride + (size_t)y * hp + kz;
            outB[gi] = res2[li * ld + t];
            spec[gi] = res2[(nb + li) * ld + t];
        }
    }
}

// ------------------------------------------------------------------------------ Z inverse

// c2r of three components: A -> x, Cc -> y, Bz * (-i k_z) -> z, written as OutT: float force grids
// as (x, y, z) per mesh point (the interpolation gathers one record per corner), double grids (the
// judge and the ICs) as three planes g[comp * m + cell].
template <typename R, typename OutT>
__global__ void __launch_bounds__(256) k_fft_z_inv3(const typename Cx<R>::T *A, const typename Cx<R>::T *Cc,
                                                    const typename Cx<R>::T *Bz, OutT *g, size_t m, PlanArg pl,
                                                    const typename Cx<R>::T *tw, const double *kvec, int nb,
                                                    int hp) {
    using C = typename Cx<R>::T;
    extern __shared__ __align__(16) unsigned char smem[];
    const int n = pl.n;
    const int ld = n + 1;
    C *a = reinterpret_cast<C *>(smem);
    C *b = a + 3 * nb * ld;
    const size_t nlines = (size_t)n * n;
    const size_t l0 = (size_t)blockIdx.x * nb;
    for (int e = threadIdx.x; e < 3 * nb * hp; e += blockDim.x) {
        const int row = e / hp, kz = e - row * hp;  // row = comp * nb + li
        const int comp = row / nb, li = row - comp * nb;
        const size_t l = l0 + li;
        C v;
        v.x = R(0);
        v.y = R(0);
        if (l < nlines) {
            const size_t gi = l * hp + kz;
            if (comp == 0) {
                v = A[gi];
            } else if (comp == 1) {
                v = Cc[gi];
            } else {
                const C u = Bz[gi];
                const R kz_r = (R)kvec[kz];
                v.x = u.y * kz_r;
                v.y = -u.x * kz_r;
            }
        }
        a[row * ld + kz] = v;
        // Hermitian completion: X[n - kz] = conj(X[kz]) for the upper half of the line.
        const int kc = n - kz;
        if (kz >= 1 && kc >= hp) {
            C cv;
            cv.x = v.x;
            cv.y = -v.y;
            a[row * ld + kc] = cv;
        }
    }
    __syncthreads();
    C *res = fft_lds<R>(a, b, 3 * nb, ld, pl, tw, true);
    if (sizeof(OutT) == sizeof(float)) {
        // Consecutive threads write consecutive floats of the (x, y, z) records.
        for (int e = threadIdx.x; e < 3 * nb * n; e += blockDim.x) {
            const int li = e / (3 * n), rem = e - li * 3 * n;
            const int z = rem / 3, comp = rem - 3 * z;
            const size_t l = l0 + li;
            if (l < nlines) g[3 * (l * n + z) + comp] = (OutT)res[(comp * nb + li) * ld + z].x;
        }
    } else {
        for (int e = threadIdx.x; e < 3 * nb * n; e += blockDim.x) {
            const int row = e / n, z = e - row * n;
            const int comp = row / nb, li = row - comp * nb;
            const size_t l = l0 + li;
            if (l < nlines) g[comp * m + l * n + z] = (OutT)res[row * ld + z].x;
        }
    }
}

// ------------------------------------------------------------ 512-point wavefront transforms
//
// For n = 512 every line transform runs in one wavefront with the four-step split 512 = 8 x 64:
// lane t holds x[t + 64 q] (q = 0..7, one coalesced load per q), computes the 8-point DFT over q
// and multiplies by w512^(t k1); the 64-point DFTs over t run as 8 x 8 through the wavefront's own
// LDS line (two exchanges); lane l ends with X[l + 64 j], j = 0..7, which is again the layout a
// following transform starts from. Forward transforms use exp(-2 pi i nk / n), inverse ones
// exp(+...), both unnormalised (src/fft.rs:27-35). Mesh columns (the y and x passes) are staged
// through LDS in tiles of 8 columns so that global accesses run along kz.
namespace w512 {

constexpr int kN = 512;
constexpr int kP1 = 72;   // pitch of the [k1][t] exchange (conflict-free for 16-lane groups)
constexpr int kP2 = 9;    // pitch of the [m][k1][t_a] exchange
constexpr int kLP = 585;  // LDS floats2 per line (>= 8 kP1 and 64 kP2; 585 mod 16 spreads columns)
constexpr int kCols = 8;  // mesh columns per workgroup in the column passes (one wavefront each)

__device__ __forceinline__ float2 f2(float x, float y) { return make_float2(x, y); }
__device__ __forceinline__ float2 add(float2 a, float2 b) { return f2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 sub(float2 a, float2 b) { return f2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 mul(float2 a, float2 b) { return f2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
// a * conj(b)
__device__ __forceinline__ float2 mulc(float2 a, float2 b) { return f2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y); }

template <bool INV>
__device__ __forceinline__ float2 twm(float2 z, float2 w) {
    return INV ? mulc(z, w) : mul(z, w);
}

// U[k] = sum_n u[n] w4^(nk), w4 = s i (s = -1 forward, +1 inverse).
template <bool INV>
__device__ __forceinline__ void dft4(float2 &u0, float2 &u1, float2 &u2, float2 &u3) {
    const float2 a = add(u0, u2), b = sub(u0, u2), c = add(u1, u3), d = sub(u1, u3);
    const float2 wd = INV ? f2(-d.y, d.x) : f2(d.y, -d.x);
    u0 = add(a, c);
    u2 = sub(a, c);
    u1 = add(b, wd);
    u3 = sub(b, wd);
}

// V[k] = sum_n v[n] w8^(nk), w8 = exp(s 2 pi i / 8), in place (radix-2 split into even and odd).
template <bool INV>
__device__ __forceinline__ void dft8(float2 v[8]) {
    float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<INV>(e0, e1, e2, e3);
    dft4<INV>(o0, o1, o2, o3);
    const float c = 0.70710678118654752f;
    // o_k w8^k: w8 = (c, s c), w8^2 = s i, w8^3 = (-c, s c)
    const float2 t1 = INV ? f2(c * (o1.x - o1.y), c * (o1.y + o1.x)) : f2(c * (o1.x + o1.y), c * (o1.y - o1.x));
    const float2 t2 = INV ? f2(-o2.y, o2.x) : f2(o2.y, -o2.x);
    const float2 t3 = INV ? f2(-c * (o3.x + o3.y), c * (o3.x - o3.y)) : f2(c * (o3.y - o3.x), -c * (o3.x + o3.y));
    v[0] = add(e0, o0);
    v[4] = sub(e0, o0);
    v[1] = add(e1, t1);
    v[5] = sub(e1, t1);
    v[2] = add(e2, t2);
    v[6] = sub(e2, t2);
    v[3] = add(e3, t3);
    v[7] = sub(e3, t3);
}

// Per-lane twiddles of the forward transform: a[k] = w512^(t k1), b[k] = w64^(t_a m) (index 0: 1).
struct Tw {
    float2 a[8];
    float2 b[8];
};

__device__ __forceinline__ void load_tw(Tw &T, const float2 *__restrict__ tw, int lane) {
    T.a[0] = T.b[0] = f2(1.f, 0.f);
#pragma unroll
    for (int k = 1; k < 8; ++k) {
        T.a[k] = tw[(lane * k) & (kN - 1)];
        T.b[k] = tw[(8 * (lane & 7) * k) & (kN - 1)];
    }
}

__device__ __forceinline__ void wsync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// v[q] = x[lane + 64 q] in, v[j] = X[lane + 64 j] out; `line` is this wavefront's LDS scratch.
template <bool INV>
__device__ __forceinline__ void fft(float2 v[8], float2 *line, const Tw &T) {
    const int lane = threadIdx.x & 63;
    dft8<INV>(v);
#pragma unroll
    for (int k = 1; k < 8; ++k) v[k] = twm<INV>(v[k], T.a[k]);
    wsync();
#pragma unroll
    for (int k = 0; k < 8; ++k) line[k * kP1 + lane] = v[k];
    wsync();
    {
        const int ta = lane & 7, k1 = lane >> 3;
#pragma unroll
        for (int tb = 0; tb < 8; ++tb) v[tb] = line[k1 * kP1 + ta + 8 * tb];
        dft8<INV>(v);
#pragma unroll
        for (int m = 1; m < 8; ++m) v[m] = twm<INV>(v[m], T.b[m]);
        wsync();
#pragma unroll
        for (int m = 0; m < 8; ++m) line[(8 * m + k1) * kP2 + ta] = v[m];
    }
    wsync();
    {
        const int k1 = lane & 7, m = lane >> 3;
#pragma unroll
        for (int ta = 0; ta < 8; ++ta) v[ta] = line[(8 * m + k1) * kP2 + ta];
        dft8<INV>(v);
    }
    wsync();
}

// z pass, r2c: four lines per workgroup, one per wavefront.
__global__ void __launch_bounds__(256) k_z_fwd(ZSrc<float> src, float2 *__restrict__ spec,
                                               const float2 *__restrict__ tw, size_t nlines) {
    __shared__ float2 lds[4 * kLP];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const size_t l = (size_t)xcd_block(blockIdx.x, gridDim.x) * 4 + wave;
    if (l >= nlines) return;  // whole wavefronts; no workgroup barrier follows
    Tw T;
    load_tw(T, tw, lane);
    float2 v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const size_t gi = l * kN + lane + 64 * q;
        const float x = zsrc_value<float>(src, gi);
        if (src.delta_out) src.delta_out[gi] = x;
        v[q] = f2(x, 0.f);
    }
    fft<false>(v, lds + wave * kLP, T);
    float2 *out = spec + l * (kN / 2 + 1);
#pragma unroll
    for (int j = 0; j < 4; ++j) out[lane + 64 * j] = v[j];
    if (lane == 0) out[kN / 2] = v[4];
}

// Column passes over the half spectrum (hp = 257 values of kz): AXIS 0 = y lines
// (S[(x n + t) hp + kz] for fixed x), AXIS 1 = x lines (S[(t n + y) hp + kz] for fixed y).
// MODE 0: forward in place; 1: inverse in place; 2: S0 <- iFFT(S0), S1 <- iFFT(S0 (-i k_t));
// 3: forward, Green's function, then S1 <- iFFT(phi), S0 <- iFFT(phi (-i k_t)).
template <int AXIS>
__device__ __forceinline__ size_t col_index(int outer, int t, int kz, int hp) {
    return AXIS == 0 ? ((size_t)outer * kN + t) * hp + kz : ((size_t)t * kN + outer) * hp + kz;
}

template <int AXIS>
__device__ __forceinline__ void tile_store(float2 *__restrict__ S, const float2 *lds, int outer, int kz0, int hp) {
    for (int i = threadIdx.x; i < kN * kCols; i += blockDim.x) {
        const int c = i & (kCols - 1), t = i / kCols, kz = kz0 + c;
        if (kz < hp) S[col_index<AXIS>(outer, t, kz, hp)] = lds[c * kLP + t];
    }
}

template <int AXIS, int MODE>
__global__ void __launch_bounds__(64 * kCols) k_col(float2 *__restrict__ S0, float2 *__restrict__ S1,
                                                     const float2 *__restrict__ tw, const double *__restrict__ kvec,
                                                     int hp, int ntiles, GreenArgs green) {
    __shared__ float2 lds[kCols * kLP];
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x);
    const int tile = (int)(b % (uint32_t)ntiles), outer = (int)(b / (uint32_t)ntiles);
    const int kz0 = tile * kCols;
    for (int i = threadIdx.x; i < kN * kCols; i += blockDim.x) {
        const int c = i & (kCols - 1), t = i / kCols, kz = kz0 + c;
        lds[c * kLP + t] = kz < hp ? S0[col_index<AXIS>(outer, t, kz, hp)] : f2(0.f, 0.f);
    }
    __syncthreads();
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    float2 *line = lds + w * kLP;
    Tw T;
    load_tw(T, tw, lane);
    float2 v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = line[lane + 64 * q];
    if (MODE == 0 || MODE == 1) {
        fft<MODE == 1>(v, line, T);
#pragma unroll
        for (int j = 0; j < 8; ++j) line[lane + 64 * j] = v[j];
        __syncthreads();
        tile_store<AXIS>(S0, lds, outer, kz0, hp);
        return;
    }
    float2 m[8];
    if (MODE == 2) {
        // (v) (0 - i k_t): re = v.y k, im = -v.x k, for t = lane + 64 q
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const float k = (float)kvec[lane + 64 * q];
            m[q] = f2(v[q].y * k, -v[q].x * k);
        }
    } else {
        fft<false>(v, line, T);
        const int kz = kz0 + w;
        const double ko = kvec[outer];
        const double kzv = kz < hp ? kvec[kz] : 0.0;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const double kt = kvec[lane + 64 * j];
            const double k2 = kt * kt + ko * ko + kzv * kzv;
            float2 phi = f2(0.f, 0.f);
            if (k2 != 0.0 && kz < hp) {
                const float g = (float)green_factor(green, k2);
                phi = f2(v[j].x * g, v[j].y * g);
            }
            const float ktf = (float)kt;
            v[j] = phi;
            m[j] = f2(phi.y * ktf, -phi.x * ktf);
        }
    }
    fft<true>(v, line, T);
#pragma unroll
    for (int j = 0; j < 8; ++j) line[lane + 64 * j] = v[j];
    __syncthreads();
    tile_store<AXIS>(MODE == 2 ? S0 : S1, lds, outer, kz0, hp);
    __syncthreads();
    fft<true>(m, line, T);
#pragma unroll
    for (int j = 0; j < 8; ++j) line[lane + 64 * j] = m[j];
    __syncthreads();
    tile_store<AXIS>(MODE == 2 ? S1 : S0, lds, outer, kz0, hp);
}

// z pass, c2r of three components: A -> x, Cc -> y, Bz (-i k_z) -> z, the upper half of each line
// by Hermitian completion; one line per wavefront, written as (x, y, z) per mesh point so that the
// interpolation gathers one 12-byte record per corner.
__global__ void __launch_bounds__(256) k_z_inv3(const float2 *__restrict__ A, const float2 *__restrict__ Cc,
                                                const float2 *__restrict__ Bz, float *__restrict__ g,
                                                const float2 *__restrict__ tw, const double *__restrict__ kvec,
                                                size_t nlines) {
    __shared__ float2 lds[4 * kLP];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const size_t l = (size_t)xcd_block(blockIdx.x, gridDim.x) * 4 + wave;
    if (l >= nlines) return;  // whole wavefronts; no workgroup barrier follows
    Tw T;
    load_tw(T, tw, lane);
    float r[3][8];
#pragma unroll
    for (int comp = 0; comp < 3; ++comp) {
        const float2 *S = (comp == 0 ? A : comp == 1 ? Cc : Bz) + l * (kN / 2 + 1);
        float2 v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int i = lane + 64 * q;
            const int kz = i <= kN / 2 ? i : kN - i;
            float2 u = S[kz];
            if (comp == 2) {
                const float k = (float)kvec[kz];
                u = f2(u.y * k, -u.x * k);
            }
            v[q] = i <= kN / 2 ? u : f2(u.x, -u.y);
        }
        fft<true>(v, lds + wave * kLP, T);
#pragma unroll
        for (int j = 0; j < 8; ++j) r[comp][j] = v[j].x;
    }
    float *out = g + 3 * (l * kN);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const size_t c = (size_t)(lane + 64 * j);
        out[3 * c] = r[0][j];
        out[3 * c + 1] = r[1][j];
        out[3 * c + 2] = r[2][j];
    }
}

}  // namespace w512

// ----------------------------------------------------------------------------- host side

static int factor_n(int n, int *radix) {
    int s = 0;
    while (n % 4 == 0) {
        radix[s++] = 4;
        n /= 4;
    }
    while (n % 2 == 0) {
        radix[s++] = 2;
        n /= 2;
    }
    for (int p = 3; n > 1; p += 2)
        while (n % p == 0) {
            radix[s++] = p;
            n /= p;
        }
    return s;
}

// src/fft.rs:86-95, evaluated in double.
static double wavenumber64(int i, int n, double box) {
    const double tpl = 2.0 * M_PI / box;
    if (n % 2 == 0 && i == n / 2) return 0.0;
    if (i <= n / 2) return (double)i * tpl;
    return ((double)i - (double)n) * tpl;
}

void make_plan(Ctx &c) {
    const int n = c.n;
    c.plan.n = n;
    c.plan.nst = factor_n(n, c.plan.radix);
    std::vector<float2> t32(n);
    std::vector<double2> t64(n);
    std::vector<double> kv(n);
    for (int t = 0; t < n; ++t) {
        const double ang = -2.0 * M_PI * (double)t / (double)n;
        t64[t] = make_double2(cos(ang), sin(ang));
        t32[t] = make_float2((float)cos(ang), (float)sin(ang));
        kv[t] = wavenumber64(t, n, c.p.box_size);
    }
    TPM_HIP(hipMemcpy(c.tw32, t32.data(), n * sizeof(float2), hipMemcpyHostToDevice));
    TPM_HIP(hipMemcpy(c.tw64, t64.data(), n * sizeof(double2), hipMemcpyHostToDevice));
    TPM_HIP(hipMemcpy(c.kvec, kv.data(), n * sizeof(double), hipMemcpyHostToDevice));
}

static PlanArg plan_arg(const Ctx &c) {
    PlanArg p{};
    p.n = c.plan.n;
    p.nst = c.plan.nst;
    for (int i = 0; i < 24; ++i) p.radix[i] = c.plan.radix[i];
    return p;
}

// Lines per workgroup so that `buffers` LDS buffers of mult*nb lines fit in 64 KiB.
static int lines_per_block(int n, size_t elem, int mult, int buffers, int cap) {
    const size_t per_line = (size_t)(n + 1) * elem * mult * buffers;
    int nb = (int)(65536 / per_line);
    if (nb > cap) nb = cap;
    if (nb < 1) nb = 1;
    return nb;
}

template <typename R>
static void pm_pipeline(Ctx &c, const ZSrc<R> &src, typename Cx<R>::T *SA, typename Cx<R>::T *SB,
                        typename Cx<R>::T *SC, const typename Cx<R>::T *tw, void *gout, const GreenArgs &green) {
    using C = typename Cx<R>::T;
    const int n = c.n, hp = c.hp;
    const PlanArg pl = plan_arg(c);
    const size_t es = sizeof(C);
    const int ld = n + 1;
    TPM_REQUIRE(2 * (size_t)ld * es <= 65536, "mesh too large for the LDS line transform");

    if constexpr (sizeof(R) == 4) {
        if (n == w512::kN && c.fft_path == 0) {
            // Wavefront transforms: z r2c, y forward, x forward + Green + 2 inverse, y inverse
            // (A; B -> B, C), z c2r of the three components.
            const size_t lines = (size_t)n * n;
            const int ntiles = (hp + w512::kCols - 1) / w512::kCols;
            const dim3 gcol((unsigned)(ntiles * n)), bcol(64 * w512::kCols);
            const double *kv = c.kvec;
            hipLaunchKernelGGL(w512::k_z_fwd, dim3((unsigned)(lines / 4)), dim3(256), 0, c.stream, src, SA, tw, lines);
            hipLaunchKernelGGL((w512::k_col<0, 0>), gcol, bcol, 0, c.stream, SA, (float2 *)nullptr, tw, kv, hp, ntiles,
                               green);
            hipLaunchKernelGGL((w512::k_col<1, 3>), gcol, bcol, 0, c.stream, SA, SB, tw, kv, hp, ntiles, green);
            hipLaunchKernelGGL((w512::k_col<0, 1>), gcol, bcol, 0, c.stream, SA, (float2 *)nullptr, tw, kv, hp, ntiles,
                               green);
            hipLaunchKernelGGL((w512::k_col<0, 2>), gcol, bcol, 0, c.stream, SB, SC, tw, kv, hp, ntiles, green);
            hipLaunchKernelGGL(w512::k_z_inv3, dim3((unsigned)(lines / 4)), dim3(256), 0, c.stream, SA, SC, SB,
                               (float *)gout, tw, kv, lines);
            TPM_HIP(hipGetLastError());
            return;
        }
    }

    // Z forward
    {
        const int nb = lines_per_block(n, es, 1, 2, c.fft_cap);
        const size_t lines = (size_t)n * n;
        dim3 grid((unsigned)((lines + nb - 1) / nb));
        const size_t sh = 2 * (size_t)nb * ld * es;
        hipLaunchKernelGGL(k_fft_z_fwd<R>, grid, dim3(256), sh, c.stream, src, SA, pl, tw, nb, hp);
        TPM_HIP(hipGetLastError());
    }
    // Y forward
    {
        const int nb = lines_per_block(n, es, 1, 2, c.fft_cap);
        dim3 grid((unsigned)((hp + nb - 1) / nb), (unsigned)n);
        const size_t sh = 2 * (size_t)nb * ld * es;
        hipLaunchKernelGGL(k_fft_y<R>, grid, dim3(256), sh, c.stream, SA, (C *)nullptr, pl, tw,
                           (const double *)c.kvec, nb, hp, 0);
        TPM_HIP(hipGetLastError());
    }
    // X forward + Green + 2 inverse
    {
        const int nb = lines_per_block(n, es, 2, 2, c.fft_cap);
        dim3 grid((unsigned)((hp + nb - 1) / nb), (unsigned)n);
        const size_t sh = 4 * (size_t)nb * ld * es;
        TPM_REQUIRE(sh <= 65536, "mesh too large for the fused x pass");
        hipLaunchKernelGGL(k_fft_x_green<R>, grid, dim3(256), sh, c.stream, SA, SB, pl, tw, (const double *)c.kvec,
                           nb, hp, green);
        TPM_HIP(hipGetLastError());
    }
    // Y inverse: A in place; B -> B, C
    {
        const int nb = lines_per_block(n, es, 1, 2, c.fft_cap);
        dim3 grid((unsigned)((hp + nb - 1) / nb), (unsigned)n);
        const size_t sh = 2 * (size_t)nb * ld * es;
        hipLaunchKernelGGL(k_fft_y<R>, grid, dim3(256), sh, c.stream, SA, (C *)nullptr, pl, tw,
                           (const double *)c.kvec, nb, hp, 1);
        TPM_HIP(hipGetLastError());
        const int nb2 = lines_per_block(n, es, 2, 2, c.fft_cap);
        dim3 grid2((unsigned)((hp + nb2 - 1) / nb2), (unsigned)n);
        const size_t sh2 = 4 * (size_t)nb2 * ld * es;
        TPM_REQUIRE(sh2 <= 65536, "mesh too large for the split y pass");
        hipLaunchKernelGGL(k_fft_y<R>, grid2, dim3(256), sh2, c.stream, SB, SC, pl, tw, (const double *)c.kvec,
                           nb2, hp, 2);
        TPM_HIP(hipGetLastError());
    }
    // Z inverse (c2r) of the three components
    {
        const int nb = lines_per_block(n, es, 3, 2, c.fft_cap);
        const size_t sh = 6 * (size_t)nb * ld * es;
        TPM_REQUIRE(sh <= 65536, "mesh too large for the c2r pass");
        const size_t lines = (size_t)n * n;
        dim3 grid((unsigned)((lines + nb - 1) / nb));
        if (sizeof(R) == 4) {
            hipLaunchKernelGGL((k_fft_z_inv3<R, float>), grid, dim3(256), sh, c.stream, SA, SC, SB, (float *)gout,
                               c.m, pl, tw, (const double *)c.kvec, nb, hp);
        } else {
            hipLaunchKernelGGL((k_fft_z_inv3<R, double>), grid, dim3(256), sh, c.stream, SA, SC, SB, (double *)gout,
                               c.m, pl, tw, (const double *)c.kvec, nb, hp);
        }
        TPM_HIP(hipGetLastError());
    }
}

static GreenArgs gravity(const Ctx &c, double a, double r_split) {
    GreenArgs g;
    g.mode = 0;
    g.c0 = 1.5 * c.p.omega_m / a / ((double)c.n * c.n * c.n);
    g.c1 = r_split * r_split;
    g.c2 = 0.0;
    return g;
}

void pm_from_grid(Ctx &c, size_t n_particles, double a, double r_split, float *delta_out) {
    ZSrc<float> src{};
    src.grid = c.grid;
    src.delta = nullptr;
    src.mean32 = (float)n_particles / (float)c.m;
    src.mean64 = (double)n_particles / (double)c.m;
    src.delta_out = delta_out;
    pm_pipeline<float>(c, src, c.specA, c.specB, c.specC, c.tw32, c.force, gravity(c, a, r_split));
}

void pm_from_delta(Ctx &c, const float *d_delta, double a, double r_split) {
    ZSrc<float> src{};
    src.grid = nullptr;
    src.delta = d_delta;
    src.delta_out = nullptr;
    pm_pipeline<float>(c, src, c.specA, c.specB, c.specC, c.tw32, c.force, gravity(c, a, r_split));
}

void pm_from_grid_f64(Ctx &c, size_t n_particles, double a, double r_split) {
    if (!c.dspecA) {
        TPM_HIP(hipMalloc(&c.dspecA, c.hs * sizeof(double2)));
        TPM_HIP(hipMalloc(&c.dspecB, c.hs * sizeof(double2)));
        TPM_HIP(hipMalloc(&c.dspecC, c.hs * sizeof(double2)));
        TPM_HIP(hipMalloc(&c.dforce, 3 * c.m * sizeof(double)));
    }
    ZSrc<double> src{};
    src.grid = c.grid;
    src.delta = nullptr;
    src.mean32 = (float)n_particles / (float)c.m;
    src.mean64 = (double)n_particles / (double)c.m;
    src.delta_out = nullptr;
    pm_pipeline<double>(c, src, c.dspecA, c.dspecB, c.dspecC, c.tw64, c.dforce, gravity(c, a, r_split));
}

// Zel'dovich displacement grids from a white-noise grid on the device, through the FP64
// pipeline. Writes c.dforce = (psi_x, psi_y, psi_z) on the mesh points.
void displacement_from_noise(Ctx &c, const float *d_noise, const DisplacementSpec &spec) {
    if (!c.dspecA) {
        TPM_HIP(hipMalloc(&c.dspecA, c.hs * sizeof(double2)));
        TPM_HIP(hipMalloc(&c.dspecB, c.hs * sizeof(double2)));
        TPM_HIP(hipMalloc(&c.dspecC, c.hs * sizeof(double2)));
        TPM_HIP(hipMalloc(&c.dforce, 3 * c.m * sizeof(double)));
    }
    ZSrc<double> src{};
    src.grid = nullptr;
    src.delta = d_noise;
    src.delta_out = nullptr;
    GreenArgs g;
    const double m = (double)c.n * c.n * c.n;
    const double vol = c.p.box_size * c.p.box_size * c.p.box_size;
    g.c0 = sqrt(m / vol) / m;
    if (spec.bbks) {
        g.mode = 2;
        g.c0 *= sqrt(spec.amplitude);
        g.c1 = spec.n_s;
        g.c2 = spec.gamma;
    } else {
        g.mode = 1;
        g.c1 = spec.k_c;
        g.c2 = 0.0;
    }
    pm_pipeline<double>(c, src, c.dspecA, c.dspecB, c.dspecC, c.tw64, c.dforce, g);
}

}  // namespace tpm
