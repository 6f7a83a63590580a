// Zel'dovich initial conditions on a particle lattice (one particle per mesh point).
//
// Follows the structure of the reference's zeldovich (src/ics.rs:119-182): white noise on the
// grid, forward transform, colour each mode, displacement Psi_k = i k F_k / k^2 with the
// reference's wavenumber convention, three inverse transforms, x = q + Psi(q) on the lattice
// q = i h (src/ics.rs:167), wrapped into [0, L), and p = (a^2 f E) Psi (src/ics.rs:157-160).
// Differences, by design of the workloads this solver is measured on: the spectrum is the
// power law P(k) = k^-2 exp(-k / k_c), and its amplitude is set by rescaling Psi to a target
// 3D displacement RMS. The noise is a counter-based hash (SplitMix64 finaliser, Box-Muller),
// so a grid point's value depends only on (seed, index) and can be restated off the GPU.
#include <algorithm>
#include <cmath>
#include <utility>
#include <vector>

#include "tpm_internal.h"

namespace tpm {

__host__ __device__ inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Standard normal deviate for grid point i (FP64 Box-Muller, rounded to float).
__global__ void k_noise(float *__restrict__ out, size_t m, uint64_t seed) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const uint64_t x = splitmix64(seed ^ splitmix64((uint64_t)i));
    const uint64_t y = splitmix64(x);
    const double u1 = ((double)(x >> 11) + 1.0) * 0x1p-53;  // (0, 1]
    const double u2 = (double)(y >> 11) * 0x1p-53;          // [0, 1)
    out[i] = (float)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
}

// Per-block sums of psi_x^2 + psi_y^2 + psi_z^2 in a fixed order (deterministic).
__global__ void __launch_bounds__(256) k_sumsq(const double *__restrict__ g, size_t m, double *__restrict__ part) {
    __shared__ double sh[256];
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < m; i += (size_t)gridDim.x * 256) {
        const double a = g[i], b = g[m + i], c = g[2 * m + i];
        s += a * a + b * b + c * c;
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__device__ inline double rem_euclid(double x, double l) {
    double r = fmod(x, l);
    if (r < 0.0) r += l;
    if (r >= l) r = 0.0;  // x slightly below 0 can round up to exactly l
    return r;
}

__global__ void k_place(const double *__restrict__ g, int n, double box, double scale, double mom_factor,
                        double *__restrict__ pos, double *__restrict__ mom) {
    const size_t m = (size_t)n * n * n;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int iz = (int)(i % n);
    const int iy = (int)((i / n) % n);
    const int ix = (int)(i / ((size_t)n * n));
    const double h = box / (double)n;
    const double q[3] = {ix * h, iy * h, iz * h};
    for (int a = 0; a < 3; ++a) {
        const double d = scale * g[a * m + i];
        pos[3 * i + a] = rem_euclid(q[a] + d, box);
        if (mom) mom[3 * i + a] = mom_factor * d;
    }
}

static inline unsigned blocks(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

// ---------------------------------------------------------------- BBKS / sigma_8 spectrum

// BBKS transfer function, dark matter only (src/cosmology.rs:93-103): q = k / (Omega_m h).
static double bbks_transfer_host(double k, double gamma) {
    const double q = k / gamma;
    if (q == 0.0) return 1.0;
    const double a = 16.1 * q, b = 5.46 * q, c = 6.71 * q;
    const double poly = 1.0 + 3.89 * q + a * a + b * b * b + (c * c) * (c * c);
    return std::log1p(2.34 * q) / (2.34 * q) * std::pow(poly, -0.25);
}

// Spherical top-hat window W(x) = 3 (sin x - x cos x) / x^3, 1 below x = 1e-4 (src/ics.rs:67-74).
static double tophat(double x) {
    if (x < 1e-4) return 1.0;
    return 3.0 * (std::sin(x) - x * std::cos(x)) / (x * x * x);
}

double PowerSpec::power(double k) const {
    if (k == 0.0) return 0.0;
    const double t = bbks_transfer_host(k, gamma);
    return amplitude * std::pow(k, n_s) * t * t;
}

// sigma(R)^2 = (1 / 2 pi^2) int dk k^2 P(k) W(kR)^2, Simpson's rule over ln k in [1e-4, 1e3] with
// 8192 intervals (src/ics.rs:35-52), in FP64.
double PowerSpec::sigma(double radius) const {
    const int steps = 8192;
    const double lo = std::log(1e-4), hi = std::log(1e3);
    const double h = (hi - lo) / steps;
    auto f = [&](double lk) {
        const double k = std::exp(lk);
        const double w = tophat(k * radius);
        return k * k * k * power(k) * w * w;
    };
    double sum = f(lo) + f(hi);
    for (int i = 1; i < steps; ++i) sum += (i % 2 == 1 ? 4.0 : 2.0) * f(lo + i * h);
    const double integral = sum * h / 3.0;
    return std::sqrt(integral / (2.0 * M_PI * M_PI));
}

// PowerSpectrum::normalized (src/ics.rs:54-64): the amplitude that gives sigma(radius) = sigma8.
PowerSpec normalized_power(const tpm_cosmology &cosmo, double h0, double n_s, double sigma8, double radius) {
    PowerSpec p;
    p.amplitude = 1.0;
    p.n_s = n_s;
    p.gamma = cosmo.omega_m * h0 / 100.0;
    const double s = p.sigma(radius);
    p.amplitude = (sigma8 / s) * (sigma8 / s);
    return p;
}

void zeldovich_bbks(Ctx &c, const tpm_cosmology &cosmo, const PowerSpec &ps, uint64_t seed, double a_init,
                    double *d_pos, double *d_mom) {
    TPM_REQUIRE(c.m <= c.max_n, "context capacity must hold mesh_n^3 particles");
    TPM_REQUIRE(a_init > 0 && std::isfinite(a_init), "a_init must be positive");
    hipLaunchKernelGGL(k_noise, dim3(blocks(c.m, 256)), dim3(256), 0, c.stream, c.delta, c.m, seed);
    TPM_HIP(hipGetLastError());
    DisplacementSpec spec;
    spec.bbks = true;
    spec.amplitude = ps.amplitude;
    spec.n_s = ps.n_s;
    spec.gamma = ps.gamma;
    displacement_from_noise(c, c.delta, spec);
    // x = q + D Psi, p = a^2 f E D Psi (src/ics.rs:155-177): the displacement at a_init times a^2 f E.
    double d = 0.0, f = 0.0;
    growth(cosmo, a_init, &d, &f);
    const double e = hubble_e(cosmo, a_init);
    hipLaunchKernelGGL(k_place, dim3(blocks(c.m, 256)), dim3(256), 0, c.stream, c.dforce, c.n, c.p.box_size, d,
                       a_init * a_init * f * e, d_pos, d_mom);
    TPM_HIP(hipGetLastError());
}

void zeldovich(Ctx &c, uint64_t seed, double kc, double disp_rms, double mom_factor, double *d_pos, double *d_mom) {
    TPM_REQUIRE(c.m <= c.max_n, "context capacity must hold mesh_n^3 particles");
    TPM_REQUIRE(kc > 0 && disp_rms >= 0, "k_c must be positive and the displacement RMS >= 0");
    // Noise lives in the delta grid buffer (n^3 floats).
    hipLaunchKernelGGL(k_noise, dim3(blocks(c.m, 256)), dim3(256), 0, c.stream, c.delta, c.m, seed);
    TPM_HIP(hipGetLastError());
    DisplacementSpec spec;
    spec.k_c = kc;
    displacement_from_noise(c, c.delta, spec);
    const unsigned nb = 1024;
    double *part = reinterpret_cast<double *>(c.judge_buf);
    if (c.judge_cap < nb) {
        if (c.judge_buf) (void)hipFree(c.judge_buf);
        TPM_HIP(hipMalloc(&c.judge_buf, nb * (sizeof(uint32_t) * 2 + 6 * sizeof(double)) + 256));
        c.judge_cap = nb;
        part = reinterpret_cast<double *>(c.judge_buf);
    }
    hipLaunchKernelGGL(k_sumsq, dim3(nb), dim3(256), 0, c.stream, c.dforce, c.m, part);
    TPM_HIP(hipGetLastError());
    std::vector<double> hpart(nb);
    TPM_HIP(hipMemcpyAsync(hpart.data(), part, nb * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    TPM_HIP(hipStreamSynchronize(c.stream));
    double s = 0.0;
    for (double v : hpart) s += v;
    const double rms = sqrt(s / (double)c.m);
    const double scale = rms > 0 ? disp_rms / rms : 0.0;
    hipLaunchKernelGGL(k_place, dim3(blocks(c.m, 256)), dim3(256), 0, c.stream, c.dforce, c.n, c.p.box_size, scale,
                       mom_factor, d_pos, d_mom);
    TPM_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ clustered NFW particles

// One halo of the clustered workload: centre, r200, NFW concentration and first particle.
struct HaloDev {
    double cx, cy, cz, r200, conc;
    uint64_t first;
};

__device__ inline double u01(uint64_t seed, uint64_t idx, uint64_t stream) {
    const uint64_t x = splitmix64(seed ^ splitmix64(idx * 4 + stream));
    return ((double)(x >> 11) + 0.5) * 0x1p-53;  // (0, 1)
}

// NFW enclosed-mass shape m(x) = ln(1 + x) - x / (1 + x), x = r / r_s.
__device__ inline double nfw_m(double x) { return log1p(x) - x / (1.0 + x); }

// Particle i: halo particles (i < n_halo_part) sample r from the NFW mass profile truncated at
// r200 (bisection on m(x) = u m(c)) in an isotropic direction; the rest are uniform.
__global__ void k_nfw_place(const HaloDev *__restrict__ halos, int nh, size_t n_halo_part, size_t n, double box,
                            uint64_t seed, double *__restrict__ pos) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double p[3];
    if (i < n_halo_part) {
        int lo = 0, hi = nh - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (halos[mid].first <= i)
                lo = mid;
            else
                hi = mid - 1;
        }
        const HaloDev h = halos[lo];
        const double target = u01(seed, i, 0) * nfw_m(h.conc);
        double a = 0.0, b = h.conc;
        for (int it = 0; it < 60; ++it) {
            const double mid = 0.5 * (a + b);
            if (nfw_m(mid) < target)
                a = mid;
            else
                b = mid;
        }
        const double r = 0.5 * (a + b) * h.r200 / h.conc;
        const double cth = 2.0 * u01(seed, i, 1) - 1.0;
        const double sth = sqrt(fmax(0.0, 1.0 - cth * cth));
        const double ph = 6.283185307179586 * u01(seed, i, 2);
        p[0] = h.cx + r * sth * cos(ph);
        p[1] = h.cy + r * sth * sin(ph);
        p[2] = h.cz + r * cth;
    } else {
        for (int a = 0; a < 3; ++a) p[a] = box * u01(seed, i, (uint64_t)a);
    }
    for (int a = 0; a < 3; ++a) pos[3 * i + a] = rem_euclid(p[a], box);
}

// Host-side halo catalogue (counter-based, so it is reproducible from the seed alone).
static double host_u01(uint64_t seed, uint64_t idx, uint64_t stream) {
    const uint64_t x = splitmix64(seed ^ splitmix64(idx * 4 + stream));
    return ((double)(x >> 11) + 0.5) * 0x1p-53;
}

void nfw_halos(Ctx &c, const NfwSpec &sp, size_t n, double *d_pos) {
    TPM_REQUIRE(n <= c.max_n, "more particles than the context capacity");
    TPM_REQUIRE(sp.n_halos > 0 && sp.halo_fraction >= 0 && sp.halo_fraction <= 1, "bad halo count or fraction");
    TPM_REQUIRE(sp.mass_min > 0 && sp.mass_max > sp.mass_min && sp.slope != 1.0, "bad halo mass range or slope");
    TPM_REQUIRE(sp.conc_min > 0 && sp.conc_max >= sp.conc_min, "bad concentration range");
    const double box = c.p.box_size;
    const int nh = sp.n_halos;
    const uint64_t hseed = splitmix64(sp.seed ^ 0xA5A5A5A5A5A5A5A5ULL);
    // Masses from dN/dM ~ M^-slope on [mass_min, mass_max] by inverse transform.
    const double e = 1.0 - sp.slope;
    const double a0 = std::pow(sp.mass_min, e), a1 = std::pow(sp.mass_max, e);
    std::vector<double> mass(nh);
    double mtot = 0.0;
    for (int h = 0; h < nh; ++h) {
        mass[h] = std::pow(a0 + host_u01(hseed, h, 0) * (a1 - a0), 1.0 / e);
        mtot += mass[h];
    }
    // Particles per halo in proportion to mass (largest remainders get the leftovers).
    const size_t nhp = (size_t)std::llround(sp.halo_fraction * (double)n);
    std::vector<size_t> cnt(nh);
    std::vector<std::pair<double, int>> rem(nh);
    size_t used = 0;
    for (int h = 0; h < nh; ++h) {
        const double want = mass[h] / mtot * (double)nhp;
        cnt[h] = (size_t)std::floor(want);
        used += cnt[h];
        rem[h] = {want - std::floor(want), h};
    }
    std::sort(rem.begin(), rem.end(), [](const std::pair<double, int> &x, const std::pair<double, int> &y) {
        return x.first > y.first || (x.first == y.first && x.second < y.second);
    });
    for (size_t k = 0; used < nhp; ++k, ++used) cnt[rem[k % nh].second]++;
    // r200 = (3 M / (4 pi 200 rho_crit))^(1/3), rho_crit = 2.775e11 (M_sun/h) / (Mpc/h)^3.
    const double rho_crit = 2.77536627e11;
    std::vector<HaloDev> hd(nh);
    uint64_t first = 0;
    for (int h = 0; h < nh; ++h) {
        hd[h].cx = box * host_u01(hseed, h, 1);
        hd[h].cy = box * host_u01(hseed, h, 2);
        hd[h].cz = box * host_u01(hseed, h, 3);
        hd[h].r200 = std::cbrt(3.0 * mass[h] / (4.0 * M_PI * 200.0 * rho_crit));
        hd[h].conc = sp.conc_min + (sp.conc_max - sp.conc_min) * host_u01(hseed + 1, h, 0);
        hd[h].first = first;
        first += cnt[h];
    }
    HaloDev *d_h = nullptr;
    TPM_HIP(hipMalloc(&d_h, nh * sizeof(HaloDev)));
    try {
        TPM_HIP(hipMemcpyAsync(d_h, hd.data(), nh * sizeof(HaloDev), hipMemcpyHostToDevice, c.stream));
        if (n) {
            hipLaunchKernelGGL(k_nfw_place, dim3(blocks(n, 256)), dim3(256), 0, c.stream, d_h, nh, nhp, n, box,
                               splitmix64(sp.seed), d_pos);
            TPM_HIP(hipGetLastError());
        }
        TPM_HIP(hipStreamSynchronize(c.stream));
    } catch (...) {
        (void)hipFree(d_h);
        throw;
    }
    (void)hipFree(d_h);
}

}  // namespace tpm
