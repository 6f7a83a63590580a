// Kick-drift-kick stepping on the device with FP64 positions and momenta.
//
// Replaces Simulation::step_to / kick / drift (src/integrator.rs:174-186, 300-316). The step
// factors (Simpson kick/drift integrals and the time midpoint, src/integrator.rs:262-297) are
// scalar host work, evaluated here in FP64 with the reference's 64 Simpson intervals and 60
// bisections; E(a) is src/cosmology.rs:32-34.
#include <cmath>

#include "tpm_internal.h"

namespace tpm {

__global__ void k_kick(double *__restrict__ mom, const float *__restrict__ acc, size_t count, double f) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) mom[i] += (double)acc[i] * f;
}

// x = (x + p f).rem_euclid(L) (src/integrator.rs:310-316).
__global__ void k_drift(double *__restrict__ pos, const double *__restrict__ mom, size_t count, double f,
                        double box) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    double r = fmod(pos[i] + mom[i] * f, box);
    if (r < 0.0) r += box;
    pos[i] = r;
}

static double cosmo_e(const tpm_cosmology &c, double a) {
    return std::sqrt(c.omega_r / (a * a * a * a) + c.omega_m / (a * a * a) + c.omega_l);
}

// Simpson's rule with 64 intervals; kind 0: da/(a E), 1: da/(a^3 E).
static double integrate(const tpm_cosmology &c, double lo, double hi, int kind) {
    auto f = [&](double a) { return kind == 0 ? 1.0 / (a * cosmo_e(c, a)) : 1.0 / (a * a * a * cosmo_e(c, a)); };
    const int n = 64;
    const double h = (hi - lo) / n;
    double sum = f(lo) + f(hi);
    for (int i = 1; i < n; ++i) sum += (i % 2 == 1 ? 4.0 : 2.0) * f(lo + i * h);
    return sum * h / 3.0;
}

KdkFactors kdk_factors(const tpm_cosmology &c, double a0, double a1) {
    KdkFactors k{};
    const double half = integrate(c, a0, a1, 0) / 2.0;
    double lo = a0, hi = a1;
    for (int it = 0; it < 60; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (integrate(c, a0, mid, 0) < half)
            lo = mid;
        else
            hi = mid;
    }
    k.am = 0.5 * (lo + hi);
    k.k1 = integrate(c, a0, k.am, 0);
    k.d = integrate(c, a0, a1, 1);
    k.k2 = integrate(c, k.am, a1, 0);
    return k;
}

// Linear growth state (delta, d delta / d ln a) at a_end: the growth ODE in ln a,
//   delta'' + (2 + dlnE/dlna) delta' = 1.5 Om delta / (a^3 E^2),
// RK4 with 4000 steps from a = 1e-3, started on the Meszaros growing mode 1 + 1.5 y (y = a Om/Or)
// with radiation, or delta = a without (Cosmology::growth_state, src/cosmology.rs:41-79), in FP64.
static void growth_state(const tpm_cosmology &c, double a_end, double *d_out, double *dl_out) {
    const double a_start = 1e-3;
    double d, dl;
    if (c.omega_r > 0.0) {
        const double y = a_start * c.omega_m / c.omega_r;
        d = 1.0 + 1.5 * y;
        dl = 1.5 * y;
    } else {
        d = a_start;
        dl = a_start;
    }
    auto rhs = [&](double la, double x, double xl, double *fx, double *fxl) {
        const double a = std::exp(la);
        const double a3 = a * a * a, a4 = a3 * a;
        const double e2 = c.omega_r / a4 + c.omega_m / a3 + c.omega_l;
        const double dln_e = 0.5 * (-4.0 * c.omega_r / a4 - 3.0 * c.omega_m / a3) / e2;
        *fx = xl;
        *fxl = -(2.0 + dln_e) * xl + 1.5 * c.omega_m / (a3 * e2) * x;
    };
    const int steps = 4000;
    const double la0 = std::log(a_start);
    const double h = (std::log(a_end) - la0) / steps;
    for (int s = 0; s < steps; ++s) {
        const double la = la0 + s * h;
        double k1, l1, k2, l2, k3, l3, k4, l4;
        rhs(la, d, dl, &k1, &l1);
        rhs(la + 0.5 * h, d + 0.5 * h * k1, dl + 0.5 * h * l1, &k2, &l2);
        rhs(la + 0.5 * h, d + 0.5 * h * k2, dl + 0.5 * h * l2, &k3, &l3);
        rhs(la + h, d + h * k3, dl + h * l3, &k4, &l4);
        d += h / 6.0 * (k1 + 2.0 * k2 + 2.0 * k3 + k4);
        dl += h / 6.0 * (l1 + 2.0 * l2 + 2.0 * l3 + l4);
    }
    *d_out = d;
    *dl_out = dl;
}

void growth(const tpm_cosmology &c, double a, double *d, double *f) {
    double d1, dl1, da, dla;
    growth_state(c, 1.0, &d1, &dl1);
    growth_state(c, a, &da, &dla);
    if (d) *d = da / d1;
    if (f) *f = dla / da;
}

double hubble_e(const tpm_cosmology &c, double a) { return cosmo_e(c, a); }

static inline unsigned blocks(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

void kick(Ctx &c, double *d_mom, const float *d_acc, size_t n, double f) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_kick, dim3(blocks(3 * n, 256)), dim3(256), 0, c.stream, d_mom, d_acc, 3 * n, f);
    TPM_HIP(hipGetLastError());
}

void drift(Ctx &c, double *d_pos, const double *d_mom, size_t n, double f) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_drift, dim3(blocks(3 * n, 256)), dim3(256), 0, c.stream, d_pos, d_mom, 3 * n, f,
                       c.p.box_size);
    TPM_HIP(hipGetLastError());
}

}  // namespace tpm
