// Headless runner: RunSettings and the worker loop of src/runner.rs (the `--no-viz` path of
// src/main.rs:34-37) on the GPU.
//
// init_simulation (src/runner.rs:186-206): BBKS/sigma_8 Zel'dovich ICs for particles_per_side^3
// particles at z_start, a mesh of twice the particle count per side, ForceSplit::with when TreePM
// is on. run_simulation (src/runner.rs:208-260): steps evenly spaced in ln a, one KDK step each,
// a snapshot per step and the maximum peculiar speed as the blow-up check. The frame goes to a
// caller callback instead of a channel; a non-zero return from it stops the run like the
// control's stop flag.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "tpm_internal.h"

namespace tpm {

// Largest non-negative float of an array, by its bits (non-negative floats order like their bits).
__global__ void k_max_nonneg(const float *__restrict__ v, size_t n, unsigned int *__restrict__ out) {
    unsigned int m = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        m = max(m, __float_as_uint(fmaxf(v[i], 0.0f)));
    for (int d = 32; d >= 1; d >>= 1) m = max(m, (unsigned int)__shfl_xor((int)m, d, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

}  // namespace tpm

tpm::Ctx &tpm_ctx_internal(tpm_ctx *ctx);

namespace {

// Cosmology::lcdm_default (src/cosmology.rs:21-30) in the reference's float values.
tpm_cosmology lcdm_default() {
    tpm_cosmology c;
    c.omega_r = (double)9.14e-5f;
    c.omega_m = (double)0.315f;
    c.omega_l = (double)(1.0f - 0.315f - 9.14e-5f);
    return c;
}

template <typename T>
T clampv(T v, T lo, T hi) {
    return std::min(std::max(v, lo), hi);
}

struct DevBuf {
    void *p = nullptr;
    explicit DevBuf(size_t bytes) { TPM_HIP(hipMalloc(&p, bytes ? bytes : 16)); }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    template <typename T>
    T *as() const {
        return reinterpret_cast<T *>(p);
    }
};

struct CtxHandle {
    tpm_ctx *h = nullptr;
    ~CtxHandle() {
        if (h) tpm_destroy(h);
    }
};

}  // namespace

extern "C" {

// RunSettings::default (src/runner.rs:49-63).
void tpm_run_settings_default(tpm_run_settings *s) {
    if (!s) return;
    s->z_start = (double)500.0f;
    s->z_end = (double)-0.9f;
    s->steps = 512;
    s->particles_per_side = 40;
    s->seed = 1;
    s->treepm = 1;
    s->theta = (double)0.5f;
    s->softening_scale = (double)0.05f;
    s->pacing_ms = 0;
}

// RunSettings::sanitized (src/runner.rs:66-82), FP32 clamps as the reference's.
int tpm_run_settings_sanitized(const tpm_run_settings *in, tpm_run_settings *out) {
    if (!in || !out) return TPM_E_INVALID;
    if (!std::isfinite(in->z_start) || !std::isfinite(in->z_end) || !std::isfinite(in->theta) ||
        !std::isfinite(in->softening_scale))
        return TPM_E_INVALID;
    tpm_run_settings s = *in;
    const float z_start = clampv((float)in->z_start, 10.0f, 2000.0f);
    s.z_start = z_start;
    s.z_end = clampv((float)in->z_end, -0.95f, z_start - 1.0f);
    s.steps = clampv(in->steps, 16, 5000);
    s.particles_per_side = clampv(in->particles_per_side, 5, 200);
    s.theta = clampv((float)in->theta, 0.1f, 1.0f);
    s.softening_scale = clampv((float)in->softening_scale, 0.0f, 0.5f);
    s.pacing_ms = std::min(in->pacing_ms, 2000u);
    *out = s;
    return TPM_OK;
}

int tpm_headless_run(const tpm_run_settings *settings, tpm_frame_fn on_frame, void *user, double *final_pos,
                     double *final_mom, size_t capacity) {
    if (!settings) return TPM_E_INVALID;
    tpm_run_settings s;
    int rc = tpm_run_settings_sanitized(settings, &s);
    if (rc != TPM_OK) return rc;
    const tpm_cosmology cosmo = lcdm_default();
    const float box = 100.0f;  // src/runner.rs:22
    const int n = s.particles_per_side;
    const int mesh_n = 2 * n;
    const size_t np = (size_t)n * n * n;
    if ((final_pos || final_mom) && capacity < np) return TPM_E_INVALID;
    const double a_start = 1.0 / (1.0 + (double)(float)s.z_start);
    const double a_end = 1.0 / (1.0 + (double)(float)s.z_end);
    try {
        DevBuf pos(3 * np * sizeof(double)), mom(3 * np * sizeof(double)), acc(3 * np * sizeof(float));
        DevBuf delta((size_t)mesh_n * mesh_n * mesh_n * sizeof(float));
        DevBuf pos32(3 * np * sizeof(float)), speed(np * sizeof(float)), dens(np * sizeof(float)),
            rsd(3 * np * sizeof(float));
        DevBuf maxv(sizeof(unsigned int));
        {
            // ICs on an n^3 mesh (the generator's lattice), src/runner.rs:193-194.
            tpm_params p{};
            p.mesh_n = n;
            p.box_size = box;
            p.omega_m = cosmo.omega_m;
            CtxHandle ic;
            rc = tpm_create(&p, np, &ic.h);
            if (rc != TPM_OK) return rc;
            rc = tpm_zeldovich_bbks_dev(ic.h, &cosmo, (double)67.4f, (double)0.965f, (double)0.811f, s.seed, a_start,
                                        pos.as<double>(), mom.as<double>(), nullptr);
            if (rc != TPM_OK) return rc;
            TPM_HIP(hipDeviceSynchronize());
        }
        tpm_params p{};
        p.mesh_n = mesh_n;
        p.box_size = box;
        p.omega_m = cosmo.omega_m;
        if (s.treepm) {
            // ForceSplit::with(BOX, mesh_n, n, softening_scale, theta) (src/tree.rs:40-50), in FP32.
            const float cell = box / (float)mesh_n;
            const float r_split = 1.25f * cell;
            p.flags = TPM_FLAG_TREEPM;
            p.r_split = r_split;
            p.r_cut = 4.5f * r_split;
            p.softening = (float)s.softening_scale * (box / (float)n);
            p.theta = (float)s.theta;
        }
        CtxHandle sim;
        rc = tpm_create(&p, np, &sim.h);
        if (rc != TPM_OK) return rc;
        tpm::Ctx &c = tpm_ctx_internal(sim.h);
        // The run's parameters for binary snapshots of its frames.
        tpm_snapshot_header snap{};
        snap.n = np;
        snap.box_size = box;
        snap.omega_m = cosmo.omega_m;
        snap.omega_l = cosmo.omega_l;
        snap.omega_r = cosmo.omega_r;
        snap.h0 = (double)67.4f;
        snap.softening = p.softening;
        snap.r_split = p.r_split;
        snap.r_cut = p.r_cut;
        snap.theta = p.theta;
        snap.mesh_n = mesh_n;
        snap.seed = s.seed;
        const auto t0 = std::chrono::steady_clock::now();
        // Simulation::new: the force at a_start (src/integrator.rs:54-79).
        tpm::evaluate(c, pos.as<double>(), np, a_start, false, acc.as<float>(), delta.as<float>());
        double a = a_start;
        for (int i = 1; i <= s.steps; ++i) {
            const double a1 = a_start * std::pow(a_end / a_start, (double)i / (double)s.steps);
            rc = tpm_step_to_dev(sim.h, &cosmo, pos.as<double>(), mom.as<double>(), np, a, a1, acc.as<float>(),
                                 delta.as<float>(), nullptr);
            if (rc != TPM_OK) return rc;
            a = a1;
            const double e = tpm::hubble_e(cosmo, a);
            tpm::snapshot(c, pos.as<double>(), mom.as<double>(), np, a, e, delta.as<float>(), pos32.as<float>(),
                          speed.as<float>(), dens.as<float>(), rsd.as<float>());
            tpm::zero_fill(c, maxv.p, sizeof(unsigned int));
            hipLaunchKernelGGL(tpm::k_max_nonneg, dim3(256), dim3(256), 0, c.stream, speed.as<float>(), np,
                               maxv.as<unsigned int>());
            TPM_HIP(hipGetLastError());
            unsigned int mbits = 0;
            TPM_HIP(hipMemcpyAsync(&mbits, maxv.p, sizeof(mbits), hipMemcpyDeviceToHost, c.stream));
            TPM_HIP(hipStreamSynchronize(c.stream));
            tpm::finish_checks(c);
            float max_speed;
            std::memcpy(&max_speed, &mbits, sizeof(float));
            if (on_frame) {
                tpm_frame f{};
                f.step = i;
                f.steps = s.steps;
                f.redshift = 1.0 / a - 1.0;
                f.max_speed = max_speed;
                f.elapsed_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                f.n = np;
                f.d_pos = pos32.as<float>();
                f.d_speed = speed.as<float>();
                f.d_density = dens.as<float>();
                f.d_rsd = rsd.as<float>();
                f.d_delta = delta.as<float>();
                f.mesh_n = mesh_n;
                f.a = a;
                f.d_state_pos = pos.as<double>();
                f.d_state_mom = mom.as<double>();
                snap.a = a;
                snap.step = i;
                f.snapshot = &snap;
                if (on_frame(user, &f) != 0) break;
            }
        }
        if (final_pos) TPM_HIP(hipMemcpy(final_pos, pos.p, 3 * np * sizeof(double), hipMemcpyDeviceToHost));
        if (final_mom) TPM_HIP(hipMemcpy(final_mom, mom.p, 3 * np * sizeof(double), hipMemcpyDeviceToHost));
    } catch (const tpm::Error &x) {
        return x.code;
    }
    return TPM_OK;
}

}  // extern "C"
