/*
 * CPU oracle for the TreePM force evaluation: a restatement of the reference Rust crate's
 * algorithm (src/mesh.rs, src/fft.rs, src/poisson.rs, src/tree.rs, src/integrator.rs:117-161).
 * It is test infrastructure: only the tests, the smoke check and the CPU-baseline leg of the
 * benchmark load it, and only as the checker. The product library never links or calls it.
 *
 * Exported symbols carry a _f32 suffix (the reference's FP32 arithmetic) or a _f64 suffix
 * (the same algorithm in double, the FP64 twin). Build with -ffp-contract=off.
 */
#include <math.h>
#include <tgmath.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define REAL float
#define SFX _f32
#include "oracle_impl.h"
#undef REAL
#undef SFX

#define REAL double
#define SFX _f64
#include "oracle_impl.h"
#undef REAL
#undef SFX

/* Thread count the OpenMP regions use (the CPU-baseline leg reports it). */
#ifdef _OPENMP
#include <omp.h>
int ora_threads(void) { return omp_get_max_threads(); }
void ora_set_threads(int t) { omp_set_num_threads(t); }
#else
int ora_threads(void) { return 1; }
void ora_set_threads(int t) { (void)t; }
#endif

/* frame.rs:33-52 downsample_grid: 1 + delta averaged over f x f x f blocks, f = ceil(n / 128) (at
 * least 1), m = n / f cells a side (a partial last block is dropped), output x fastest
 * ((z m + y) m + x), the block sum in the reference's loop order, all in float. Returns m; `out`
 * may be NULL to query it. */
int ora_downsample_grid(const float *delta, int n, float *out) {
    int f = (n + 127) / 128;
    if (f < 1) f = 1;
    const int m = n / f;
    if (!out) return m;
    const float norm = 1.0f / (float)(f * f * f);
    for (int z = 0; z < m; ++z)
        for (int y = 0; y < m; ++y)
            for (int x = 0; x < m; ++x) {
                float sum = 0.0f;
                for (int dx = 0; dx < f; ++dx)
                    for (int dy = 0; dy < f; ++dy)
                        for (int dz = 0; dz < f; ++dz) {
                            const int64_t ix = x * f + dx, iy = y * f + dy, iz = z * f + dz;
                            sum += delta[(ix * n + iy) * n + iz];
                        }
                out[((int64_t)z * m + y) * m + x] = 1.0f + sum * norm;
            }
    return m;
}

/* runner.rs:262-284 snapshot for FP64 state: positions rounded to float; speed |p| / a; density
 * max(1 + delta(x), 0) with the CIC weights of the FP64 stencil rounded to float and summed in
 * float in stencil order (the solver's FP64 interpolation convention); redshift-space shift
 * p / (a^2 E) (frame.rs:56-64). Any output may be NULL. */
void ora_snapshot_f64(const double *pos, const double *mom, int64_t np, double a, double e, const float *delta, int n,
                      double box_size, float *opos, float *ospeed, float *odens, float *orsd) {
    const double inv_a2e = 1.0 / (a * a * e);
    for (int64_t i = 0; i < np; ++i) {
        const double *x = pos + 3 * i, *p = mom + 3 * i;
        if (opos)
            for (int k = 0; k < 3; ++k) opos[3 * i + k] = (float)x[k];
        if (ospeed) ospeed[i] = (float)(sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]) / a);
        if (odens) {
            int64_t idx[8];
            double w[8];
            stencil__f64(n, box_size, x, idx, w);
            float s = 0.0f;
            for (int c = 0; c < 8; ++c) s = s + delta[idx[c]] * (float)w[c];
            const float d = 1.0f + s;
            odens[i] = d > 0.0f ? d : 0.0f;
        }
        if (orsd)
            for (int k = 0; k < 3; ++k) orsd[3 * i + k] = (float)(p[k] * inv_a2e);
    }
}
