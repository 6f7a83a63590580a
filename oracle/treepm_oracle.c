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
