/*
 * Type-generic body of the CPU oracle. Included twice by treepm_oracle.c, once with
 * REAL = float (the reference's precision, src/real.rs:2) and once with REAL = double
 * (the FP64 twin used to judge force error). Test infrastructure only: nothing in the
 * product library links this code.
 *
 * Every function restates the reference's arithmetic in the reference's order, so the FP32
 * instantiation reproduces the reference's rounding wherever the reference's own operations
 * are deterministic (compile with -ffp-contract=off so no FMA is formed that Rust would not
 * form). Two parts cannot be bit-identical because the reference delegates them to
 * third-party code: the 1D FFT (rustfft 6.4.1, src/fft.rs:19-23) and erfcf (libm 0.2.16,
 * src/tree.rs:56). Here the FFT is a mixed-radix Stockham transform with twiddles rounded
 * from double, and erfc is the C library's.
 */

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define F(name) CAT(name, SFX)

typedef struct { REAL re, im; } F(cplx_);

/* mesh.rs:81-83 axis_weight */
static inline REAL F(axis_weight_)(REAL f, int d) { return d == 0 ? (REAL)1.0 - f : f; }

/* i64 rem_euclid as used in mesh.rs:51-53 */
static inline int64_t F(rem_euclid_i64_)(int64_t a, int64_t n) {
    int64_t r = a % n;
    return r < 0 ? r + n : r;
}

/* mesh.rs:29-59 stencil: the 8 (cell index, weight) pairs of a position, z fastest. */
static void F(stencil_)(int n, REAL box_size, const REAL x[3], int64_t idx[8], REAL w[8]) {
    REAL h = box_size / (REAL)n;
    int64_t base[3];
    REAL frac[3];
    for (int k = 0; k < 3; ++k) {
        REAL s = x[k] / h - (REAL)0.5;
        REAL i0 = floor(s);
        base[k] = (int64_t)i0;
        frac[k] = s - i0;
    }
    int c = 0;
    for (int dx = 0; dx < 2; ++dx)
        for (int dy = 0; dy < 2; ++dy)
            for (int dz = 0; dz < 2; ++dz) {
                REAL ww = F(axis_weight_)(frac[0], dx) * F(axis_weight_)(frac[1], dy);
                ww = ww * F(axis_weight_)(frac[2], dz);
                int64_t ix = F(rem_euclid_i64_)(base[0] + dx, n);
                int64_t iy = F(rem_euclid_i64_)(base[1] + dy, n);
                int64_t iz = F(rem_euclid_i64_)(base[2] + dz, n);
                idx[c] = (ix * n + iy) * n + iz;
                w[c] = ww;
                ++c;
            }
}

/* mesh.rs:62-69 deposit: clears the grid, then adds mass*w per corner, particle order. */
void F(ora_deposit)(const REAL *pos, int64_t np, int n, REAL box_size, REAL mass, REAL *cells) {
    int64_t m = (int64_t)n * n * n;
    for (int64_t i = 0; i < m; ++i) cells[i] = 0;
    int64_t idx[8];
    REAL w[8];
    for (int64_t p = 0; p < np; ++p) {
        F(stencil_)(n, box_size, pos + 3 * p, idx, w);
        for (int c = 0; c < 8; ++c) cells[idx[c]] += mass * w[c];
    }
}

/* integrator.rs:117-126 density_contrast: unit-mass deposit, then delta = c/mean - 1. */
void F(ora_density_contrast)(const REAL *pos, int64_t np, int n, REAL box_size, REAL *delta) {
    F(ora_deposit)(pos, np, n, box_size, (REAL)1.0, delta);
    int64_t m = (int64_t)n * n * n;
    REAL cells = (REAL)m;
    REAL mean = (REAL)np / cells;
    for (int64_t i = 0; i < m; ++i) delta[i] = delta[i] / mean - (REAL)1.0;
}

/* mesh.rs:72-77 interpolate, applied per particle (integrator.rs:144-148 runs it in parallel). */
void F(ora_interpolate)(const REAL *cells, int n, REAL box_size, const REAL *pos, int64_t np, REAL *out) {
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < np; ++p) {
        int64_t idx[8];
        REAL w[8];
        F(stencil_)(n, box_size, pos + 3 * p, idx, w);
        REAL acc = (REAL)0.0;
        for (int c = 0; c < 8; ++c) acc = acc + cells[idx[c]] * w[c];
        out[p] = acc;
    }
}

/* fft.rs:86-95 wavenumber: the Nyquist mode of an even grid maps to 0. */
REAL F(ora_wavenumber)(int64_t i, int64_t n, REAL box_size) {
    REAL two_pi_over_l = (REAL)2.0 * (REAL)3.14159265358979323846 / box_size;
    if (n % 2 == 0 && i == n / 2) return (REAL)0.0;
    if (i <= n / 2) return (REAL)i * two_pi_over_l;
    return ((REAL)i - (REAL)n) * two_pi_over_l;
}

/*
 * One unnormalised 1D transform of a contiguous line: X_k = sum_x f_x e^{sign*2*pi*i*k*x/n}
 * (fft.rs:27-35: forward sign -1, inverse +1). Mixed-radix Stockham autosort: each stage of
 * radix R maps in[j + r*n/R] through the twiddle e^{sign*2*pi*i*(j mod Ns)*r/(Ns*R)} and an
 * R-point DFT to out[(j/Ns)*Ns*R + (j mod Ns) + q*Ns]. The FFT tests check this routine against
 * a direct O(n^2) DFT in double and against numpy, so a slip in the formulation shows there.
 * Twiddles come from one table w[t] = e^{sign*2*pi*i*t/n}, computed in double and rounded.
 */
typedef struct {
    int n, sign, nstages;
    int radix[40];
    F(cplx_) *w;
} F(fftplan_);

static int F(factor_)(int n, int *radix) {
    int s = 0;
    while (n % 4 == 0) { radix[s++] = 4; n /= 4; }
    while (n % 2 == 0) { radix[s++] = 2; n /= 2; }
    for (int p = 3; n > 1; p += 2)
        while (n % p == 0) { radix[s++] = p; n /= p; }
    return s;
}

static void F(plan_init_)(F(fftplan_) *p, int n, int sign) {
    p->n = n;
    p->sign = sign;
    p->nstages = F(factor_)(n, p->radix);
    p->w = (F(cplx_) *)malloc(sizeof(F(cplx_)) * n);
    for (int t = 0; t < n; ++t) {
        double ang = sign * 2.0 * 3.14159265358979323846 * (double)t / (double)n;
        p->w[t].re = (REAL)cos(ang);
        p->w[t].im = (REAL)sin(ang);
    }
}

static void F(plan_free_)(F(fftplan_) *p) { free(p->w); }

static inline F(cplx_) F(cmul_)(F(cplx_) a, F(cplx_) b) {
    F(cplx_) t = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return t;
}

static void F(fft_line_)(const F(fftplan_) *p, F(cplx_) *data, F(cplx_) *work, F(cplx_) *v, F(cplx_) *o) {
    const int n = p->n;
    F(cplx_) *src = data, *dst = work;
    int ns = 1;
    const REAL s = (REAL)p->sign;
    for (int st = 0; st < p->nstages; ++st) {
        const int R = p->radix[st];
        const int stride = n / R;
        const int tw_step = n / (ns * R);
        for (int j = 0; j < stride; ++j) {
            const int k = j % ns;
            for (int r = 0; r < R; ++r) {
                F(cplx_) a = src[j + r * stride];
                v[r] = r == 0 ? a : F(cmul_)(a, p->w[(int)(((int64_t)k * r * tw_step) % n)]);
            }
            if (R == 2) {
                o[0].re = v[0].re + v[1].re; o[0].im = v[0].im + v[1].im;
                o[1].re = v[0].re - v[1].re; o[1].im = v[0].im - v[1].im;
            } else if (R == 4) {
                F(cplx_) t0 = {v[0].re + v[2].re, v[0].im + v[2].im};
                F(cplx_) t1 = {v[0].re - v[2].re, v[0].im - v[2].im};
                F(cplx_) t2 = {v[1].re + v[3].re, v[1].im + v[3].im};
                F(cplx_) t3 = {v[1].re - v[3].re, v[1].im - v[3].im};
                /* s*i*t3 */
                F(cplx_) it3 = {-s * t3.im, s * t3.re};
                o[0].re = t0.re + t2.re; o[0].im = t0.im + t2.im;
                o[2].re = t0.re - t2.re; o[2].im = t0.im - t2.im;
                o[1].re = t1.re + it3.re; o[1].im = t1.im + it3.im;
                o[3].re = t1.re - it3.re; o[3].im = t1.im - it3.im;
            } else {
                const int wr = n / R;
                for (int q = 0; q < R; ++q) {
                    REAL sre = 0, sim = 0;
                    for (int r = 0; r < R; ++r) {
                        F(cplx_) t = F(cmul_)(v[r], p->w[(int)(((int64_t)r * q) % R) * wr]);
                        sre += t.re;
                        sim += t.im;
                    }
                    o[q].re = sre;
                    o[q].im = sim;
                }
            }
            const int base = (j / ns) * ns * R + k;
            for (int q = 0; q < R; ++q) dst[base + q * ns] = o[q];
        }
        ns *= R;
        F(cplx_) *t = src; src = dst; dst = t;
    }
    if (src != data) memcpy(data, src, sizeof(F(cplx_)) * n);
}

void F(ora_fft1d)(REAL *line_interleaved, int n, int sign) {
    F(fftplan_) p;
    F(plan_init_)(&p, n, sign);
    F(cplx_) *work = (F(cplx_) *)malloc(sizeof(F(cplx_)) * 3 * n);
    F(fft_line_)(&p, (F(cplx_) *)line_interleaved, work, work + n, work + 2 * n);
    free(work);
    F(plan_free_)(&p);
}

/* fft.rs:76-82 flat_index */
static inline int64_t F(flat_index_)(int64_t n, int axis, int64_t a, int64_t b, int64_t t) {
    if (axis == 0) return (t * n + a) * n + b;
    if (axis == 1) return (a * n + t) * n + b;
    return (a * n + b) * n + t;
}

/* fft.rs:40-72 transform: z lines in place, then y, then x lines gathered, transformed, scattered. */
void F(ora_fft3)(REAL *data_interleaved, int n, int sign) {
    F(cplx_) *data = (F(cplx_) *)data_interleaved;
    F(fftplan_) p;
    F(plan_init_)(&p, n, sign);
    int64_t nn = (int64_t)n * n;
#pragma omp parallel
    {
        F(cplx_) *work = (F(cplx_) *)malloc(sizeof(F(cplx_)) * 4 * n);
        F(cplx_) *line = work + n, *v = work + 2 * n, *o = work + 3 * n;
#pragma omp for schedule(static)
        for (int64_t ab = 0; ab < nn; ++ab) F(fft_line_)(&p, data + ab * n, work, v, o);
        for (int axis = 1; axis >= 0; --axis) {
#pragma omp for schedule(static)
            for (int64_t ab = 0; ab < nn; ++ab) {
                int64_t a = ab / n, b = ab % n;
                for (int64_t t = 0; t < n; ++t) line[t] = data[F(flat_index_)(n, axis, a, b, t)];
                F(fft_line_)(&p, line, work, v, o);
                for (int64_t t = 0; t < n; ++t) data[F(flat_index_)(n, axis, a, b, t)] = line[t];
            }
        }
        free(work);
    }
    F(plan_free_)(&p);
}

/*
 * poisson.rs:33-88 PoissonSolver::acceleration: forward FFT of delta, then for each mode
 * phi_k = delta_k * (-(1.5*omega_m/a)/k^2 * exp(-k^2 r_s^2)) and g_k[axis] = phi_k * (0 - i k_axis),
 * skipping k^2 == 0; three inverse FFTs scaled by 1/n^3.
 */
void F(ora_poisson_accel)(const REAL *delta, int n, REAL box_size, REAL omega_m, REAL a, REAL r_split,
                          REAL *gx, REAL *gy, REAL *gz) {
    int64_t m = (int64_t)n * n * n;
    F(cplx_) *dk = (F(cplx_) *)malloc(sizeof(F(cplx_)) * m);
    for (int64_t i = 0; i < m; ++i) { dk[i].re = delta[i]; dk[i].im = (REAL)0.0; }
    F(ora_fft3)((REAL *)dk, n, -1);
    REAL prefactor = (REAL)1.5 * omega_m / a;
    F(cplx_) *ak[3];
    for (int c = 0; c < 3; ++c) ak[c] = (F(cplx_) *)calloc(m, sizeof(F(cplx_)));
    for (int64_t ix = 0; ix < n; ++ix)
        for (int64_t iy = 0; iy < n; ++iy)
            for (int64_t iz = 0; iz < n; ++iz) {
                int64_t idx = (ix * n + iy) * n + iz;
                REAL k[3] = {F(ora_wavenumber)(ix, n, box_size), F(ora_wavenumber)(iy, n, box_size),
                             F(ora_wavenumber)(iz, n, box_size)};
                REAL k2 = k[0] * k[0];
                k2 = k2 + k[1] * k[1];
                k2 = k2 + k[2] * k[2];
                if (k2 == (REAL)0.0) continue;
                REAL filter = exp(-k2 * r_split * r_split);
                REAL sc = -prefactor / k2 * filter;
                F(cplx_) phi = {dk[idx].re * sc, dk[idx].im * sc};
                for (int c = 0; c < 3; ++c) {
                    REAL kr = (REAL)0.0, ki = -k[c];
                    ak[c][idx].re = phi.re * kr - phi.im * ki;
                    ak[c][idx].im = phi.re * ki + phi.im * kr;
                }
            }
    REAL scale = (REAL)1.0 / (REAL)m;
    REAL *outs[3] = {gx, gy, gz};
    for (int c = 0; c < 3; ++c) {
        F(ora_fft3)((REAL *)ak[c], n, +1);
        for (int64_t i = 0; i < m; ++i) outs[c][i] = ak[c][i].re * scale;
        free(ak[c]);
    }
    free(dk);
}

/* ---------------------------------------------------------------- short-range tree (tree.rs) */

#define NO_CHILD_ 0xFFFFFFFFu

typedef struct {
    REAL center[3];
    REAL half;
    REAL com[3];
    uint32_t count;
    uint32_t children[8];
    uint32_t start, end;
} F(node_);

typedef struct {
    const REAL *pos;
    REAL box_size;
    uint32_t *order;
    F(node_) *nodes;
    int64_t nnodes, cap;
    uint32_t *sorted;
} F(tree_);

/* tree.rs:54-57 short_range_factor: S(r) = erfc(u) + (2/sqrt(pi)) u e^{-u^2}, u = r/(2 r_s). */
REAL F(ora_short_range_factor)(REAL r, REAL r_split) {
    const REAL two_over_sqrt_pi = (REAL)1.1283791670955126;
    REAL u = r / ((REAL)2.0 * r_split);
    REAL e = exp(-u * u);
    return erfc(u) + two_over_sqrt_pi * u * e;
}

static inline REAL F(norm2_)(const REAL v[3]) { return v[0] * v[0] + v[1] * v[1] + v[2] * v[2]; }

/* tree.rs:61-66 pair */
static inline void F(pair_add_)(const REAL d[3], REAL mass, REAL k, REAL r_split, REAL soft, REAL acc[3]) {
    REAL r2 = F(norm2_)(d);
    REAL s = r2 + soft * soft;
    REAL f = -k * mass * F(ora_short_range_factor)(sqrt(r2), r_split) / (s * sqrt(s));
    for (int c = 0; c < 3; ++c) acc[c] += d[c] * f;
}

/* tree.rs:253-261 min_image: a - b wrapped with round-half-away-from-zero. */
static inline void F(min_image_)(REAL l, const REAL a[3], const REAL b[3], REAL d[3]) {
    for (int k = 0; k < 3; ++k) {
        REAL x = a[k] - b[k];
        d[k] = x - l * round(x / l);
    }
}

/* tree.rs:186-250 build_node (recursive, counting sort into octants, leaf if <= 8 or depth >= 40). */
static uint32_t F(build_node_)(F(tree_) *t, int64_t start, int64_t end, const REAL center[3], REAL half, uint32_t depth) {
    if (t->nnodes == t->cap) {
        t->cap = t->cap * 2 + 16;
        t->nodes = (F(node_) *)realloc(t->nodes, sizeof(F(node_)) * t->cap);
    }
    uint32_t idx = (uint32_t)t->nnodes++;
    REAL com[3] = {0, 0, 0};
    for (int64_t q = start; q < end; ++q) {
        const REAL *p = t->pos + 3 * (int64_t)t->order[q];
        for (int k = 0; k < 3; ++k) com[k] += p[k];
    }
    REAL count = (REAL)(end - start);
    for (int k = 0; k < 3; ++k) com[k] /= count;
    F(node_) *nd = &t->nodes[idx];
    for (int k = 0; k < 3; ++k) { nd->center[k] = center[k]; nd->com[k] = com[k]; }
    nd->half = half;
    nd->count = (uint32_t)(end - start);
    for (int o = 0; o < 8; ++o) nd->children[o] = NO_CHILD_;
    nd->start = (uint32_t)start;
    nd->end = (uint32_t)end;
    if (end - start <= 8 || depth >= 40) return idx;

    int64_t counts[8] = {0}, offsets[8], next[8];
#define OCT_(p) ((int)((p)[0] >= center[0]) | ((int)((p)[1] >= center[1]) << 1) | ((int)((p)[2] >= center[2]) << 2))
    for (int64_t q = start; q < end; ++q) counts[OCT_(t->pos + 3 * (int64_t)t->order[q])]++;
    int64_t running = start;
    for (int o = 0; o < 8; ++o) { offsets[o] = running; running += counts[o]; }
    for (int o = 0; o < 8; ++o) next[o] = offsets[o];
    for (int64_t q = start; q < end; ++q) {
        uint32_t j = t->order[q];
        int o = OCT_(t->pos + 3 * (int64_t)j);
        t->sorted[next[o] - start] = j;
        next[o]++;
    }
#undef OCT_
    memcpy(t->order + start, t->sorted, sizeof(uint32_t) * (end - start));
    REAL quarter = half / (REAL)2.0;
    for (int o = 0; o < 8; ++o) {
        if (counts[o] == 0) continue;
        REAL cc[3];
        for (int k = 0; k < 3; ++k) cc[k] = center[k] + (((o >> k) & 1) ? quarter : -quarter);
        uint32_t child = F(build_node_)(t, offsets[o], offsets[o] + counts[o], cc, quarter, depth + 1);
        t->nodes[idx].children[o] = child;
    }
    return idx;
}

/* tree.rs:171-183 Tree::build */
static void F(tree_build_)(F(tree_) *t, const REAL *pos, int64_t np, REAL box_size) {
    t->pos = pos;
    t->box_size = box_size;
    t->order = (uint32_t *)malloc(sizeof(uint32_t) * (np > 0 ? np : 1));
    t->sorted = (uint32_t *)malloc(sizeof(uint32_t) * (np > 0 ? np : 1));
    for (int64_t i = 0; i < np; ++i) t->order[i] = (uint32_t)i;
    t->cap = 2 * np / 8 + 1;
    t->nodes = (F(node_) *)malloc(sizeof(F(node_)) * t->cap);
    t->nnodes = 0;
    REAL half = box_size / (REAL)2.0;
    REAL c[3] = {half, half, half};
    F(build_node_)(t, 0, np, c, half, 0);
}

static void F(tree_free_)(F(tree_) *t) {
    free(t->order);
    free(t->sorted);
    free(t->nodes);
}

/*
 * tree.rs:264-310 accel_on: LIFO walk from the root with children pushed 0..7. A cube is
 * skipped when |c - x|^2 > (r_cut + sqrt(3) half)^2; a node whose children[0] is absent is
 * walked as a leaf (tree.rs:280); an internal node is accepted when 2 half < theta |com - x|.
 * Returns the number of interactions evaluated in *ninter when non-null.
 */
static void F(accel_on_)(const F(tree_) *t, int64_t i, REAL r_split, REAL r_cut, REAL soft, REAL theta, REAL k,
                         uint32_t *stack, REAL acc[3], int64_t *ninter) {
    const REAL sqrt_3 = (REAL)1.7320508075688772;
    const REAL *xi = t->pos + 3 * i;
    REAL cut2 = r_cut * r_cut;
    acc[0] = acc[1] = acc[2] = 0;
    int64_t sp = 0, cnt = 0;
    stack[sp++] = 0;
    while (sp > 0) {
        uint32_t n = stack[--sp];
        const F(node_) *nd = &t->nodes[n];
        REAL to_centre[3];
        F(min_image_)(t->box_size, nd->center, xi, to_centre);
        REAL max_r = r_cut + sqrt_3 * nd->half;
        if (F(norm2_)(to_centre) > max_r * max_r) continue;
        if (nd->children[0] == NO_CHILD_) {
            for (uint32_t q = nd->start; q < nd->end; ++q) {
                int64_t j = t->order[q];
                if (j == i) continue;
                REAL d[3];
                F(min_image_)(t->box_size, xi, t->pos + 3 * j, d);
                if (F(norm2_)(d) > cut2) continue;
                F(pair_add_)(d, (REAL)1.0, k, r_split, soft, acc);
                ++cnt;
            }
        } else {
            REAL to_com[3];
            F(min_image_)(t->box_size, nd->com, xi, to_com);
            REAL r = sqrt(F(norm2_)(to_com));
            if ((REAL)2.0 * nd->half < theta * r) {
                REAL d[3] = {-to_com[0], -to_com[1], -to_com[2]};
                F(pair_add_)(d, (REAL)nd->count, k, r_split, soft, acc);
                ++cnt;
            } else {
                for (int o = 0; o < 8; ++o)
                    if (nd->children[o] != NO_CHILD_) stack[sp++] = nd->children[o];
            }
        }
    }
    if (ninter) *ninter = cnt;
}

/* tree.rs:314-323 accelerations: build, then walk every particle (rayon-parallel there, OpenMP here). */
int64_t F(ora_tree_accel)(const REAL *pos, int64_t np, REAL box_size, REAL r_split, REAL r_cut, REAL soft,
                          REAL theta, REAL k, REAL *acc) {
    F(tree_) t;
    F(tree_build_)(&t, pos, np, box_size);
    int64_t total = 0;
#pragma omp parallel reduction(+ : total)
    {
        uint32_t *stack = (uint32_t *)malloc(sizeof(uint32_t) * (8 * 64 + 64));
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < np; ++i) {
            int64_t c = 0;
            F(accel_on_)(&t, i, r_split, r_cut, soft, theta, k, stack, acc + 3 * i, &c);
            total += c;
        }
        free(stack);
    }
    F(tree_free_)(&t);
    return total;
}

/* Walks only the listed targets (used to judge a subsample); same tree and walk as above. */
void F(ora_tree_accel_subset)(const REAL *pos, int64_t np, const uint32_t *targets, int64_t nt, REAL box_size,
                              REAL r_split, REAL r_cut, REAL soft, REAL theta, REAL k, REAL *acc) {
    F(tree_) t;
    F(tree_build_)(&t, pos, np, box_size);
#pragma omp parallel
    {
        uint32_t *stack = (uint32_t *)malloc(sizeof(uint32_t) * (8 * 64 + 64));
#pragma omp for schedule(dynamic, 64)
        for (int64_t q = 0; q < nt; ++q) F(accel_on_)(&t, targets[q], r_split, r_cut, soft, theta, k, stack, acc + 3 * q, NULL);
        free(stack);
    }
    F(tree_free_)(&t);
}

/* tree.rs:351-370 brute_force: O(N^2) pair sum with the same kernel, cutoff and min image. */
void F(ora_brute_force)(const REAL *pos, int64_t np, REAL box_size, REAL r_split, REAL r_cut, REAL soft, REAL k,
                        REAL *acc) {
    REAL cut2 = r_cut * r_cut;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < np; ++i) {
        REAL a[3] = {0, 0, 0};
        for (int64_t j = 0; j < np; ++j) {
            if (j == i) continue;
            REAL d[3];
            F(min_image_)(box_size, pos + 3 * i, pos + 3 * j, d);
            if (F(norm2_)(d) > cut2) continue;
            F(pair_add_)(d, (REAL)1.0, k, r_split, soft, a);
        }
        acc[3 * i] = a[0];
        acc[3 * i + 1] = a[1];
        acc[3 * i + 2] = a[2];
    }
}

/* integrator.rs:150-152: K = 1.5 Omega_m L^3 / (4 pi a N). */
REAL F(ora_coupling)(REAL omega_m, REAL box_size, REAL a, int64_t np) {
    REAL l3 = box_size * box_size * box_size;
    REAL pi = (REAL)3.14159265358979323846;
    return (REAL)1.5 * omega_m * l3 / ((REAL)4.0 * pi * a * (REAL)np);
}

/*
 * integrator.rs:140-161 Simulation::accelerations: delta -> PM mesh force -> CIC interpolation
 * per particle, plus the tree's short-range force when treepm != 0. Writes acc (3N) and, if
 * non-null, delta (n^3). Returns the interaction count of the walk.
 */
int64_t F(ora_accelerations)(const REAL *pos, int64_t np, int n, REAL box_size, REAL omega_m, REAL a, int treepm,
                             REAL r_split, REAL r_cut, REAL soft, REAL theta, REAL *acc, REAL *delta_out) {
    int64_t m = (int64_t)n * n * n;
    REAL *delta = delta_out ? delta_out : (REAL *)malloc(sizeof(REAL) * m);
    F(ora_density_contrast)(pos, np, n, box_size, delta);
    REAL *g = (REAL *)malloc(sizeof(REAL) * 3 * m);
    F(ora_poisson_accel)(delta, n, box_size, omega_m, a, treepm ? r_split : (REAL)0.0, g, g + m, g + 2 * m);
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < np; ++p) {
        int64_t idx[8];
        REAL w[8];
        F(stencil_)(n, box_size, pos + 3 * p, idx, w);
        for (int c = 0; c < 3; ++c) {
            REAL s = (REAL)0.0;
            for (int q = 0; q < 8; ++q) s = s + g[c * m + idx[q]] * w[q];
            acc[3 * p + c] = s;
        }
    }
    int64_t inter = 0;
    if (treepm) {
        REAL k = F(ora_coupling)(omega_m, box_size, a, np);
        REAL *sr = (REAL *)malloc(sizeof(REAL) * 3 * (np > 0 ? np : 1));
        inter = F(ora_tree_accel)(pos, np, box_size, r_split, r_cut, soft, theta, k, sr);
        for (int64_t i = 0; i < 3 * np; ++i) acc[i] += sr[i];
        free(sr);
    }
    free(g);
    if (!delta_out) free(delta);
    return inter;
}

/* ------------------------------------------------- cosmology and KDK (cosmology.rs, integrator.rs) */

/* cosmology.rs:32-34 E(a) = sqrt(Or a^-4 + Om a^-3 + OL) */
REAL F(ora_cosmo_e)(REAL omr, REAL omm, REAL oml, REAL a) {
    return sqrt(omr / (a * a * a * a) + omm / (a * a * a) + oml);
}

/* cosmology.rs:47-80 growth_state: RK4 in ln a, 4000 steps from a = 1e-3 (Meszaros start). */
static void F(growth_state_)(REAL omr, REAL omm, REAL oml, REAL a_end, REAL *d_out, REAL *dl_out) {
    const REAL a_start = (REAL)1e-3;
    REAL d, dl;
    if (omr > (REAL)0.0) {
        REAL y = a_start * omm / omr;
        d = (REAL)1.0 + (REAL)1.5 * y;
        dl = (REAL)1.5 * y;
    } else {
        d = a_start;
        dl = a_start;
    }
#define RHS_(la_, d_, dl_, od_, odl_)                                                                   \
    do {                                                                                                \
        REAL a_ = exp(la_);                                                                             \
        REAL e2_ = omr / (a_ * a_ * a_ * a_) + omm / (a_ * a_ * a_) + oml;                              \
        REAL dln_e_ = (REAL)0.5 * ((REAL)-4.0 * omr / (a_ * a_ * a_ * a_) - (REAL)3.0 * omm / (a_ * a_ * a_)) / e2_; \
        REAL src_ = (REAL)1.5 * omm / (a_ * a_ * a_ * e2_) * (d_);                                      \
        od_ = (dl_);                                                                                    \
        odl_ = -((REAL)2.0 + dln_e_) * (dl_) + src_;                                                    \
    } while (0)
    const int steps = 4000;
    REAL la0 = log(a_start);
    REAL h = (log(a_end) - la0) / (REAL)steps;
    for (int step = 0; step < steps; ++step) {
        REAL la = la0 + (REAL)step * h;
        REAL k1d, k1l, k2d, k2l, k3d, k3l, k4d, k4l;
        RHS_(la, d, dl, k1d, k1l);
        RHS_(la + (REAL)0.5 * h, d + (REAL)0.5 * h * k1d, dl + (REAL)0.5 * h * k1l, k2d, k2l);
        RHS_(la + (REAL)0.5 * h, d + (REAL)0.5 * h * k2d, dl + (REAL)0.5 * h * k2l, k3d, k3l);
        RHS_(la + h, d + h * k3d, dl + h * k3l, k4d, k4l);
        d += h / (REAL)6.0 * (k1d + (REAL)2.0 * k2d + (REAL)2.0 * k3d + k4d);
        dl += h / (REAL)6.0 * (k1l + (REAL)2.0 * k2l + (REAL)2.0 * k3l + k4l);
    }
#undef RHS_
    *d_out = d;
    *dl_out = dl;
}

/* cosmology.rs:82-91 growth (D(1) = 1) and growth_rate f = dlnD/dlna. */
REAL F(ora_growth)(REAL omr, REAL omm, REAL oml, REAL a) {
    REAL d, dl, d1, dl1;
    F(growth_state_)(omr, omm, oml, a, &d, &dl);
    F(growth_state_)(omr, omm, oml, (REAL)1.0, &d1, &dl1);
    return d / d1;
}

REAL F(ora_growth_rate)(REAL omr, REAL omm, REAL oml, REAL a) {
    REAL d, dl;
    F(growth_state_)(omr, omm, oml, a, &d, &dl);
    return dl / d;
}

/* integrator.rs:287-297 Simpson's rule with 64 intervals; kind 0: 1/(a E) (kick), 1: 1/(a^3 E) (drift). */
static REAL F(integrate_)(REAL omr, REAL omm, REAL oml, REAL lo, REAL hi, int kind) {
#define FN_(a_) (kind == 0 ? (REAL)1.0 / ((a_) * F(ora_cosmo_e)(omr, omm, oml, (a_))) \
                           : (REAL)1.0 / ((a_) * (a_) * (a_) * F(ora_cosmo_e)(omr, omm, oml, (a_))))
    const int n = 64;
    REAL h = (hi - lo) / (REAL)n;
    REAL sum = FN_(lo) + FN_(hi);
    for (int i = 1; i < n; ++i) {
        REAL weight = (i % 2 == 1) ? (REAL)4.0 : (REAL)2.0;
        sum += weight * FN_(lo + (REAL)i * h);
    }
#undef FN_
    return sum * h / (REAL)3.0;
}

REAL F(ora_kick_factor)(REAL omr, REAL omm, REAL oml, REAL a0, REAL a1) { return F(integrate_)(omr, omm, oml, a0, a1, 0); }
REAL F(ora_drift_factor)(REAL omr, REAL omm, REAL oml, REAL a0, REAL a1) { return F(integrate_)(omr, omm, oml, a0, a1, 1); }

/* integrator.rs:273-285 midpoint: 60 bisections for half the kick factor. */
REAL F(ora_midpoint)(REAL omr, REAL omm, REAL oml, REAL a0, REAL a1) {
    REAL half = F(ora_kick_factor)(omr, omm, oml, a0, a1) / (REAL)2.0;
    REAL lo = a0, hi = a1;
    for (int it = 0; it < 60; ++it) {
        REAL mid = (REAL)0.5 * (lo + hi);
        if (F(ora_kick_factor)(omr, omm, oml, a0, mid) < half)
            lo = mid;
        else
            hi = mid;
    }
    return (REAL)0.5 * (lo + hi);
}

/* Rust's f32/f64 rem_euclid. */
static inline REAL F(rem_euclid_)(REAL x, REAL l) {
    REAL r = fmod(x, l);
    return r < (REAL)0.0 ? r + fabs(l) : r;
}

/*
 * integrator.rs:54-79 (Simulation::new computes the initial accelerations) and :165-186
 * (run / step_to): `steps` KDK steps from a_start to a_end evenly spaced in ln a. pos and mom are
 * updated in place. treepm / split as in ora_accelerations.
 */
void F(ora_sim_run)(REAL *pos, REAL *mom, int64_t np, int n, REAL box_size, REAL omr, REAL omm, REAL oml,
                    REAL a_start, REAL a_end, int steps, int treepm, REAL r_split, REAL r_cut, REAL soft,
                    REAL theta) {
    REAL *acc = (REAL *)malloc(sizeof(REAL) * 3 * (np > 0 ? np : 1));
    F(ora_accelerations)(pos, np, n, box_size, omm, a_start, treepm, r_split, r_cut, soft, theta, acc, NULL);
    REAL a = a_start;
    for (int i = 0; i < steps; ++i) {
        REAL a1 = a_start * pow(a_end / a_start, (REAL)(i + 1) / (REAL)steps);
        REAL am = F(ora_midpoint)(omr, omm, oml, a, a1);
        REAL k1 = F(ora_kick_factor)(omr, omm, oml, a, am);
        for (int64_t q = 0; q < 3 * np; ++q) mom[q] += acc[q] * k1;
        REAL d = F(ora_drift_factor)(omr, omm, oml, a, a1);
        for (int64_t q = 0; q < 3 * np; ++q) pos[q] = F(rem_euclid_)(pos[q] + mom[q] * d, box_size);
        F(ora_accelerations)(pos, np, n, box_size, omm, a1, treepm, r_split, r_cut, soft, theta, acc, NULL);
        REAL k2 = F(ora_kick_factor)(omr, omm, oml, am, a1);
        for (int64_t q = 0; q < 3 * np; ++q) mom[q] += acc[q] * k2;
        a = a1;
    }
    free(acc);
}

#undef F
#undef CAT
#undef CAT2
