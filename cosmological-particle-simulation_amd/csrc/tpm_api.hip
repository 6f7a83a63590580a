// C ABI of libtpm (declared in include/tpm.h): context lifecycle, the full force evaluation
// that replaces Simulation::accelerations (src/integrator.rs:140-161), and the component
// entry points that mirror the reference's public functions.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>

#include "device_scan.h"

#include "tpm_internal.h"

#include <algorithm>

namespace tpm {

// Message of the last failure of a call made without a context (tpm_create, the allocation-free
// deposit); a fixed buffer, so recording it allocates nothing.
static char g_create_error[256] = "";

void set_global_error(const char *msg) {
    std::snprintf(g_create_error, sizeof g_create_error, "%s", msg ? msg : "");
}

__global__ void k_inverse_perm(const uint32_t *__restrict__ idx, size_t n, uint32_t *__restrict__ inv) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) inv[idx[i]] = (uint32_t)i;
}

__global__ void k_map_targets(const uint32_t *__restrict__ inv, const uint32_t *__restrict__ caller, size_t m,
                              size_t n, uint32_t *__restrict__ out, int *bad) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    const uint32_t c = caller[t];
    if (c >= n) {
        *bad = 1;
        out[t] = 0;
        return;
    }
    out[t] = inv[c];
}

__global__ void k_add3(double *__restrict__ a, const double *__restrict__ b, size_t count) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) a[i] += b[i];
}

static inline unsigned blocks(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

// One thread: ORs the failure flags of the evaluation just enqueued into the sticky status word.
__global__ void k_fold_status(const unsigned long long *__restrict__ walk_overflow, const uint32_t *__restrict__ tree_full,
                              const int *__restrict__ nonfinite, uint32_t *__restrict__ status) {
    uint32_t s = 0;
    if (walk_overflow && *walk_overflow) s |= TPM_STATUS_WALK_OVERFLOW;
    if (tree_full && *tree_full) s |= TPM_STATUS_TREE_CAPACITY;
    if (nonfinite && *nonfinite) s |= TPM_STATUS_NONFINITE;
    if (s) atomicOr(status, s);
}

void fold_status(Ctx &c, bool tree_built) {
    const uint32_t *tree_full = tree_built && c.tree_info ? c.tree_info + 1 : nullptr;
    const int *nonfinite = (c.p.flags & TPM_FLAG_CHECK_FINITE) ? c.flag : nullptr;
    hipLaunchKernelGGL(k_fold_status, dim3(1), dim3(1), 0, c.stream, c.counter + 1, tree_full, nonfinite, c.status);
    TPM_HIP(hipGetLastError());
    c.status_stream = c.stream;
}

static void validate(const tpm_params &p) {
    TPM_REQUIRE(p.mesh_n >= 2 && p.mesh_n <= 1024, "mesh_n must be in [2, 1024]");
    TPM_REQUIRE(std::isfinite(p.box_size) && p.box_size > 0, "box_size must be positive");
    TPM_REQUIRE(std::isfinite(p.omega_m), "omega_m must be finite");
    TPM_REQUIRE(std::isfinite(p.r_split) && p.r_split >= 0, "r_split must be >= 0");
    if (p.flags & TPM_FLAG_TREEPM) {
        TPM_REQUIRE(p.r_split > 0, "TreePM needs r_split > 0");
        TPM_REQUIRE(std::isfinite(p.r_cut) && p.r_cut > 0, "TreePM needs r_cut > 0");
        TPM_REQUIRE(std::isfinite(p.softening) && p.softening >= 0, "softening must be >= 0");
        TPM_REQUIRE(std::isfinite(p.theta) && p.theta >= 0, "theta must be >= 0");
    }
}

static void free_all(Ctx &c) {
    for (hipEvent_t &e : c.walk_ev)
        if (e) (void)hipEventDestroy(e);
    void *ptrs[] = {c.tw32, c.tw64, c.kvec, c.pos, c.acc, c.keys, c.keys_alt, c.idx, c.idx_alt, c.radix_table, c.scan_part, c.spos,
                    c.counter, c.flag, c.status, c.scan_part2, c.grid, c.specA, c.specB, c.specC, c.force, c.delta, c.dspecA, c.dspecB,
                    c.dspecC, c.dforce, c.tree.range, c.tree.link, c.tree.box, c.tree.com, c.tree_scan,
                    c.ngroups_dev, c.walk_ctr, c.walk_goff, c.dep_start, c.inv, c.judge_buf, c.gstart, c.gscratch, c.stab,
                    c.acc_sr, c.members, c.reject, c.mlp.params, c.mlp.grads, c.mlp.m1, c.mlp.m2};
    for (int l = 0; l < 4; ++l) {
        if (c.mlp.wt16[l]) (void)hipFree(c.mlp.wt16[l]);
        if (c.mlp.w16[l]) (void)hipFree(c.mlp.w16[l]);
    }
    for (void *p : c.mlp.work)
        if (p) (void)hipFree(p);
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    for (hipEvent_t e : {c.ev_fork, c.ev_join, c.tree_ev[0], c.tree_ev[1]})
        if (e) (void)hipEventDestroy(e);
    if (c.side) (void)hipStreamDestroy(c.side);
    if (c.stream) (void)hipStreamDestroy(c.stream);
}

static void create(Ctx &c, const tpm_params &p, size_t max_n) {
    validate(p);
    TPM_REQUIRE(max_n < (size_t)0xFFFFFFF0u, "max_particles must fit 32-bit indices");
    c.p = p;
    c.max_n = max_n;
    c.n = p.mesh_n;
    c.hp = c.n / 2 + 1;
    c.m = (size_t)c.n * c.n * c.n;
    c.hs = (size_t)c.n * c.n * c.hp;
    const size_t mn = max_n ? max_n : 1;
    TPM_HIP(hipStreamCreate(&c.stream));
    TPM_HIP(hipMalloc(&c.tw32, c.n * sizeof(float2)));
    TPM_HIP(hipMalloc(&c.tw64, c.n * sizeof(double2)));
    TPM_HIP(hipMalloc(&c.kvec, c.n * sizeof(double)));
    TPM_HIP(hipMalloc(&c.pos, 3 * mn * sizeof(double)));
    TPM_HIP(hipMalloc(&c.acc, 3 * mn * sizeof(float)));
    TPM_HIP(hipMalloc(&c.keys, mn * sizeof(uint64_t)));
    TPM_HIP(hipMalloc(&c.keys_alt, mn * sizeof(uint64_t)));
    TPM_HIP(hipMalloc(&c.idx, mn * sizeof(uint32_t)));
    TPM_HIP(hipMalloc(&c.idx_alt, mn * sizeof(uint32_t)));
    c.radix_table_words = radix_tiles(mn) * 256;
    TPM_HIP(hipMalloc(&c.radix_table, c.radix_table_words * sizeof(uint32_t)));
    TPM_HIP(hipMalloc(&c.scan_part, kLbSlotWords * sizeof(unsigned long long)));
    TPM_HIP(hipMalloc(&c.scan_part2, kLbSlotWords * sizeof(unsigned long long)));
    TPM_HIP(hipMemset(c.scan_part, 0, kLbSlotWords * sizeof(unsigned long long)));
    TPM_HIP(hipMemset(c.scan_part2, 0, kLbSlotWords * sizeof(unsigned long long)));
    TPM_HIP(hipMalloc(&c.spos, 3 * mn * sizeof(double)));
    TPM_HIP(hipMalloc(&c.gstart, (mn + 1) * sizeof(uint32_t)));
    TPM_HIP(hipMalloc(&c.gscratch, 2 * mn * sizeof(uint32_t)));
    TPM_HIP(hipMalloc(&c.ngroups_dev, 16));
    TPM_HIP(hipMalloc(&c.walk_ctr, 8 * 1024 * sizeof(uint32_t)));
    TPM_HIP(hipMalloc(&c.walk_goff, 64 * sizeof(uint32_t)));
    c.dep_level = 0;
    while (((c.n + (1 << c.dep_level) - 1) >> c.dep_level) > kDepSide) ++c.dep_level;
    {
        const size_t nodes1 = (size_t(1) << (3 * c.dep_level)) + 1;
        TPM_HIP(hipMalloc(&c.dep_start, 3 * nodes1 * sizeof(uint32_t)));
    }
    TPM_HIP(hipMalloc(&c.counter, 8 * sizeof(unsigned long long)));
    TPM_HIP(hipMalloc(&c.flag, sizeof(int) * 4));
    TPM_HIP(hipMalloc(&c.status, sizeof(uint32_t)));
    TPM_HIP(hipMemset(c.status, 0, sizeof(uint32_t)));
    TPM_HIP(hipMalloc(&c.grid, c.m * sizeof(unsigned long long)));
    TPM_HIP(hipMalloc(&c.specA, c.hs * sizeof(float2)));
    TPM_HIP(hipMalloc(&c.specB, c.hs * sizeof(float2)));
    TPM_HIP(hipMalloc(&c.specC, c.hs * sizeof(float2)));
    TPM_HIP(hipMalloc(&c.force, 3 * c.m * sizeof(float)));
    TPM_HIP(hipMalloc(&c.delta, c.m * sizeof(float)));
    TPM_HIP(hipMalloc(&c.stab, 1040 * sizeof(float2)));
    int dev = 0;
    TPM_HIP(hipGetDevice(&dev));
    TPM_HIP(hipDeviceGetAttribute(&c.num_cu, hipDeviceAttributeMultiprocessorCount, dev));
    make_plan(c);
    make_short_range_table(c);
}

// Stage timer: records an event on the context stream when timing is on.
struct Timer {
    Ctx &c;
    std::vector<hipEvent_t> ev;
    explicit Timer(Ctx &cc) : c(cc) {}
    void mark() {
        if (!c.timing) return;
        hipEvent_t e;
        TPM_HIP(hipEventCreate(&e));
        TPM_HIP(hipEventRecord(e, c.stream));
        ev.push_back(e);
    }
    double ms(size_t a, size_t b) {
        if (!c.timing || b >= ev.size()) return 0.0;
        float t = 0;
        TPM_HIP(hipEventElapsedTime(&t, ev[a], ev[b]));
        return t;
    }
    ~Timer() {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
};

// integrator.rs:150-152: K = (3/2) Omega_m L^3 / (4 pi a N); in FP32 for the FP32 interface.
static double coupling(const Ctx &c, double a, size_t n, bool f32) {
    if (f32) {
        const float l = (float)c.p.box_size;
        const float l3 = l * l * l;
        return (double)(1.5f * (float)c.p.omega_m * l3 / (4.0f * 3.14159265358979323846f * (float)a * (float)n));
    }
    const double l = c.p.box_size;
    return 1.5 * c.p.omega_m * l * l * l / (4.0 * M_PI * a * (double)n);
}

// Node and group counts of the last build (a synchronous read; statistics only).
static void read_tree_stats(Ctx &c, bool built) {
    uint32_t info[2] = {0, 0}, ng = 0;
    if (built && c.tree_info) {
        TPM_HIP(hipMemcpy(info, c.tree_info, sizeof(info), hipMemcpyDeviceToHost));
        TPM_HIP(hipMemcpy(&ng, c.ngroups_dev, sizeof(ng), hipMemcpyDeviceToHost));
        if (info[1]) throw Error{TPM_E_CAPACITY, "octree node capacity exceeded"};
    }
    c.stats.nodes = info[0];
    c.stats.groups = ng;
}

// Switches the context's enqueues to its second stream, which first waits for everything enqueued
// on the main stream so far (the sort), with the second set of scan partials; leave() switches back
// (the second stream's work runs on, beside what the main stream does next) and join() makes the
// main stream wait for it. The octree and the walk groups need only the sort, as do the deposit,
// FFT and interpolation, so the two chains overlap and share no scratch buffer.
struct SideStream {
    Ctx &c;
    hipStream_t main;
    unsigned long long *main_scan;
    bool on = false;
    explicit SideStream(Ctx &cc) : c(cc), main(cc.stream), main_scan(cc.scan_part) {
        if (!c.side) {
            // Made on the context's first large evaluation (the ensemble warms its boxes eagerly
            // before a capture, so none is made during one).
            TPM_HIP(hipStreamCreate(&c.side));
            TPM_HIP(hipEventCreateWithFlags(&c.ev_fork, hipEventDisableTiming));
            TPM_HIP(hipEventCreateWithFlags(&c.ev_join, hipEventDisableTiming));
        }
        TPM_HIP(hipEventRecord(c.ev_fork, main));
        TPM_HIP(hipStreamWaitEvent(c.side, c.ev_fork, 0));
        c.stream = c.side;
        c.scan_part = c.scan_part2;
        on = true;
    }
    bool joined = false;
    void leave() {
        TPM_HIP(hipEventRecord(c.ev_join, c.side));
        c.stream = main;
        c.scan_part = main_scan;
        on = false;
    }
    void join() {
        TPM_HIP(hipStreamWaitEvent(main, c.ev_join, 0));
        joined = true;
    }
    // On an error between fork and join the main stream still waits for the second stream's work,
    // so the next evaluation's sort cannot overwrite what an unfinished octree build reads.
    ~SideStream() {
        if (on) {
            (void)hipEventRecord(c.ev_join, c.side);
            c.stream = main;
            c.scan_part = main_scan;
        }
        if (!joined) (void)hipStreamWaitEvent(main, c.ev_join, 0);
    }
};

// Evaluations of more particles than this overlap the octree with the PM part. Smaller ones run
// everything on one stream: there the kernels are short and latency-bound, and in an ensemble of
// boxes, whose streams already overlap one another, the second streams made a 96-box run slower
// (2.41 -> 2.61 s eager, 2.72 -> 3.50 s as a graph), while at 512^3 the overlap saves 1.2 ms.
constexpr size_t kOverlapMin = size_t(1) << 22;

// Octree and walk groups (the group walk's arguments), on whatever stream the context is on.
static WalkArgs tree_chain(Ctx &c, size_t n, double k, float *d_acc, bool group_walk) {
    if (c.timing) {
        for (auto &e : c.tree_ev)
            if (!e) TPM_HIP(hipEventCreate(&e));
        TPM_HIP(hipEventRecord(c.tree_ev[0], c.stream));
    }
    build_tree(c, n, group_walk);
    WalkArgs w{};
    if (group_walk) w = walk_prepare(c, n, k, d_acc);
    if (c.timing) TPM_HIP(hipEventRecord(c.tree_ev[1], c.stream));
    return w;
}

// The force evaluation on device-resident positions (caller order, FP64).
void evaluate(Ctx &c, const double *d_pos, size_t n, double a, bool f32w, float *d_acc, float *d_delta) {
    TPM_REQUIRE(n <= c.max_n, "more particles than the context capacity");
    TPM_REQUIRE(a > 0 && std::isfinite(a), "scale factor must be positive");
    const bool treepm = (c.p.flags & TPM_FLAG_TREEPM) != 0;
    const bool walk = treepm && n > 0;
    const bool group_walk = walk && !(c.p.flags & TPM_FLAG_REF_WALK);
    const double k = walk ? coupling(c, a, n, f32w) : 0.0;
    Timer tm(c);
    zero_fill(c, c.counter, 8 * sizeof(unsigned long long));
    tm.mark();
    sort_particles(c, d_pos, n);
    tm.mark();
    WalkArgs wa{};
    std::unique_ptr<SideStream> side;
    if (walk && n > kOverlapMin && c.overlap) {
        side.reset(new SideStream(c));
        wa = tree_chain(c, n, k, d_acc, group_walk);
        side->leave();
    }
    deposit_fixed(c, n, 1.0, f32w, c.spos, true);
    tm.mark();
    if (n > 0) {
        pm_from_grid(c, n, a, treepm ? c.p.r_split : 0.0, d_delta);
    } else {
        if (d_delta) zero_fill(c, d_delta, c.m * sizeof(float));
    }
    tm.mark();
    interp_force(c, n, f32w, d_acc);
    tm.mark();
    if (side)
        side->join();
    else if (walk)
        wa = tree_chain(c, n, k, d_acc, group_walk);
    tm.mark();
    if (group_walk)
        walk_launch(c, wa, n);
    else if (walk)
        walk_short_range(c, n, k, d_acc);
    tm.mark();
    if (c.p.flags & TPM_FLAG_CHECK_FINITE) check_finite(c, d_acc, 3 * n);
    fold_status(c, treepm && n > 0);
    if (c.timing) {
        TPM_HIP(hipStreamSynchronize(c.stream));
        c.stats.ms_sort = tm.ms(0, 1);
        c.stats.ms_deposit = tm.ms(1, 2);
        c.stats.ms_fft = tm.ms(2, 3);
        c.stats.ms_interp = tm.ms(3, 4);
        // The octree and walk groups (on the second stream beside the three stages above for large
        // evaluations); the walk stage is the walk's launch.
        c.stats.ms_tree = 0.0;
        if (walk) {
            float t = 0.f;
            TPM_HIP(hipEventElapsedTime(&t, c.tree_ev[0], c.tree_ev[1]));
            c.stats.ms_tree = t;
        }
        c.stats.ms_walk = tm.ms(5, 6);
        c.stats.ms_total = tm.ms(0, 6);
        c.stats.ms_walk_kernel = 0.0;
        if (treepm && n > 0 && c.walk_ev[1]) {
            float t = 0.f;
            TPM_HIP(hipEventElapsedTime(&t, c.walk_ev[0], c.walk_ev[1]));
            c.stats.ms_walk_kernel = t;
        }
        unsigned long long cnt[4] = {0, 0, 0, 0};
        TPM_HIP(hipMemcpy(cnt, c.counter, sizeof(cnt), hipMemcpyDeviceToHost));
        c.stats.interactions = cnt[0];
        read_tree_stats(c, treepm && n > 0);
        c.stats.listed = cnt[2];
        c.stats.kept = cnt[3];
        if (cnt[1]) throw Error{TPM_E_DEVICE, "walk stack overflow"};
    }
}

// evaluate() for FP64 positions, split around the short-range walk's launch so that the walks of
// several boxes can share one launch: the same kernels on the same streams (sort; deposit, PM and
// interpolation into d_acc; octree and walk groups, beside them on the second stream above
// kOverlapMin particles), then the prepared walk's arguments.
WalkArgs evaluate_pre_walk(Ctx &c, const double *d_pos, size_t n, double a, float *d_acc) {
    TPM_REQUIRE(n > 0 && n <= c.max_n, "bad particle count for a batched evaluation");
    TPM_REQUIRE(a > 0 && std::isfinite(a), "scale factor must be positive");
    TPM_REQUIRE((c.p.flags & TPM_FLAG_TREEPM) && !(c.p.flags & TPM_FLAG_REF_WALK),
                "a batched evaluation needs the TreePM split and the group walk");
    zero_fill(c, c.counter, 8 * sizeof(unsigned long long));
    sort_particles(c, d_pos, n);
    std::unique_ptr<SideStream> side;
    WalkArgs w{};
    if (n > kOverlapMin && c.overlap) {
        side.reset(new SideStream(c));
        w = tree_chain(c, n, coupling(c, a, n, false), d_acc, true);
        side->leave();
    }
    deposit_fixed(c, n, 1.0, false, c.spos, true);
    pm_from_grid(c, n, a, c.p.r_split, nullptr);
    interp_force(c, n, false, d_acc);
    if (side)
        side->join();
    else
        w = tree_chain(c, n, coupling(c, a, n, false), d_acc, true);
    return w;
}

// The end of evaluate() after the walk: the finiteness check when asked and the status fold.
void evaluate_post_walk(Ctx &c, float *d_acc, size_t n) {
    if (c.p.flags & TPM_FLAG_CHECK_FINITE) check_finite(c, d_acc, 3 * n);
    fold_status(c, true);
}

void finish_checks(Ctx &c) {
    unsigned long long ovf = 0;
    uint32_t info[2] = {0, 0};
    TPM_HIP(hipMemcpyAsync(&ovf, c.counter + 1, sizeof(ovf), hipMemcpyDeviceToHost, c.stream));
    if (c.tree_info)
        TPM_HIP(hipMemcpyAsync(info, c.tree_info, sizeof(info), hipMemcpyDeviceToHost, c.stream));
    TPM_HIP(hipStreamSynchronize(c.stream));
    if (ovf) throw Error{TPM_E_DEVICE, "walk stack overflow"};
    if (info[1]) throw Error{TPM_E_CAPACITY, "octree node capacity exceeded"};
    if (c.p.flags & TPM_FLAG_CHECK_FINITE) {
        int bad = 0;
        TPM_HIP(hipMemcpyAsync(&bad, c.flag, sizeof(int), hipMemcpyDeviceToHost, c.stream));
        TPM_HIP(hipStreamSynchronize(c.stream));
        if (bad) throw Error{TPM_E_NONFINITE, "non-finite acceleration"};
    }
}

// Copies caller positions into c.pos as FP64 (f32 inputs are widened exactly).
static void upload_positions(Ctx &c, const void *pos, size_t n, bool f32) {
    TPM_REQUIRE(n <= c.max_n, "more particles than the context capacity");
    if (n == 0) return;
    TPM_REQUIRE(pos != nullptr, "positions pointer is null");
    if (f32) {
        TPM_HIP(hipMemcpyAsync(c.acc, pos, 3 * n * sizeof(float), hipMemcpyHostToDevice, c.stream));
        convert_f32_to_f64(c, c.acc, c.pos, 3 * n);
    } else {
        TPM_HIP(hipMemcpyAsync(c.pos, pos, 3 * n * sizeof(double), hipMemcpyHostToDevice, c.stream));
    }
}

static int run_host_accel(Ctx &c, const void *pos, size_t n, double a, float *acc, float *delta, bool f32) {
    TPM_REQUIRE(n == 0 || acc != nullptr, "acc pointer is null");
    upload_positions(c, pos, n, f32);
    evaluate(c, c.pos, n, a, f32, c.acc, delta ? c.delta : nullptr);
    finish_checks(c);
    if (n) TPM_HIP(hipMemcpyAsync(acc, c.acc, 3 * n * sizeof(float), hipMemcpyDeviceToHost, c.stream));
    if (delta) TPM_HIP(hipMemcpyAsync(delta, c.delta, c.m * sizeof(float), hipMemcpyDeviceToHost, c.stream));
    TPM_HIP(hipStreamSynchronize(c.stream));
    return TPM_OK;
}

// Judge scratch: m caller targets (u32) at *caller, then sorted targets and 6 m doubles (PM and
// short-range parts); *out receives the 3 m doubles of the result.
void judge_buffers(Ctx &c, size_t m, uint32_t **caller, double **out) {
    if (!c.inv) TPM_HIP(hipMalloc(&c.inv, (c.max_n ? c.max_n : 1) * sizeof(uint32_t)));
    if (c.judge_cap < m) {
        if (c.judge_buf) (void)hipFree(c.judge_buf);
        c.judge_buf = nullptr;
        TPM_HIP(hipMalloc(&c.judge_buf, m * (sizeof(uint32_t) * 2 + 6 * sizeof(double)) + 256));
        c.judge_cap = m;
    }
    *caller = (uint32_t *)c.judge_buf;
    *out = (double *)((char *)c.judge_buf + ((m * 2 * sizeof(uint32_t) + 15) / 16) * 16);
}

// The FP64 judge on device positions for the caller targets d_caller (judge_buffers' layout), result
// into d_out (3 m doubles, = judge_buffers' out). Synchronises: a target out of range throws.
void judge_dev(Ctx &c, const double *d_pos, size_t n, const uint32_t *d_caller, size_t m, double a, double theta,
               double *d_out) {
    TPM_REQUIRE(n > 0 && n <= c.max_n, "judge needs 1..capacity particles");
    TPM_REQUIRE(a > 0 && std::isfinite(a), "scale factor must be positive");
    TPM_REQUIRE(theta >= 0 && std::isfinite(theta), "theta must be >= 0");
    TPM_REQUIRE(m <= c.judge_cap, "judge buffers too small");
    // The judge's octree is the full-depth one whatever the tuned depth limit (the yardstick must
    // not depend on tuning); the product's next evaluation sorts and builds again.
    struct DepthScope {
        Ctx &c;
        int saved;
        ~DepthScope() { c.tree_depth = saved; }
    } depth_scope{c, c.tree_depth};
    c.tree_depth = kMaxDepth;
    sort_particles(c, d_pos, n);
    deposit_fixed(c, n, 1.0, false, c.spos, true);
    const bool treepm = (c.p.flags & TPM_FLAG_TREEPM) != 0;
    pm_from_grid_f64(c, n, a, treepm ? c.p.r_split : 0.0);
    uint32_t *d_sorted = const_cast<uint32_t *>(d_caller) + m;
    double *d_pm = d_out;
    double *d_sr = d_pm + 3 * m;
    hipLaunchKernelGGL(k_inverse_perm, dim3(blocks(n, 256)), dim3(256), 0, c.stream, c.idx, n, c.inv);
    zero_fill(c, c.flag, sizeof(int));
    hipLaunchKernelGGL(k_map_targets, dim3(blocks(m, 256)), dim3(256), 0, c.stream, c.inv, d_caller, m, n, d_sorted,
                       c.flag);
    TPM_HIP(hipGetLastError());
    interp_judge_f64(c, d_sorted, m, d_pm);
    if (treepm) {
        build_tree(c, n);
        const double k = 1.5 * c.p.omega_m * c.p.box_size * c.p.box_size * c.p.box_size / (4.0 * M_PI * a * (double)n);
        walk_judge_f64(c, n, d_sorted, m, k, theta, d_sr);
        hipLaunchKernelGGL(k_add3, dim3(blocks(3 * m, 256)), dim3(256), 0, c.stream, d_pm, d_sr, 3 * m);
        TPM_HIP(hipGetLastError());
    }
    int bad = 0;
    TPM_HIP(hipMemcpyAsync(&bad, c.flag, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    TPM_HIP(hipStreamSynchronize(c.stream));
    TPM_REQUIRE(!bad, "judge target index out of range");
}

}  // namespace tpm

using tpm::Ctx;

struct tpm_ctx {
    Ctx c;
};

tpm::Ctx &tpm_ctx_internal(tpm_ctx *ctx) { return ctx->c; }

template <typename F>
static int guarded(tpm_ctx *ctx, F &&f) {
    if (!ctx) return TPM_E_INVALID;
    try {
        return f(ctx->c);
    } catch (const tpm::Error &e) {
        ctx->c.err = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        ctx->c.err = "host allocation failed";
        return TPM_E_DEVICE;
    }
}

extern "C" {

const char *tpm_version(void) { return "tpm 0.1 (gfx950)"; }

int tpm_create(const tpm_params *params, size_t max_particles, tpm_ctx **out) {
    if (!params || !out) {
        tpm::set_global_error("null argument");
        return TPM_E_INVALID;
    }
    *out = nullptr;
    tpm_ctx *ctx = new (std::nothrow) tpm_ctx();
    if (!ctx) {
        tpm::set_global_error("host allocation failed");
        return TPM_E_DEVICE;
    }
    try {
        tpm::create(ctx->c, *params, max_particles);
    } catch (const tpm::Error &e) {
        tpm::set_global_error(e.msg.c_str());
        tpm::free_all(ctx->c);
        delete ctx;
        return e.code;
    }
    *out = ctx;
    return TPM_OK;
}

void tpm_destroy(tpm_ctx *ctx) {
    if (!ctx) return;
    (void)hipStreamSynchronize(ctx->c.stream);
    tpm::free_all(ctx->c);
    delete ctx;
}

const char *tpm_last_error(const tpm_ctx *ctx) {
    if (!ctx) return tpm::g_create_error;
    return ctx->c.err.c_str();
}

int tpm_set_params(tpm_ctx *ctx, const tpm_params *params) {
    return guarded(ctx, [&](Ctx &c) {
        TPM_REQUIRE(params != nullptr, "null params");
        tpm::validate(*params);
        TPM_REQUIRE(params->mesh_n == c.p.mesh_n, "mesh_n cannot change on a context");
        const bool box_changed = params->box_size != c.p.box_size;
        c.p = *params;
        if (box_changed) tpm::make_plan(c);
        tpm::make_short_range_table(c);
        return TPM_OK;
    });
}

int tpm_set_timing(tpm_ctx *ctx, int enabled) {
    return guarded(ctx, [&](Ctx &c) {
        c.timing = enabled != 0;
        return TPM_OK;
    });
}

int tpm_set_counting(tpm_ctx *ctx, int enabled) {
    return guarded(ctx, [&](Ctx &c) {
        c.count_pairs = enabled != 0;
        return TPM_OK;
    });
}

int tpm_set_overlap(tpm_ctx *ctx, int enabled) {
    return guarded(ctx, [&](Ctx &c) {
        c.overlap = enabled != 0;
        return TPM_OK;
    });
}

int tpm_get_stats(const tpm_ctx *ctx, tpm_stats *out) {
    if (!ctx || !out) return TPM_E_INVALID;
    *out = ctx->c.stats;
    return TPM_OK;
}

int tpm_accelerations_f32(tpm_ctx *ctx, const float *pos, size_t n, double a, float *acc, float *delta) {
    return guarded(ctx, [&](Ctx &c) { return tpm::run_host_accel(c, pos, n, a, acc, delta, true); });
}

int tpm_accelerations_f64(tpm_ctx *ctx, const double *pos, size_t n, double a, float *acc, float *delta) {
    return guarded(ctx, [&](Ctx &c) { return tpm::run_host_accel(c, pos, n, a, acc, delta, false); });
}

int tpm_accelerations_dev_f64(tpm_ctx *ctx, const double *d_pos, size_t n, double a, float *d_acc, float *d_delta,
                              void *stream) {
    return guarded(ctx, [&](Ctx &c) {
        TPM_REQUIRE(n == 0 || (d_pos && d_acc), "null device pointer");
        hipStream_t saved = c.stream;
        if (stream) c.stream = (hipStream_t)stream;
        try {
            tpm::evaluate(c, d_pos, n, a, false, d_acc, d_delta);
        } catch (...) {
            c.stream = saved;
            throw;
        }
        c.stream = saved;
        return TPM_OK;
    });
}

int tpm_deposit_f32(tpm_ctx *ctx, const float *pos, size_t n, float mass, float *cells) {
    return guarded(ctx, [&](Ctx &c) {
        TPM_REQUIRE(cells != nullptr, "cells pointer is null");
        tpm::upload_positions(c, pos, n, true);
        tpm::deposit_fixed(c, n, (double)mass, true, c.pos, false);
        tpm::grid_to_float(c, n ? n : 1, nullptr, c.delta);
        TPM_HIP(hipMemcpyAsync(cells, c.delta, c.m * sizeof(float), hipMemcpyDeviceToHost, c.stream));
        TPM_HIP(hipStreamSynchronize(c.stream));
        return TPM_OK;
    });
}

int tpm_deposit_f64(tpm_ctx *ctx, const double *pos, size_t n, double mass, float *cells) {
    return guarded(ctx, [&](Ctx &c) {
        TPM_REQUIRE(cells != nullptr, "cells pointer is null");
        tpm::upload_positions(c, pos, n, false);
        tpm::deposit_fixed(c, n, mass, false, c.pos, false);
        tpm::grid_to_float(c, n ? n : 1, nullptr, c.delta);
        TPM_HIP(hipMemcpyAsync(cells, c.delta, c.m * sizeof(float), hipMemcpyDeviceToHost, c.stream));
        TPM_HIP(hipStreamSynchronize(c.stream));
        return TPM_OK;
    });
}

int tpm_density_contrast_f32(tpm_ctx *ctx, const float *pos, size_t n, float *delta) {
    return guarded(ctx, [&](Ctx &c) {
        TPM_REQUIRE(delta != nullptr, "delta pointer is null");
        TPM_REQUIRE(n > 0, "density contrast needs at least one particle");
        tpm::upload_positions(c, pos, n, true);
        tpm::deposit_fixed(c, n, 1.0, true, c.pos, false);
        tpm::grid_to_float(c, n, c.delta, nullptr);
        TPM_HIP(hipMemcpyAsync(delta, c.delta, c.m * sizeof(float), hipMemcpyDeviceToHost, c.stream));
        TPM_HIP(hipStreamSynchronize(c.stream));
        return TPM_OK;
    });
}

int tpm_pm_accel(tpm_ctx *ctx, const float *delta, double a, double r_split, float *gx, float *gy, float *gz) {
    return guarded(ctx, [&](Ctx &c) {
        TPM_REQUIRE(delta && gx && gy && gz, "null grid pointer");
        TPM_REQUIRE(a > 0 && std::isfinite(a), "scale factor must be positive");
        TPM_REQUIRE(r_split >= 0 && std::isfinite(r_split), "r_split must be >= 0");
        TPM_HIP(hipMemcpyAsync(c.delta, delta, c.m * sizeof(float), hipMemcpyHostToDevice, c.stream));
        tpm::pm_from_delta(c, c.delta, a, r_split);
        // The device grid holds (x, y, z) per mesh point; the caller gets three planes.
        std::vector<float> g(3 * c.m);
        TPM_HIP(hipMemcpyAsync(g.data(), c.force, 3 * c.m * sizeof(float), hipMemcpyDeviceToHost, c.stream));
        TPM_HIP(hipStreamSynchronize(c.stream));
        for (size_t i = 0; i < c.m; ++i) {
            gx[i] = g[3 * i];
            gy[i] = g[3 * i + 1];
            gz[i] = g[3 * i + 2];
        }
        return TPM_OK;
    });
}

int tpm_interpolate_f32(tpm_ctx *ctx, const float *cells, const float *pos, size_t n, float *out) {
    return guarded(ctx, [&](Ctx &c) {
        TPM_REQUIRE(cells && (n == 0 || out), "null pointer");
        TPM_HIP(hipMemcpyAsync(c.delta, cells, c.m * sizeof(float), hipMemcpyHostToDevice, c.stream));
        tpm::upload_positions(c, pos, n, true);
        tpm::interp_one(c, c.delta, c.pos, n, true, c.acc);
        if (n) TPM_HIP(hipMemcpyAsync(out, c.acc, n * sizeof(float), hipMemcpyDeviceToHost, c.stream));
        TPM_HIP(hipStreamSynchronize(c.stream));
        return TPM_OK;
    });
}

static int short_range(tpm_ctx *ctx, const void *pos, size_t n, double k, float *acc, bool f32) {
    return guarded(ctx, [&](Ctx &c) {
        TPM_REQUIRE(n == 0 || acc != nullptr, "acc pointer is null");
        TPM_REQUIRE(c.p.r_split > 0 && c.p.r_cut > 0, "short-range force needs r_split > 0 and r_cut > 0");
        tpm::upload_positions(c, pos, n, f32);
        if (n == 0) return TPM_OK;
        tpm::zero_fill(c, c.counter, 8 * sizeof(unsigned long long));
        tpm::sort_particles(c, c.pos, n);
        tpm::zero_fill(c, c.acc, 3 * n * sizeof(float));
        tpm::build_tree(c, n, !(c.p.flags & TPM_FLAG_REF_WALK));
        tpm::walk_short_range(c, n, k, c.acc);
        TPM_HIP(hipMemcpyAsync(acc, c.acc, 3 * n * sizeof(float), hipMemcpyDeviceToHost, c.stream));
        unsigned long long cnt = 0;
        TPM_HIP(hipMemcpyAsync(&cnt, c.counter, sizeof(cnt), hipMemcpyDeviceToHost, c.stream));
        TPM_HIP(hipStreamSynchronize(c.stream));
        c.stats.interactions = cnt;
        tpm::read_tree_stats(c, true);
        return TPM_OK;
    });
}

int tpm_short_range_f32(tpm_ctx *ctx, const float *pos, size_t n, double k, float *acc) {
    return short_range(ctx, pos, n, k, acc, true);
}

int tpm_short_range_f64(tpm_ctx *ctx, const double *pos, size_t n, double k, float *acc) {
    return short_range(ctx, pos, n, k, acc, false);
}

int tpm_judge_f64(tpm_ctx *ctx, const double *pos, size_t n, const uint32_t *idx, size_t m, double a, double theta,
                  double *acc) {
    return guarded(ctx, [&](Ctx &c) {
        TPM_REQUIRE(n > 0 && pos && idx && acc, "judge needs particles, targets and an output");
        tpm::upload_positions(c, pos, n, false);
        uint32_t *d_caller = nullptr;
        double *d_out = nullptr;
        tpm::judge_buffers(c, m, &d_caller, &d_out);
        TPM_HIP(hipMemcpyAsync(d_caller, idx, m * sizeof(uint32_t), hipMemcpyHostToDevice, c.stream));
        tpm::judge_dev(c, c.pos, n, d_caller, m, a, theta, d_out);
        TPM_HIP(hipMemcpyAsync(acc, d_out, 3 * m * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        TPM_HIP(hipStreamSynchronize(c.stream));
        return TPM_OK;
    });
}

int tpm_morton_order(tpm_ctx *ctx, const double *pos, size_t n, uint32_t *order, uint64_t *keys) {
    return guarded(ctx, [&](Ctx &c) {
        TPM_REQUIRE(n == 0 || order != nullptr, "order pointer is null");
        tpm::upload_positions(c, pos, n, false);
        tpm::sort_particles(c, c.pos, n);
        if (n) {
            TPM_HIP(hipMemcpyAsync(order, c.idx, n * sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
            if (keys) TPM_HIP(hipMemcpyAsync(keys, c.keys, n * sizeof(uint64_t), hipMemcpyDeviceToHost, c.stream));
        }
        TPM_HIP(hipStreamSynchronize(c.stream));
        return TPM_OK;
    });
}

int tpm_get_device_status(tpm_ctx *ctx, uint32_t *status, int clear) {
    return guarded(ctx, [&](Ctx &c) {
        TPM_REQUIRE(status != nullptr, "status pointer is null");
        if (c.status_stream && c.status_stream != c.stream) TPM_HIP(hipStreamSynchronize(c.status_stream));
        TPM_HIP(hipStreamSynchronize(c.stream));
        uint32_t s = 0;
        TPM_HIP(hipMemcpy(&s, c.status, sizeof(s), hipMemcpyDeviceToHost));
        if (clear) TPM_HIP(hipMemset(c.status, 0, sizeof(uint32_t)));
        *status = s;
        return TPM_OK;
    });
}

int tpm_device_alloc(size_t bytes, void **out) {
    if (!out) return TPM_E_INVALID;
    *out = nullptr;
    if (bytes == 0) return TPM_OK;
    return hipMalloc(out, bytes) == hipSuccess ? TPM_OK : TPM_E_DEVICE;
}

int tpm_device_free(void *ptr) {
    if (!ptr) return TPM_OK;
    return hipFree(ptr) == hipSuccess ? TPM_OK : TPM_E_DEVICE;
}

int tpm_memcpy(void *dst, const void *src, size_t bytes, int kind) {
    if (bytes == 0) return TPM_OK;
    if (!dst || !src || kind < 0 || kind > 2) return TPM_E_INVALID;
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost
                                                                          : hipMemcpyDeviceToDevice;
    return hipMemcpy(dst, src, bytes, k) == hipSuccess ? TPM_OK : TPM_E_DEVICE;
}

int tpm_sync(tpm_ctx *ctx) {
    return guarded(ctx, [&](Ctx &c) {
        TPM_HIP(hipStreamSynchronize(c.stream));
        return TPM_OK;
    });
}

size_t tpm_mlp_param_count(void) { return tpm::mlp_param_count(); }

int tpm_mlp_init(tpm_ctx *ctx, uint64_t seed) {
    return guarded(ctx, [&](Ctx &c) {
        tpm::mlp_init(c, seed);
        c.mlp.trained = false;
        TPM_HIP(hipStreamSynchronize(c.stream));
        return TPM_OK;
    });
}

int tpm_mlp_get_params(tpm_ctx *ctx, float *params, size_t count) {
    return guarded(ctx, [&](Ctx &c) {
        TPM_REQUIRE(params && count == tpm::mlp_param_count(), "params must hold tpm_mlp_param_count() floats");
        TPM_REQUIRE(c.mlp.params != nullptr, "the MLP is not initialised");
        TPM_HIP(hipStreamSynchronize(c.stream));
        TPM_HIP(hipMemcpy(params, c.mlp.params, count * sizeof(float), hipMemcpyDeviceToHost));
        return TPM_OK;
    });
}

int tpm_mlp_set_params(tpm_ctx *ctx, const float *params, size_t count) {
    return guarded(ctx, [&](Ctx &c) {
        TPM_REQUIRE(params && count == tpm::mlp_param_count(), "params must hold tpm_mlp_param_count() floats");
        if (!c.mlp.params) tpm::mlp_init(c, 1);
        TPM_HIP(hipMemcpy(c.mlp.params, params, count * sizeof(float), hipMemcpyHostToDevice));
        tpm::mlp_refresh_bf16(c);
        c.mlp.trained = true;
        TPM_HIP(hipStreamSynchronize(c.stream));
        return TPM_OK;
    });
}

static size_t round32(size_t r) { return (r + 31) / 32 * 32; }

int tpm_mlp_gradients(tpm_ctx *ctx, const float *x, const float *t, size_t rows, float *grads, double *loss) {
    return guarded(ctx, [&](Ctx &c) {
        TPM_REQUIRE(x && t && grads && rows > 0, "null argument");
        TPM_REQUIRE(c.mlp.params != nullptr, "the MLP is not initialised");
        const size_t rp = round32(rows);
        tpm::mlp_prepare(c, rp);
        float *dx = nullptr;
        TPM_HIP(hipMalloc(&dx, rows * 36 * sizeof(float)));
        try {
            TPM_HIP(hipMemcpyAsync(dx, x, rows * 33 * sizeof(float), hipMemcpyHostToDevice, c.stream));
            TPM_HIP(hipMemcpyAsync(dx + rows * 33, t, rows * 3 * sizeof(float), hipMemcpyHostToDevice, c.stream));
            tpm::mlp_pack_rows(c, dx, dx + rows * 33, rows, rp);
            const double l = tpm::mlp_forward_backward(c, rp, rows, true);
            if (loss) *loss = l;
            TPM_HIP(hipMemcpy(grads, c.mlp.grads, tpm::mlp_param_count() * sizeof(float), hipMemcpyDeviceToHost));
        } catch (...) {
            (void)hipFree(dx);
            throw;
        }
        (void)hipFree(dx);
        return TPM_OK;
    });
}

int tpm_mlp_predict(tpm_ctx *ctx, const float *x, size_t rows, float *out) {
    return guarded(ctx, [&](Ctx &c) {
        TPM_REQUIRE(x && out && rows > 0, "null argument");
        TPM_REQUIRE(c.mlp.params != nullptr, "the MLP is not initialised");
        const size_t rp = round32(rows);
        tpm::mlp_prepare(c, rp);
        float *dx = nullptr;
        TPM_HIP(hipMalloc(&dx, rows * 33 * sizeof(float)));
        try {
            TPM_HIP(hipMemcpyAsync(dx, x, rows * 33 * sizeof(float), hipMemcpyHostToDevice, c.stream));
            tpm::mlp_pack_rows(c, dx, nullptr, rows, rp);
            // The inference kernel's network (the same arithmetic as the training forward pass).
            tpm::mlp_predict_rows(c, c.mlp.X, rows, c.mlp.Zout);
            TPM_HIP(hipStreamSynchronize(c.stream));
            TPM_HIP(hipMemcpy(out, c.mlp.Zout, rows * 4 * sizeof(float), hipMemcpyDeviceToHost));
        } catch (...) {
            (void)hipFree(dx);
            throw;
        }
        (void)hipFree(dx);
        return TPM_OK;
    });
}

// K / h^2: the force scale of the learned model's inputs and outputs.
static double learned_fscale(const Ctx &c, double a, size_t n) {
    const double k = 1.5 * c.p.omega_m * c.p.box_size * c.p.box_size * c.p.box_size / (4.0 * M_PI * a * (double)n);
    const double h = c.p.box_size / c.n;
    return k / (h * h);
}

// PM part of an evaluation with the density contrast kept in c.delta: sort, deposit, r2c ... c2r,
// interpolation into d_acc.
static void pm_part(Ctx &c, const double *d_pos, size_t n, double a, float *d_acc) {
    tpm::sort_particles(c, d_pos, n);
    tpm::deposit_fixed(c, n, 1.0, false, c.spos, true);
    tpm::pm_from_grid(c, n, a, c.p.r_split, c.delta);
    tpm::interp_force(c, n, false, d_acc);
}

int tpm_mlp_train_dev(tpm_ctx *ctx, const double *d_pos, size_t n, double a, size_t n_samples, int epochs,
                      size_t batch, float lr, uint64_t seed, double *loss_history) {
    return guarded(ctx, [&](Ctx &c) {
        TPM_REQUIRE(d_pos && n > 0 && n <= c.max_n, "bad particle set");
        TPM_REQUIRE((c.p.flags & TPM_FLAG_TREEPM) != 0, "training needs the TreePM split (targets come from the walk)");
        TPM_REQUIRE(epochs >= 0 && batch >= 32 && lr > 0, "bad training parameters");
        if (!c.mlp.params) tpm::mlp_init(c, seed);
        if (!c.acc_sr) TPM_HIP(hipMalloc(&c.acc_sr, 3 * c.max_n * sizeof(float)));
        const double fscale = learned_fscale(c, a, n);
        const double k = 1.5 * c.p.omega_m * c.p.box_size * c.p.box_size * c.p.box_size / (4.0 * M_PI * a * (double)n);
        pm_part(c, d_pos, n, a, c.acc);
        tpm::zero_fill(c, c.acc_sr, 3 * n * sizeof(float));
        tpm::build_tree(c, n, !(c.p.flags & TPM_FLAG_REF_WALK));
        tpm::walk_short_range(c, n, k, c.acc_sr);
        const size_t ns = std::min(n_samples, n);
        // Evenly spaced particles of the Morton order (every particle equally likely: density weighted).
        std::vector<uint32_t> sel(ns);
        const uint64_t off = seed % std::max<size_t>(1, n / ns);
        for (size_t t = 0; t < ns; ++t) sel[t] = (uint32_t)std::min<uint64_t>(n - 1, t * (uint64_t)n / ns + off);
        uint32_t *d_sel = nullptr, *d_perm = nullptr;
        uint16_t *Xall = nullptr;
        float *Tall = nullptr;
        auto cleanup = [&]() {
            (void)hipFree(d_sel);
            (void)hipFree(d_perm);
            (void)hipFree(Xall);
            (void)hipFree(Tall);
        };
        try {
            TPM_HIP(hipMalloc(&d_sel, ns * sizeof(uint32_t)));
            TPM_HIP(hipMalloc(&d_perm, ns * sizeof(uint32_t)));
            TPM_HIP(hipMalloc(&Xall, ns * 64 * sizeof(uint16_t)));
            TPM_HIP(hipMalloc(&Tall, ns * 3 * sizeof(float)));
            TPM_HIP(hipMemcpyAsync(d_sel, sel.data(), ns * sizeof(uint32_t), hipMemcpyHostToDevice, c.stream));
            tpm::mlp_features(c, d_sel, 0, ns, Xall, (float)(1.0 / fscale));
            tpm::mlp_targets(c, d_sel, ns, c.acc_sr, (float)(1.0 / fscale), Tall);
            const size_t rows = round32(batch);
            tpm::mlp_prepare(c, rows);
            std::vector<uint32_t> perm(ns);
            uint64_t st = seed * 0x9E3779B97F4A7C15ULL + 7;
            auto next = [&]() {
                st ^= st << 13;
                st ^= st >> 7;
                st ^= st << 17;
                return st;
            };
            const size_t steps_total = (size_t)epochs * ((ns + rows - 1) / rows);
            size_t step = 0;
            for (int e = 0; e < epochs; ++e) {
                for (size_t t = 0; t < ns; ++t) perm[t] = (uint32_t)t;
                for (size_t t = ns - 1; t > 0; --t) std::swap(perm[t], perm[next() % (t + 1)]);
                TPM_HIP(hipMemcpyAsync(d_perm, perm.data(), ns * sizeof(uint32_t), hipMemcpyHostToDevice, c.stream));
                double last = 0.0;
                for (size_t b0 = 0; b0 < ns; b0 += rows) {
                    const size_t cnt = std::min(rows, ns - b0);
                    tpm::mlp_gather_batch(c, Xall, Tall, d_perm, b0, ns, rows);
                    const bool want = b0 + rows >= ns;
                    last = tpm::mlp_forward_backward(c, rows, cnt, want);
                    // Linear decay of the learning rate to a tenth over the run.
                    const float f = 1.0f - 0.9f * (float)step / (float)std::max<size_t>(1, steps_total);
                    tpm::mlp_adam_step(c, lr * f);
                    ++step;
                }
                if (loss_history) loss_history[e] = last;
            }
            // Calibration of the gate on up to 65,536 particles midway between consecutive training samples
            // (held out when the samples are at least two apart; the training samples themselves otherwise).
            const size_t gap = n / ns;
            const size_t mcal = std::min<size_t>(65536, ns);
            std::vector<uint32_t> cal(mcal);
            for (size_t t = 0; t < mcal; ++t)
                cal[t] = (uint32_t)std::min<uint64_t>(n - 1, sel[t * ns / mcal] + (gap >= 2 ? gap / 2 : 0));
            TPM_HIP(hipMemcpyAsync(d_sel, cal.data(), mcal * sizeof(uint32_t), hipMemcpyHostToDevice, c.stream));
            tpm::mlp_calibrate(c, d_sel, mcal, (float)fscale, c.acc_sr, c.acc);
            TPM_HIP(hipStreamSynchronize(c.stream));
        } catch (...) {
            cleanup();
            throw;
        }
        cleanup();
        c.mlp.trained = true;
        return TPM_OK;
    });
}

__global__ void k_compact(const uint32_t *__restrict__ flag, const uint32_t *__restrict__ pos, size_t n,
                          uint32_t *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && flag[i]) out[pos[i]] = (uint32_t)i;
}

// walked[idx[i]] = reject[i]: the fallback mask in caller order.
__global__ void k_walked_mask(const uint32_t *__restrict__ reject, const uint32_t *__restrict__ order, size_t n,
                              uint32_t *__restrict__ walked) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) walked[order[i]] = reject[i];
}

int tpm_mlp_error_scale(tpm_ctx *ctx, double tol, double *kappa) {
    return guarded(ctx, [&](Ctx &c) {
        TPM_REQUIRE(kappa != nullptr, "null output");
        *kappa = tpm::mlp_kappa(c, tol);
        return TPM_OK;
    });
}

int tpm_accelerations_learned_dev(tpm_ctx *ctx, const double *d_pos, size_t n, double a, double tol, float *d_acc,
                                  uint64_t *n_fallback, uint32_t *d_walked) {
    return guarded(ctx, [&](Ctx &c) {
        TPM_REQUIRE(d_pos && d_acc && n > 0 && n <= c.max_n, "bad particle set");
        TPM_REQUIRE((c.p.flags & TPM_FLAG_TREEPM) != 0, "the learned term replaces the TreePM short-range force");
        TPM_REQUIRE(c.mlp.params != nullptr, "the MLP is not initialised");
        TPM_REQUIRE(tol >= 0, "tol must be >= 0");
        if (!c.members) TPM_HIP(hipMalloc(&c.members, c.max_n * sizeof(uint32_t)));
        if (!c.reject) TPM_HIP(hipMalloc(&c.reject, c.max_n * sizeof(uint32_t)));
        const double fscale = learned_fscale(c, a, n);
        const double k = 1.5 * c.p.omega_m * c.p.box_size * c.p.box_size * c.p.box_size / (4.0 * M_PI * a * (double)n);
        tpm::Timer tm(c);
        tm.mark();
        pm_part(c, d_pos, n, a, d_acc);
        tm.mark();
        const double kappa = tpm::mlp_kappa(c, tol);
        tpm::mlp_infer_fused(c, nullptr, 0, n, (float)fscale, (float)tol, std::isfinite(kappa) ? (float)kappa : INFINITY,
                             d_acc, c.reject, nullptr);
        if (d_walked) {
            hipLaunchKernelGGL(k_walked_mask, dim3(tpm::blocks(n, 256)), dim3(256), 0, c.stream, c.reject, c.idx, n,
                               d_walked);
            TPM_HIP(hipGetLastError());
        }
        tm.mark();
        // Members = rejected sorted indices (exclusive scan + scatter).
        tpm::exclusive_sum_u32(c, c.reject, c.gscratch, n);
        hipLaunchKernelGGL(k_compact, dim3(tpm::blocks(n, 256)), dim3(256), 0, c.stream, c.reject, c.gscratch, n,
                           c.members);
        uint32_t last[2];
        TPM_HIP(hipMemcpyAsync(&last[0], c.gscratch + n - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
        TPM_HIP(hipMemcpyAsync(&last[1], c.reject + n - 1, sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream));
        TPM_HIP(hipStreamSynchronize(c.stream));
        const size_t nrej = (size_t)last[0] + last[1];
        if (n_fallback) *n_fallback = nrej;
        if (nrej > 0) {
            tpm::zero_fill(c, c.counter, 8 * sizeof(unsigned long long));
            tpm::build_tree(c, n, !(c.p.flags & TPM_FLAG_REF_WALK));
            tpm::walk_short_range(c, n, k, d_acc, c.members, nrej);
            tpm::fold_status(c, true);
        }
        tm.mark();
        if (c.timing) {
            TPM_HIP(hipStreamSynchronize(c.stream));
            c.stats.ms_fft = tm.ms(0, 1);
            c.stats.ms_interp = tm.ms(1, 2);
            c.stats.ms_walk = tm.ms(2, 3);
            c.stats.ms_total = tm.ms(0, 3);
            c.stats.ms_walk_kernel = 0.0;
            if (nrej > 0 && c.walk_ev[1]) {
                float t = 0.f;
                TPM_HIP(hipEventElapsedTime(&t, c.walk_ev[0], c.walk_ev[1]));
                c.stats.ms_walk_kernel = t;
            }
        }
        return TPM_OK;
    });
}

int tpm_kdk_factors(const tpm_cosmology *cosmo, double a0, double a1, double *am, double *k1, double *d, double *k2) {
    if (!cosmo || !(a0 > 0) || !(a1 > 0) || !std::isfinite(a0) || !std::isfinite(a1)) return TPM_E_INVALID;
    const tpm::KdkFactors f = tpm::kdk_factors(*cosmo, a0, a1);
    if (am) *am = f.am;
    if (k1) *k1 = f.k1;
    if (d) *d = f.d;
    if (k2) *k2 = f.k2;
    return TPM_OK;
}

int tpm_snapshot_dev(tpm_ctx *ctx, const double *d_pos, const double *d_mom, size_t n, double a, double e,
                     const float *d_delta, float *d_pos32, float *d_speed, float *d_density, float *d_rsd,
                     void *stream) {
    return guarded(ctx, [&](Ctx &c) {
        TPM_REQUIRE(a > 0 && e > 0 && std::isfinite(a) && std::isfinite(e), "a and E must be positive");
        TPM_REQUIRE(n == 0 || d_pos != nullptr, "null positions");
        TPM_REQUIRE(!(d_speed || d_rsd) || d_mom != nullptr, "speeds and shifts need momenta");
        TPM_REQUIRE(!d_density || d_delta != nullptr, "densities need the density-contrast mesh");
        hipStream_t saved = c.stream;
        if (stream) c.stream = (hipStream_t)stream;
        try {
            tpm::snapshot(c, d_pos, d_mom, n, a, e, d_delta, d_pos32, d_speed, d_density, d_rsd);
        } catch (...) {
            c.stream = saved;
            throw;
        }
        c.stream = saved;
        return TPM_OK;
    });
}

int tpm_downsample_grid_dev(tpm_ctx *ctx, const float *d_delta, float *d_out, int32_t *side, void *stream) {
    return guarded(ctx, [&](Ctx &c) {
        const int m = tpm::downsample_side(c.n, nullptr);
        if (side) *side = m;
        if (!d_out) return TPM_OK;
        TPM_REQUIRE(d_delta != nullptr, "null density-contrast mesh");
        hipStream_t saved = c.stream;
        if (stream) c.stream = (hipStream_t)stream;
        try {
            tpm::downsample_grid(c, d_delta, d_out);
        } catch (...) {
            c.stream = saved;
            throw;
        }
        c.stream = saved;
        return TPM_OK;
    });
}

int tpm_power_spectrum(const tpm_cosmology *cosmo, double h0, double n_s, double sigma8, double k, double *p,
                       double *amplitude) {
    if (!cosmo || !(h0 > 0) || !(sigma8 > 0) || !(k >= 0) || !std::isfinite(k)) return TPM_E_INVALID;
    const tpm::PowerSpec ps = tpm::normalized_power(*cosmo, h0, n_s, sigma8, 8.0);
    if (p) *p = ps.power(k);
    if (amplitude) *amplitude = ps.amplitude;
    return TPM_OK;
}

int tpm_power_sigma(const tpm_cosmology *cosmo, double h0, double n_s, double sigma8, double radius, double *sigma) {
    if (!cosmo || !(h0 > 0) || !(sigma8 > 0) || !(radius > 0) || !sigma) return TPM_E_INVALID;
    *sigma = tpm::normalized_power(*cosmo, h0, n_s, sigma8, 8.0).sigma(radius);
    return TPM_OK;
}

int tpm_zeldovich_bbks_dev(tpm_ctx *ctx, const tpm_cosmology *cosmo, double h0, double n_s, double sigma8,
                           uint64_t seed, double a_init, double *d_pos, double *d_mom, void *stream) {
    return guarded(ctx, [&](Ctx &c) {
        TPM_REQUIRE(cosmo != nullptr && d_pos != nullptr, "null argument");
        TPM_REQUIRE(h0 > 0 && sigma8 > 0, "h0 and sigma8 must be positive");
        const tpm::PowerSpec ps = tpm::normalized_power(*cosmo, h0, n_s, sigma8, 8.0);
        hipStream_t saved = c.stream;
        if (stream) c.stream = (hipStream_t)stream;
        try {
            tpm::zeldovich_bbks(c, *cosmo, ps, seed, a_init, d_pos, d_mom);
        } catch (...) {
            c.stream = saved;
            throw;
        }
        c.stream = saved;
        return TPM_OK;
    });
}

int tpm_zeldovich_bbks(tpm_ctx *ctx, const tpm_cosmology *cosmo, double h0, double n_s, double sigma8, uint64_t seed,
                       double a_init, double *pos, double *mom) {
    return guarded(ctx, [&](Ctx &c) {
        TPM_REQUIRE(cosmo != nullptr && pos != nullptr, "null argument");
        TPM_REQUIRE(h0 > 0 && sigma8 > 0, "h0 and sigma8 must be positive");
        TPM_REQUIRE(c.m <= c.max_n, "context capacity must hold mesh_n^3 particles");
        const tpm::PowerSpec ps = tpm::normalized_power(*cosmo, h0, n_s, sigma8, 8.0);
        double *d_mom = nullptr;
        TPM_HIP(hipMalloc(&d_mom, 3 * c.m * sizeof(double)));
        try {
            tpm::zeldovich_bbks(c, *cosmo, ps, seed, a_init, c.pos, d_mom);
            TPM_HIP(hipMemcpyAsync(pos, c.pos, 3 * c.m * sizeof(double), hipMemcpyDeviceToHost, c.stream));
            if (mom) TPM_HIP(hipMemcpyAsync(mom, d_mom, 3 * c.m * sizeof(double), hipMemcpyDeviceToHost, c.stream));
            TPM_HIP(hipStreamSynchronize(c.stream));
        } catch (...) {
            (void)hipFree(d_mom);
            throw;
        }
        (void)hipFree(d_mom);
        return TPM_OK;
    });
}

int tpm_growth(const tpm_cosmology *cosmo, double a, double *d, double *f, double *e) {
    if (!cosmo || !(a > 0) || !std::isfinite(a)) return TPM_E_INVALID;
    tpm::growth(*cosmo, a, d, f);
    if (e) *e = tpm::hubble_e(*cosmo, a);
    return TPM_OK;
}

int tpm_step_to_dev(tpm_ctx *ctx, const tpm_cosmology *cosmo, double *d_pos, double *d_mom, size_t n, double a0,
                    double a1, float *d_acc, float *d_delta, void *stream) {
    return guarded(ctx, [&](Ctx &c) {
        TPM_REQUIRE(cosmo != nullptr, "null cosmology");
        TPM_REQUIRE(n == 0 || (d_pos && d_mom && d_acc), "null device pointer");
        TPM_REQUIRE(a0 > 0 && a1 > 0 && std::isfinite(a0) && std::isfinite(a1), "scale factors must be positive");
        hipStream_t saved = c.stream;
        if (stream) c.stream = (hipStream_t)stream;
        try {
            const tpm::KdkFactors f = tpm::kdk_factors(*cosmo, a0, a1);
            tpm::kick(c, d_mom, d_acc, n, f.k1);
            tpm::drift(c, d_pos, d_mom, n, f.d);
            tpm::evaluate(c, d_pos, n, a1, false, d_acc, d_delta);
            tpm::kick(c, d_mom, d_acc, n, f.k2);
        } catch (...) {
            c.stream = saved;
            throw;
        }
        c.stream = saved;
        return TPM_OK;
    });
}

int tpm_run(tpm_ctx *ctx, const tpm_cosmology *cosmo, double *pos, double *mom, size_t n, double a_start,
            double a_end, int steps) {
    return guarded(ctx, [&](Ctx &c) {
        TPM_REQUIRE(cosmo != nullptr, "null cosmology");
        TPM_REQUIRE(n == 0 || (pos && mom), "null host pointer");
        TPM_REQUIRE(steps >= 0, "steps must be >= 0");
        TPM_REQUIRE(a_start > 0 && a_end > 0, "scale factors must be positive");
        TPM_REQUIRE(n <= c.max_n, "more particles than the context capacity");
        if (n == 0) return TPM_OK;
        double *d_mom = nullptr;
        TPM_HIP(hipMalloc(&d_mom, 3 * n * sizeof(double)));
        try {
            TPM_HIP(hipMemcpyAsync(c.pos, pos, 3 * n * sizeof(double), hipMemcpyHostToDevice, c.stream));
            TPM_HIP(hipMemcpyAsync(d_mom, mom, 3 * n * sizeof(double), hipMemcpyHostToDevice, c.stream));
            tpm::evaluate(c, c.pos, n, a_start, false, c.acc, nullptr);
            double a = a_start;
            for (int i = 0; i < steps; ++i) {
                const double a1 = a_start * std::pow(a_end / a_start, (double)(i + 1) / (double)steps);
                const tpm::KdkFactors f = tpm::kdk_factors(*cosmo, a, a1);
                tpm::kick(c, d_mom, c.acc, n, f.k1);
                tpm::drift(c, c.pos, d_mom, n, f.d);
                tpm::evaluate(c, c.pos, n, a1, false, c.acc, nullptr);
                tpm::kick(c, d_mom, c.acc, n, f.k2);
                a = a1;
            }
            tpm::finish_checks(c);
            TPM_HIP(hipMemcpyAsync(pos, c.pos, 3 * n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
            TPM_HIP(hipMemcpyAsync(mom, d_mom, 3 * n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
            TPM_HIP(hipStreamSynchronize(c.stream));
        } catch (...) {
            (void)hipFree(d_mom);
            throw;
        }
        (void)hipFree(d_mom);
        return TPM_OK;
    });
}

int tpm_nfw_halos_dev(tpm_ctx *ctx, const tpm_nfw_spec *spec, size_t n, double *d_pos, void *stream) {
    return guarded(ctx, [&](Ctx &c) {
        TPM_REQUIRE(spec != nullptr && (n == 0 || d_pos != nullptr), "null argument");
        hipStream_t saved = c.stream;
        if (stream) c.stream = (hipStream_t)stream;
        try {
            tpm::nfw_halos(c, *spec, n, d_pos);
        } catch (...) {
            c.stream = saved;
            throw;
        }
        c.stream = saved;
        return TPM_OK;
    });
}

int tpm_nfw_halos(tpm_ctx *ctx, const tpm_nfw_spec *spec, size_t n, double *pos) {
    return guarded(ctx, [&](Ctx &c) {
        TPM_REQUIRE(spec != nullptr && (n == 0 || pos != nullptr), "null argument");
        tpm::nfw_halos(c, *spec, n, c.pos);
        if (n) TPM_HIP(hipMemcpy(pos, c.pos, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
        return TPM_OK;
    });
}

int tpm_zeldovich_dev(tpm_ctx *ctx, uint64_t seed, double k_c, double disp_rms, double mom_factor, double *d_pos,
                      double *d_mom, void *stream) {
    return guarded(ctx, [&](Ctx &c) {
        TPM_REQUIRE(d_pos != nullptr, "null positions pointer");
        hipStream_t saved = c.stream;
        if (stream) c.stream = (hipStream_t)stream;
        try {
            tpm::zeldovich(c, seed, k_c, disp_rms, mom_factor, d_pos, d_mom);
        } catch (...) {
            c.stream = saved;
            throw;
        }
        c.stream = saved;
        return TPM_OK;
    });
}

int tpm_zeldovich(tpm_ctx *ctx, uint64_t seed, double k_c, double disp_rms, double mom_factor, double *pos,
                  double *mom) {
    return guarded(ctx, [&](Ctx &c) {
        TPM_REQUIRE(pos != nullptr, "null positions pointer");
        TPM_REQUIRE(c.m <= c.max_n, "context capacity must hold mesh_n^3 particles");
        double *d_mom = nullptr;
        TPM_HIP(hipMalloc(&d_mom, 3 * c.m * sizeof(double)));
        try {
            tpm::zeldovich(c, seed, k_c, disp_rms, mom_factor, c.pos, d_mom);
            TPM_HIP(hipMemcpyAsync(pos, c.pos, 3 * c.m * sizeof(double), hipMemcpyDeviceToHost, c.stream));
            if (mom) TPM_HIP(hipMemcpyAsync(mom, d_mom, 3 * c.m * sizeof(double), hipMemcpyDeviceToHost, c.stream));
            TPM_HIP(hipStreamSynchronize(c.stream));
        } catch (...) {
            (void)hipFree(d_mom);
            throw;
        }
        (void)hipFree(d_mom);
        return TPM_OK;
    });
}

}  // extern "C"
