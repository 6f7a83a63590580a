// Internal declarations of libtpm: the context, device helpers shared by the kernel files,
// and the launchers each kernel file exports to the C ABI layer (tpm_api.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "tpm.h"

namespace tpm {

// Workgroups are dealt to the 8 XCDs round-robin; this renumbering gives each XCD a contiguous
// range of logical blocks, so neighbouring blocks (neighbouring particles in Morton order,
// neighbouring mesh columns) share its L2. It is a bijection of [0, g).
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t g) {
    const uint32_t q = g / 8, r = g % 8, x = b % 8;
    return x * q + min(x, r) + b / 8;
}

// Fixed-point scale of the deterministic mass grid: one unit of mass is 2^32 counts.
constexpr int kFixedShift = 32;
// Morton key resolution per axis (3*21 = 63-bit keys); also the deepest octree level.
constexpr int kKeyBits = 21;
// Octree leaf size and depth limit (src/tree.rs:12-13 uses 8 and 40; the key resolution
// bounds the depth here at 21 levels, below which particles share a leaf).
constexpr int kLeafSize = 8;
constexpr int kMaxDepth = kKeyBits;
// The group walk never opens a node of at most kWalkLeaf particles: it takes it as a monopole or
// lists its particles. Children of such nodes are built only for the per-particle walks.
constexpr uint32_t kWalkLeaf = 16;
// The deposit's Morton cubes are the largest aligned cubes of at most kDepSide mesh cells a side.
constexpr int kDepSide = 16;
// Particles per walk group: one wavefront.
constexpr int kGroup = 64;

struct Error {
    int code;
    std::string msg;
};

#define TPM_HIP(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            throw ::tpm::Error{TPM_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)}; \
    } while (0)

#define TPM_REQUIRE(cond, msg)                                                             \
    do {                                                                                   \
        if (!(cond)) throw ::tpm::Error{TPM_E_INVALID, (msg)};                             \
    } while (0)

// Plan of one line FFT length: Stockham stages and the forward twiddle table
// w[t] = exp(-2 pi i t / n), computed in double and stored in both precisions.
struct FftPlan {
    int n = 0;
    int nst = 0;
    int radix[24] = {0};
};

// Octree in structure-of-arrays form; nodes are stored level by level (breadth first) and
// the children of a node are contiguous. TreeView is the plain part passed to kernels.
struct TreeView {
    uint2 *range;    // [start, end) into the sorted particle order
    uint4 *link;     // x: first child, y: child count, z: child octant mask, w: level | walk-leaf<<8
    float4 *box;     // centre (FP32, the reference's arithmetic) and half side
    double4 *com;    // centre of mass (FP64) and particle count
};

struct Tree {
    size_t cap = 0;
    size_t count = 0;
    uint2 *range = nullptr;
    uint4 *link = nullptr;
    float4 *box = nullptr;
    double4 *com = nullptr;
    TreeView view() const { return TreeView{range, link, box, com}; }
};

// MLP layer sizes: 64 (33 features, zero padded) -> 128 -> 128 -> 128 -> 32 (4 used).
constexpr int kMlpIn[4] = {64, 128, 128, 128};
constexpr int kMlpOut[4] = {128, 128, 128, 32};

// Learned short-range correction state (mlp_kernels.hip).
struct Mlp {
    float *params = nullptr, *grads = nullptr, *m1 = nullptr, *m2 = nullptr;  // FP32 master + Adam
    uint16_t *wt16[4] = {nullptr, nullptr, nullptr, nullptr};  // BF16 Wt [out x in]
    uint16_t *w16[4] = {nullptr, nullptr, nullptr, nullptr};   // BF16 W [in x out]
    int step = 0;
    bool trained = false;
    double kappa = 5.196152422706632;  // error-estimate multiplier of the gate until calibrated (3 sqrt(3))
    std::vector<float> cal_est, cal_err;  // calibration pairs (sigma / |a|, error / |a|), sorted by estimate
    std::map<double, double> kappa_cache;  // tol -> multiplier
    size_t cap = 0;  // rows of the work buffers
    uint16_t *X = nullptr, *Xt = nullptr, *Y[3] = {nullptr, nullptr, nullptr}, *Yt[3] = {nullptr, nullptr, nullptr};
    float *Z[3] = {nullptr, nullptr, nullptr}, *C = nullptr, *Zout = nullptr, *dY = nullptr, *T = nullptr,
          *Wrow = nullptr, *loss_part = nullptr;
    uint16_t *dZ = nullptr, *dZt = nullptr;
    std::vector<void *> work;
};

struct Ctx {
    tpm_params p{};
    size_t max_n = 0;
    int n = 0;          // mesh cells per side
    int hp = 0;         // n/2 + 1, the half-spectrum length along z
    size_t m = 0;       // n^3
    size_t hs = 0;      // n * n * hp
    hipStream_t stream = nullptr;
    // Second stream (made on first use): the octree and walk groups of a large evaluation run there
    // beside the deposit, FFT and interpolation (both need only the sort), with a look-back slot of
    // their own.
    hipStream_t side = nullptr;
    unsigned long long *scan_part2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipEvent_t tree_ev[2] = {nullptr, nullptr};  // around the tree chain when timing
    bool timing = false;
    bool count_pairs = true;  // walk interactions counted (tpm_set_counting)
    bool overlap = true;      // octree beside the PM part on the second stream (tpm_set_overlap)
    std::string err;
    tpm_stats stats{};

    FftPlan plan;
    float2 *tw32 = nullptr;   // forward twiddles, FP32
    double2 *tw64 = nullptr;  // forward twiddles, FP64
    double *kvec = nullptr;   // wavenumber(i) for i < n (src/fft.rs:86-95), FP64

    // Particle buffers (capacity max_n).
    double *pos = nullptr;     // 3 max_n staging positions (caller order)
    float *acc = nullptr;      // 3 max_n accelerations (caller order)
    uint64_t *keys = nullptr, *keys_alt = nullptr;
    uint32_t *idx = nullptr, *idx_alt = nullptr;
    uint32_t *radix_table = nullptr;  // per-tile digit counts of a radix pass (256 per 4096 keys)
    size_t radix_table_words = 0;
    // Look-back slot of the single-pass scans and octree levels on the context's stream
    // (kLbSlotWords words, zeroed at creation).
    unsigned long long *scan_part = nullptr;
    double *spos = nullptr;    // 3 max_n positions in Morton order
    unsigned long long *counter = nullptr;  // interaction counter + scratch words
    int *flag = nullptr;       // non-finite flag
    uint32_t *status = nullptr;  // sticky TPM_STATUS_* bits of every evaluation since the last clear
    hipStream_t status_stream = nullptr;  // stream of the last evaluation that folded into `status`

    // Mesh buffers.
    unsigned long long *grid = nullptr;  // fixed-point mass grid, m words
    float2 *specA = nullptr, *specB = nullptr, *specC = nullptr;  // half spectra, hs each
    float *force = nullptr;    // (gx, gy, gz) per mesh point (3 m)
    float *delta = nullptr;    // density contrast grid (m)

    // FP64 judge buffers (allocated on first use).
    double2 *dspecA = nullptr, *dspecB = nullptr, *dspecC = nullptr;
    double *dforce = nullptr;
    uint32_t *inv = nullptr;      // inverse of the Morton permutation
    void *judge_buf = nullptr;    // judge targets and FP64 outputs
    size_t judge_cap = 0;

    Tree tree;
    Mlp mlp;
    float *acc_sr = nullptr;     // short-range force (caller order) while training
    uint32_t *members = nullptr; // walk members (sorted indices) for partial walks
    uint32_t *reject = nullptr;  // per sorted particle: 1 = learned estimate rejected
    float2 *stab = nullptr;     // S(u) table of the walk (kSTab + 1 entries)
    double stab_scale = 0.0;
    int num_cu = 0;
    unsigned walk_blocks = 0;
    hipEvent_t walk_ev[2] = {nullptr, nullptr};  // around the walk kernel when timing
    // Walk groups: particles [gstart[g], gstart[g+1]) of the Morton order.
    uint32_t *gstart = nullptr;     // max_n + 1
    uint32_t *gscratch = nullptr;   // 2 max_n words
    int group_level = 2;
    int group_level_shift = 0;  // autotuned offset of the group cube level
    int fft_cap = 16;           // autotuned cap on FFT lines per workgroup (LDS line kernels)
    int tree_depth = kMaxDepth; // octree depth limit (autotuned): deeper nodes stay leaves (see sort_depth)
    int small_direct = 0;       // autotuned: small nodes passing the opening criterion summed directly
    int fft_path = 0;           // 0: wavefront transforms where the mesh side has them (512), 1: LDS line kernels
    uint32_t *tree_scan = nullptr;  // per-level child counts, offsets, scan partials, level table
    uint32_t *tree_info = nullptr;  // [0] node count, [1] capacity exceeded
    // First sorted index of every level-`tree_tab_level` cube, from the last octree build over
    // `tree_tab_n` particles (0: none).
    const uint32_t *tree_tab = nullptr;
    int tree_tab_level = 0;
    size_t tree_tab_n = 0;
    uint32_t *ngroups_dev = nullptr;  // walk group count (device)
    uint32_t *walk_ctr = nullptr;     // walk group counters, one per XCD range, 1024 words apart (8192 words)
    uint32_t *walk_goff = nullptr;    // first global group of each box of a batched walk (kWalkBatch + 1)
    uint32_t *dep_start = nullptr;    // per deposit cube: first sorted index, item count, first item (3 (8^dep_level + 1))
    int dep_level = 0;                // Morton level of the deposit cubes (at most kDepSide cells a side)
};

// ---------------------------------------------------------------------------------------
// Launchers (all enqueue on ctx.stream).

// Morton keys of positions (FP64, caller order) and a stable radix sort of (key, index) over the
// key bits of sort_depth(c, n) levels; then spos = positions gathered into key order.
void sort_particles(Ctx &c, const double *d_pos, size_t n);
// Morton level of the walk groups' cubes for n particles (the tuned shift applied).
int walk_group_level(const Ctx &c, size_t n);
// Levels the sort orders and the octree splits: the tuned depth limit, raised to the finest cube
// level that is looked up in the sorted keys (the deposit cubes and the walk groups).
int sort_depth(const Ctx &c, size_t n);

// Fixed-point CIC deposit of positions into c.grid. f32_weights: compute the stencil in FP32
// from the positions rounded to float (the reference's arithmetic). sorted: positions are in
// Morton order (enables the LDS-tiled kernel; the grid is bit-identical either way).
void deposit_fixed(Ctx &c, size_t n, double mass, bool f32_weights, const double *pos_src, bool sorted);
// grid -> density contrast (FP32, the reference's c/mean - 1) and/or mass grid.
void grid_to_float(Ctx &c, size_t n_particles, float *delta_out, float *mass_out);

// PM force grids from the fixed-point grid (forward r2c fused with the density contrast,
// Green's function, gradient and inverse passes). Writes c.force as (gx, gy, gz) per mesh point.
void pm_from_grid(Ctx &c, size_t n_particles, double a, double r_split, float *delta_out);
// PM force grids from a density-contrast grid given on the device.
void pm_from_delta(Ctx &c, const float *d_delta, double a, double r_split);
// FP64 variant for the judge: writes c.dforce.
void pm_from_grid_f64(Ctx &c, size_t n_particles, double a, double r_split);

// CIC interpolation of the three force grids at sorted positions; writes acc[idx[i]].
void interp_force(Ctx &c, size_t n, bool f32_weights, float *d_acc);
// Interpolation of one grid at caller-order positions.
void interp_one(Ctx &c, const float *d_cells, const double *d_pos, size_t n, bool f32_weights, float *d_out);

// Octree over the sorted particles (levels, ranges, boxes, centres of mass).
void build_tree(Ctx &c, size_t n, bool group_walk_only = false);

// Walk groups: runs of at most 64 sorted particles that never cross a cube of octree level
// c.group_level (sets c.gstart, c.ngroups).
void build_groups(Ctx &c, size_t n, const uint64_t *keys = nullptr);
// Exclusive sum of n u32 values on c.stream (in == out allowed).
void exclusive_sum_u32(Ctx &c, const uint32_t *in, uint32_t *out, size_t n);
// Stable Morton sort of positions by the key digits of levels 1..depth (sort_kernels.hip): c.idx
// (sorted slot -> caller index), c.spos (positions in sorted order) and c.keys (their keys; for
// depth <= 10 with the levels below 10 cleared).
size_t radix_tiles(size_t n);
void morton_sort_positions(Ctx &c, const double *d_pos, size_t n, int depth);

// S(u) table of the group walk for the context's r_split and r_cut.
void make_short_range_table(Ctx &c);

// Arguments of the group walk (walk_kernels.hip).
struct WalkArgs {
    const double *spos;
    const uint32_t *idx;
    const uint32_t *gstart;
    const uint32_t *members;  // sorted indices of the walked particles (null: all, in order)
    size_t n;
    TreeView t;
    double box;
    double inv_box;
    double rcut;
    double theta2;
    int small_direct;  // nodes of at most kWalkLeaf particles are listed even where they could be accepted
    // The FP32 pair arithmetic works in coordinates scaled by c = sqrt(kSTab / s_max), where the S
    // table, uniform in s = r^2 + eps^2 on [0, s_max], has unit spacing: s' = c^2 s is the table
    // coordinate itself. All FP32 lengths below are scaled; forces come out unscaled because the
    // couplings carry c^2 (f' dx' = K c^2 S / (c^3 rho^3) c dx = K S dx / rho^3).
    double scale;  // c
    float boxf;
    float invboxf;
    float cut2;
    float cut2s;  // r_cut^2 + eps^2 (the pair cutoff on s2 = r^2 + eps^2)
    float eps2;
    float rcutf;  // r_cut, scaled
    float scalef;
    float inv2rs;
    float negk;  // -K c^2
    float tabk;  // -K c^2
    const float2 *stab;
    uint32_t *next_group;
    const uint32_t *ngroups;  // device word written by the group builder
    float *acc;
    unsigned long long *counter;
};

// One box of a batched walk: the per-box arrays of a prepared walk (the scalars are shared).
struct WalkBox {
    const double *spos;
    const uint32_t *idx;
    const uint32_t *gstart;
    const uint32_t *ngroups;
    TreeView t;
    float *acc;
    unsigned long long *counter;
};
// Boxes per batched launch: the table travels in the kernel arguments (at most 4 KiB).
constexpr int kWalkBatch = 40;

// The group walk up to its launch: walk groups, counters and the launch arguments (group walk
// only; the reference walk has no batched form).
WalkArgs walk_prepare(Ctx &c, size_t n, double k, float *d_acc);
// Launches a prepared walk of c's particles on c.stream.
void walk_launch(Ctx &c, const WalkArgs &w, size_t n);
// One launch (per kWalkBatch boxes) walking the prepared walks of boxes that share every scalar
// of their arguments (same split, softening, coupling, opening angle and small-node rule), on
// lead.stream; each box's forces are those of its own walk_launch.
void walk_launch_batch(Ctx &lead, const std::vector<WalkArgs> &boxes);

// Short-range force of every particle added into acc[idx[i]] (group walk or reference walk).
// With members != null only the `count` listed sorted particles are walked (all n are sources).
void walk_short_range(Ctx &c, size_t n, double k, float *d_acc, const uint32_t *members = nullptr,
                      size_t count = 0);
// FP64 reference walk for listed targets (caller indices); writes 3 doubles per target.
void walk_judge_f64(Ctx &c, size_t n, const uint32_t *d_targets_sorted, size_t m, double k, double theta,
                    double *d_out);

// FP64 interpolation of c.dforce at listed sorted indices.
void interp_judge_f64(Ctx &c, const uint32_t *d_targets_sorted, size_t m, double *d_out);

// Per-step outputs (mesh_kernels.hip): snapshot of FP64 state and the downsampled 1 + delta grid.
void snapshot(Ctx &c, const double *pos, const double *mom, size_t n, double a, double e, const float *delta,
              float *opos, float *ospeed, float *odens, float *orsd);
int downsample_side(int n, int *f_out);
void downsample_grid(Ctx &c, const float *delta, float *out);

// Force evaluation on device-resident FP64 positions (tpm_api.hip); no host synchronisation
// unless timing is on.
void evaluate(Ctx &c, const double *d_pos, size_t n, double a, bool f32w, float *d_acc, float *d_delta);
// evaluate() (FP64 positions, group walk) split around the walk's launch, for batched walks:
// everything before the launch, returning its arguments; and what follows the walk.
WalkArgs evaluate_pre_walk(Ctx &c, const double *d_pos, size_t n, double a, float *d_acc);
void evaluate_post_walk(Ctx &c, float *d_acc, size_t n);
// Enqueues the fold of this evaluation's failure flags (walk stack overflow, octree capacity when
// the tree was built, non-finite accelerations when checked) into the sticky c.status word.
void fold_status(Ctx &c, bool tree_built);
// FP64 TreePM judge (FP64 PM + per-particle FP64 walk at theta on the full-depth octree) of the
// caller-order targets d_caller (from judge_buffers) on device positions; 3 m doubles into d_out.
void judge_buffers(Ctx &c, size_t m, uint32_t **caller, double **out);
void judge_dev(Ctx &c, const double *d_pos, size_t n, const uint32_t *d_caller, size_t m, double a, double theta,
               double *d_out);
// Synchronises c.stream and throws on a walk stack overflow, octree capacity overflow or (with
// TPM_FLAG_CHECK_FINITE) a non-finite acceleration of the last evaluation.
void finish_checks(Ctx &c);

// Misc.
// Records the message tpm_last_error(NULL) returns (fixed buffer, no allocation).
void set_global_error(const char *msg);
// Zero `bytes` (a multiple of 4) at ptr with a fill kernel (graph-capture safe, unlike memset).
void zero_fill(Ctx &c, void *ptr, size_t bytes);
size_t tree_capacity(size_t n);
void convert_f32_to_f64(Ctx &c, const float *src, double *dst, size_t count);
void check_finite(Ctx &c, const float *d_acc, size_t count);

// Zel'dovich displacement grids from device white noise (n^3 floats): c.dforce.
// Spectrum of the Zel'dovich displacement field: the power law k^-2 exp(-k / k_c) (rescaled to a
// displacement RMS afterwards) or amplitude k^n_s T_BBKS(k / Gamma)^2 (src/ics.rs:20-33).
struct DisplacementSpec {
    bool bbks = false;
    double k_c = 0.0;
    double amplitude = 0.0, n_s = 0.0, gamma = 0.0;
};
void displacement_from_noise(Ctx &c, const float *d_noise, const DisplacementSpec &spec);

// PowerSpectrum with the BBKS transfer (src/ics.rs:20-64), FP64 on the host.
struct PowerSpec {
    double amplitude = 1.0, n_s = 1.0, gamma = 0.2;
    double power(double k) const;
    double sigma(double radius) const;
};
PowerSpec normalized_power(const tpm_cosmology &cosmo, double h0, double n_s, double sigma8, double radius);
// Lattice Zel'dovich ICs from the BBKS spectrum at a_init (src/ics.rs:111-182), mesh_n^3 particles.
void zeldovich_bbks(Ctx &c, const tpm_cosmology &cosmo, const PowerSpec &ps, uint64_t seed, double a_init,
                    double *d_pos, double *d_mom);

// Zel'dovich lattice ICs (ic_kernels.hip): positions and momenta (3 n^3 doubles each, lattice
// order) on the device.
void zeldovich(Ctx &c, uint64_t seed, double kc, double disp_rms, double mom_factor, double *d_pos, double *d_mom);

// KDK (kdk_kernels.hip): step factors in FP64 and the device kick / drift.
struct KdkFactors {
    double am, k1, d, k2;
};
KdkFactors kdk_factors(const tpm_cosmology &c, double a0, double a1);
// Cosmology::growth (D, normalised to D(1) = 1) and growth_rate f = dlnD/dlna, FP64.
void growth(const tpm_cosmology &c, double a, double *d, double *f);
// E(a) = H(a)/H0.
double hubble_e(const tpm_cosmology &c, double a);
void kick(Ctx &c, double *d_mom, const float *d_acc, size_t n, double f);
void drift(Ctx &c, double *d_pos, const double *d_mom, size_t n, double f);

// Clustered workload (ic_kernels.hip): halo particles from NFW profiles plus a uniform
// background, positions written to d_pos (3 n doubles).
void nfw_halos(Ctx &c, const tpm_nfw_spec &sp, size_t n, double *d_pos);
using NfwSpec = tpm_nfw_spec;

// MLP (mlp_kernels.hip).
size_t mlp_param_count();
void mlp_init(Ctx &c, uint64_t seed);
void mlp_refresh_bf16(Ctx &c);
void mlp_prepare(Ctx &c, size_t rows);
void mlp_forward(Ctx &c, size_t rows, bool train);
double mlp_forward_backward(Ctx &c, size_t rows, size_t count, bool want_loss);
void mlp_pack_rows(Ctx &c, const float *d_x, const float *d_t, size_t count, size_t rows);
size_t mlp_weight_offset(int l);
void mlp_adam_step(Ctx &c, float lr);
void mlp_features(Ctx &c, const uint32_t *d_sel, size_t first, size_t count, uint16_t *X, float inv_fscale);
void mlp_targets(Ctx &c, const uint32_t *d_sel, size_t count, const float *acc_sr, float inv_fscale, float *T);
void mlp_gather_batch(Ctx &c, const uint16_t *Xall, const float *Tall, const uint32_t *perm, size_t b0,
                      size_t count, size_t rows);
void mlp_apply(Ctx &c, size_t count, size_t first, float fscale, float tol, float *d_acc, uint32_t *reject);
void mlp_infer_fused(Ctx &c, const uint32_t *sel, size_t first, size_t count, float fscale, float tol, float kappa,
                     float *d_acc, uint32_t *reject, float *pred);
void mlp_predict_rows(Ctx &c, const uint16_t *rows, size_t count, float *zout);
void mlp_calibrate(Ctx &c, const uint32_t *d_sel, size_t count, float fscale, const float *acc_sr,
                   const float *acc_pm);
double mlp_kappa(Ctx &c, double tol);

// Plans and tables (fft_kernels.hip).
void make_plan(Ctx &c);

}  // namespace tpm
