// Short-range TreePM force: the Barnes-Hut walk with the Gaussian-split kernel.
//
// Replaces Tree::accel_on, pair and short_range_factor (src/tree.rs:54-66, 253-310).
//
// Group walk (default). One wavefront owns 64 consecutive particles in Morton order. The
// wave walks the octree once for the whole group, against the group's bounding box:
//   - a node is skipped when even the nearest point of the box is beyond r_cut + sqrt(3) half
//     (the reference's per-particle cube test, src/tree.rs:274-278, made conservative);
//   - a walk-leaf, or a node of at most kWalkLeaf particles that would be opened (with the
//     small_direct knob, also one that would be accepted), contributes its particle range
//     (direct sums), if its cube comes within r_cut of the box;
//   - a split node is taken as a monopole only when 2 half < theta * d for the nearest point
//     of the box, so the reference's criterion (src/tree.rs:297) holds for every particle of
//     the group; otherwise its children are pushed.
// The lists are then evaluated in LDS tiles of 64 sources, each lane applying the reference's
// per-particle tests (pair cutoff r^2 <= r_cut^2, node cube test) and pair force. Where the
// reference would accept a node for one particle but the group opens it, the particle gets
// the node's children instead: the result is never less accurate than the reference's.
//
// Pair force (src/tree.rs:61-66): a = -K m S(r) d / (r^2 + eps^2)^{3/2},
// S(r) = erfc(u) + (2/sqrt(pi)) u exp(-u^2), u = r / (2 r_s). In the FP32 walk, particle pairs
// (r <= r_cut) interpolate S from an LDS table (pair_tab); for nodes, whose centre of mass may
// lie beyond r_cut, erfc(u) is the Abramowitz & Stegun 7.1.26 rational approximation (Hastings),
//   erfc(u) = t (a1 + t (a2 + t (a3 + t (a4 + t a5)))) exp(-u^2),  t = 1 / (1 + p u),
// with absolute error below 1.5e-7, which shares exp(-u^2) with the second term of S.
//
// Reference walk (TPM_FLAG_REF_WALK, and the FP64 judge): one thread per particle walking the
// same octree with the reference's LIFO order and per-particle tests, in T = float or double.
#include <algorithm>
#include <cmath>
#include <vector>

#include "tpm_internal.h"

namespace tpm {

constexpr int kSTab = 1024;  // intervals of the S table on [0, s_max]


constexpr float kErfP = 0.3275911f;
constexpr float kErfA1 = 0.254829592f;
constexpr float kErfA2 = -0.284496736f;
constexpr float kErfA3 = 1.421413741f;
constexpr float kErfA4 = -1.453152027f;
constexpr float kErfA5 = 1.061405429f;
constexpr float kTwoOverSqrtPi = 1.12837916709551257f;
constexpr float kLog2e = 1.44269504088896341f;

// S(r) for r = sqrt(r2), u = r * inv2rs. The hardware's single-instruction exp2 and
// reciprocal (about 1 ulp) are used: exp2's argument stays above -40 for u < 5.
__device__ __forceinline__ float short_range_factor_f32(float r, float inv2rs) {
    const float u = r * inv2rs;
    const float e = __builtin_amdgcn_exp2f(-(u * u) * kLog2e);
    const float t = __builtin_amdgcn_rcpf(fmaf(kErfP, u, 1.0f));
    const float poly = t * fmaf(t, fmaf(t, fmaf(t, fmaf(t, kErfA5, kErfA4), kErfA3), kErfA2), kErfA1);
    return e * fmaf(kTwoOverSqrtPi, u, poly);
}

// Pair force with single-instruction rsqrt/sqrt (about 1 ulp). s = r^2 + eps^2 > 0 whenever the
// self pair is excluded or eps > 0.
__device__ __forceinline__ void pair_f32(float dx, float dy, float dz, float r2, float mass, const WalkArgs &w,
                                         float &ax, float &ay, float &az) {
    const float s = r2 + w.eps2;
    const float rs = __builtin_amdgcn_rsqf(s);
    const float inv3 = rs * rs * rs;
    const float r = __builtin_amdgcn_sqrtf(r2);
    const float f = w.negk * mass * short_range_factor_f32(r, w.inv2rs) * inv3;
    ax = fmaf(dx, f, ax);
    ay = fmaf(dy, f, ay);
    az = fmaf(dz, f, az);
}

// Exclusive prefix sum over the 64 lanes (all lanes must be active) with DPP row shifts and row
// broadcasts, which move data between lanes without the LDS crossbar: Hillis-Steele within each
// row of 16 lanes (shifts 1, 2, 4, 8), then lane 15 of each row is added to the next row and
// lane 31 to the upper half. Lanes whose source is outside the row, or in a row the row mask
// leaves out, receive 0.
__device__ __forceinline__ int wave_excl_scan(int v, int &total) {
    int x = v;
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15 into rows 1, 3
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31 into rows 2, 3
    total = __builtin_amdgcn_readlane(x, 63);
    return x - v;
}

// Position of this lane among the lanes whose bit is set in m (m from a ballot), and their count.
__device__ __forceinline__ int mask_rank(uint64_t m, int &total) {
    total = __popcll(m);
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// One DPP move of a double (both halves) with row control CTRL and row mask ROWS; lanes with no
// source lane, or in rows the mask leaves out, receive `fill`.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_f64(double v, double fill) {
    const uint64_t b = __double_as_longlong(v), f = __double_as_longlong(fill);
    const int lo = __builtin_amdgcn_update_dpp((int)(uint32_t)f, (int)(uint32_t)b, CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(uint32_t)(f >> 32), (int)(uint32_t)(b >> 32), CTRL, ROWS, 0xf, false);
    return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

// Minimum and maximum over the 64 lanes, returned to every lane: the DPP steps of the prefix sum
// (row shifts 1, 2, 4, 8, row broadcasts 15 and 31) leave the reduction in lane 63.
__device__ __forceinline__ double wave_min(double v) {
    const double inf = __builtin_inf();
    v = fmin(v, dpp_f64<0x111, 0xf>(v, inf));
    v = fmin(v, dpp_f64<0x112, 0xf>(v, inf));
    v = fmin(v, dpp_f64<0x114, 0xf>(v, inf));
    v = fmin(v, dpp_f64<0x118, 0xf>(v, inf));
    v = fmin(v, dpp_f64<0x142, 0xa>(v, inf));
    v = fmin(v, dpp_f64<0x143, 0xc>(v, inf));
    const uint64_t b = __double_as_longlong(v);
    const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)b, 63), hi = __builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), 63);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

__device__ __forceinline__ double wave_max(double v) {
    const double ninf = -__builtin_inf();
    v = fmax(v, dpp_f64<0x111, 0xf>(v, ninf));
    v = fmax(v, dpp_f64<0x112, 0xf>(v, ninf));
    v = fmax(v, dpp_f64<0x114, 0xf>(v, ninf));
    v = fmax(v, dpp_f64<0x118, 0xf>(v, ninf));
    v = fmax(v, dpp_f64<0x142, 0xa>(v, ninf));
    v = fmax(v, dpp_f64<0x143, 0xc>(v, ninf));
    const uint64_t b = __double_as_longlong(v);
    const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)b, 63), hi = __builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), 63);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

__device__ __forceinline__ double min_image64(double d, double box, double inv_box) { return d - box * rint(d * inv_box); }

// Squared distance from point offset d (already minimum-imaged, relative to the box centre)
// to an axis-aligned box of half extents h.
__device__ __forceinline__ double box_dist2(double dx, double dy, double dz, double hx, double hy, double hz) {
    const double ex = fmax(fabs(dx) - hx, 0.0);
    const double ey = fmax(fabs(dy) - hy, 0.0);
    const double ez = fmax(fabs(dz) - hz, 0.0);
    return ex * ex + ey * ey + ez * ez;
}

// Wavefronts per workgroup; each walks its own groups. Twelve share one copy of the S table, so
// two workgroups (six wavefronts per SIMD) fit in the CU's LDS with the largest stacks.
constexpr int kWaves = 12;
// Pops are limited so the stack stays below kStackCap + 7 * depth, within kStackAlloc for any tree.
// The cap bounds how many nodes one traversal iteration may pop ((kStackCap - size) / 7), so a
// larger stack means fewer iterations per group (the traversal's instructions are a fifth of the
// walk's); 472 is what the LDS holds at two workgroups per CU.
constexpr int kStackCap = 472;
constexpr int kStackAlloc = kStackCap + 192;
constexpr int kRangeCap = 96;
constexpr int kNodeCap = 96;
constexpr int kStage = 128;  // staged sources (kept ones accumulate here, evaluated 64 at a time)
// Nodes of at most kWalkLeaf particles that the opening criterion would open are listed as particle
// ranges instead: direct sums over their particles rather than one more level of traversal, whose
// dependent loads cost more than the extra candidates.

// Per-wavefront LDS state. The node tiles and the source stage are never live at the same time.
struct WaveShared {
    uint32_t stack[kStackAlloc];
    uint32_t r_start[kRangeCap];
    uint32_t r_off[kRangeCap + 1];
    uint32_t nlist[kNodeCap];
    uint32_t mark[64];  // range starts within a batch of 64 candidates, tagged by batch
    union {
        float4 stage[kStage];
        struct {
            float4 tileA[64];
            float4 tileB[64];
        };
    };
};

struct BlockShared {
    float2 stab[kSTab + 1];  // S table: intercept and slope per interval
    WaveShared wave[kWaves];
};

// Orders a wavefront's LDS writes before its later reads (one wavefront executes in lockstep, so
// no hardware barrier is needed; this keeps the compiler from moving memory operations across).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Pair force with -K S interpolated from the LDS table, linear in s = r^2 + eps^2 over kSTab
// intervals of [0, s_max], s_max = (r_cut (1 + 1e-4))^2 + eps^2 (relative interpolation error of
// the force below 5e-5, largest for the closest pairs). Interval k holds the line through its end
// points as intercept a_k and slope b_k in the table coordinate s', so no fractional part is
// needed: the force factor (a_k + b_k s') / s'^{3/2} is (a_k rs^2 + b_k) rs with rs = 1/sqrt(s'),
// one multiply-add between two products (the terms are at most about 5 times their sum, so the
// rounding stays at the level of the FP32 coordinate). The table index
// does not depend on the reciprocal square root, which only gives the softened
// 1 / (r^2 + eps^2)^{3/2}, so the two chains overlap. Used for particle pairs, which satisfy
// r <= r_cut, so the interval index is at most kSTab - 1 and needs no clamp.
__device__ __forceinline__ void pair_tab(float dx, float dy, float dz, float s2, const float2 *stab,
                                         const WalkArgs &w, float &ax, float &ay, float &az) {
    const float x = s2;  // scaled: the table coordinate
    const float2 t = stab[(int)x];
    const float rs = __builtin_amdgcn_rsqf(s2);
    // (a + b s') / s'^{3/2} = (a rs^2 + b) rs, with rs^2 = 1 / s'.
    const float f = fmaf(t.x, rs * rs, t.y) * rs;
    ax = fmaf(dx, f, ax);
    ay = fmaf(dy, f, ay);
    az = fmaf(dz, f, az);
}

template <bool WRAP, bool SELF, bool CNT>
__device__ __forceinline__ void stage_one(const float4 src, const float2 *stab, const WalkArgs &w, uint32_t self,
                                          float xi, float yi, float zi, bool active, float &ax, float &ay,
                                          float &az, uint32_t &ninter) {
    float dx = xi - src.x;
    float dy = yi - src.y;
    float dz = zi - src.z;
    if (WRAP) {
        dx -= w.boxf * rintf(dx * w.invboxf);
        dy -= w.boxf * rintf(dy * w.invboxf);
        dz -= w.boxf * rintf(dz * w.invboxf);
    }
    // s2 = r^2 + eps^2; the cutoff r^2 <= r_cut^2 is tested as s2 <= r_cut^2 + eps^2.
    const float s2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, w.eps2)));
    // Inactive lanes hold NaN coordinates, so every comparison is false for them.
    bool take = s2 <= w.cut2s;
    if (SELF) take = take && __float_as_uint(src.w) != self;
    if (take) {
        pair_tab(dx, dy, dz, s2, stab, w, ax, ay, az);
        if (CNT) ++ninter;
    }
}

// Pairs of every active lane with the first `cnt` staged sources (broadcast LDS reads).
template <bool WRAP, bool SELF, bool CNT>
__device__ __forceinline__ void eval_stage(const WaveShared &sh, const float2 *stab, int cnt, const WalkArgs &w,
                                           uint32_t self, float xi, float yi, float zi, bool active, float &ax,
                                           float &ay, float &az, uint32_t &ninter) {
    if (cnt == 64) {
        // Four sources are read ahead into registers so their LDS latency overlaps.
        for (int c0 = 0; c0 < 64; c0 += 4) {
            float4 src[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) src[j] = sh.stage[c0 + j];
#pragma unroll
            for (int j = 0; j < 4; ++j)
                stage_one<WRAP, SELF, CNT>(src[j], stab, w, self, xi, yi, zi, active, ax, ay, az, ninter);
        }
    } else {
        for (int s = 0; s < cnt; ++s)
            stage_one<WRAP, SELF, CNT>(sh.stage[s], stab, w, self, xi, yi, zi, active, ax, ay, az, ninter);
    }
}

// Direct sums over the particles of the listed leaf ranges. Each source is first tested
// against the group's bounding box (half extents hx, hy, hz about the group centre): one
// farther than r_cut from the whole box cannot be within r_cut of any particle of the group,
// so it is dropped before the 64-lane evaluation. Kept sources are compacted into the stage
// and evaluated 64 at a time.
//
// Candidate e of the listed ranges is found without a search: a batch of 64 candidates
// [base, base + 64) lies in range rcur (the one holding candidate base) and the ranges that start
// inside the batch, at most 64 after rcur. Those mark their start position in sh.mark with the
// batch's tag; a ballot over the marks then gives, by a masked population count, the range of
// every lane's candidate.
template <bool WRAP, bool SELF, bool CNT>
__device__ __forceinline__ void eval_particles(WaveShared &sh, const float2 *stab, int nr, int tot,
                                               const WalkArgs &w, double gx, double gy, double gz, float hx,
                                               float hy, float hz, uint32_t self, float xi, float yi, float zi,
                                               bool active, float &ax, float &ay, float &az, uint32_t &ninter,
                                               uint32_t &kept, uint32_t &tag) {
    const int lane = threadIdx.x & 63;
    if (lane == 0) sh.r_off[nr] = (uint32_t)tot;
    wave_sync();
    const float cut2 = w.cut2 * (1.0f + 1e-5f);
    // |d| below this (float) bound cannot need a periodic shift: min_image(d) = d for |d| < L/2.
    const float no_wrap = 0.49f * w.boxf;
    const double sc = w.scale;
    int cnt = 0;
    int rcur = 0;
    for (int base = 0; base < tot; base += 64, ++tag) {
        const int r = rcur + 1 + lane;
        if (r < nr) {
            const int pos = (int)sh.r_off[r] - base;  // >= 0: range rcur holds candidate base
            if (pos < 64) sh.mark[pos] = tag;
        }
        wave_sync();
        const uint64_t starts = __builtin_amdgcn_ballot_w64(sh.mark[lane] == tag);
        const int mine = rcur + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(starts >> 32),
                                                               __builtin_amdgcn_mbcnt_lo((uint32_t)starts, 0u)) +
                         (int)((starts >> lane) & 1ull);
        rcur += __popcll(starts);
        const int e = base + lane;
        float d2 = __builtin_inff();
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (e < tot) {
            const uint32_t j = sh.r_start[mine] + (uint32_t)(e - (int)sh.r_off[mine]);
            double dx = w.spos[3 * (size_t)j] - gx;
            double dy = w.spos[3 * (size_t)j + 1] - gy;
            double dz = w.spos[3 * (size_t)j + 2] - gz;
            v = make_float4((float)(dx * sc), (float)(dy * sc), (float)(dz * sc), __uint_as_float(j));
            if (!WRAP) {
                const bool far = fmaxf(fabsf(v.x), fmaxf(fabsf(v.y), fabsf(v.z))) >= no_wrap;
                if (far) {
                    dx = min_image64(dx, w.box, w.inv_box);
                    dy = min_image64(dy, w.box, w.inv_box);
                    dz = min_image64(dz, w.box, w.inv_box);
                    v = make_float4((float)(dx * sc), (float)(dy * sc), (float)(dz * sc), v.w);
                }
                const float ex = fmaxf(fabsf(v.x) - hx, 0.f);
                const float ey = fmaxf(fabsf(v.y) - hy, 0.f);
                const float ez = fmaxf(fabsf(v.z) - hz, 0.f);
                d2 = fmaf(ex, ex, fmaf(ey, ey, ez * ez));
            } else {
                d2 = 0.f;
            }
        }
        const bool keep = d2 <= cut2;
        const uint64_t m = __builtin_amdgcn_ballot_w64(keep);
        const int at = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        if (keep) sh.stage[cnt + at] = v;
        cnt += __popcll(m);
        wave_sync();
        if (cnt >= 64) {
            eval_stage<WRAP, SELF, CNT>(sh, stab, 64, w, self, xi, yi, zi, active, ax, ay, az, ninter);
            kept += 64;
            const int rem = cnt - 64;
            float4 t = make_float4(0.f, 0.f, 0.f, 0.f);
            if (lane < rem) t = sh.stage[64 + lane];
            wave_sync();
            if (lane < rem) sh.stage[lane] = t;
            wave_sync();
            cnt = rem;
        }
    }
    if (cnt > 0) {
        eval_stage<WRAP, SELF, CNT>(sh, stab, cnt, w, self, xi, yi, zi, active, ax, ay, az, ninter);
        kept += cnt;
        wave_sync();
    }
}

template <bool WRAP>
__device__ __forceinline__ void eval_nodes(WaveShared &sh, const float2 *stab, int nn, const WalkArgs &w, double gx, double gy,
                                           double gz, float xi, float yi, float zi, bool active, float &ax,
                                           float &ay, float &az, uint32_t &ninter) {
    const int lane = threadIdx.x & 63;
    const float sqrt3 = 1.7320508075688772f;
    for (int base = 0; base < nn; base += 64) {
        const int e = base + lane;
        if (e < nn) {
            const uint32_t q = sh.nlist[e];
            const double4 cm = w.t.com[q];
            const float4 bx = w.t.box[q];
            double cx = cm.x - gx, cy = cm.y - gy, cz = cm.z - gz;
            double bxx = (double)bx.x - gx, byy = (double)bx.y - gy, bzz = (double)bx.z - gz;
            if (!WRAP) {
                cx = min_image64(cx, w.box, w.inv_box);
                cy = min_image64(cy, w.box, w.inv_box);
                cz = min_image64(cz, w.box, w.inv_box);
                bxx = min_image64(bxx, w.box, w.inv_box);
                byy = min_image64(byy, w.box, w.inv_box);
                bzz = min_image64(bzz, w.box, w.inv_box);
            }
            // src/tree.rs:275: max_r = r_cut + SQRT_3 * half, in FP32.
            const float max_r = w.rcutf + sqrt3 * (bx.w * w.scalef);
            const double sc = w.scale;
            sh.tileA[lane] = make_float4((float)(cx * sc), (float)(cy * sc), (float)(cz * sc), (float)cm.w);
            sh.tileB[lane] = make_float4((float)(bxx * sc), (float)(byy * sc), (float)(bzz * sc), max_r * max_r);
        }
        wave_sync();
        const int cnt = min(64, nn - base);
        for (int s = 0; s < cnt; ++s) {
            const float4 bb = sh.tileB[s];
            float ux = bb.x - xi, uy = bb.y - yi, uz = bb.z - zi;
            if (WRAP) {
                ux -= w.boxf * rintf(ux * w.invboxf);
                uy -= w.boxf * rintf(uy * w.invboxf);
                uz -= w.boxf * rintf(uz * w.invboxf);
            }
            const bool inside = fmaf(ux, ux, fmaf(uy, uy, uz * uz)) <= bb.w;  // false for NaN (inactive) lanes
            if (!inside) continue;
            ++ninter;
            const float4 src = sh.tileA[s];
            float dx = xi - src.x, dy = yi - src.y, dz = zi - src.z;
            if (WRAP) {
                dx -= w.boxf * rintf(dx * w.invboxf);
                dy -= w.boxf * rintf(dy * w.invboxf);
                dz -= w.boxf * rintf(dz * w.invboxf);
            }
            // Within r_cut the pair table serves, scaled by the node's mass; a node's centre of mass
            // may lie beyond r_cut, outside the table: there the analytic S(r).
            const float s2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, w.eps2)));
            if (s2 <= w.cut2s) {
                const float2 t = stab[(int)s2];
                const float rs = __builtin_amdgcn_rsqf(s2);
                const float f = fmaf(t.x, rs * rs, t.y) * rs * src.w;
                ax = fmaf(dx, f, ax);
                ay = fmaf(dy, f, ay);
                az = fmaf(dz, f, az);
            } else {
                const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                pair_f32(dx, dy, dz, r2, src.w, w, ax, ay, az);
            }
        }
        wave_sync();
    }
}

template <bool WRAP, bool SELF, bool CNT>
__device__ __forceinline__ void group_walk_body(WaveShared &sh, const float2 *stab, const WalkArgs &w, double gx, double gy,
                                double gz, double hx, double hy, double hz, uint32_t self, float xi, float yi,
                                float zi, bool active, float &ax, float &ay, float &az, uint32_t &ninter,
                                uint32_t &listed, uint32_t &kept, uint32_t &tag) {
    const float hxf = (float)(hx * w.scale) * (1.0f + 1e-6f), hyf = (float)(hy * w.scale) * (1.0f + 1e-6f),
                hzf = (float)(hz * w.scale) * (1.0f + 1e-6f);
    const int lane = threadIdx.x & 63;
    const double sqrt3 = 1.7320508075688772;
    int size = 1;
    if (lane == 0) sh.stack[0] = 0;
    int nr = 0, tot = 0, nn = 0;
    wave_sync();
    while (size > 0) {
        int room = (kStackCap - size) / 7;
        int k = min(64, size);
        if (k > room) k = room > 1 ? room : 1;
        uint32_t node = 0xFFFFFFFFu;
        if (lane < k) node = sh.stack[size - k + lane];
        wave_sync();
        size -= k;
        int kind = 0;  // 0 none, 1 leaf range, 2 accept, 3 open
        uint2 rng = make_uint2(0, 0);
        uint4 lk = make_uint4(0, 0, 0, 0);
        if (node != 0xFFFFFFFFu) {
            // All four records are requested at once (the centre of mass is needed only for split
            // nodes, but waiting for the link first would add a second round trip).
            lk = w.t.link[node];
            rng = w.t.range[node];
            const float4 bx = w.t.box[node];
            const double4 cm = w.t.com[node];
            double dx = (double)bx.x - gx, dy = (double)bx.y - gy, dz = (double)bx.z - gz;
            if (!WRAP) {
                dx = min_image64(dx, w.box, w.inv_box);
                dy = min_image64(dy, w.box, w.inv_box);
                dz = min_image64(dz, w.box, w.inv_box);
            }
            const double maxr = w.rcut + sqrt3 * (double)bx.w;
            const double d2 = box_dist2(dx, dy, dz, hx, hy, hz);
            if (WRAP || d2 <= maxr * maxr * (1.0 + 1e-6)) {
                bool leaf = (lk.w >> 8) != 0;
                if (!leaf) {
                    double cx = cm.x - gx, cy = cm.y - gy, cz = cm.z - gz;
                    if (!WRAP) {
                        cx = min_image64(cx, w.box, w.inv_box);
                        cy = min_image64(cy, w.box, w.inv_box);
                        cz = min_image64(cz, w.box, w.inv_box);
                    }
                    const double dc2 = WRAP ? 0.0 : box_dist2(cx, cy, cz, hx, hy, hz);
                    const double side = 2.0 * (double)bx.w;
                    const bool small = rng.y - rng.x <= kWalkLeaf;
                    if (small && w.small_direct)
                        leaf = true;
                    else if (side * side < w.theta2 * dc2 * (1.0 - 1e-6))
                        kind = 2;
                    else if (small)
                        leaf = true;  // a small node that would be opened
                    else
                        kind = 3;
                }
                if (leaf) {
                    // Listed only if the node's cube comes within r_cut of the group's box: no
                    // particle farther away can be within r_cut of a member.
                    const double hw = (double)bx.w;
                    const double dc = box_dist2(dx, dy, dz, hx + hw, hy + hw, hz + hw);
                    kind = (WRAP || dc <= w.rcut * w.rcut * (1.0 + 1e-6)) ? 1 : 0;
                }
            }
        }
        // Appends (wave-wide prefix sums).
        int t_open, t_leaf, t_leafcnt, t_acc;
        const int o_open = wave_excl_scan(kind == 3 ? (int)lk.y : 0, t_open);
        const int o_leaf = mask_rank(__builtin_amdgcn_ballot_w64(kind == 1), t_leaf);
        const int o_leafcnt = wave_excl_scan(kind == 1 ? (int)(rng.y - rng.x) : 0, t_leafcnt);
        const int o_acc = mask_rank(__builtin_amdgcn_ballot_w64(kind == 2), t_acc);
        if (kind == 3)
            for (uint32_t cix = 0; cix < lk.y; ++cix) {
                const int at = size + o_open + (int)cix;
                if (at < kStackAlloc) sh.stack[at] = lk.x + cix;
            }
        if (kind == 1) {
            sh.r_start[nr + o_leaf] = rng.x;
            sh.r_off[nr + o_leaf] = (uint32_t)(tot + o_leafcnt);
        }
        if (kind == 2) sh.nlist[nn + o_acc] = node;
        size += t_open;
        if (size > kStackAlloc) {
            // Cannot happen with the pop limit above (bounded by kStackCap + 7 * depth); recorded
            // as an error instead of writing past the stack.
            if (lane == 0 && w.counter) atomicAdd(w.counter + 1, 1ull);
            size = kStackAlloc;
        }
        nr += t_leaf;
        tot += t_leafcnt;
        nn += t_acc;
        wave_sync();
        if (nr > kRangeCap - 64) {
            eval_particles<WRAP, SELF, CNT>(sh, stab, nr, tot, w, gx, gy, gz, hxf, hyf, hzf, self, xi, yi, zi, active,
                                       ax, ay, az, ninter, kept, tag);
            listed += (uint32_t)tot;
            nr = 0;
            tot = 0;
        }
        if (nn > kNodeCap - 64) {
            eval_nodes<WRAP>(sh, stab, nn, w, gx, gy, gz, xi, yi, zi, active, ax, ay, az, ninter);
            nn = 0;
        }
    }
    if (nr > 0) {
        eval_particles<WRAP, SELF, CNT>(sh, stab, nr, tot, w, gx, gy, gz, hxf, hyf, hzf, self, xi, yi, zi, active, ax,
                                   ay, az, ninter, kept, tag);
        listed += (uint32_t)tot;
    }
    if (nn > 0) eval_nodes<WRAP>(sh, stab, nn, w, gx, gy, gz, xi, yi, zi, active, ax, ay, az, ninter);
}

// Statistics of one wavefront's groups, summed in scalar registers and added to the global
// counters once, when the wavefront exits (per-group atomics on three shared words serialise),
// and the wavefront's candidate-batch tag (WaveShared::mark).
struct WaveTotals {
    unsigned long long inter = 0, listed = 0, kept = 0;
    uint32_t tag = 0;
};

template <bool CNT>
__device__ __forceinline__ void walk_group(WaveShared &sh, const float2 *stab, const WalkArgs &w, uint32_t g, WaveTotals &tot) {
    const int lane = threadIdx.x & 63;
    const size_t g0 = w.gstart[g];
    const size_t g1 = w.gstart[g + 1];
    const bool active = g0 + lane < g1;
    // Sorted index of this lane's particle (through the member list for a partial walk).
    const size_t i = w.members ? (size_t)w.members[active ? g0 + lane : g0] : (active ? g0 + lane : g0);
    const size_t ii = i;
    const double px = w.spos[3 * ii], py = w.spos[3 * ii + 1], pz = w.spos[3 * ii + 2];
    const double mnx = wave_min(px), mny = wave_min(py), mnz = wave_min(pz);
    const double mxx = wave_max(px), mxy = wave_max(py), mxz = wave_max(pz);
    const double gx = 0.5 * (mnx + mxx), gy = 0.5 * (mny + mxy), gz = 0.5 * (mnz + mxz);
    // Half extents, widened by a relative 1e-12 of the box so the box surely contains the group.
    const double pad = 1e-12 * w.box;
    const double hx = 0.5 * (mxx - mnx) + pad, hy = 0.5 * (mxy - mny) + pad, hz = 0.5 * (mxz - mnz) + pad;
    // Inactive lanes get NaN coordinates: all their distance tests fail, so they take no pair.
    const float nan = __builtin_nanf("");
    const double sc = w.scale;
    const float xi = active ? (float)((px - gx) * sc) : nan, yi = active ? (float)((py - gy) * sc) : nan,
                zi = active ? (float)((pz - gz) * sc) : nan;
    float ax = 0.f, ay = 0.f, az = 0.f;
    uint32_t ninter = 0, listed = 0, kept = 0;
    // Sources are placed at their image nearest the group centre; that image is the one nearest
    // each particle whenever half extent + r_cut < L/2 on every axis. Otherwise every pair and
    // node distance is wrapped individually.
    const double lim = 0.5 * w.box * (1.0 - 1e-9);
    const bool wrap = (hx + w.rcut >= lim) || (hy + w.rcut >= lim) || (hz + w.rcut >= lim);
    // With softening > 0 a particle's pair with itself has d = 0 and a finite kernel, so it adds
    // exactly zero (as the reference's j == i skip does); the index test is needed only at eps = 0.
    const bool self_test = !(w.eps2 > 0.0f);
    const uint32_t self = (uint32_t)i;
    if (wrap) {
        if (self_test)
            group_walk_body<true, true, CNT>(sh, stab, w, gx, gy, gz, hx, hy, hz, self, xi, yi, zi, active, ax, ay, az,
                                        ninter, listed, kept, tot.tag);
        else
            group_walk_body<true, false, CNT>(sh, stab, w, gx, gy, gz, hx, hy, hz, self, xi, yi, zi, active, ax, ay, az,
                                         ninter, listed, kept, tot.tag);
    } else {
        if (self_test)
            group_walk_body<false, true, CNT>(sh, stab, w, gx, gy, gz, hx, hy, hz, self, xi, yi, zi, active, ax, ay,
                                         az, ninter, listed, kept, tot.tag);
        else
            group_walk_body<false, false, CNT>(sh, stab, w, gx, gy, gz, hx, hy, hz, self, xi, yi, zi, active, ax, ay,
                                          az, ninter, listed, kept, tot.tag);
    }
    if (active) {
        const size_t o = w.idx[i];
        w.acc[3 * o] += ax;
        w.acc[3 * o + 1] += ay;
        w.acc[3 * o + 2] += az;
    }
    int ntot;
    wave_excl_scan(active ? (int)ninter : 0, ntot);
    tot.inter += (uint32_t)ntot;
    tot.listed += listed;
    tot.kept += kept;
}

// Persistent workgroups of kWaves wavefronts. The groups (Morton order) are split into
// kSlots contiguous ranges, one per XCD: workgroups are dealt to the XCDs round-robin, so
// workgroup b starts on range b mod kSlots and its XCD's L2 keeps serving neighbouring groups,
// whose source regions overlap. A wavefront takes the next group of its range from that range's
// counter; when the range is exhausted it moves on to the next one, so every wavefront reaches
// the exit after visiting each range once and the load stays balanced.
constexpr int kSlots = 8;
// The ranges' counters sit 4 KiB apart, in lines that different memory channels serve, so their
// atomics proceed in parallel: every wavefront passes every counter once on its way out, which on
// a small box costs more than the walk when the counters share one line.
constexpr int kSlotStride = 1024;

// Six wavefronts per SIMD: the register budget is held to 80 VGPRs. CNT: count the interactions
// (one vector instruction per pair); without it the interaction total is not reported.
template <bool CNT>
__global__ void __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(6, 8))) k_group_walk(WalkArgs w) {
    __shared__ BlockShared bs;
    // The table holds -K S(u) for this evaluation's coupling K.
    for (int i = threadIdx.x; i <= kSTab; i += blockDim.x) {
        const float2 t = w.stab[i];
        bs.stab[i] = make_float2(w.tabk * t.x, w.tabk * t.y);
    }
    __syncthreads();
    WaveShared &sh = bs.wave[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    sh.mark[lane] = 0xFFFFFFFFu;  // never a tag
    wave_sync();
    const uint32_t ng = *w.ngroups;
    int slot = (int)(blockIdx.x % kSlots);
    WaveTotals tot;
    for (int visited = 0; visited < kSlots;) {
        const uint32_t r0 = (uint32_t)(((uint64_t)ng * slot) / kSlots);
        const uint32_t r1 = (uint32_t)(((uint64_t)ng * (slot + 1)) / kSlots);
        uint32_t g = 0;
        if (lane == 0) g = atomicAdd(w.next_group + slot * kSlotStride, 1u);
        g = __builtin_amdgcn_readfirstlane(g);
        if (g >= r1 - r0) {
            slot = (slot + 1) % kSlots;
            ++visited;
            continue;
        }
        walk_group<CNT>(sh, bs.stab, w, r0 + g, tot);
    }
    // Statistics; a wavefront that walked no group adds nothing (on a small box most of them).
    if (lane == 0 && w.counter) {
        if (CNT && tot.inter) atomicAdd(w.counter, tot.inter);
        if (tot.listed) atomicAdd(w.counter + 2, tot.listed);
        if (tot.kept) atomicAdd(w.counter + 3, tot.kept);
    }
}

// The boxes of a batched walk, passed by value in the kernel arguments; goff[b] is the first global
// group of box b (goff[nbox]: all groups), written by k_group_offsets before the walk.
struct WalkBoxArgs {
    WalkBox b[kWalkBatch];
    uint32_t nbox;
    const uint32_t *goff;
};

__global__ void k_group_offsets(WalkBoxArgs bx, uint32_t *__restrict__ goff) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint32_t s = 0;
    goff[0] = 0;
    for (uint32_t b = 0; b < bx.nbox; ++b) {
        s += *bx.b[b].ngroups;
        goff[b + 1] = s;
    }
}

// k_group_walk over the groups of several boxes: the global group list is the boxes' lists one
// after another, split into the same XCD ranges; a group's box is found by a binary search of
// goff and the box's arrays replace those of the shared arguments. Each group is walked exactly
// as k_group_walk walks it in its own box.
template <bool CNT>
__global__ void __launch_bounds__(64 * kWaves) __attribute__((amdgpu_waves_per_eu(6, 8)))
k_group_walk_batch(WalkArgs w, WalkBoxArgs bx) {
    __shared__ BlockShared bs;
    for (int i = threadIdx.x; i <= kSTab; i += blockDim.x) {
        const float2 t = w.stab[i];
        bs.stab[i] = make_float2(w.tabk * t.x, w.tabk * t.y);
    }
    __syncthreads();
    WaveShared &sh = bs.wave[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    sh.mark[lane] = 0xFFFFFFFFu;  // never a tag
    wave_sync();
    const uint32_t ng = bx.goff[bx.nbox];
    int slot = (int)(blockIdx.x % kSlots);
    WaveTotals tot;
    for (int visited = 0; visited < kSlots;) {
        const uint32_t r0 = (uint32_t)(((uint64_t)ng * slot) / kSlots);
        const uint32_t r1 = (uint32_t)(((uint64_t)ng * (slot + 1)) / kSlots);
        uint32_t g = 0;
        if (lane == 0) g = atomicAdd(w.next_group + slot * kSlotStride, 1u);
        g = __builtin_amdgcn_readfirstlane(g);
        if (g >= r1 - r0) {
            slot = (slot + 1) % kSlots;
            ++visited;
            continue;
        }
        const uint32_t gg = r0 + g;
        uint32_t lo = 0, hi = bx.nbox - 1;  // last box with goff <= gg
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (bx.goff[mid] <= gg)
                lo = mid;
            else
                hi = mid - 1;
        }
        WalkArgs wb = w;
        wb.spos = bx.b[lo].spos;
        wb.idx = bx.b[lo].idx;
        wb.gstart = bx.b[lo].gstart;
        wb.t = bx.b[lo].t;
        wb.acc = bx.b[lo].acc;
        wb.counter = bx.b[lo].counter;
        walk_group<CNT>(sh, bs.stab, wb, gg - bx.goff[lo], tot);
    }
    if (lane == 0 && w.counter) {
        if (CNT && tot.inter) atomicAdd(w.counter, tot.inter);
        if (tot.listed) atomicAdd(w.counter + 2, tot.listed);
        if (tot.kept) atomicAdd(w.counter + 3, tot.kept);
    }
}

// -------------------------------------------------------------------- reference walk

template <typename T>
struct RefWalkArgs {
    const double *spos;
    const uint32_t *idx;
    const uint32_t *targets;  // sorted indices to walk, or null = all particles in order
    size_t m;
    TreeView t;
    T box, r_split, rcut, eps, theta, k;
    T *out_f64;    // 3 per target (judge), or null
    float *acc;    // += into caller order, or null
    unsigned long long *counter;
};

template <typename T>
__device__ __forceinline__ T sr_factor_ref(T r, T r_split) {
    const T u = r / ((T)2 * r_split);
    return erfc(u) + (T)1.12837916709551257 * u * exp(-u * u);
}

template <typename T>
__device__ __forceinline__ void min_image_ref(T l, const T a[3], const T b[3], T d[3]) {
    for (int q = 0; q < 3; ++q) {
        const T x = a[q] - b[q];
        d[q] = x - l * round(x / l);
    }
}

template <typename T>
__device__ __forceinline__ void pair_ref(const T d[3], T mass, const RefWalkArgs<T> &w, T acc[3]) {
    const T r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const T s = r2 + w.eps * w.eps;
    const T f = -w.k * mass * sr_factor_ref<T>(sqrt(r2), w.r_split) / (s * sqrt(s));
    for (int q = 0; q < 3; ++q) acc[q] += d[q] * f;
}

template <typename T>
__global__ void __launch_bounds__(128) k_ref_walk(RefWalkArgs<T> w) {
    const size_t tix = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tix >= w.m) return;
    const size_t i = w.targets ? w.targets[tix] : tix;
    const T xi[3] = {(T)w.spos[3 * i], (T)w.spos[3 * i + 1], (T)w.spos[3 * i + 2]};
    const T cut2 = w.rcut * w.rcut;
    const T sqrt3 = (T)1.7320508075688772;
    T acc[3] = {0, 0, 0};
    uint32_t stack[8 * (kMaxDepth + 2)];
    int sp = 0;
    stack[sp++] = 0;
    unsigned long long cnt = 0;
    while (sp > 0) {
        const uint32_t nd = stack[--sp];
        const float4 bx = w.t.box[nd];
        const uint4 lk = w.t.link[nd];
        const T ctr[3] = {(T)bx.x, (T)bx.y, (T)bx.z};
        T tc[3];
        min_image_ref<T>(w.box, ctr, xi, tc);
        const T max_r = w.rcut + sqrt3 * (T)bx.w;
        if (tc[0] * tc[0] + tc[1] * tc[1] + tc[2] * tc[2] > max_r * max_r) continue;
        if (lk.w >> 8) {
            const uint2 r = w.t.range[nd];
            for (uint32_t j = r.x; j < r.y; ++j) {
                if (j == i) continue;
                const T xj[3] = {(T)w.spos[3 * (size_t)j], (T)w.spos[3 * (size_t)j + 1], (T)w.spos[3 * (size_t)j + 2]};
                T d[3];
                min_image_ref<T>(w.box, xi, xj, d);
                if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] > cut2) continue;
                pair_ref<T>(d, (T)1, w, acc);
                ++cnt;
            }
        } else {
            const double4 cm = w.t.com[nd];
            const T com[3] = {(T)cm.x, (T)cm.y, (T)cm.z};
            T tcm[3];
            min_image_ref<T>(w.box, com, xi, tcm);
            const T r = sqrt(tcm[0] * tcm[0] + tcm[1] * tcm[1] + tcm[2] * tcm[2]);
            if ((T)2 * (T)bx.w < w.theta * r) {
                const T d[3] = {-tcm[0], -tcm[1], -tcm[2]};
                pair_ref<T>(d, (T)cm.w, w, acc);
                ++cnt;
            } else {
                for (uint32_t cix = 0; cix < lk.y; ++cix) stack[sp++] = lk.x + cix;
            }
        }
    }
    if (w.out_f64) {
        w.out_f64[3 * tix] = acc[0];
        w.out_f64[3 * tix + 1] = acc[1];
        w.out_f64[3 * tix + 2] = acc[2];
    }
    if (w.acc) {
        const size_t o = w.idx[i];
        w.acc[3 * o] += (float)acc[0];
        w.acc[3 * o + 1] += (float)acc[1];
        w.acc[3 * o + 2] += (float)acc[2];
    }
    if (w.counter) atomicAdd(w.counter, cnt);
}

static inline unsigned blocks(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

__global__ void k_gather_keys(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ members, size_t count,
                              uint64_t *__restrict__ out) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) out[t] = keys[members[t]];
}

WalkArgs walk_prepare(Ctx &c, size_t n, double k, float *d_acc, const uint32_t *members, size_t count);

void walk_short_range(Ctx &c, size_t n, double k, float *d_acc, const uint32_t *members, size_t count) {
    if (n == 0) return;
    if (members && count == 0) return;
    if (c.p.flags & TPM_FLAG_REF_WALK) {
        RefWalkArgs<float> w{};
        w.spos = c.spos;
        w.idx = c.idx;
        w.targets = members;  // the per-particle walk of each listed particle (or of all)
        w.m = members ? count : n;
        w.t = c.tree.view();
        w.box = (float)c.p.box_size;
        w.r_split = (float)c.p.r_split;
        w.rcut = (float)c.p.r_cut;
        w.eps = (float)c.p.softening;
        w.theta = (float)c.p.theta;
        w.k = (float)k;
        w.out_f64 = nullptr;
        w.acc = d_acc;
        w.counter = c.counter;
        hipLaunchKernelGGL(k_ref_walk<float>, dim3(blocks(w.m, 128)), dim3(128), 0, c.stream, w);
        TPM_HIP(hipGetLastError());
        return;
    }
    const WalkArgs w = walk_prepare(c, n, k, d_acc, members, count);
    walk_launch(c, w, members ? count : n);
}

WalkArgs walk_prepare(Ctx &c, size_t n, double k, float *d_acc, const uint32_t *members, size_t count) {
    WalkArgs w{};
    w.spos = c.spos;
    w.idx = c.idx;
    w.n = n;
    w.t = c.tree.view();
    w.box = c.p.box_size;
    w.inv_box = 1.0 / c.p.box_size;
    w.rcut = c.p.r_cut;
    w.theta2 = c.p.theta * c.p.theta;
    w.small_direct = c.small_direct;
    const double c2 = c.stab_scale > 0 ? c.stab_scale : 1.0;  // scale^2
    const double sc = std::sqrt(c2);
    w.scale = sc;
    w.scalef = (float)sc;
    w.boxf = (float)(c.p.box_size * sc);
    w.invboxf = (float)(1.0 / (c.p.box_size * sc));
    w.cut2 = (float)(c.p.r_cut * c.p.r_cut * c2);
    w.eps2 = (float)(c.p.softening * c.p.softening * c2);
    w.cut2s = (float)((c.p.r_cut * c.p.r_cut + c.p.softening * c.p.softening) * c2);
    w.rcutf = (float)(c.p.r_cut * sc);
    w.inv2rs = (float)(1.0 / (2.0 * c.p.r_split * sc));
    w.negk = (float)(-k * c2);
    w.tabk = (float)(-k * c2);
    w.stab = c.stab;
    w.acc = d_acc;
    w.counter = c.counter;
    if (members) {
        // Keys of the members, in member order, for the group segmentation.
        hipLaunchKernelGGL(k_gather_keys, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, c.stream, c.keys,
                           members, count, c.keys_alt);
        TPM_HIP(hipGetLastError());
        build_groups(c, count, c.keys_alt);
    } else {
        build_groups(c, n);
    }
    w.gstart = c.gstart;
    w.members = members;
    w.ngroups = c.ngroups_dev;
    w.next_group = c.walk_ctr;
    zero_fill(c, w.next_group, kSlots * kSlotStride * sizeof(uint32_t));
    return w;
}

WalkArgs walk_prepare(Ctx &c, size_t n, double k, float *d_acc) { return walk_prepare(c, n, k, d_acc, nullptr, 0); }

static unsigned walk_grid(Ctx &c, size_t walked) {
    if (c.walk_blocks == 0) {
        int per_cu = 0;
        TPM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_group_walk<true>, 64 * kWaves, 0));
        c.walk_blocks = (unsigned)std::max(1, per_cu) * (unsigned)c.num_cu;
    }
    // Persistent grid; workgroups without a group left exit at once. No more workgroups than one
    // wavefront per 32 particles (groups hold 64 at most and about 50 at mean density; with fewer
    // wavefronts than groups each wavefront takes several): idle wavefronts cost their exit
    // protocol, which on a small box outweighs the walk.
    const size_t bound = (walked + kWaves * 32 - 1) / (kWaves * 32);
    return (unsigned)std::max<size_t>(1, std::min<size_t>(c.walk_blocks, bound));
}

void walk_launch(Ctx &c, const WalkArgs &w, size_t walked) {
    const unsigned nblk = walk_grid(c, walked);
    if (c.timing) {
        if (!c.walk_ev[0]) TPM_HIP(hipEventCreate(&c.walk_ev[0]));
        if (!c.walk_ev[1]) TPM_HIP(hipEventCreate(&c.walk_ev[1]));
        TPM_HIP(hipEventRecord(c.walk_ev[0], c.stream));
    }
    if (c.count_pairs)
        hipLaunchKernelGGL(k_group_walk<true>, dim3(nblk), dim3(64 * kWaves), 0, c.stream, w);
    else
        hipLaunchKernelGGL(k_group_walk<false>, dim3(nblk), dim3(64 * kWaves), 0, c.stream, w);
    TPM_HIP(hipGetLastError());
    if (c.timing) TPM_HIP(hipEventRecord(c.walk_ev[1], c.stream));
}

void walk_launch_batch(Ctx &lead, const std::vector<WalkArgs> &boxes) {
    static_assert(sizeof(WalkArgs) + sizeof(WalkBoxArgs) <= 4096, "walk batch arguments exceed 4 KiB");
    TPM_REQUIRE(!boxes.empty(), "empty walk batch");
    for (const WalkArgs &a : boxes) {
        const WalkArgs &l = boxes[0];
        TPM_REQUIRE(a.box == l.box && a.rcut == l.rcut && a.theta2 == l.theta2 && a.small_direct == l.small_direct &&
                        a.scale == l.scale && a.eps2 == l.eps2 && a.cut2s == l.cut2s && a.inv2rs == l.inv2rs &&
                        a.negk == l.negk && a.tabk == l.tabk,
                    "the boxes of a batched walk must share every scalar of the walk");
    }
    for (size_t b0 = 0; b0 < boxes.size(); b0 += kWalkBatch) {
        const size_t nb = std::min<size_t>(kWalkBatch, boxes.size() - b0);
        WalkBoxArgs bx{};
        bx.nbox = (uint32_t)nb;
        bx.goff = lead.walk_goff;
        size_t walked = 0;
        for (size_t q = 0; q < nb; ++q) {
            const WalkArgs &a = boxes[b0 + q];
            TPM_REQUIRE(a.members == nullptr, "a batched walk takes whole boxes");
            bx.b[q] = WalkBox{a.spos, a.idx, a.gstart, a.ngroups, a.t, a.acc, a.counter};
            walked += a.n;
        }
        WalkArgs w = boxes[b0];  // the shared scalars, the S table and the lead's counters
        w.next_group = lead.walk_ctr;
        w.counter = lead.counter;
        if (b0 > 0) zero_fill(lead, lead.walk_ctr, kSlots * kSlotStride * sizeof(uint32_t));
        hipLaunchKernelGGL(k_group_offsets, dim3(1), dim3(64), 0, lead.stream, bx, lead.walk_goff);
        TPM_HIP(hipGetLastError());
        const unsigned nblk = walk_grid(lead, walked);
        if (lead.count_pairs)
            hipLaunchKernelGGL(k_group_walk_batch<true>, dim3(nblk), dim3(64 * kWaves), 0, lead.stream, w, bx);
        else
            hipLaunchKernelGGL(k_group_walk_batch<false>, dim3(nblk), dim3(64 * kWaves), 0, lead.stream, w, bx);
        TPM_HIP(hipGetLastError());
    }
}

// S(u) = erfc(u) + (2/sqrt(pi)) u exp(-u^2) (src/tree.rs:54-57), u = r / (2 r_s), tabulated in FP64
// against s = r^2 + eps^2 on [0, s_max], s_max = (r_cut (1 + 1e-4))^2 + eps^2: kSTab + 1 points,
// the line through each interval's end points as intercept and slope in the table coordinate
// (entries below s = eps^2, never read, are flat at S(0)).
void make_short_range_table(Ctx &c) {
    if (!(c.p.r_split > 0) || !(c.p.r_cut > 0)) {
        c.stab_scale = 0.0;
        return;
    }
    const double eps2 = c.p.softening * c.p.softening;
    const double rc = c.p.r_cut * (1.0 + 1e-4);
    const double s_max = rc * rc + eps2;
    const double ds = s_max / kSTab;
    auto S = [&](double s) {
        const double u = std::sqrt(std::max(s - eps2, 0.0)) / (2.0 * c.p.r_split);
        return std::erfc(u) + 1.1283791670955126 * u * std::exp(-u * u);
    };
    std::vector<float2> t(kSTab + 1);
    for (int i = 0; i <= kSTab; ++i) {
        // Interval i spans [i, i + 1) in the table coordinate: S ~ (v0 - i (v1 - v0)) + s' (v1 - v0).
        const double v0 = S(i * ds), v1 = S((i + 1) * ds);
        t[i] = make_float2((float)(v0 - i * (v1 - v0)), (float)(v1 - v0));
    }
    TPM_HIP(hipMemcpy(c.stab, t.data(), t.size() * sizeof(float2), hipMemcpyHostToDevice));
    c.stab_scale = 1.0 / ds;
}

void walk_judge_f64(Ctx &c, size_t n, const uint32_t *d_targets_sorted, size_t m, double k, double theta,
                    double *d_out) {
    (void)n;
    if (m == 0) return;
    RefWalkArgs<double> w{};
    w.spos = c.spos;
    w.idx = c.idx;
    w.targets = d_targets_sorted;
    w.m = m;
    w.t = c.tree.view();
    w.box = c.p.box_size;
    w.r_split = c.p.r_split;
    w.rcut = c.p.r_cut;
    w.eps = c.p.softening;
    w.theta = theta;
    w.k = k;
    w.out_f64 = d_out;
    w.acc = nullptr;
    w.counter = nullptr;
    hipLaunchKernelGGL(k_ref_walk<double>, dim3(blocks(m, 128)), dim3(128), 0, c.stream, w);
    TPM_HIP(hipGetLastError());
}

}  // namespace tpm
