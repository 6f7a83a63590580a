// Morton keys of positions: the octant digits of Tree::build (src/tree.rs:216-218) for every level.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "tpm_internal.h"

namespace tpm {

__device__ __forceinline__ uint64_t spread3(uint32_t v) {
    uint64_t x = v & 0x1fffffu;
    x = (x | x << 32) & 0x1f00000000ffffULL;
    x = (x | x << 16) & 0x1f0000ff0000ffULL;
    x = (x | x << 8) & 0x100f00f00f00f00fULL;
    x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
    x = (x | x << 2) & 0x1249249249249249ULL;
    return x;
}

// 63-bit Morton key of a position: floor(x / L * 2^21) per axis, the division first, so a position
// exactly on a dyadic plane of the box maps to the upper cell, like the reference's `p >= centre`
// octant test.
__device__ __forceinline__ uint64_t morton_key(const double *p, double box) {
    const double q0 = floor((p[0] / box) * (double)(1u << kKeyBits));
    const double q1 = floor((p[1] / box) * (double)(1u << kKeyBits));
    const double q2 = floor((p[2] / box) * (double)(1u << kKeyBits));
    const double qmax = (double)((1u << kKeyBits) - 1);
    const uint32_t a = (uint32_t)fmin(fmax(q0, 0.0), qmax);
    const uint32_t b = (uint32_t)fmin(fmax(q1, 0.0), qmax);
    const uint32_t c = (uint32_t)fmin(fmax(q2, 0.0), qmax);
    return spread3(a) | (spread3(b) << 1) | (spread3(c) << 2);
}

// The top 30 bits of the key (octree levels 1-10), for a sort of 32-bit keys; widened back by
// this shift they are the full key with the levels below 10 cleared.
constexpr int kHiShift = 3 * kKeyBits - 30;

}  // namespace tpm
