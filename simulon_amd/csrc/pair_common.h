// Shared pieces of the full-list pair kernels (LJ, EAM, MLP descriptors).
#pragma once
#include "sim_common.h"

// Box constants held in registers for the whole neighbor loop.
struct BoxRegs {
    double h[9];
    double hi[9];
};

__device__ __forceinline__ BoxRegs load_box(const sim_replica_t& P) {
    BoxRegs b;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        b.h[k] = P.h[k];
        b.hi[k] = P.hinv[k];
    }
    return b;
}

// float64 displacement xi - xj folded to the minimum image, rounded once to
// float32. ORTHO: multiplies by 1/L; otherwise fractional rounding with H^-1.
template <bool ORTHO>
__device__ __forceinline__ void displacement(const double4& xi, const double4& xj, const BoxRegs& B,
                                             float& rx, float& ry, float& rz) {
    double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
    if (ORTHO) {
        dx = fma(-B.h[0], rint(dx * B.hi[0]), dx);
        dy = fma(-B.h[4], rint(dy * B.hi[4]), dy);
        dz = fma(-B.h[8], rint(dz * B.hi[8]), dz);
    } else {
        const double s0 = rint(fma(dz, B.hi[6], fma(dy, B.hi[3], dx * B.hi[0])));
        const double s1 = rint(fma(dz, B.hi[7], fma(dy, B.hi[4], dx * B.hi[1])));
        const double s2 = rint(fma(dz, B.hi[8], fma(dy, B.hi[5], dx * B.hi[2])));
        dx = fma(-s2, B.h[6], fma(-s1, B.h[3], fma(-s0, B.h[0], dx)));
        dy = fma(-s2, B.h[7], fma(-s1, B.h[4], fma(-s0, B.h[1], dy)));
        dz = fma(-s2, B.h[8], fma(-s1, B.h[5], fma(-s0, B.h[2], dz)));
    }
    rx = (float)dx;
    ry = (float)dy;
    rz = (float)dz;
}

__device__ __forceinline__ float sq3(float x, float y, float z) {
    return __builtin_fmaf(z, z, __builtin_fmaf(y, y, x * x));
}

// Smallest float x with sqrtf(x) >= rc: r2 < thr  <=>  sqrtf(r2) < rc, so the
// squared test reproduces a `r < rc` mask exactly.
static inline float sqrt_threshold(float rc) {
    float x = rc * rc;
    while (sqrtf(x) < rc) x = nextafterf(x, 1e30f);
    while (true) {
        float y = nextafterf(x, 0.f);
        if (sqrtf(y) >= rc) x = y;
        else break;
    }
    return x;
}

// Neighbors processed per batch: all loads of a batch are issued before any
// of them is used, so their latencies overlap.
#define SIM_NB_BATCH 4
