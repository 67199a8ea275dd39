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

// ---- fractional fixed-point position records --------------------------------
// The global-list kernels read every atom as one 16-byte record (int4): its
// fractional coordinates s = x H^-1 (the row-vector convention of
// core/box.py:127-142) wrapped into [0, 1) and stored as round(s 2^32) mod
// 2^32, plus a 32-bit payload (the atom type; F'(rho) of the atom in the
// record the EAM density pass writes for the force pass). A neighbour then
// costs one 16-byte gather instead of a 32-byte float64 one.
//
// The wrapped 32-bit difference of two records is s_i - s_j - round(s_i - s_j)
// (the fractional minimum image of core/box.py:109-130) exactly, away from a
// tie at half a box vector. It is converted to float32 (one rounding, 2^-24
// relative) and mapped through H * 2^-32 held in float32 (one rounding of
// each entry, a common scale error below 6e-8 of every distance): the
// displacement comes within about one float32 ulp of the float64 value, at
// two VALU instructions per component of an orthogonal box instead of three
// in float64. Quantisation moves an atom by at most 2^-33 of each box vector
// (4e-8 A in a 340 A box), far below that rounding.
// H scaled by 2^-32, rounded to float32: maps a record difference to Cartesian.
struct QBox {
    float h[9];
};

__device__ __forceinline__ QBox load_qbox(const sim_replica_t& P) {
    QBox b;
#pragma unroll
    for (int k = 0; k < 9; ++k) b.h[k] = (float)(P.h[k] * SIM_Q_INV);
    return b;
}

// Box forms of the displacement: 0 general rows, 1 orthogonal, 2 lower
// triangular rows (a along x, b in the xy plane: h01 = h02 = h12 = 0, the
// LAMMPS form). Form 2 leaves out products with those zeros, which the
// general form adds exactly, so both give the same displacement.
template <int GEO>
__device__ __forceinline__ void qdisp(const int4& a, const int4& b, const QBox& B, float& rx,
                                      float& ry, float& rz) {
    const float s0 = (float)(int)((unsigned)a.x - (unsigned)b.x);
    const float s1 = (float)(int)((unsigned)a.y - (unsigned)b.y);
    const float s2 = (float)(int)((unsigned)a.z - (unsigned)b.z);
    if (GEO == 1) {
        rx = s0 * B.h[0];
        ry = s1 * B.h[4];
        rz = s2 * B.h[8];
    } else if (GEO == 2) {
        rx = __builtin_fmaf(s2, B.h[6], __builtin_fmaf(s1, B.h[3], s0 * B.h[0]));
        ry = __builtin_fmaf(s2, B.h[7], s1 * B.h[4]);
        rz = s2 * B.h[8];
    } else {
        rx = __builtin_fmaf(s2, B.h[6], __builtin_fmaf(s1, B.h[3], s0 * B.h[0]));
        ry = __builtin_fmaf(s2, B.h[7], __builtin_fmaf(s1, B.h[4], s0 * B.h[1]));
        rz = __builtin_fmaf(s2, B.h[8], __builtin_fmaf(s1, B.h[5], s0 * B.h[2]));
    }
}

// The same displacement formed in float64 and rounded once to float32 (three
// VALU instructions per component of an orthogonal box). The LJ kernels use
// it: their virial sums pair terms of both signs that cancel to a few parts in
// a hundred, which magnifies the float32 form's common scale error (the
// rounded H * 2^-32) about two hundred times; on a 1,000-atom argon cell the
// float32 form puts the virial 7.6e-6 (relative) from float64, this one 5e-8.
struct QBoxD {
    double h[9];
};

__device__ __forceinline__ QBoxD load_qbox_d(const sim_replica_t& P) {
    QBoxD b;
#pragma unroll
    for (int k = 0; k < 9; ++k) b.h[k] = P.h[k] * SIM_Q_INV;
    return b;
}

template <int GEO>
__device__ __forceinline__ void qdisp_d(const int4& a, const int4& b, const QBoxD& B, float& rx,
                                        float& ry, float& rz) {
    const double s0 = (double)(int)((unsigned)a.x - (unsigned)b.x);
    const double s1 = (double)(int)((unsigned)a.y - (unsigned)b.y);
    const double s2 = (double)(int)((unsigned)a.z - (unsigned)b.z);
    if (GEO == 1) {
        rx = (float)(s0 * B.h[0]);
        ry = (float)(s1 * B.h[4]);
        rz = (float)(s2 * B.h[8]);
    } else if (GEO == 2) {
        rx = (float)fma(s2, B.h[6], fma(s1, B.h[3], s0 * B.h[0]));
        ry = (float)fma(s2, B.h[7], s1 * B.h[4]);
        rz = (float)(s2 * B.h[8]);
    } else {
        rx = (float)fma(s2, B.h[6], fma(s1, B.h[3], s0 * B.h[0]));
        ry = (float)fma(s2, B.h[7], fma(s1, B.h[4], s0 * B.h[1]));
        rz = (float)fma(s2, B.h[8], fma(s1, B.h[5], s0 * B.h[2]));
    }
}

// Box form of a replica for qdisp (uniform per replica).
__device__ __forceinline__ int box_geo(const sim_replica_t& P) {
    return P.ortho ? 1 : (P.h[1] == 0.0 && P.h[2] == 0.0 && P.h[5] == 0.0 ? 2 : 0);
}

// block_sum of one 256-thread group g of a larger workgroup (same order, so
// the same value as block_sum); lds holds [groups][4 * NV]; every thread of
// the workgroup must call it.
template <int NV>
__device__ __forceinline__ void group_sum(double (&v)[NV], double* lds, int g, int lt) {
    const int lane = lt & 63, wid = lt >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) lds[(g * 4 + wid) * NV + k] = v[k];
    }
    __syncthreads();
    if (lt == 0) {
        const double* p = lds + g * 4 * NV;
#pragma unroll
        for (int k = 0; k < NV; ++k) v[k] = (p[k] + p[NV + k]) + (p[2 * NV + k] + p[3 * NV + k]);
    }
}

// An int the caller knows to be equal on every lane of the wave (such as a
// tile index derived from threadIdx.x / 256), moved to a scalar register: the
// tile record, the replica record and every field read through them then
// come by scalar loads instead of one vector load per lane.
__device__ __forceinline__ int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Wave-uniform extrema of an int over the 64 lanes (all lanes active): DPP
// within rows of 16, then the four row results (VALU only, as wave_sum).
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}

__device__ __forceinline__ int wave_max_int(int v) {
    v = max(v, dpp_i<0xB1>(v));
    v = max(v, dpp_i<0x4E>(v));
    v = max(v, dpp_i<0x141>(v));
    v = max(v, dpp_i<0x140>(v));
    return max(max(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
               max(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

__device__ __forceinline__ int wave_min_int(int v) {
    v = min(v, dpp_i<0xB1>(v));
    v = min(v, dpp_i<0x4E>(v));
    v = min(v, dpp_i<0x141>(v));
    v = min(v, dpp_i<0x140>(v));
    return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
               min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

// Raw buffer view of device memory (gfx9 descriptor, 32-bit data format):
// loads take a 32-bit per-lane byte offset from a base held in scalar
// registers, so no per-lane 64-bit address arithmetic is spent; bytes is the
// range checked by the hardware (< 2^32).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                             0x00020000);
}

// Entry k of atom i's column-major list, or i itself at or past the lane's
// end cnt. k and cmax (the longest list of the wave, <= max_nbr) are
// wave-uniform: the row is clamped to cmax - 1 by a scalar operation and its
// base goes into the buffer descriptor, the lane's offset i*4 is loop
// invariant; rows between the lane's end and cmax, stale, are discarded by
// the select.
__device__ __forceinline__ int nbr_row(const int* __restrict__ nbr, int ns, int i, int k, int cnt,
                                       int cmax) {
    const int kk = min(k, max(cmax - 1, 0));
    const int j = __builtin_amdgcn_raw_buffer_load_b32(
        buf_rsrc(nbr + (size_t)kk * ns, (unsigned)ns * 4u), i * 4, 0, 0);
    return k < cnt ? j : i;
}

// Record j of a 16-byte record array (j < 2^28: 32-bit byte offsets).
__device__ __forceinline__ int4 rec_at(const int4* __restrict__ p, int j) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(buf_rsrc(p, 0xFFFFFFFFu), j * 16, 0, 0);
    return make_int4((int)v[0], (int)v[1], (int)v[2], (int)v[3]);
}

// Square root of a squared pair distance with the hardware instruction
// (within 1 ulp) instead of the correctly rounded expansion, about ten
// instructions more per pair; r only selects a table interval and its
// fraction, and enters forces through them.
__device__ __forceinline__ float pair_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

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
// Batch of the full-list kernels (thread per atom): the last batch of a list
// is padded with masked self entries instead of a serial tail. Measured on
// gfx950, 16 in flight per thread beats a higher occupancy with fewer.
#ifndef SIM_LIST_BATCH
#define SIM_LIST_BATCH 16
#endif

// The integrator's second half fused into a force pass (INTEG; the EAM force
// pass and k_lj): with the force of atom i in registers the pass finishes the
// step for it (O + B, the arithmetic of k_second) and writes the kinetic
// columns of its tile's partial row, instead of a separate kernel re-reading
// forces and v_half. With pos set (EAM), the pass also does the next step's
// first half for the replicas first_fusable admits (force and new velocity
// in registers): v_half, position, record, and the displacement maximum.
struct SecondHalf {
    float* vh;
    float* vel;
    const float* noise;
    const long long* key;
    unsigned long long seed;
    const unsigned long long* step_dev;
    sim_types_t ty;
    double4* pos;
    const double4* pos_last;
    unsigned long long* maxd2;
    int4* posq;
};

