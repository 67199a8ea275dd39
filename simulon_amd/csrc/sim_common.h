// Device helpers shared by the MD-step kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/simulon_hip.h"

#define SIM_CHECK_LAUNCH()                                   \
    do {                                                     \
        if (hipGetLastError() != hipSuccess) return SIM_E_LAUNCH; \
    } while (0)

static inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Scale of the fractional fixed-point position records (sim_pos_quantize).
constexpr double SIM_Q_SCALE = 4294967296.0;                // 2^32
constexpr double SIM_Q_INV = 2.3283064365386962890625e-10;  // 2^-32

// gfx950 dispatches workgroup b of a launch to XCD b % 8, and each XCD has
// its own L2. Tiles are in cell order, so neighbouring tiles share most of
// their neighbour positions: give each XCD a contiguous run of tiles
// (q or q + 1 of them) instead of every eighth one. Bijective on [0, nb).
#define SIM_NXCD 8
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int q = nb / SIM_NXCD, r = nb % SIM_NXCD;
    const int x = b % SIM_NXCD, k = b / SIM_NXCD;
    return x * q + (x < r ? x : r) + k;
}

// Fractional fixed-point record of position x in replica P: s = x H^-1
// (fixed-order products) wrapped into [0, 1), stored as round(s 2^32) mod
// 2^32 per component; w = atom type.
__device__ __forceinline__ int4 quant_record(const sim_replica_t& P, const double4& x) {
    unsigned u[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        double s = fma(x.z, P.hinv[6 + m], fma(x.y, P.hinv[3 + m], x.x * P.hinv[m]));
        s -= floor(s);
        u[m] = (unsigned)(unsigned long long)rint(s * SIM_Q_SCALE);   // 2^32 wraps to 0
    }
    return make_int4((int)u[0], (int)u[1], (int)u[2], (int)x.w);
}

// Tile table entry: (replica, first global atom index).
struct SimTile {
    int rep;
    int start;
};

// Lane l' of v within its row of 16 lanes by a DPP lane pattern (ctrl), for
// both 32-bit halves of a double (VALU only: no LDS crossbar traffic).
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double readlane_d(double v, int lane) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sum over the 64 lanes (all active), the same value in every lane: pairs,
// quads, half rows and rows of 16 by DPP (each step adds two equal-order
// partial sums, so all lanes of a row agree bit for bit), then the four row
// sums in a fixed order. Deterministic; no LDS traffic.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_d<0xB1>(v);    // quad_perm [1, 0, 3, 2]
    v += dpp_d<0x4E>(v);    // quad_perm [2, 3, 0, 1]
    v += dpp_d<0x141>(v);   // row_half_mirror: the other quad of the half row
    v += dpp_d<0x140>(v);   // row_mirror: the other half row
    return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Sum of NV doubles per thread over a 256-thread workgroup; thread 0 of the
// workgroup receives the totals in v[]. Tree order is fixed (deterministic).
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* lds /* [4*NV] */) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) lds[wid * NV + k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k)
            v[k] = (lds[k] + lds[NV + k]) + (lds[2 * NV + k] + lds[3 * NV + k]);
    }
}

// Minimum image of a displacement with the reference's arithmetic
// (core/box.py:122-130), in T = float or double, contraction-free (the
// library is compiled with -ffp-contract=off).
template <typename T>
__device__ __forceinline__ void min_image_ref(T& dx, T& dy, T& dz, const sim_replica_t& R) {
    if (R.ortho) {
        const T lx = (T)R.h[0], ly = (T)R.h[4], lz = (T)R.h[8];
        dx = dx - lx * rint(dx / lx);
        dy = dy - ly * rint(dy / ly);
        dz = dz - lz * rint(dz / lz);
    } else {
        const double* hi = R.hinv;
        const double* h = R.h;
        T s0 = (dx * (T)hi[0] + dy * (T)hi[3]) + dz * (T)hi[6];
        T s1 = (dx * (T)hi[1] + dy * (T)hi[4]) + dz * (T)hi[7];
        T s2 = (dx * (T)hi[2] + dy * (T)hi[5]) + dz * (T)hi[8];
        s0 = s0 - rint(s0);
        s1 = s1 - rint(s1);
        s2 = s2 - rint(s2);
        dx = (s0 * (T)h[0] + s1 * (T)h[3]) + s2 * (T)h[6];
        dy = (s0 * (T)h[1] + s1 * (T)h[4]) + s2 * (T)h[7];
        dz = (s0 * (T)h[2] + s1 * (T)h[5]) + s2 * (T)h[8];
    }
}

// Minimum image for force evaluation in float64: multiplies by 1/L (and by
// H^-1) instead of dividing; any image choice that differs from the
// reference's can only happen at an exact half-box separation, far beyond
// the cutoff.
__device__ __forceinline__ void min_image_fast(double& dx, double& dy, double& dz,
                                               const sim_replica_t& R) {
    if (R.ortho) {
        dx = fma(-R.h[0], rint(dx * R.hinv[0]), dx);
        dy = fma(-R.h[4], rint(dy * R.hinv[4]), dy);
        dz = fma(-R.h[8], rint(dz * R.hinv[8]), dz);
    } else {
        const double* hi = R.hinv;
        const double* h = R.h;
        double s0 = fma(dz, hi[6], fma(dy, hi[3], dx * hi[0]));
        double s1 = fma(dz, hi[7], fma(dy, hi[4], dx * hi[1]));
        double s2 = fma(dz, hi[8], fma(dy, hi[5], dx * hi[2]));
        s0 = rint(s0);
        s1 = rint(s1);
        s2 = rint(s2);
        dx = fma(-s2, h[6], fma(-s1, h[3], fma(-s0, h[0], dx)));
        dy = fma(-s2, h[7], fma(-s1, h[4], fma(-s0, h[1], dy)));
        dz = fma(-s2, h[8], fma(-s1, h[5], fma(-s0, h[2], dz)));
    }
}

// ---- Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) --------------------
struct u32x4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t hi0 = __umulhi(M0, c.x), lo0 = M0 * c.x;
        const uint32_t hi1 = __umulhi(M1, c.z), lo1 = M1 * c.z;
        c = u32x4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// Uniform in (0, 1] from 32 random bits.
__device__ __forceinline__ float u01(uint32_t b) {
    return ((float)(b >> 8) + 1.0f) * (1.0f / 16777216.0f);
}

// Three N(0,1) draws for (atom, step) by Box-Muller on one Philox block.
// tag separates independent streams (atoms: the default; barostat: its own).
// The transcendentals are the hardware's: v_log_f32 (log2), v_sqrt_f32 and
// v_sin_f32 / v_cos_f32, which take the angle in turns, so sin(2 pi u) is one
// instruction on u (their errors, about one float32 ulp for the logarithm
// and square root and below 1e-6 absolute for sine and cosine, are far
// below what a sample of N(0,1) noise can show); the library's logf, sqrtf
// and sincospif took about 120 VALU instructions per atom.
__device__ __forceinline__ void gauss3(unsigned long long seed, unsigned long long step,
                                       uint32_t atom, float& g0, float& g1, float& g2,
                                       uint32_t tag = 0x5eedu) {
    u32x4 c{atom, (uint32_t)step, (uint32_t)(step >> 32), tag};
    u32x4 r = philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const float m2ln2 = -1.3862943611198906f;   // -2 ln 2: -2 ln u = -2 ln2 log2 u
    const float ra = __builtin_amdgcn_sqrtf(m2ln2 * __builtin_amdgcn_logf(u01(r.x)));
    const float rb = __builtin_amdgcn_sqrtf(m2ln2 * __builtin_amdgcn_logf(u01(r.z)));
    const float ta = u01(r.y), tb = u01(r.w);
    g0 = ra * __builtin_amdgcn_cosf(ta);
    g1 = ra * __builtin_amdgcn_sinf(ta);
    g2 = rb * __builtin_amdgcn_cosf(tb);
}

// Second half of the velocity-Verlet step for atom i (core/integrator/
// integrator.py:102-124): v = vh damp + xi rff (Langevin replicas), then
// v += (0.5 F/m) dt, all float32; writes vel and returns the kinetic terms
// (1/2 m v^2, m v_a v_b for xx yy zz xy xz yz) in float64. Shared by
// k_second and the EAM force pass that fuses it, so both give the same bits.
// noise: optional float[N][3] of N(0,1) draws (else Philox keyed by the
// replica's key, the atom's index within its replica and the step).
// Split in two: second_half_pre (the O step: it needs no force, so the fused
// force pass does it before its neighbour walk and keeps the result in
// registers) and second_half_post (the B step, velocity and kinetic terms).
__device__ __forceinline__ float3 second_half_pre(const sim_replica_t& P, int i, int ty_i,
                                                  const float* vh, const float* noise,
                                                  const long long* key, unsigned long long seed,
                                                  unsigned long long step) {
    const size_t o = (size_t)i * 3;
    float v0 = vh[o], v1 = vh[o + 1], v2 = vh[o + 2];
    if (P.langevin) {
        float g0, g1, g2;
        if (noise) {
            g0 = noise[o];
            g1 = noise[o + 1];
            g2 = noise[o + 2];
        } else {
            const uint32_t id = key ? (uint32_t)(key[i] - P.atom_base) : (uint32_t)i;
            gauss3(seed ^ P.noise_key, step, id, g0, g1, g2);
        }
        const float s = P.rff[ty_i];
        v0 = v0 * P.damp + g0 * s;
        v1 = v1 * P.damp + g1 * s;
        v2 = v2 * P.damp + g2 * s;
    }
    return make_float3(v0, v1, v2);
}

// m: the atom's mass. Returns the new velocity (the values written to vel).
__device__ __forceinline__ float3 second_half_post(const sim_replica_t& P, int i, float m,
                                                   float3 w, float f0, float f1, float f2,
                                                   float* vel, double (&k)[7]) {
    const float dt = P.dtf;
    const size_t o = (size_t)i * 3;
    const float v0 = w.x + (0.5f * (f0 / m)) * dt;
    const float v1 = w.y + (0.5f * (f1 / m)) * dt;
    const float v2 = w.z + (0.5f * (f2 / m)) * dt;
    vel[o] = v0;
    vel[o + 1] = v1;
    vel[o + 2] = v2;
    const double md = (double)m, a = v0, b = v1, c = v2;
    k[0] = 0.5 * md * ((a * a + b * b) + c * c);
    k[1] = md * a * a;
    k[2] = md * b * b;
    k[3] = md * c * c;
    k[4] = md * a * b;
    k[5] = md * a * c;
    k[6] = md * b * c;
    return make_float3(v0, v1, v2);
}

__device__ __forceinline__ void second_half_atom(const sim_replica_t& P, int i, int ty_i,
                                                 const sim_types_t& ty, const float* vh,
                                                 float f0, float f1, float f2, float* vel,
                                                 const float* noise, const long long* key,
                                                 unsigned long long seed,
                                                 unsigned long long step, double (&k)[7]) {
    const float3 w = second_half_pre(P, i, ty_i, vh, noise, key, seed, step);
    (void)second_half_post(P, i, ty.mass[ty_i], w, f0, f1, f2, vel, k);
}

// First half of the velocity-Verlet step (core/integrator/integrator.py:91-99)
// for one atom of mass m (float32 velocities with the reference's
// arithmetic, float64 positions): h = v + (0.5 F/m) dt, x += double(h) dt.
// Returns |x - l|^2, the displacement since the last list build. Shared by
// k_first and the EAM force pass that does the next step's first half, so
// both give the same bits.
__device__ __forceinline__ double first_half_atom(const sim_replica_t& P, float m, float3 v,
                                                  float f0, float f1, float f2, double4& x,
                                                  const double4& l, float3& h) {
    const float dt = P.dtf;
    h.x = v.x + (0.5f * (f0 / m)) * dt;
    h.y = v.y + (0.5f * (f1 / m)) * dt;
    h.z = v.z + (0.5f * (f2 / m)) * dt;
    x.x += (double)h.x * P.dt;
    x.y += (double)h.y * P.dt;
    x.z += (double)h.z * P.dt;
    const double ex = x.x - l.x, ey = x.y - l.y, ez = x.z - l.z;
    return (ex * ex + ey * ey) + ez * ez;
}

// Skin trigger of one replica (io_utils/reader.py:136-149): rebuild when the
// largest displacement since the last build, m (bits of a non-negative
// double), exceeds what the list's skin still covers. Pairs closer than r_l
// at the build stay inside the list while
//   2 max|dx| <= skin - max(0, 1 - mu_acc) r_l   (isotropic box scaling).
__device__ __forceinline__ bool skin_trip(const sim_replica_t& P, unsigned long long m,
                                          int force_all) {
    const double d2 = __longlong_as_double((long long)m);
    const double allow = 0.5 * (P.skin - fmax(0.0, 1.0 - P.mu_acc) * P.rl);
    const bool trip = allow <= 0.0 || d2 > (P.mu_acc == 1.0 ? P.trig2 : allow * allow);
    return force_all || trip;
}

// Replicas whose next first half the fused EAM force pass may do: no
// barostat and no loading step between this step's end and the next step's
// first half.
__device__ __forceinline__ bool first_fusable(const sim_replica_t& P) {
    return P.barostat == 0 && P.loading == 0;
}
