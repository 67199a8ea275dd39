// Fused velocity-Verlet / Langevin / Berendsen update, skin trigger and the
// float64 per-replica reductions (gfx950).
//
// core/integrator/integrator.py:91-124 splits a step into B+A (first half)
// and O+B (second half); each half is one kernel here. Velocities stay float32
// with the reference's arithmetic (a = F/m, v + (0.5 a) dt); positions are
// float64 and advanced as x += double(v_half) dt. The first half also measures
// the displacement since the last list build (io_utils/reader.py:136-149) and
// folds it into a per-replica maximum; the Langevin kick uses Philox4x32-10
// with Box-Muller: key = seed XOR the replica's noise_key, counter = (atom
// index within the replica, step), so replicas with different keys draw
// independent kicks.
#include "sim_common.h"

namespace {

__device__ __forceinline__ SimTile tile_of(const SimTile* tiles) { return tiles[blockIdx.x]; }

__global__ void __launch_bounds__(SIM_TILE)
    k_first(const sim_replica_t* __restrict__ reps, const SimTile* __restrict__ tiles,
            sim_types_t ty, double4* __restrict__ pos, const float* __restrict__ vel,
            const float* __restrict__ force, float* __restrict__ vh_out,
            double4* __restrict__ pos_last, unsigned long long* __restrict__ maxd2,
            int4* __restrict__ posq, int flags, float* __restrict__ vel_w) {
    __shared__ double lds[4];
    const SimTile t = tile_of(tiles);
    const sim_replica_t& P = reps[t.rep];
    // the previous step's EAM force pass did this replica's first half
    if ((flags & SIM_FIRST_UNFUSED_ONLY) && first_fusable(P)) return;   // uniform over the tile
    // the previous step's Berendsen update left its scaling of x, the build
    // positions and v to this launch (k_baro_scale's arithmetic, then B + A)
    const bool scale = (flags & SIM_FIRST_BERENDSEN_SCALE) && P.barostat == 1;
    const int i = t.start + threadIdx.x;
    double d2 = 0.0;
    if (i < P.atom_base + P.natoms) {
        double4 x = pos[i];
        double4 l = pos_last[i];
        const size_t o = (size_t)i * 3;
        float3 v = make_float3(vel[o], vel[o + 1], vel[o + 2]);
        if (scale) {
            const double mu = P.mu;
            const float muf = (float)mu;
            x.x *= mu;
            x.y *= mu;
            x.z *= mu;
            l.x *= mu;
            l.y *= mu;
            l.z *= mu;
            pos_last[i] = l;
            v.x *= muf;
            v.y *= muf;
            v.z *= muf;
            vel_w[o] = v.x;
            vel_w[o + 1] = v.y;
            vel_w[o + 2] = v.z;
        }
        const float m = ty.mass[(int)x.w];
        float3 h;
        d2 = first_half_atom(P, m, v, force[o], force[o + 1], force[o + 2], x, l, h);
        vh_out[o] = h.x;
        vh_out[o + 1] = h.y;
        vh_out[o + 2] = h.z;
        pos[i] = x;
        if (posq) posq[i] = quant_record(P, x);   // records for this step's forces
    }
    d2 = wave_max(d2);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = d2;
    __syncthreads();
    if (threadIdx.x == 0) {
        const double m = fmax(fmax(lds[0], lds[1]), fmax(lds[2], lds[3]));
        // non-negative doubles order like their bit patterns; the plain read
        // first skips the atomic (contended when one replica has thousands of
        // tiles) whenever it could not raise the maximum (a stale read only
        // costs an extra atomic)
        const unsigned long long b = (unsigned long long)__double_as_longlong(m);
        if (b > maxd2[t.rep]) atomicMax(&maxd2[t.rep], b);
    }
}

__global__ void __launch_bounds__(SIM_TILE)
    k_second(const sim_replica_t* __restrict__ reps, const SimTile* __restrict__ tiles,
             sim_types_t ty, const double4* __restrict__ pos, const float* __restrict__ vh,
             const float* __restrict__ force, float* __restrict__ vel,
             const float* __restrict__ noise, const long long* __restrict__ key,
             unsigned long long seed, unsigned long long step0,
             const unsigned long long* __restrict__ step_dev, double* __restrict__ partial) {
    const unsigned long long step = step0 + (step_dev ? step_dev[0] : 0ull);
    __shared__ double lds[4 * 7];
    const SimTile t = tile_of(tiles);
    const sim_replica_t& P = reps[t.rep];
    const int i = t.start + threadIdx.x;
    double v[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (i < P.atom_base + P.natoms) {
        // one type: no need to read the position record for it
        const int ty_i = ty.ntypes == 1 ? 0 : (int)pos[i].w;
        const size_t o = (size_t)i * 3;
        second_half_atom(P, i, ty_i, ty, vh, force[o], force[o + 1], force[o + 2], vel, noise,
                         key, seed, step, v);
    }
    block_sum<7>(v, lds);
    if (threadIdx.x == 0) {
        double* p = partial + (size_t)blockIdx.x * SIM_NPART;
#pragma unroll
        for (int k = 0; k < 7; ++k) p[SIM_T_KE + k] = v[k];
    }
}

__global__ void k_step_advance(unsigned long long* c) { c[0] += 1ull; }

__global__ void k_trigger(int R, sim_replica_t* reps, unsigned long long* maxd2, int* flag,
                          int* nbuild, int force_all) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    const int f = skin_trip(reps[r], maxd2[r], force_all) ? 1 : 0;
    flag[r] = f;
    if (nbuild) nbuild[r] += f;
    maxd2[r] = 0ull;
}

// Per-replica sums of partial columns [c0, c1) in two levels: workgroup
// (r, k) sums chunk k of replica r's rows with a compensated (Neumaier)
// running sum per column and a fixed float64 tree across its 256 threads;
// one workgroup per replica then adds the chunk sums in order. A chunk holds
// at least 256 rows (one per thread), so a small replica uses few chunks and
// the other workgroups of its column return at once. Both levels have a fixed
// order, so the result is deterministic.
#define SIM_RED_CHUNKS 64
#ifndef SIM_RED_GX
#define SIM_RED_GX 8      // blocks per replica in the chunk pass
#endif

// Rows per chunk of a replica with n rows, and the number of chunks used.
__device__ __forceinline__ int red_len(int n) {
    return max(256, (n + SIM_RED_CHUNKS - 1) / SIM_RED_CHUNKS);
}

template <int NC>
__global__ void __launch_bounds__(256)
    k_reduce_chunks(const int* __restrict__ rep_rows, const double* __restrict__ partial, int c0,
                    int c1, double* __restrict__ chunks) {
    __shared__ double lds[4 * NC];
    const int r = blockIdx.y;
    const int b0 = rep_rows[r], e0 = rep_rows[r + 1];
    const int len = red_len(e0 - b0);
    // a few blocks per replica walk its chunks (most replicas need one or
    // two), instead of one block per possible chunk
    for (int ck = blockIdx.x; ck < SIM_RED_CHUNKS; ck += gridDim.x) {
        const int b = b0 + ck * len, e = min(e0, b + len);
        if (ck > 0 && b >= e) break;   // unused chunk (chunk 0 always writes)
        double s[NC], comp[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) s[c] = comp[c] = 0.0;
        for (int k = b + (int)threadIdx.x; k < e; k += 256) {
            const double* row = partial + (size_t)k * SIM_NPART;
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const double x = c0 + c < c1 ? row[c0 + c] : 0.0;
                const double tt = s[c] + x;
                comp[c] += fabs(s[c]) >= fabs(x) ? (s[c] - tt) + x : (x - tt) + s[c];
                s[c] = tt;
            }
        }
        double v[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) v[c] = s[c] + comp[c];
        block_sum<NC>(v, lds);
        if (threadIdx.x == 0) {
            double* out = chunks + ((size_t)r * SIM_RED_CHUNKS + ck) * SIM_NPART;
#pragma unroll
            for (int c = 0; c < NC; ++c)
                if (c0 + c < c1) out[c0 + c] = v[c];
        }
        __syncthreads();   // lds is reused by the next chunk
    }
}

__global__ void k_reduce_final(const int* __restrict__ rep_rows, int c0, int c1,
                               const double* __restrict__ chunks, double* __restrict__ thermo,
                               unsigned long long* __restrict__ step_dev) {
    const int r = blockIdx.x, c = c0 + (int)threadIdx.x;
    // the step counter of the noise advances here (after the second half
    // read it) instead of in a launch of its own
    if (step_dev != nullptr && r == 0 && threadIdx.x == 0) step_dev[0] += 1ull;
    if (c >= c1) return;
    const int n = rep_rows[r + 1] - rep_rows[r];
    const int nck = max(1, (n + red_len(n) - 1) / red_len(n));
    double s = 0.0, comp = 0.0;
    for (int k = 0; k < nck; ++k) {
        const double x = chunks[((size_t)r * SIM_RED_CHUNKS + k) * SIM_NPART + c];
        const double tt = s + x;
        comp += fabs(s) >= fabs(x) ? (s - tt) + x : (x - tt) + s;
        s = tt;
    }
    thermo[(size_t)r * SIM_NPART + c] = s + comp;
}

// Berendsen, core/barostat.py:46-66 (pressure from K + W, mu^3 clamp, box and
// coordinates scaled); pos_last is scaled too so that the rebuild trigger keeps
// measuring displacements in the co-moving frame.
__global__ void k_baro_box(int R, sim_replica_t* reps, const double* thermo) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    sim_replica_t& P = reps[r];
    if (P.barostat != 1) {
        P.mu = 1.0;
        return;
    }
    const double* th = thermo + (size_t)r * SIM_NPART;
    const double tr = (th[SIM_T_K] + th[SIM_T_K + 1] + th[SIM_T_K + 2]) +
                      (th[SIM_T_W] + th[SIM_T_W + 1] + th[SIM_T_W + 2]);
    const double Pbar = tr / P.volume / 3.0 * 1602176.6208;
    const double dt = P.baro_dt > 0.0 ? P.baro_dt : P.dt;
    double mu3 = 1.0 + P.kappa * (dt / P.tau_p) * (Pbar - P.target_p);
    const double lo = (1.0 - P.mu_max) * (1.0 - P.mu_max) * (1.0 - P.mu_max);
    const double hi = (1.0 + P.mu_max) * (1.0 + P.mu_max) * (1.0 + P.mu_max);
    mu3 = fmin(fmax(mu3, lo), hi);
    const double mu = pow(mu3, 1.0 / 3.0);
    for (int k = 0; k < 9; ++k) {
        P.h[k] *= mu;
        P.hinv[k] /= mu;
    }
    P.volume *= mu3 == 0.0 ? 0.0 : mu * mu * mu;
    P.mu = mu;
    P.mu_acc *= mu;
    P.press = Pbar;
}

__global__ void __launch_bounds__(SIM_TILE)
    k_baro_scale(const sim_replica_t* __restrict__ reps, const SimTile* __restrict__ tiles,
                 double4* __restrict__ pos, float* __restrict__ vel, double4* __restrict__ pos_last) {
    const SimTile t = tile_of(tiles);
    const sim_replica_t& P = reps[t.rep];
    if (P.barostat != 1) return;
    const int i = t.start + threadIdx.x;
    if (i >= P.atom_base + P.natoms) return;
    const double mu = P.mu;
    const float muf = (float)mu;
    double4 x = pos[i];
    x.x *= mu;
    x.y *= mu;
    x.z *= mu;
    pos[i] = x;
    double4 l = pos_last[i];
    l.x *= mu;
    l.y *= mu;
    l.z *= mu;
    pos_last[i] = l;
    const size_t o = (size_t)i * 3;
    vel[o] *= muf;
    vel[o + 1] *= muf;
    vel[o + 2] *= muf;
}

// ---- lattice-row scaling: anisotropic barostat and uniaxial loading ----------
//
// Both scale the lattice rows, H' = diag(s) H, and carry the atoms with the
// box: x' = x H^-1 H' (fractional coordinates unchanged), pos_last too so that
// the rebuild trigger keeps measuring displacements in the co-moving frame.

// Smallest eigenvalue of the symmetric 3x3 G (xx yy zz xy xz yz) by the
// trigonometric closed form (O. K. Smith, CACM 4(4):168, 1961).
__device__ double sym3_min_eig(double a, double b, double c, double d, double e, double f) {
    const double p1 = d * d + e * e + f * f;
    if (p1 == 0.0) return fmin(a, fmin(b, c));
    const double q = (a + b + c) / 3.0;
    const double p2 = (a - q) * (a - q) + (b - q) * (b - q) + (c - q) * (c - q) + 2.0 * p1;
    const double p = sqrt(p2 / 6.0);
    const double ba = (a - q) / p, bb = (b - q) / p, bc = (c - q) / p;
    const double bd = d / p, be = e / p, bf = f / p;
    const double det = ba * (bb * bc - bf * bf) - bd * (bd * bc - bf * be) + be * (bd * bf - bb * be);
    const double r = fmin(1.0, fmax(-1.0, 0.5 * det));
    const double phi = acos(r) / 3.0;
    return q + 2.0 * p * cos(phi + 2.0943951023931957);     // + 2 pi / 3
}

// mu_acc = smallest singular value of F = Hb^-1 H (rows map as d' = d F), the
// factor by which the box change since the last build can shrink a
// separation, lowered by 1e-12 to stay a bound under rounding.
__device__ void update_mu_acc(sim_replica_t& P) {
    double F[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            F[3 * i + j] = P.hb_inv[3 * i] * P.h[j] + P.hb_inv[3 * i + 1] * P.h[3 + j] +
                           P.hb_inv[3 * i + 2] * P.h[6 + j];
    auto dot = [&](int i, int k) {
        return F[3 * i] * F[3 * k] + F[3 * i + 1] * F[3 * k + 1] + F[3 * i + 2] * F[3 * k + 2];
    };
    const double lam = sym3_min_eig(dot(0, 0), dot(1, 1), dot(2, 2), dot(0, 1), dot(0, 2), dot(1, 2));
    P.mu_acc = sqrt(fmax(lam, 0.0)) - 1e-12;
}

// H' = diag(s) H; amap = H^-1 H'; hinv' = H^-1 diag(1/s); volume scales by
// s0 s1 s2.
__device__ void scale_rows(sim_replica_t& P, const double s[3]) {
    double hn[9];
    for (int a = 0; a < 3; ++a)
        for (int j = 0; j < 3; ++j) hn[3 * a + j] = s[a] * P.h[3 * a + j];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            P.amap[3 * i + j] = P.hinv[3 * i] * hn[j] + P.hinv[3 * i + 1] * hn[3 + j] +
                                P.hinv[3 * i + 2] * hn[6 + j];
    for (int k = 0; k < 9; ++k) P.h[k] = hn[k];
    for (int i = 0; i < 3; ++i)
        for (int a = 0; a < 3; ++a) P.hinv[3 * i + a] /= s[a];
    P.volume *= s[0] * s[1] * s[2];
    update_mu_acc(P);
}

// core/barostat.py:131-208 for replicas with barostat = 2.
__global__ void k_aniso_box(int R, sim_replica_t* reps, const double* thermo,
                            unsigned long long seed, const unsigned long long* step_dev) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    sim_replica_t& P = reps[r];
    if (P.barostat != 2) return;
    const double* th = thermo + (size_t)r * SIM_NPART;
    const double bar = 1602176.6208;
    // (K + W) / V, symmetric, in eV/A^3
    double S[9];
    const int ia[6] = {0, 4, 8, 1, 2, 5}, ib[6] = {0, 4, 8, 3, 6, 7};
    for (int k = 0; k < 6; ++k) {
        const double v = (th[SIM_T_K + k] + th[SIM_T_W + k]) / P.volume;
        S[ia[k]] = v;
        S[ib[k]] = v;
    }
    const double dt = P.baro_dt > 0.0 ? P.baro_dt : P.dt;
    float gn[3] = {0.f, 0.f, 0.f};
    if (P.stochastic) {
        const unsigned long long step = step_dev ? step_dev[0] : 0ull;
        // a replica with its own key draws the same stream wherever it is packed
        gauss3(seed ^ P.noise_key, step, P.noise_key ? 0u : (uint32_t)r, gn[0], gn[1], gn[2],
               0xba20u);
    }
    const double tau2 = fmax(P.tau_p * P.tau_p, 1e-12);
    bool within[3];
    bool any_within = false;
    double accel[3];
    for (int a = 0; a < 3; ++a) {
        const double* u = P.h + 3 * a;
        const double n = fmax(sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]), 1e-12);
        const double e0 = u[0] / n, e1 = u[1] / n, e2 = u[2] / n;
        const double pa = e0 * (S[0] * e0 + S[1] * e1 + S[2] * e2) +
                          e1 * (S[3] * e0 + S[4] * e1 + S[5] * e2) +
                          e2 * (S[6] * e0 + S[7] * e1 + S[8] * e2);
        const double pbar = pa * bar;
        P.p_axis[a] = pbar;
        const double dp = pbar - P.p_target3[a];
        within[a] = fabs(dp) <= P.p_tol;
        any_within |= within[a];
        accel[a] = P.kappa * (within[a] ? 0.0 : dp) / tau2;
        if (P.gamma_p > 0.0) accel[a] -= P.gamma_p * P.eta_dot[a];
        if (P.stochastic)
            accel[a] += (double)gn[a] * (P.kappa * 8.617333262e-5 * fmax(P.temp_k, 1.0)) /
                        fmax(P.tau_p, 1e-12);
    }
    const double hold = fmax(P.gamma_p, 5.0);
    double s[3];
    for (int a = 0; a < 3; ++a) {
        if (any_within && within[a]) accel[a] = -hold * P.eta_dot[a];
        const bool ctl = (P.axes >> a) & 1;
        double e = P.eta_dot[a] + dt * accel[a];
        e = fmin(fmax(e, -P.max_eta_dot), P.max_eta_dot);
        if (!ctl) e = 0.0;
        if (within[a] && fabs(e) < 1e-7) e = 0.0;
        P.eta_dot[a] = e;
        s[a] = ctl ? exp(dt * e) : 1.0;
        P.vscale[a] = ctl ? (float)exp(-dt * e) : 1.0f;
    }
    scale_rows(P, s);
    P.press = (P.p_axis[0] + P.p_axis[1] + P.p_axis[2]) / 3.0;
}

// core/mechanics/loading.py:47-78 for replicas with loading = 1.
__global__ void k_deform_box(int R, sim_replica_t* reps) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    sim_replica_t& P = reps[r];
    if (!P.loading) return;
    const double s[3] = {P.deform[0], P.deform[1], P.deform[2]};
    scale_rows(P, s);
}

// x' = x amap for pos and pos_last; with VEL, v_a *= vscale[a]. MODE 2 serves
// the replicas with barostat = 2, MODE 3 those with loading = 1.
template <int MODE, bool VEL>
__global__ void __launch_bounds__(SIM_TILE)
    k_affine_map(const sim_replica_t* __restrict__ reps, const SimTile* __restrict__ tiles,
                 double4* __restrict__ pos, float* __restrict__ vel,
                 double4* __restrict__ pos_last) {
    const SimTile t = tile_of(tiles);
    const sim_replica_t& P = reps[t.rep];
    if (MODE == 2 ? P.barostat != 2 : !P.loading) return;
    const int i = t.start + threadIdx.x;
    if (i >= P.atom_base + P.natoms) return;
    const double* M = P.amap;
    double4 x = pos[i];
    pos[i] = double4{x.x * M[0] + x.y * M[3] + x.z * M[6], x.x * M[1] + x.y * M[4] + x.z * M[7],
                     x.x * M[2] + x.y * M[5] + x.z * M[8], x.w};
    double4 l = pos_last[i];
    pos_last[i] = double4{l.x * M[0] + l.y * M[3] + l.z * M[6], l.x * M[1] + l.y * M[4] + l.z * M[7],
                          l.x * M[2] + l.y * M[5] + l.z * M[8], l.w};
    if (VEL) {
        const size_t o = (size_t)i * 3;
        vel[o] *= P.vscale[0];
        vel[o + 1] *= P.vscale[1];
        vel[o + 2] *= P.vscale[2];
    }
}

template <typename T>
__global__ void k_export(const double4* pos, int n, T* xyz) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 x = pos[i];
    xyz[(size_t)i * 3] = (T)x.x;
    xyz[(size_t)i * 3 + 1] = (T)x.y;
    xyz[(size_t)i * 3 + 2] = (T)x.z;
}

template <typename T>
__global__ void k_import(double4* pos, int n, const T* xyz) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double4 x = pos[i];
    x.x = (double)xyz[(size_t)i * 3];
    x.y = (double)xyz[(size_t)i * 3 + 1];
    x.z = (double)xyz[(size_t)i * 3 + 2];
    pos[i] = x;
}

// Fractional fixed-point record of every atom: s = x H^-1 (row vectors,
// core/box.py:132-142), wrapped into [0, 1), stored as round(s 2^32) mod 2^32;
// w = atom type.
__global__ void __launch_bounds__(SIM_TILE)
    k_quantize(const sim_replica_t* __restrict__ reps, const SimTile* __restrict__ tiles,
               const double4* __restrict__ pos, int4* __restrict__ q) {
    const SimTile t = tile_of(tiles);
    const sim_replica_t& P = reps[t.rep];
    const int i = t.start + threadIdx.x;
    if (i >= P.atom_base + P.natoms) return;
    q[i] = quant_record(P, pos[i]);
}

}  // namespace

extern "C" int sim_pos_quantize(const void* reps, const int* tiles, int ntiles, const double* pos4,
                                int* posq, void* stream) {
    if (!reps || !tiles || ntiles <= 0 || !pos4 || !posq) return SIM_E_ARG;
    hipLaunchKernelGGL(k_quantize, dim3(ntiles), dim3(SIM_TILE), 0, as_stream(stream),
                       static_cast<const sim_replica_t*>(reps),
                       reinterpret_cast<const SimTile*>(tiles),
                       reinterpret_cast<const double4*>(pos4), reinterpret_cast<int4*>(posq));
    SIM_CHECK_LAUNCH();
    return SIM_OK;
}

static int first_impl(const void* reps, const int* tiles, int ntiles, sim_types_t types,
                      double* pos4, float* vel, const float* force, float* vel_half,
                      double* pos_last, double* maxd2, int* posq, int flags,
                      void* stream) {
    if (!reps || !tiles || ntiles <= 0 || !pos4 || !vel || !force || !vel_half || !pos_last ||
        !maxd2 || types.ntypes < 1 || types.ntypes > SIM_MAX_TYPES ||
        (flags & ~(SIM_FIRST_UNFUSED_ONLY | SIM_FIRST_BERENDSEN_SCALE)) != 0)
        return SIM_E_ARG;
    hipLaunchKernelGGL(k_first, dim3(ntiles), dim3(SIM_TILE), 0, as_stream(stream),
                       static_cast<const sim_replica_t*>(reps),
                       reinterpret_cast<const SimTile*>(tiles), types,
                       reinterpret_cast<double4*>(pos4), vel, force, vel_half,
                       reinterpret_cast<double4*>(pos_last),
                       reinterpret_cast<unsigned long long*>(maxd2), reinterpret_cast<int4*>(posq),
                       flags, vel);
    SIM_CHECK_LAUNCH();
    return SIM_OK;
}

extern "C" int sim_integrate_first(const void* reps, const int* tiles, int ntiles,
                                   sim_types_t types, double* pos4, const float* vel,
                                   const float* force, float* vel_half, const double* pos_last,
                                   double* maxd2, void* stream) {
    // flags 0: vel and pos_last are only read
    return first_impl(reps, tiles, ntiles, types, pos4, const_cast<float*>(vel), force, vel_half,
                      const_cast<double*>(pos_last), maxd2, nullptr, 0, stream);
}

extern "C" int sim_integrate_first_q(const void* reps, const int* tiles, int ntiles,
                                     sim_types_t types, double* pos4, float* vel,
                                     const float* force, float* vel_half, double* pos_last,
                                     double* maxd2, int* posq, int flags, void* stream) {
    if (!posq) return SIM_E_ARG;
    return first_impl(reps, tiles, ntiles, types, pos4, vel, force, vel_half, pos_last, maxd2,
                      posq, flags, stream);
}

extern "C" int sim_integrate_second(const void* reps, const int* tiles, int ntiles,
                                    sim_types_t types, const double* pos4, const float* vel_half,
                                    const float* force, float* vel, const float* noise,
                                    const long long* key,
                                    unsigned long long seed, unsigned long long step,
                                    const unsigned long long* step_dev, double* partial,
                                    void* stream) {
    if (!reps || !tiles || ntiles <= 0 || !pos4 || !vel_half || !force || !vel || !partial ||
        types.ntypes < 1 || types.ntypes > SIM_MAX_TYPES)
        return SIM_E_ARG;
    hipLaunchKernelGGL(k_second, dim3(ntiles), dim3(SIM_TILE), 0, as_stream(stream),
                       static_cast<const sim_replica_t*>(reps),
                       reinterpret_cast<const SimTile*>(tiles), types,
                       reinterpret_cast<const double4*>(pos4), vel_half, force, vel, noise, key, seed,
                       step, step_dev, partial);
    SIM_CHECK_LAUNCH();
    return SIM_OK;
}

extern "C" int sim_step_advance(unsigned long long* step_dev, void* stream) {
    if (!step_dev) return SIM_E_ARG;
    hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(1), 0, as_stream(stream), step_dev);
    SIM_CHECK_LAUNCH();
    return SIM_OK;
}

extern "C" int sim_nl_trigger(int R, void* reps, double* maxd2, int* flag, int* nbuild,
                              int force_all, void* stream) {
    if (R <= 0 || !reps || !maxd2 || !flag) return SIM_E_ARG;
    hipLaunchKernelGGL(k_trigger, dim3((R + 63) / 64), dim3(64), 0, as_stream(stream), R,
                       static_cast<sim_replica_t*>(reps),
                       reinterpret_cast<unsigned long long*>(maxd2), flag, nbuild, force_all);
    SIM_CHECK_LAUNCH();
    return SIM_OK;
}

static int reduce_impl(int R, const int* rep_rows, const double* partial, int col0, int col1,
                          double* chunks, double* thermo, unsigned long long* step_dev, void* stream) {
    if (R <= 0 || !rep_rows || !partial || !chunks || !thermo || col0 < 0 || col1 > SIM_NPART ||
        col0 >= col1)
        return SIM_E_ARG;
    hipStream_t st = as_stream(stream);
    const dim3 g(SIM_RED_GX, R);
    if (col1 - col0 <= 8)
        hipLaunchKernelGGL(k_reduce_chunks<8>, g, dim3(256), 0, st, rep_rows, partial, col0, col1,
                           chunks);
    else
        hipLaunchKernelGGL(k_reduce_chunks<SIM_NPART>, g, dim3(256), 0, st, rep_rows, partial,
                           col0, col1, chunks);
    hipLaunchKernelGGL(k_reduce_final, dim3(R), dim3(SIM_NPART), 0, st, rep_rows, col0, col1,
                       chunks, thermo, step_dev);
    SIM_CHECK_LAUNCH();
    return SIM_OK;
}

extern "C" int sim_reduce(int R, const int* rep_rows, const double* partial, int col0, int col1,
                          double* chunks, double* thermo, void* stream) {
    return reduce_impl(R, rep_rows, partial, col0, col1, chunks, thermo, nullptr, stream);
}

extern "C" int sim_reduce_step(int R, const int* rep_rows, const double* partial, int col0,
                               int col1, double* chunks, double* thermo,
                               unsigned long long* step_dev, void* stream) {
    if (!step_dev) return SIM_E_ARG;
    return reduce_impl(R, rep_rows, partial, col0, col1, chunks, thermo, step_dev, stream);
}

extern "C" int sim_berendsen_box(int R, void* reps, const double* thermo, void* stream) {
    if (R <= 0 || !reps || !thermo) return SIM_E_ARG;
    hipLaunchKernelGGL(k_baro_box, dim3((R + 63) / 64), dim3(64), 0, as_stream(stream), R,
                       static_cast<sim_replica_t*>(reps), thermo);
    SIM_CHECK_LAUNCH();
    return SIM_OK;
}

extern "C" int sim_berendsen(int R, void* reps, const int* tiles, int ntiles, const double* thermo,
                             double* pos4, float* vel, double* pos_last, void* stream) {
    if (R <= 0 || !reps || !tiles || ntiles <= 0 || !thermo || !pos4 || !vel || !pos_last)
        return SIM_E_ARG;
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(k_baro_box, dim3((R + 63) / 64), dim3(64), 0, st, R,
                       static_cast<sim_replica_t*>(reps), thermo);
    hipLaunchKernelGGL(k_baro_scale, dim3(ntiles), dim3(SIM_TILE), 0, st,
                       static_cast<const sim_replica_t*>(reps),
                       reinterpret_cast<const SimTile*>(tiles), reinterpret_cast<double4*>(pos4), vel,
                       reinterpret_cast<double4*>(pos_last));
    SIM_CHECK_LAUNCH();
    return SIM_OK;
}

extern "C" int sim_aniso_barostat(int R, void* reps, const int* tiles, int ntiles,
                                  const double* thermo, double* pos4, float* vel,
                                  double* pos_last, unsigned long long seed,
                                  const unsigned long long* step_dev, void* stream) {
    if (R <= 0 || !reps || !tiles || ntiles <= 0 || !thermo || !pos4 || !vel || !pos_last)
        return SIM_E_ARG;
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(k_aniso_box, dim3((R + 63) / 64), dim3(64), 0, st, R,
                       static_cast<sim_replica_t*>(reps), thermo, seed, step_dev);
    hipLaunchKernelGGL((k_affine_map<2, true>), dim3(ntiles), dim3(SIM_TILE), 0, st,
                       static_cast<const sim_replica_t*>(reps),
                       reinterpret_cast<const SimTile*>(tiles), reinterpret_cast<double4*>(pos4), vel,
                       reinterpret_cast<double4*>(pos_last));
    SIM_CHECK_LAUNCH();
    return SIM_OK;
}

extern "C" int sim_box_deform(int R, void* reps, const int* tiles, int ntiles, double* pos4,
                              double* pos_last, void* stream) {
    if (R <= 0 || !reps || !tiles || ntiles <= 0 || !pos4 || !pos_last) return SIM_E_ARG;
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(k_deform_box, dim3((R + 63) / 64), dim3(64), 0, st, R,
                       static_cast<sim_replica_t*>(reps));
    hipLaunchKernelGGL((k_affine_map<3, false>), dim3(ntiles), dim3(SIM_TILE), 0, st,
                       static_cast<const sim_replica_t*>(reps),
                       reinterpret_cast<const SimTile*>(tiles), reinterpret_cast<double4*>(pos4),
                       nullptr, reinterpret_cast<double4*>(pos_last));
    SIM_CHECK_LAUNCH();
    return SIM_OK;
}

extern "C" int sim_pos_export(const double* pos4, int n, void* xyz, int dtype, void* stream) {
    if (!pos4 || n <= 0 || !xyz) return SIM_E_ARG;
    hipStream_t st = as_stream(stream);
    auto* X = reinterpret_cast<const double4*>(pos4);
    if (dtype == 0)
        hipLaunchKernelGGL(k_export<float>, dim3((n + 255) / 256), dim3(256), 0, st, X, n,
                           static_cast<float*>(xyz));
    else if (dtype == 1)
        hipLaunchKernelGGL(k_export<double>, dim3((n + 255) / 256), dim3(256), 0, st, X, n,
                           static_cast<double*>(xyz));
    else
        return SIM_E_ARG;
    SIM_CHECK_LAUNCH();
    return SIM_OK;
}

extern "C" int sim_pos_import(double* pos4, int n, const void* xyz, int dtype, void* stream) {
    if (!pos4 || n <= 0 || !xyz) return SIM_E_ARG;
    hipStream_t st = as_stream(stream);
    auto* X = reinterpret_cast<double4*>(pos4);
    if (dtype == 0)
        hipLaunchKernelGGL(k_import<float>, dim3((n + 255) / 256), dim3(256), 0, st, X, n,
                           static_cast<const float*>(xyz));
    else if (dtype == 1)
        hipLaunchKernelGGL(k_import<double>, dim3((n + 255) / 256), dim3(256), 0, st, X, n,
                           static_cast<const double*>(xyz));
    else
        return SIM_E_ARG;
    SIM_CHECK_LAUNCH();
    return SIM_OK;
}
