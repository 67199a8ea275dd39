// Cell-tiled pair kernels (gfx950): LJ, EAM density and EAM force with the
// 27-cell neighbourhood of each cell staged in LDS.
//
// After a rebuild the atoms of every cell are contiguous (sim_nl_build with a
// reorder), positions are wrapped into the primary cell, and each atom's
// Verlet list also exists as 16-bit column slots (stencil column x-1, x, x+1
// and the offset within that column's 9 cells). The halo concatenates the 27
// stencil cells column by column, so a slot maps to a halo index with one
// add. These per-cell kernels suit small systems (many short workgroups);
// pair_rows.hip walks rows of cells for large ones. One workgroup takes one
// cell:
//   1. the 27 stencil cells' ranges, periodic image translations and LDS
//      offsets go to LDS;
//   2. all stencil atoms are copied into LDS with coalesced loads, as float64
//      positions already translated to the image adjacent to the cell (so a
//      pair displacement is a plain float64 subtraction), plus one auxiliary
//      float (atom type, or F'(rho) for the EAM force pass);
//   3. LANES lanes share each home atom and walk its list with SIM_NB_BATCH
//      neighbours in flight per lane; neighbour data come from LDS, only the
//      EAM table lookups go to global memory (L2-resident tables);
//   4. per-atom sums are combined across the LANES lanes with wave shuffles;
//      energy and virial leave as one float64 partial per cell.
// A cell whose stencil holds more atoms than the LDS tile (a dense
// fluctuation in a hot system) reads its neighbours from global memory
// through the same stencil table instead, with the same arithmetic; overflow
// bit 2 reports that this happened (slower, not wrong).
// Physics and masks are those of pair_lj.hip / pair_eam.hip
// (core/force/lennard_jones_force.py:15-95, core/force/eam_force_cu.py:461-579).
#include "pair_common.h"

#define SIM_CELL_THREADS 256

namespace {

template <int CAP>
struct Halo {
    double x[CAP], y[CAP], z[CAP];
    float aux[CAP];
    int start[27];
    int count[27];
    int off[28];
    double sh[27][3];
    double red[4 * 8];
};

// Stencil metadata and halo positions for `cell`. The stencil cells are the
// 27 neighbours in fixed (x, z, y) order, i.e. column by column, matching the
// column slots of the lists; their periodic translation is
// floor((c + o) / n) lattice vectors. Returns HALO_SKIP when the grid has
// fewer than 3 cells in a dimension (overflow bit 4), HALO_LDS when the
// stencil atoms are in LDS, HALO_GLOBAL when the stencil holds more than CAP
// atoms (overflow bit 2, informational): then only the stencil table is in
// LDS and the atoms are read from global memory through it.
enum { HALO_SKIP = 0, HALO_LDS = 1, HALO_GLOBAL = 2 };

template <int CAP, int AUX>
__device__ __forceinline__ int load_halo(Halo<CAP>& H, const sim_replica_t& P, int cell,
                                          const int* __restrict__ cell_start,
                                          const int* __restrict__ cell_count,
                                          const double4* __restrict__ pos,
                                          const float* __restrict__ aux, int* overflow, int& total) {
    const int nx = P.ncell[0], ny = P.ncell[1], nz = P.ncell[2];
    const int local = cell - P.cell_base;
    const int tid = threadIdx.x;
    if (nx < 3 || ny < 3 || nz < 3) {   // the 27 stencil cells must be distinct
        if (tid == 0) atomicOr(overflow, 4);
        return HALO_SKIP;
    }
    if (tid < 27) {
        const int cx = local % nx, cy = (local / nx) % ny, cz = local / (nx * ny);
        const int ox = tid / 9 - 1, oz = (tid / 3) % 3 - 1, oy = tid % 3 - 1;
        int xx = cx + ox, yy = cy + oy, zz = cz + oz;
        const int fx = xx < 0 ? -1 : (xx >= nx ? 1 : 0);
        const int fy = yy < 0 ? -1 : (yy >= ny ? 1 : 0);
        const int fz = zz < 0 ? -1 : (zz >= nz ? 1 : 0);
        xx -= fx * nx;
        yy -= fy * ny;
        zz -= fz * nz;
        const int c = P.cell_base + (zz * ny + yy) * nx + xx;
        H.start[tid] = cell_start[c];
        H.count[tid] = cell_count[c];
        for (int k = 0; k < 3; ++k)
            H.sh[tid][k] = fx * P.h[k] + fy * P.h[3 + k] + fz * P.h[6 + k];
    }
    __syncthreads();
    if (tid == 0) {
        int o = 0;
        for (int s = 0; s < 27; ++s) {
            H.off[s] = o;
            o += H.count[s];
        }
        H.off[27] = o;
    }
    __syncthreads();
    total = H.off[27];
    if (total > CAP) {
        if (tid == 0) atomicOr(overflow, 2);
        return HALO_GLOBAL;
    }
    for (int f = tid; f < total; f += SIM_CELL_THREADS) {
        int lo = 0, hi = 26;   // stencil cell holding flat slot f
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (H.off[mid] <= f) lo = mid;
            else hi = mid - 1;
        }
        const int j = H.start[lo] + (f - H.off[lo]);
        const double4 xj = pos[j];
        H.x[f] = xj.x + H.sh[lo][0];
        H.y[f] = xj.y + H.sh[lo][1];
        H.z[f] = xj.z + H.sh[lo][2];
        H.aux[f] = AUX ? aux[j] : (float)xj.w;
    }
    __syncthreads();
    return HALO_LDS;
}

// Position (translated to the image next to the cell) and auxiliary float of
// stencil slot s: from LDS, or from global memory through the stencil table
// with the same float64 translation, so both give the same numbers.
template <bool FROM_LDS, int AUX, int CAP>
__device__ __forceinline__ void slot_get(const Halo<CAP>& H, const double4* __restrict__ pos,
                                         const float* __restrict__ aux, int s, double& x,
                                         double& y, double& z, float& a) {
    if (FROM_LDS) {
        x = H.x[s];
        y = H.y[s];
        z = H.z[s];
        a = H.aux[s];
        return;
    }
    int lo = 0, hi = 26;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (H.off[mid] <= s) lo = mid;
        else hi = mid - 1;
    }
    const int j = H.start[lo] + (s - H.off[lo]);
    const double4 p = pos[j];
    x = p.x + H.sh[lo][0];
    y = p.y + H.sh[lo][1];
    z = p.z + H.sh[lo][2];
    a = AUX ? aux[j] : (float)p.w;
}

// Column slot (stencil column << 14 | offset within the column's 9 cells,
// sim_nl_build) -> flat index into the halo, whose cells are stored column by
// column in the same (z, y) order.
template <int CAP>
__device__ __forceinline__ int flat_slot(const Halo<CAP>& H, int s16) {
    return H.off[(s16 >> 14) * 9] + (s16 & 0x3fff);
}

template <int LANES>
__device__ __forceinline__ float lane_sum(float v) {
#pragma unroll
    for (int o = 1; o < LANES; o <<= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// xi - x(slot s) rounded to float32, and the slot's auxiliary float.
template <bool FROM_LDS, int AUX, int CAP>
__device__ __forceinline__ void disp_slot(const Halo<CAP>& H, const double4* __restrict__ pos,
                                          const float* __restrict__ aux, double xi, double yi,
                                          double zi, int s, float& rx, float& ry, float& rz,
                                          float& a) {
    double x, y, z;
    slot_get<FROM_LDS, AUX>(H, pos, aux, s, x, y, z, a);
    rx = (float)(xi - x);
    ry = (float)(yi - y);
    rz = (float)(zi - z);
}

// ---- EAM ------------------------------------------------------------------------
struct LerpT {
    int idx;
    float frac;
};

__device__ __forceinline__ LerpT lerp_idx(float r, const sim_eam_params_t& E) {
    const float rc = fminf(fmaxf(r, 0.0f), E.r_clamp);
    const float xf = rc * E.inv_dr;
    const int idx = (int)xf;
    return LerpT{idx, xf - (float)idx};
}

__device__ __forceinline__ float lp(float a, float b, float t) { return a + t * (b - a); }

// Density pass over the home atoms of one cell; returns the cell's energy.
template <bool FROM_LDS, int CAP, int LANES>
__device__ __forceinline__ double eam_density_home(const Halo<CAP>& H,
                                                   const double4* __restrict__ pos,
                                                   const unsigned short* __restrict__ nbr16,
                                                   const int* __restrict__ nnbr, int max_nbr,
                                                   const sim_eam_params_t& E,
                                                   float* __restrict__ rho_out,
                                                   float* __restrict__ fp_out, float thr) {
    const float4* __restrict__ T1 = static_cast<const float4*>(E.pass1);
    const int nt = E.ntypes;
    const int n_home = H.count[13], home0 = H.off[13], g0 = H.start[13];
    const int l = threadIdx.x % LANES;
    double etot = 0.0;
    for (int b = 0; b < n_home; b += SIM_CELL_THREADS / LANES) {
        const int a = b + threadIdx.x / LANES;
        float rho = 0.f, ep = 0.f;
        int ti = 0;
        if (a < n_home) {
            const int i = g0 + a;
            double xi, yi, zi;
            float ai;
            slot_get<FROM_LDS, 0>(H, pos, nullptr, home0 + a, xi, yi, zi, ai);
            ti = (int)ai;
            const int cnt = nnbr[i];
            const unsigned short* row = nbr16 + (size_t)i * max_nbr;
            constexpr int U = SIM_NB_BATCH;
            int k = l;
            for (; k + (U - 1) * LANES < cnt; k += U * LANES) {
                float r2[U];
                LerpT L[U];
                float4 v[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int s = flat_slot(H, row[k + u * LANES]);
                    float rx, ry, rz, aj;
                    disp_slot<FROM_LDS, 0>(H, pos, nullptr, xi, yi, zi, s, rx, ry, rz, aj);
                    r2[u] = sq3(rx, ry, rz);
                    L[u] = lerp_idx(pair_sqrt(r2[u]), E);
                    const int tj = nt == 1 ? 0 : (int)aj;
                    v[u] = T1[(size_t)(ti * nt + tj) * E.n_r + L[u].idx];
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const float w = r2[u] < thr ? 1.0f : 0.0f;
                    rho += w * lp(v[u].x, v[u].y, L[u].frac);
                    ep += w * lp(v[u].z, v[u].w, L[u].frac);
                }
            }
            for (; k < cnt; k += LANES) {
                const int s = flat_slot(H, row[k]);
                float rx, ry, rz, aj;
                disp_slot<FROM_LDS, 0>(H, pos, nullptr, xi, yi, zi, s, rx, ry, rz, aj);
                const float r2 = sq3(rx, ry, rz);
                const LerpT L = lerp_idx(pair_sqrt(r2), E);
                const int tj = nt == 1 ? 0 : (int)aj;
                const float4 v = T1[(size_t)(ti * nt + tj) * E.n_r + L.idx];
                const float w = r2 < thr ? 1.0f : 0.0f;
                rho += w * lp(v.x, v.y, L.frac);
                ep += w * lp(v.z, v.w, L.frac);
            }
        }
        rho = lane_sum<LANES>(rho);
        ep = lane_sum<LANES>(ep);
        if (a < n_home && l == 0) {
            const int i = g0 + a;
            const float lo = E.rho_lo[ti];
            const float rcl = fminf(fmaxf(rho, lo), E.rho_hi[ti]);
            const float xf = (rcl - lo) * E.inv_drho[ti];
            const int idx = (int)xf;
            const float fr = xf - (float)idx;
            const float2 F = reinterpret_cast<const float2*>(E.emb)[(size_t)ti * E.n_rho + idx];
            const float2 dF = reinterpret_cast<const float2*>(E.demb)[(size_t)ti * E.n_rho + idx];
            rho_out[i] = rho;
            fp_out[i] = lp(dF.x, dF.y, fr);
            etot += (double)lp(F.x, F.y, fr) + 0.5 * (double)ep;
        }
    }
    return etot;
}

template <int CAP, int LANES>
__global__ void __launch_bounds__(SIM_CELL_THREADS)
    k_eam_density_cell(const sim_replica_t* __restrict__ reps, const int* __restrict__ cell_rep,
                       const int* __restrict__ cell_start, const int* __restrict__ cell_count,
                       const double4* __restrict__ pos, const unsigned short* __restrict__ nbr16,
                       const int* __restrict__ nnbr, int max_nbr, sim_eam_params_t E,
                       float* __restrict__ rho_out, float* __restrict__ fp_out,
                       double* __restrict__ partial, int* overflow, float thr) {
    __shared__ Halo<CAP> H;
    const int cell = blockIdx.x;
    const sim_replica_t& P = reps[cell_rep[cell]];
    double etot = 0.0;
    const int local = cell - P.cell_base;
    const bool live = local < P.ncell[0] * P.ncell[1] * P.ncell[2] && cell_count[cell] > 0;
    int total = 0;
    const int st = live ? load_halo<CAP, 0>(H, P, cell, cell_start, cell_count, pos, nullptr,
                                            overflow, total)
                        : HALO_SKIP;
    if (st == HALO_LDS)
        etot = eam_density_home<true, CAP, LANES>(H, pos, nbr16, nnbr, max_nbr, E, rho_out, fp_out,
                                                  thr);
    else if (st == HALO_GLOBAL)
        etot = eam_density_home<false, CAP, LANES>(H, pos, nbr16, nnbr, max_nbr, E, rho_out,
                                                   fp_out, thr);
    double v[1] = {etot};
    block_sum<1>(v, H.red);
    if (threadIdx.x == 0) partial[(size_t)cell * SIM_NPART + SIM_T_EPOT] = v[0];
}

// Force pass over the home atoms of one cell; acc receives the cell's
// virial columns (halved).
template <bool FROM_LDS, int CAP, int LANES, bool VIRIAL>
__device__ __forceinline__ void eam_force_home(double (&acc)[7], const Halo<CAP>& H,
                                               const double4* __restrict__ pos,
                                               const unsigned short* __restrict__ nbr16,
                                               const int* __restrict__ nnbr, int max_nbr,
                                               const sim_eam_params_t& E,
                                               const float* __restrict__ fp,
                                               float* __restrict__ force, float thr) {
    const float4* __restrict__ T2 = static_cast<const float4*>(E.pass2);
    const int n_home = H.count[13], home0 = H.off[13], g0 = H.start[13];
    const int l = threadIdx.x % LANES;
    for (int b = 0; b < n_home; b += SIM_CELL_THREADS / LANES) {
        const int a = b + threadIdx.x / LANES;
        float fx = 0.f, fy = 0.f, fz = 0.f;
        float w7[7] = {0, 0, 0, 0, 0, 0, 0};  // virial, xx yy zz xy xz yz
        if (a < n_home) {
            const int i = g0 + a;
            double xi, yi, zi;
            float fpi;
            slot_get<FROM_LDS, 1>(H, pos, fp, home0 + a, xi, yi, zi, fpi);
            const int cnt = nnbr[i];
            const unsigned short* row = nbr16 + (size_t)i * max_nbr;
            constexpr int U = SIM_NB_BATCH;
            int k = l;
            auto pair = [&](float rx, float ry, float rz, float r2, const LerpT& L,
                            const float4& va, float fpj) {
                const float r = pair_sqrt(r2);
                const float w = r2 < thr ? 1.0f : 0.0f;
                const float dd = lp(va.x, va.y, L.frac);
                const float s = w * (fpi * dd + fpj * dd + lp(va.z, va.w, L.frac));
                const float g = -s * __builtin_amdgcn_rcpf(r + 1e-8f);
                fx = __builtin_fmaf(g, rx, fx);
                fy = __builtin_fmaf(g, ry, fy);
                fz = __builtin_fmaf(g, rz, fz);
                if (VIRIAL) {
                    w7[0] = __builtin_fmaf(s, r, w7[0]);
                    const float gx = g * rx, gy = g * ry;
                    w7[1] = __builtin_fmaf(gx, rx, w7[1]);
                    w7[2] = __builtin_fmaf(gy, ry, w7[2]);
                    w7[3] = __builtin_fmaf(g * rz, rz, w7[3]);
                    w7[4] = __builtin_fmaf(gx, ry, w7[4]);
                    w7[5] = __builtin_fmaf(gx, rz, w7[5]);
                    w7[6] = __builtin_fmaf(gy, rz, w7[6]);
                }
            };
            for (; k + (U - 1) * LANES < cnt; k += U * LANES) {
                float rx[U], ry[U], rz[U], r2[U], fpj[U];
                LerpT L[U];
                float4 va[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int s = flat_slot(H, row[k + u * LANES]);
                    disp_slot<FROM_LDS, 1>(H, pos, fp, xi, yi, zi, s, rx[u], ry[u], rz[u], fpj[u]);
                    r2[u] = sq3(rx[u], ry[u], rz[u]);
                    L[u] = lerp_idx(pair_sqrt(r2[u]), E);
                    va[u] = T2[L[u].idx];   // one element type (checked at launch)
                }
#pragma unroll
                for (int u = 0; u < U; ++u) pair(rx[u], ry[u], rz[u], r2[u], L[u], va[u], fpj[u]);
            }
            for (; k < cnt; k += LANES) {
                const int s = flat_slot(H, row[k]);
                float rx, ry, rz, fpj;
                disp_slot<FROM_LDS, 1>(H, pos, fp, xi, yi, zi, s, rx, ry, rz, fpj);
                const float r2 = sq3(rx, ry, rz);
                const LerpT L = lerp_idx(pair_sqrt(r2), E);
                pair(rx, ry, rz, r2, L, T2[L.idx], fpj);
            }
        }
        fx = lane_sum<LANES>(fx);
        fy = lane_sum<LANES>(fy);
        fz = lane_sum<LANES>(fz);
        if (VIRIAL) {
#pragma unroll
            for (int c = 0; c < 7; ++c) acc[c] += 0.5 * (double)w7[c];
        }
        if (a < n_home && l == 0) {
            const int i = g0 + a;
            force[(size_t)i * 3 + 0] = fx;
            force[(size_t)i * 3 + 1] = fy;
            force[(size_t)i * 3 + 2] = fz;
        }
    }
}

template <int CAP, int LANES, bool VIRIAL>
__global__ void __launch_bounds__(SIM_CELL_THREADS)
    k_eam_force_cell(const sim_replica_t* __restrict__ reps, const int* __restrict__ cell_rep,
                     const int* __restrict__ cell_start, const int* __restrict__ cell_count,
                     const double4* __restrict__ pos, const unsigned short* __restrict__ nbr16,
                     const int* __restrict__ nnbr, int max_nbr, sim_eam_params_t E,
                     const float* __restrict__ fp, float* __restrict__ force,
                     double* __restrict__ partial, int* overflow, float thr) {
    __shared__ Halo<CAP> H;
    const int cell = blockIdx.x;
    const sim_replica_t& P = reps[cell_rep[cell]];
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    const int local = cell - P.cell_base;
    const bool live = local < P.ncell[0] * P.ncell[1] * P.ncell[2] && cell_count[cell] > 0;
    int total = 0;
    const int st = live ? load_halo<CAP, 1>(H, P, cell, cell_start, cell_count, pos, fp, overflow,
                                            total)
                        : HALO_SKIP;
    if (st == HALO_LDS)
        eam_force_home<true, CAP, LANES, VIRIAL>(acc, H, pos, nbr16, nnbr, max_nbr, E, fp, force,
                                                 thr);
    else if (st == HALO_GLOBAL)
        eam_force_home<false, CAP, LANES, VIRIAL>(acc, H, pos, nbr16, nnbr, max_nbr, E, fp, force,
                                                  thr);
    block_sum<7>(acc, H.red);
    if (threadIdx.x == 0) {
        double* p = partial + (size_t)cell * SIM_NPART;
#pragma unroll
        for (int c = 0; c < 7; ++c) p[SIM_T_VIRIAL + c] = acc[c];
    }
}

// ---- LJ -------------------------------------------------------------------------
template <bool FROM_LDS, int CAP, int LANES, bool VIRIAL>
__device__ __forceinline__ void lj_home(double (&acc)[8], const Halo<CAP>& H,
                                        const double4* __restrict__ pos,
                                        const unsigned short* __restrict__ nbr16,
                                        const int* __restrict__ nnbr, int max_nbr,
                                        const sim_lj_params_t& lj, float* __restrict__ force,
                                        float thr) {
    const int n_home = H.count[13], home0 = H.off[13], g0 = H.start[13];
    const int l = threadIdx.x % LANES;
    for (int b = 0; b < n_home; b += SIM_CELL_THREADS / LANES) {
        const int a = b + threadIdx.x / LANES;
        float fx = 0.f, fy = 0.f, fz = 0.f, e = 0.f;
        float w7[7] = {0, 0, 0, 0, 0, 0, 0};
        if (a < n_home) {
            const int i = g0 + a;
            double xi, yi, zi;
            float ai;
            slot_get<FROM_LDS, 0>(H, pos, nullptr, home0 + a, xi, yi, zi, ai);
            const int ti = (int)ai;
            const int cnt = nnbr[i];
            const unsigned short* row = nbr16 + (size_t)i * max_nbr;
            for (int k = l; k < cnt; k += LANES) {
                const int s = flat_slot(H, row[k]);
                float rx, ry, rz, aj;
                disp_slot<FROM_LDS, 0>(H, pos, nullptr, xi, yi, zi, s, rx, ry, rz, aj);
                const float r2 = sq3(rx, ry, rz);
                if (!(r2 < thr)) continue;
                const int tj = (int)aj;
                const float eps = lj.eps[ti * SIM_MAX_TYPES + tj];
                const float sig = lj.sig[ti * SIM_MAX_TYPES + tj];
                const float inv_r2 = __builtin_amdgcn_rcpf(r2);   // 1 ulp
                const float s2 = sig * sig * inv_r2;
                const float s6 = s2 * s2 * s2;
                const float s12 = s6 * s6;
                float U = 4.0f * eps * (s12 - s6);
                float fr = 24.0f * eps * inv_r2 * (2.0f * s12 - s6);
                if (lj.mode == 1) {
                    const float rc = lj.rc, r = sqrtf(r2);
                    const float src = sig / rc;
                    const float q6 = src * src * src * src * src * src;
                    const float q12 = q6 * q6;
                    const float Frc = 24.0f * eps * (1.0f / rc) * (2.0f * q12 - q6);
                    U = U - 4.0f * eps * (q12 - q6) + (r - rc) * Frc;
                    fr = fr - Frc / r;
                } else if (lj.mode == 2) {
                    const float rc = lj.rc, rs = lj.rs, r = sqrtf(r2);
                    if (r > rs) {
                        float x = fminf(fmaxf((r - rs) / (rc - rs), 0.f), 1.f);
                        const float x2 = x * x, x3 = x2 * x;
                        const float S = 1.0f - 10.0f * x3 + 15.0f * x3 * x - 6.0f * x3 * x2;
                        const float dS = (-30.0f * x2 + 60.0f * x3 - 30.0f * x2 * x2) / (rc - rs);
                        fr = fr * S - U * dS / r;
                        U = U * S;
                    }
                }
                fx = __builtin_fmaf(fr, rx, fx);
                fy = __builtin_fmaf(fr, ry, fy);
                fz = __builtin_fmaf(fr, rz, fz);
                e += U;
                if (VIRIAL) {
                    w7[0] = __builtin_fmaf(fr, r2, w7[0]);
                    const float gx = fr * rx, gy = fr * ry;
                    w7[1] = __builtin_fmaf(gx, rx, w7[1]);
                    w7[2] = __builtin_fmaf(gy, ry, w7[2]);
                    w7[3] = __builtin_fmaf(fr * rz, rz, w7[3]);
                    w7[4] = __builtin_fmaf(gx, ry, w7[4]);
                    w7[5] = __builtin_fmaf(gx, rz, w7[5]);
                    w7[6] = __builtin_fmaf(gy, rz, w7[6]);
                }
            }
        }
        fx = lane_sum<LANES>(fx);
        fy = lane_sum<LANES>(fy);
        fz = lane_sum<LANES>(fz);
        acc[0] += 0.5 * (double)e;
        if (VIRIAL) {
#pragma unroll
            for (int c = 0; c < 7; ++c) acc[1 + c] += 0.5 * (double)w7[c];
        }
        if (a < n_home && l == 0) {
            const int i = g0 + a;
            force[(size_t)i * 3 + 0] = fx;
            force[(size_t)i * 3 + 1] = fy;
            force[(size_t)i * 3 + 2] = fz;
        }
    }
}

template <int CAP, int LANES, bool VIRIAL>
__global__ void __launch_bounds__(SIM_CELL_THREADS)
    k_lj_cell(const sim_replica_t* __restrict__ reps, const int* __restrict__ cell_rep,
              const int* __restrict__ cell_start, const int* __restrict__ cell_count,
              const double4* __restrict__ pos, const unsigned short* __restrict__ nbr16,
              const int* __restrict__ nnbr, int max_nbr, sim_lj_params_t lj,
              float* __restrict__ force, double* __restrict__ partial, int* overflow, float thr) {
    __shared__ Halo<CAP> H;
    const int cell = blockIdx.x;
    const sim_replica_t& P = reps[cell_rep[cell]];
    double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int local = cell - P.cell_base;
    const bool live = local < P.ncell[0] * P.ncell[1] * P.ncell[2] && cell_count[cell] > 0;
    int total = 0;
    const int st = live ? load_halo<CAP, 0>(H, P, cell, cell_start, cell_count, pos, nullptr,
                                            overflow, total)
                        : HALO_SKIP;
    if (st == HALO_LDS)
        lj_home<true, CAP, LANES, VIRIAL>(acc, H, pos, nbr16, nnbr, max_nbr, lj, force, thr);
    else if (st == HALO_GLOBAL)
        lj_home<false, CAP, LANES, VIRIAL>(acc, H, pos, nbr16, nnbr, max_nbr, lj, force, thr);
    block_sum<8>(acc, H.red);
    if (threadIdx.x == 0) {
        double* p = partial + (size_t)cell * SIM_NPART;
#pragma unroll
        for (int c = 0; c < 8; ++c) p[c] = acc[c];
    }
}

}  // namespace

#define SIM_CAP_DISPATCH(CAPV, CALL) \
    if (cap <= CAPV) {               \
        CALL(CAPV);                  \
    } else

extern "C" int sim_eam_density_cell(int ncells, const void* reps, const int* cell_rep,
                                    const int* cell_start, const int* cell_count,
                                    const double* pos4, const unsigned short* nbr16,
                                    const int* nnbr, int max_nbr, int cap, sim_eam_params_t eam,
                                    float* rho, float* fp, double* partial, int* overflow,
                                    void* stream) {
    if (ncells <= 0 || !reps || !cell_rep || !cell_start || !cell_count || !pos4 || !nbr16 ||
        !nnbr || max_nbr <= 0 || !rho || !fp || !partial || !overflow || !eam.pass1 ||
        eam.ntypes != 1 || cap <= 0 || cap > 2048)
        return SIM_E_ARG;
    const float thr = sqrt_threshold(eam.cutoff);
    auto* P = static_cast<const sim_replica_t*>(reps);
    auto* X = reinterpret_cast<const double4*>(pos4);
    hipStream_t st = as_stream(stream);
#define L_DENS(C)                                                                              \
    hipLaunchKernelGGL((k_eam_density_cell<C, 8>), dim3(ncells), dim3(SIM_CELL_THREADS), 0, st, \
                       P, cell_rep, cell_start, cell_count, X, nbr16, nnbr, max_nbr, eam, rho, \
                       fp, partial, overflow, thr)
    SIM_CAP_DISPATCH(512, L_DENS) SIM_CAP_DISPATCH(1024, L_DENS) SIM_CAP_DISPATCH(1280, L_DENS)
    SIM_CAP_DISPATCH(1536, L_DENS) SIM_CAP_DISPATCH(2048, L_DENS) {}
#undef L_DENS
    SIM_CHECK_LAUNCH();
    return SIM_OK;
}

extern "C" int sim_eam_force_cell(int ncells, const void* reps, const int* cell_rep,
                                  const int* cell_start, const int* cell_count, const double* pos4,
                                  const unsigned short* nbr16, const int* nnbr, int max_nbr,
                                  int cap, sim_eam_params_t eam, const float* fp, float* force,
                                  double* partial, int virial, int* overflow, void* stream) {
    if (ncells <= 0 || !reps || !cell_rep || !cell_start || !cell_count || !pos4 || !nbr16 ||
        !nnbr || max_nbr <= 0 || !fp || !force || !partial || !overflow || !eam.pass2 ||
        eam.ntypes != 1 || cap <= 0 || cap > 2048)
        return SIM_E_ARG;
    const float thr = sqrt_threshold(eam.cutoff);
    auto* P = static_cast<const sim_replica_t*>(reps);
    auto* X = reinterpret_cast<const double4*>(pos4);
    hipStream_t st = as_stream(stream);
#define L_FV(C)                                                                                  \
    hipLaunchKernelGGL((k_eam_force_cell<C, 8, true>), dim3(ncells), dim3(SIM_CELL_THREADS), 0, \
                       st, P, cell_rep, cell_start, cell_count, X, nbr16, nnbr, max_nbr, eam, fp, \
                       force, partial, overflow, thr)
#define L_F(C)                                                                                    \
    hipLaunchKernelGGL((k_eam_force_cell<C, 8, false>), dim3(ncells), dim3(SIM_CELL_THREADS), 0, \
                       st, P, cell_rep, cell_start, cell_count, X, nbr16, nnbr, max_nbr, eam, fp,  \
                       force, partial, overflow, thr)
    if (virial) {
        SIM_CAP_DISPATCH(512, L_FV) SIM_CAP_DISPATCH(1024, L_FV) SIM_CAP_DISPATCH(1280, L_FV)
        SIM_CAP_DISPATCH(1536, L_FV) SIM_CAP_DISPATCH(2048, L_FV) {}
    } else {
        SIM_CAP_DISPATCH(512, L_F) SIM_CAP_DISPATCH(1024, L_F) SIM_CAP_DISPATCH(1280, L_F)
        SIM_CAP_DISPATCH(1536, L_F) SIM_CAP_DISPATCH(2048, L_F) {}
    }
#undef L_FV
#undef L_F
    SIM_CHECK_LAUNCH();
    return SIM_OK;
}

extern "C" int sim_lj_force_cell(int ncells, const void* reps, const int* cell_rep,
                                 const int* cell_start, const int* cell_count, const double* pos4,
                                 const unsigned short* nbr16, const int* nnbr, int max_nbr,
                                 int cap, sim_lj_params_t lj, float* force, double* partial,
                                 int virial, int* overflow, void* stream) {
    if (ncells <= 0 || !reps || !cell_rep || !cell_start || !cell_count || !pos4 || !nbr16 ||
        !nnbr || max_nbr <= 0 || !force || !partial || !overflow || lj.mode < 0 || lj.mode > 2 ||
        !(lj.rc > 0.f) || cap <= 0 || cap > 2048)
        return SIM_E_ARG;
    const float thr = sqrt_threshold(lj.rc);
    auto* P = static_cast<const sim_replica_t*>(reps);
    auto* X = reinterpret_cast<const double4*>(pos4);
    hipStream_t st = as_stream(stream);
#define L_LV(C)                                                                                   \
    hipLaunchKernelGGL((k_lj_cell<C, 8, true>), dim3(ncells), dim3(SIM_CELL_THREADS), 0, st, P,  \
                       cell_rep, cell_start, cell_count, X, nbr16, nnbr, max_nbr, lj, force,     \
                       partial, overflow, thr)
#define L_L(C)                                                                                    \
    hipLaunchKernelGGL((k_lj_cell<C, 8, false>), dim3(ncells), dim3(SIM_CELL_THREADS), 0, st, P, \
                       cell_rep, cell_start, cell_count, X, nbr16, nnbr, max_nbr, lj, force,     \
                       partial, overflow, thr)
    if (virial) {
        SIM_CAP_DISPATCH(512, L_LV) SIM_CAP_DISPATCH(1024, L_LV) SIM_CAP_DISPATCH(1280, L_LV)
        SIM_CAP_DISPATCH(1536, L_LV) SIM_CAP_DISPATCH(2048, L_LV) {}
    } else {
        SIM_CAP_DISPATCH(512, L_L) SIM_CAP_DISPATCH(1024, L_L) SIM_CAP_DISPATCH(1280, L_L)
        SIM_CAP_DISPATCH(1536, L_L) SIM_CAP_DISPATCH(2048, L_L) {}
    }
#undef L_LV
#undef L_L
    SIM_CHECK_LAUNCH();
    return SIM_OK;
}
