// Row pair kernels (gfx950): LJ, EAM density and EAM force over rows of cells.
//
// After a rebuild the atoms of every cell are contiguous and in cell order
// (sim_nl_build with a reorder), wrapped into the primary cell, and each
// atom's Verlet list also exists as 16-bit column slots: (stencil column
// 0..2 for x-1, x, x+1) << 14 | offset of the neighbour within that column,
// where a column is the 9 cells (dz, dy in -1..1, z-major) at one x.
//
// One workgroup walks one row of cells along x (fixed y, z of a replica). A
// ring of 4 columns lives in LDS as float64 positions already translated to
// the image next to the row (a pair displacement is a plain float64
// subtraction), plus an auxiliary float where needed (atom type, or F'(rho)
// for the EAM force pass). Moving from cell x to x+1 needs one new column,
// so each cell costs 9 cells of halo traffic instead of 27; the next column's
// loads are issued before the current cell's pair loop and stored after it.
// LANES lanes share each home atom; per-atom sums combine by wave shuffles;
// energy and virial leave as one float64 partial per cell slot, written in a
// fixed order (deterministic).
//
// A column holding more atoms than the LDS ring slot (a dense fluctuation in
// a hot system) is not staged: cells that need it read their neighbours from
// global memory through the column table with the same float64 arithmetic
// (bit-identical results), and overflow bit 2 reports that this happened.
// Cell slots beyond a replica's current grid get zero partials.
//
// Physics and masks are those of pair_lj.hip / pair_eam.hip
// (core/force/lennard_jones_force.py:15-95, core/force/eam_force_cu.py:461-579).
#include <type_traits>

#include "pair_common.h"

#define SIM_ROW_THREADS 256
#ifndef SIM_ROW_REGS
#define SIM_ROW_REGS 0   // 1: hold the next column in registers across the pair loop
#endif

namespace {

// The 9 cells of one stencil column.
struct ColMeta {
    int start[9];
    int count[9];
    int pre[10];        // prefix of count; pre[9] = population
    double sh[9][3];    // periodic translation of each cell
};

// AUX: 0 none, 1 atom type from pos.w, 2 from the aux array (F'(rho)).
template <int COLCAP, int AUX>
struct RowLDS {
    double x[4][COLCAP], y[4][COLCAP], z[4][COLCAP];
    float aux[4][AUX ? COLCAP : 1];
    ColMeta meta[4];
    int fits[4];
    double red[4 * 8];
};

constexpr int per_thread(int COLCAP) { return (COLCAP + SIM_ROW_THREADS - 1) / SIM_ROW_THREADS; }

// Table entry of column col (any integer; wrapped periodically) cell m of
// the row (y0, z0) of replica P.
__device__ __forceinline__ void col_cell(const sim_replica_t& P, int y0, int z0, int col, int m,
                                         const int* __restrict__ cell_start,
                                         const int* __restrict__ cell_count, int& start, int& count,
                                         double (&sh)[3]) {
    const int nx = P.ncell[0], ny = P.ncell[1], nz = P.ncell[2];
    int xx = col, yy = y0 + m % 3 - 1, zz = z0 + m / 3 - 1;
    const int fx = xx < 0 ? -1 : (xx >= nx ? 1 : 0);
    const int fy = yy < 0 ? -1 : (yy >= ny ? 1 : 0);
    const int fz = zz < 0 ? -1 : (zz >= nz ? 1 : 0);
    xx -= fx * nx;
    yy -= fy * ny;
    zz -= fz * nz;
    const int c = P.cell_base + (zz * ny + yy) * nx + xx;
    start = cell_start[c];
    count = cell_count[c];
#pragma unroll
    for (int k = 0; k < 3; ++k) sh[k] = fx * P.h[k] + fy * P.h[3 + k] + fz * P.h[6 + k];
}

// Lanes 0..8 of a wave write column `col`'s table into ring slot col & 3
// (with the prefix by a shuffle scan). Every lane of the wave must call.
template <int COLCAP, int AUX>
__device__ __forceinline__ void put_meta(RowLDS<COLCAP, AUX>& L, int lane, int start, int count,
                                         const double (&sh)[3], int slot) {
    int incl = lane < 9 ? count : 0;
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
        const int t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    ColMeta& M = L.meta[slot];
    if (lane < 9) {
        M.start[lane] = start;
        M.count[lane] = count;
        M.pre[lane] = incl - count;
        M.sh[lane][0] = sh[0];
        M.sh[lane][1] = sh[1];
        M.sh[lane][2] = sh[2];
    }
    if (lane == 8) {
        M.pre[9] = incl;
        L.fits[slot] = incl <= COLCAP;
    }
}

__device__ __forceinline__ int col_find(const ColMeta& M, int f) {
    int lo = 0, hi = 8;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (M.pre[mid] <= f) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Registers holding one column's atoms between their load and their store.
template <int COLCAP>
struct ColRegs {
    double4 v[per_thread(COLCAP)];
    float a[per_thread(COLCAP)];
    int m[per_thread(COLCAP)];
};

template <int COLCAP, int AUX>
__device__ __forceinline__ void col_load(const RowLDS<COLCAP, AUX>& L, int slot,
                                         const double4* __restrict__ pos,
                                         const float* __restrict__ aux, ColRegs<COLCAP>& C) {
    const ColMeta& M = L.meta[slot];
    const int total = L.fits[slot] ? M.pre[9] : 0;
#pragma unroll
    for (int q = 0; q < per_thread(COLCAP); ++q) {
        const int f = threadIdx.x + q * SIM_ROW_THREADS;
        if (f < total) {
            const int m = col_find(M, f);
            const int j = M.start[m] + (f - M.pre[m]);
            C.m[q] = m;
            C.v[q] = pos[j];
            if (AUX == 2) C.a[q] = aux[j];
        }
    }
}

template <int COLCAP, int AUX>
__device__ __forceinline__ void col_store(RowLDS<COLCAP, AUX>& L, int slot,
                                          const ColRegs<COLCAP>& C) {
    const ColMeta& M = L.meta[slot];
    const int total = L.fits[slot] ? M.pre[9] : 0;
#pragma unroll
    for (int q = 0; q < per_thread(COLCAP); ++q) {
        const int f = threadIdx.x + q * SIM_ROW_THREADS;
        if (f < total) {
            const int m = C.m[q];
            L.x[slot][f] = C.v[q].x + M.sh[m][0];
            L.y[slot][f] = C.v[q].y + M.sh[m][1];
            L.z[slot][f] = C.v[q].z + M.sh[m][2];
            if (AUX == 1) L.aux[slot][f] = (float)C.v[q].w;
            if (AUX == 2) L.aux[slot][f] = C.a[q];
        }
    }
}

// Neighbour (column slot s16 of home cell x): translated position and aux.
template <bool FROM_LDS, int COLCAP, int AUX>
__device__ __forceinline__ void nb_get(const RowLDS<COLCAP, AUX>& L, int x, int s16,
                                       const double4* __restrict__ pos,
                                       const float* __restrict__ aux, double& px, double& py,
                                       double& pz, float& a) {
    const int slot = (x - 1 + (s16 >> 14)) & 3;
    const int off = s16 & 0x3fff;
    if (FROM_LDS) {
        px = L.x[slot][off];
        py = L.y[slot][off];
        pz = L.z[slot][off];
        a = AUX ? L.aux[slot][off] : 0.f;
        return;
    }
    const ColMeta& M = L.meta[slot];
    const int m = col_find(M, off);
    const int j = M.start[m] + (off - M.pre[m]);
    const double4 p = pos[j];
    px = p.x + M.sh[m][0];
    py = p.y + M.sh[m][1];
    pz = p.z + M.sh[m][2];
    a = AUX == 2 ? aux[j] : (AUX == 1 ? (float)p.w : 0.f);
}

template <int LANES>
__device__ __forceinline__ float lane_sum(float v) {
#pragma unroll
    for (int o = 1; o < LANES; o <<= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Walk the row (y0, z0) of replica P: cell_fn(x, from_lds_tag) computes home
// cell x (all threads call it; it may synchronise).
template <int COLCAP, int AUX, typename CellFn>
__device__ __forceinline__ void walk_row(RowLDS<COLCAP, AUX>& L, const sim_replica_t& P, int y0,
                                         int z0, const int* __restrict__ cell_start,
                                         const int* __restrict__ cell_count,
                                         const double4* __restrict__ pos,
                                         const float* __restrict__ aux, int* overflow,
                                         CellFn&& cell_fn) {
    const int nx = P.ncell[0];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    // tables of columns -1, 0, 1, 2 (one wave each)
    {
        const int col = wid - 1;
        int st = 0, cn = 0;
        double sh[3] = {0.0, 0.0, 0.0};
        if (lane < 9 && col <= nx) col_cell(P, y0, z0, col, lane, cell_start, cell_count, st, cn, sh);
        put_meta(L, lane, st, cn, sh, col & 3);
    }
    __syncthreads();
    {
        ColRegs<COLCAP> C;
        for (int col = -1; col <= 1; ++col) {
            col_load(L, col & 3, pos, aux, C);
            col_store(L, col & 3, C);
        }
    }
    __syncthreads();
    for (int x = 0; x < nx; ++x) {
        // next column's atoms (its table is in LDS; its ring slot, that of
        // column x-2, is free) and the table after it
        const bool next = x + 2 <= nx;
#if SIM_ROW_REGS
        ColRegs<COLCAP> C;
        if (next) col_load(L, (x + 2) & 3, pos, aux, C);
#else
        if (next) {
            ColRegs<COLCAP> C;
            col_load(L, (x + 2) & 3, pos, aux, C);
            col_store(L, (x + 2) & 3, C);
        }
#endif
        int st = 0, cn = 0;
        double sh[3] = {0.0, 0.0, 0.0};
        const bool next2 = x + 3 <= nx;
        if (wid == 0 && next2 && lane < 9)
            col_cell(P, y0, z0, x + 3, lane, cell_start, cell_count, st, cn, sh);
        const bool lds = L.fits[(x - 1) & 3] && L.fits[x & 3] && L.fits[(x + 1) & 3];
        if (!lds && tid == 0) atomicOr(overflow, 2);
        if (lds)
            cell_fn(x, std::true_type());
        else
            cell_fn(x, std::false_type());
        __syncthreads();   // column x-1 (ring slot of x+3) no longer read
#if SIM_ROW_REGS
        if (next) col_store(L, (x + 2) & 3, C);
#endif
        if (wid == 0 && next2) put_meta(L, lane, st, cn, sh, (x + 3) & 3);
        __syncthreads();
    }
}

// Replica of global row `row` (binary search over row_base[0..R]).
__device__ __forceinline__ int row_replica(const int* __restrict__ row_base, int R, int row) {
    int lo = 0, hi = R - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (row_base[mid] <= row) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Zero columns [c0, c1) of the partial rows of cell slots outside their
// replica's current grid.
__device__ __forceinline__ void zero_slack(const sim_replica_t* __restrict__ reps,
                                           const int* __restrict__ cell_rep, int ncells,
                                           double* __restrict__ partial, int c0, int c1) {
    for (int c = blockIdx.x * SIM_ROW_THREADS + threadIdx.x; c < ncells;
         c += gridDim.x * SIM_ROW_THREADS) {
        const sim_replica_t& P = reps[cell_rep[c]];
        if (c - P.cell_base >= P.ncell[0] * P.ncell[1] * P.ncell[2])
            for (int k = c0; k < c1; ++k) partial[(size_t)c * SIM_NPART + k] = 0.0;
    }
}

// Row of this workgroup: returns false when there is none (or the grid is
// too small for three distinct columns; overflow bit 4).
__device__ __forceinline__ bool my_row(const sim_replica_t* __restrict__ reps,
                                       const int* __restrict__ row_base, int R, int* overflow,
                                       int& rep, int& y0, int& z0) {
    const int row = blockIdx.x;
    if (row >= row_base[R]) return false;
    rep = row_replica(row_base, R, row);
    const sim_replica_t& P = reps[rep];
    if (P.ncell[0] < 3 || P.ncell[1] < 3 || P.ncell[2] < 3) {
        if (threadIdx.x == 0) atomicOr(overflow, 4);
        return false;
    }
    const int q = row - row_base[rep];
    y0 = q % P.ncell[1];
    z0 = q / P.ncell[1];
    return true;
}

__device__ __forceinline__ int cell_slot(const sim_replica_t& P, int x, int y0, int z0) {
    return P.cell_base + (z0 * P.ncell[1] + y0) * P.ncell[0] + x;
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

template <int COLCAP, int LANES>
__global__ void __launch_bounds__(SIM_ROW_THREADS)
    k_eam_density_rows(const sim_replica_t* __restrict__ reps, const int* __restrict__ cell_rep,
                       int ncells, const int* __restrict__ row_base, int R,
                       const int* __restrict__ cell_start, const int* __restrict__ cell_count,
                       const double4* __restrict__ pos, const unsigned short* __restrict__ nbr16,
                       const int* __restrict__ nnbr, int max_nbr, sim_eam_params_t E,
                       float* __restrict__ rho_out, float* __restrict__ fp_out,
                       double* __restrict__ partial, int* overflow, float thr) {
    __shared__ RowLDS<COLCAP, 0> L;
    zero_slack(reps, cell_rep, ncells, partial, SIM_T_EPOT, SIM_T_EPOT + 1);
    int rep, y0, z0;
    if (!my_row(reps, row_base, R, overflow, rep, y0, z0)) return;
    const sim_replica_t& P = reps[rep];
    const float4* __restrict__ T1 = static_cast<const float4*>(E.pass1);
    const int l = threadIdx.x % LANES;
    walk_row(L, P, y0, z0, cell_start, cell_count, pos, nullptr, overflow,
             [&](int x, auto tag) {
        constexpr bool FL = decltype(tag)::value;
        const ColMeta& M = L.meta[x & 3];
        const int n_home = M.count[4], hoff = M.pre[4], g0 = M.start[4];
        double etot = 0.0;
        for (int b = 0; b < n_home; b += SIM_ROW_THREADS / LANES) {
            const int a = b + threadIdx.x / LANES;
            const int i = g0 + a;
            float rho = 0.f, ep = 0.f;
            if (a < n_home) {
                const int self = (1 << 14) | (hoff + a);
                double xi, yi, zi;
                float ai;
                nb_get<FL>(L, x, self, pos, nullptr, xi, yi, zi, ai);
                const int cnt = nnbr[i];
                const unsigned short* row = nbr16 + (size_t)i * max_nbr;
                constexpr int U = SIM_NB_BATCH;
                int sn[U];
#pragma unroll
                for (int u = 0; u < U; ++u)
                    sn[u] = l + u * LANES < cnt ? (int)row[l + u * LANES] : self;
                for (int k = l; k < cnt; k += U * LANES) {
                    int sl[U];
                    float r2[U];
                    LerpT Lt[U];
                    float4 v[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) sl[u] = sn[u];
#pragma unroll
                    for (int u = 0; u < U; ++u) {   // next batch's slots
                        const int kk = k + (U + u) * LANES;
                        sn[u] = kk < cnt ? (int)row[kk] : self;
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        double px, py, pz;
                        float aj;
                        nb_get<FL>(L, x, sl[u], pos, nullptr, px, py, pz, aj);
                        const float rx = (float)(xi - px), ry = (float)(yi - py),
                                    rz = (float)(zi - pz);
                        r2[u] = sq3(rx, ry, rz);
                        Lt[u] = lerp_idx(pair_sqrt(r2[u]), E);
                        v[u] = T1[Lt[u].idx];   // one element type (checked at launch)
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const float w = (k + u * LANES < cnt && r2[u] < thr) ? 1.0f : 0.0f;
                        rho += w * lp(v[u].x, v[u].y, Lt[u].frac);
                        ep += w * lp(v[u].z, v[u].w, Lt[u].frac);
                    }
                }
            }
            rho = lane_sum<LANES>(rho);
            ep = lane_sum<LANES>(ep);
            if (a < n_home && l == 0) {
                const float lo = E.rho_lo[0];
                const float rcl = fminf(fmaxf(rho, lo), E.rho_hi[0]);
                const float xf = (rcl - lo) * E.inv_drho[0];
                const int idx = (int)xf;
                const float fr = xf - (float)idx;
                const float2 F = reinterpret_cast<const float2*>(E.emb)[idx];
                const float2 dF = reinterpret_cast<const float2*>(E.demb)[idx];
                rho_out[i] = rho;
                fp_out[i] = lp(dF.x, dF.y, fr);
                etot += (double)lp(F.x, F.y, fr) + 0.5 * (double)ep;
            }
        }
        double v1[1] = {etot};
        block_sum<1>(v1, L.red);
        if (threadIdx.x == 0) partial[(size_t)cell_slot(P, x, y0, z0) * SIM_NPART + SIM_T_EPOT] = v1[0];
    });
}

template <int COLCAP, int LANES, bool VIRIAL>
__global__ void __launch_bounds__(SIM_ROW_THREADS)
    k_eam_force_rows(const sim_replica_t* __restrict__ reps, const int* __restrict__ cell_rep,
                     int ncells, const int* __restrict__ row_base, int R,
                     const int* __restrict__ cell_start, const int* __restrict__ cell_count,
                     const double4* __restrict__ pos, const unsigned short* __restrict__ nbr16,
                     const int* __restrict__ nnbr, int max_nbr, sim_eam_params_t E,
                     const float* __restrict__ fp, float* __restrict__ force,
                     double* __restrict__ partial, int* overflow, float thr) {
    __shared__ RowLDS<COLCAP, 2> L;
    zero_slack(reps, cell_rep, ncells, partial, SIM_T_VIRIAL, SIM_T_VIRIAL + 7);
    int rep, y0, z0;
    if (!my_row(reps, row_base, R, overflow, rep, y0, z0)) return;
    const sim_replica_t& P = reps[rep];
    const float4* __restrict__ T2 = static_cast<const float4*>(E.pass2);
    const int l = threadIdx.x % LANES;
    walk_row(L, P, y0, z0, cell_start, cell_count, pos, fp, overflow, [&](int x, auto tag) {
        constexpr bool FL = decltype(tag)::value;
        const ColMeta& M = L.meta[x & 3];
        const int n_home = M.count[4], hoff = M.pre[4], g0 = M.start[4];
        double acc[7] = {0, 0, 0, 0, 0, 0, 0};
        for (int b = 0; b < n_home; b += SIM_ROW_THREADS / LANES) {
            const int a = b + threadIdx.x / LANES;
            const int i = g0 + a;
            float fx = 0.f, fy = 0.f, fz = 0.f;
            float w7[7] = {0, 0, 0, 0, 0, 0, 0};   // virial, xx yy zz xy xz yz
            if (a < n_home) {
                const int self = (1 << 14) | (hoff + a);
                double xi, yi, zi;
                float fpi;
                nb_get<FL>(L, x, self, pos, fp, xi, yi, zi, fpi);
                const int cnt = nnbr[i];
                const unsigned short* row = nbr16 + (size_t)i * max_nbr;
                constexpr int U = SIM_NB_BATCH;
                int sn[U];
#pragma unroll
                for (int u = 0; u < U; ++u)
                    sn[u] = l + u * LANES < cnt ? (int)row[l + u * LANES] : self;
                for (int k = l; k < cnt; k += U * LANES) {
                    int sl[U];
                    float rx[U], ry[U], rz[U], r2[U], fpj[U];
                    LerpT Lt[U];
                    float4 va[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) sl[u] = sn[u];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int kk = k + (U + u) * LANES;
                        sn[u] = kk < cnt ? (int)row[kk] : self;
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        double px, py, pz;
                        nb_get<FL>(L, x, sl[u], pos, fp, px, py, pz, fpj[u]);
                        rx[u] = (float)(xi - px);
                        ry[u] = (float)(yi - py);
                        rz[u] = (float)(zi - pz);
                        r2[u] = sq3(rx[u], ry[u], rz[u]);
                        if (!(k + u * LANES < cnt)) r2[u] = 3.0e38f;   // masked: fails r2 < thr
                        Lt[u] = lerp_idx(pair_sqrt(r2[u]), E);
                        va[u] = T2[Lt[u].idx];   // one element type (checked at launch)
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const float r = pair_sqrt(r2[u]);
                        const float w = r2[u] < thr ? 1.0f : 0.0f;
                        const float dd = lp(va[u].x, va[u].y, Lt[u].frac);
                        const float s = w * (fpi * dd + fpj[u] * dd + lp(va[u].z, va[u].w, Lt[u].frac));
                        const float g = -s * __builtin_amdgcn_rcpf(r + 1e-8f);
                        fx = __builtin_fmaf(g, rx[u], fx);
                        fy = __builtin_fmaf(g, ry[u], fy);
                        fz = __builtin_fmaf(g, rz[u], fz);
                        if (VIRIAL) {
                            w7[0] = __builtin_fmaf(s, r, w7[0]);
                            const float gx = g * rx[u], gy = g * ry[u];
                            w7[1] = __builtin_fmaf(gx, rx[u], w7[1]);
                            w7[2] = __builtin_fmaf(gy, ry[u], w7[2]);
                            w7[3] = __builtin_fmaf(g * rz[u], rz[u], w7[3]);
                            w7[4] = __builtin_fmaf(gx, ry[u], w7[4]);
                            w7[5] = __builtin_fmaf(gx, rz[u], w7[5]);
                            w7[6] = __builtin_fmaf(gy, rz[u], w7[6]);
                        }
                    }
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
                force[(size_t)i * 3 + 0] = fx;
                force[(size_t)i * 3 + 1] = fy;
                force[(size_t)i * 3 + 2] = fz;
            }
        }
        block_sum<7>(acc, L.red);
        if (threadIdx.x == 0) {
            double* p = partial + (size_t)cell_slot(P, x, y0, z0) * SIM_NPART;
#pragma unroll
            for (int c = 0; c < 7; ++c) p[SIM_T_VIRIAL + c] = acc[c];
        }
    });
}

// ---- LJ -------------------------------------------------------------------------
template <int COLCAP, int LANES, bool VIRIAL, int AUX>
__global__ void __launch_bounds__(SIM_ROW_THREADS)
    k_lj_rows(const sim_replica_t* __restrict__ reps, const int* __restrict__ cell_rep, int ncells,
              const int* __restrict__ row_base, int R, const int* __restrict__ cell_start,
              const int* __restrict__ cell_count, const double4* __restrict__ pos,
              const unsigned short* __restrict__ nbr16, const int* __restrict__ nnbr, int max_nbr,
              sim_lj_params_t lj, float* __restrict__ force, double* __restrict__ partial,
              int* overflow, float thr) {
    __shared__ RowLDS<COLCAP, AUX> L;
    zero_slack(reps, cell_rep, ncells, partial, 0, 8);
    int rep, y0, z0;
    if (!my_row(reps, row_base, R, overflow, rep, y0, z0)) return;
    const sim_replica_t& P = reps[rep];
    const int l = threadIdx.x % LANES;
    walk_row(L, P, y0, z0, cell_start, cell_count, pos, nullptr, overflow, [&](int x, auto tag) {
        constexpr bool FL = decltype(tag)::value;
        const ColMeta& M = L.meta[x & 3];
        const int n_home = M.count[4], hoff = M.pre[4], g0 = M.start[4];
        double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int b = 0; b < n_home; b += SIM_ROW_THREADS / LANES) {
            const int a = b + threadIdx.x / LANES;
            const int i = g0 + a;
            float fx = 0.f, fy = 0.f, fz = 0.f, e = 0.f;
            float w7[7] = {0, 0, 0, 0, 0, 0, 0};
            if (a < n_home) {
                const int self = (1 << 14) | (hoff + a);
                double xi, yi, zi;
                float ai;
                nb_get<FL>(L, x, self, pos, nullptr, xi, yi, zi, ai);
                const int ti = (int)ai;
                const int cnt = nnbr[i];
                const unsigned short* row = nbr16 + (size_t)i * max_nbr;
                constexpr int U = SIM_NB_BATCH;
                int sn[U];
#pragma unroll
                for (int u = 0; u < U; ++u)
                    sn[u] = l + u * LANES < cnt ? (int)row[l + u * LANES] : self;
                for (int k = l; k < cnt; k += U * LANES) {
                    int sl[U];
                    float rx[U], ry[U], rz[U], tj[U];
#pragma unroll
                    for (int u = 0; u < U; ++u) sl[u] = sn[u];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const int kk = k + (U + u) * LANES;
                        sn[u] = kk < cnt ? (int)row[kk] : self;
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        double px, py, pz;
                        nb_get<FL>(L, x, sl[u], pos, nullptr, px, py, pz, tj[u]);
                        rx[u] = (float)(xi - px);
                        ry[u] = (float)(yi - py);
                        rz[u] = (float)(zi - pz);
                    }
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const float r2 = sq3(rx[u], ry[u], rz[u]);
                        if (!(k + u * LANES < cnt) || !(r2 < thr)) continue;
                        const int t2 = ti * SIM_MAX_TYPES + (int)tj[u];
                        const float eps = lj.eps[t2];
                        const float sig = lj.sig[t2];
                        const float inv_r2 = __builtin_amdgcn_rcpf(r2);   // 1 ulp
                        const float s2 = sig * sig * inv_r2;
                        const float s6 = s2 * s2 * s2;
                        const float s12 = s6 * s6;
                        float Ue = 4.0f * eps * (s12 - s6);
                        float fr = 24.0f * eps * inv_r2 * (2.0f * s12 - s6);
                        if (lj.mode == 1) {
                            const float rc = lj.rc, r = sqrtf(r2);
                            const float src = sig / rc;
                            const float q6 = src * src * src * src * src * src;
                            const float q12 = q6 * q6;
                            const float Frc = 24.0f * eps * (1.0f / rc) * (2.0f * q12 - q6);
                            Ue = Ue - 4.0f * eps * (q12 - q6) + (r - rc) * Frc;
                            fr = fr - Frc / r;
                        } else if (lj.mode == 2) {
                            const float rc = lj.rc, rs = lj.rs, r = sqrtf(r2);
                            if (r > rs) {
                                float xs = fminf(fmaxf((r - rs) / (rc - rs), 0.f), 1.f);
                                const float x2 = xs * xs, x3 = x2 * xs;
                                const float Sw = 1.0f - 10.0f * x3 + 15.0f * x3 * xs - 6.0f * x3 * x2;
                                const float dS = (-30.0f * x2 + 60.0f * x3 - 30.0f * x2 * x2) / (rc - rs);
                                fr = fr * Sw - Ue * dS / r;
                                Ue = Ue * Sw;
                            }
                        }
                        fx = __builtin_fmaf(fr, rx[u], fx);
                        fy = __builtin_fmaf(fr, ry[u], fy);
                        fz = __builtin_fmaf(fr, rz[u], fz);
                        e += Ue;
                        if (VIRIAL) {
                            w7[0] = __builtin_fmaf(fr, r2, w7[0]);
                            const float gx = fr * rx[u], gy = fr * ry[u];
                            w7[1] = __builtin_fmaf(gx, rx[u], w7[1]);
                            w7[2] = __builtin_fmaf(gy, ry[u], w7[2]);
                            w7[3] = __builtin_fmaf(fr * rz[u], rz[u], w7[3]);
                            w7[4] = __builtin_fmaf(gx, ry[u], w7[4]);
                            w7[5] = __builtin_fmaf(gx, rz[u], w7[5]);
                            w7[6] = __builtin_fmaf(gy, rz[u], w7[6]);
                        }
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
                force[(size_t)i * 3 + 0] = fx;
                force[(size_t)i * 3 + 1] = fy;
                force[(size_t)i * 3 + 2] = fz;
            }
        }
        block_sum<8>(acc, L.red);
        if (threadIdx.x == 0) {
            double* p = partial + (size_t)cell_slot(P, x, y0, z0) * SIM_NPART;
#pragma unroll
            for (int c = 0; c < 8; ++c) p[c] = acc[c];
        }
    });
}

bool rows_args_ok(int ncells, const void* reps, const int* cell_rep, const int* row_base, int R,
                  const int* cell_start, const int* cell_count, const double* pos4,
                  const unsigned short* nbr16, const int* nnbr, int max_nbr, int cap,
                  const int* overflow) {
    return ncells > 0 && reps && cell_rep && row_base && R > 0 && cell_start && cell_count &&
           pos4 && nbr16 && nnbr && max_nbr > 0 && overflow && cap > 0 && cap <= 512;
}

}  // namespace

// Launch one workgroup per possible row: with >= 3 cells along x there are at
// most ncells / 3 rows; the rest only zero slack partials.
#define SIM_ROWS_GRID dim3((ncells + 2) / 3)
#define SIM_COLCAP_DISPATCH(CAPV, CALL) \
    if (cap <= CAPV) {                  \
        CALL(CAPV);                     \
    } else

extern "C" int sim_eam_density_rows(int ncells, const void* reps, const int* cell_rep,
                                    const int* row_base, int R, const int* cell_start,
                                    const int* cell_count, const double* pos4,
                                    const unsigned short* nbr16, const int* nnbr, int max_nbr,
                                    int cap, sim_eam_params_t eam, float* rho, float* fp,
                                    double* partial, int* overflow, void* stream) {
    if (!rows_args_ok(ncells, reps, cell_rep, row_base, R, cell_start, cell_count, pos4, nbr16,
                      nnbr, max_nbr, cap, overflow) ||
        !rho || !fp || !partial || !eam.pass1 || eam.ntypes != 1)
        return SIM_E_ARG;
    const float thr = sqrt_threshold(eam.cutoff);
    auto* P = static_cast<const sim_replica_t*>(reps);
    auto* X = reinterpret_cast<const double4*>(pos4);
    hipStream_t st = as_stream(stream);
#define L_DENS(C)                                                                               \
    hipLaunchKernelGGL((k_eam_density_rows<C, 8>), SIM_ROWS_GRID, dim3(SIM_ROW_THREADS), 0, st, \
                       P, cell_rep, ncells, row_base, R, cell_start, cell_count, X, nbr16, nnbr, \
                       max_nbr, eam, rho, fp, partial, overflow, thr)
    SIM_COLCAP_DISPATCH(192, L_DENS) SIM_COLCAP_DISPATCH(256, L_DENS)
    SIM_COLCAP_DISPATCH(320, L_DENS) SIM_COLCAP_DISPATCH(384, L_DENS)
    SIM_COLCAP_DISPATCH(512, L_DENS) {}
#undef L_DENS
    SIM_CHECK_LAUNCH();
    return SIM_OK;
}

extern "C" int sim_eam_force_rows(int ncells, const void* reps, const int* cell_rep,
                                  const int* row_base, int R, const int* cell_start,
                                  const int* cell_count, const double* pos4,
                                  const unsigned short* nbr16, const int* nnbr, int max_nbr,
                                  int cap, sim_eam_params_t eam, const float* fp, float* force,
                                  double* partial, int virial, int* overflow, void* stream) {
    if (!rows_args_ok(ncells, reps, cell_rep, row_base, R, cell_start, cell_count, pos4, nbr16,
                      nnbr, max_nbr, cap, overflow) ||
        !fp || !force || !partial || !eam.pass2 || eam.ntypes != 1)
        return SIM_E_ARG;
    const float thr = sqrt_threshold(eam.cutoff);
    auto* P = static_cast<const sim_replica_t*>(reps);
    auto* X = reinterpret_cast<const double4*>(pos4);
    hipStream_t st = as_stream(stream);
#define L_FV(C)                                                                                 \
    hipLaunchKernelGGL((k_eam_force_rows<C, 8, true>), SIM_ROWS_GRID, dim3(SIM_ROW_THREADS), 0, \
                       st, P, cell_rep, ncells, row_base, R, cell_start, cell_count, X, nbr16,  \
                       nnbr, max_nbr, eam, fp, force, partial, overflow, thr)
#define L_F(C)                                                                                   \
    hipLaunchKernelGGL((k_eam_force_rows<C, 8, false>), SIM_ROWS_GRID, dim3(SIM_ROW_THREADS), 0, \
                       st, P, cell_rep, ncells, row_base, R, cell_start, cell_count, X, nbr16,   \
                       nnbr, max_nbr, eam, fp, force, partial, overflow, thr)
    if (virial) {
        SIM_COLCAP_DISPATCH(192, L_FV) SIM_COLCAP_DISPATCH(256, L_FV)
        SIM_COLCAP_DISPATCH(320, L_FV) SIM_COLCAP_DISPATCH(384, L_FV)
        SIM_COLCAP_DISPATCH(512, L_FV) {}
    } else {
        SIM_COLCAP_DISPATCH(192, L_F) SIM_COLCAP_DISPATCH(256, L_F)
        SIM_COLCAP_DISPATCH(320, L_F) SIM_COLCAP_DISPATCH(384, L_F)
        SIM_COLCAP_DISPATCH(512, L_F) {}
    }
#undef L_FV
#undef L_F
    SIM_CHECK_LAUNCH();
    return SIM_OK;
}

extern "C" int sim_lj_force_rows(int ncells, const void* reps, const int* cell_rep,
                                 const int* row_base, int R, const int* cell_start,
                                 const int* cell_count, const double* pos4,
                                 const unsigned short* nbr16, const int* nnbr, int max_nbr, int cap,
                                 int ntypes, sim_lj_params_t lj, float* force, double* partial,
                                 int virial, int* overflow, void* stream) {
    if (!rows_args_ok(ncells, reps, cell_rep, row_base, R, cell_start, cell_count, pos4, nbr16,
                      nnbr, max_nbr, cap, overflow) ||
        !force || !partial || lj.mode < 0 || lj.mode > 2 || !(lj.rc > 0.f) || ntypes < 1 ||
        ntypes > SIM_MAX_TYPES)
        return SIM_E_ARG;
    const float thr = sqrt_threshold(lj.rc);
    auto* P = static_cast<const sim_replica_t*>(reps);
    auto* X = reinterpret_cast<const double4*>(pos4);
    hipStream_t st = as_stream(stream);
#define L_LJ(C, V, A)                                                                             \
    hipLaunchKernelGGL((k_lj_rows<C, 8, V, A>), SIM_ROWS_GRID, dim3(SIM_ROW_THREADS), 0, st, P,  \
                       cell_rep, ncells, row_base, R, cell_start, cell_count, X, nbr16, nnbr,     \
                       max_nbr, lj, force, partial, overflow, thr)
#define L_LV1(C) L_LJ(C, true, 1)
#define L_LV0(C) L_LJ(C, true, 0)
#define L_L1(C) L_LJ(C, false, 1)
#define L_L0(C) L_LJ(C, false, 0)
#define L_ALL(M)                                                                                \
    SIM_COLCAP_DISPATCH(192, M) SIM_COLCAP_DISPATCH(256, M) SIM_COLCAP_DISPATCH(320, M)         \
    SIM_COLCAP_DISPATCH(384, M) SIM_COLCAP_DISPATCH(512, M) {}
    if (virial) {
        if (ntypes > 1) { L_ALL(L_LV1) } else { L_ALL(L_LV0) }
    } else {
        if (ntypes > 1) { L_ALL(L_L1) } else { L_ALL(L_L0) }
    }
#undef L_ALL
#undef L_LV1
#undef L_LV0
#undef L_L1
#undef L_L0
#undef L_LJ
    SIM_CHECK_LAUNCH();
    return SIM_OK;
}
