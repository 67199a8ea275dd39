// Two-pass EAM over the full Verlet list (gfx950).
//
// Physics of core/force/eam_force_cu.py:461-579 (tables from :334-395, clamped
// linear interpolation :437-459):
//   pass 1  rho_i = sum_j f_{ti,tj}(r_ij)            (r < cutoff)
//           E_i   = F_ti(rho_i) + 1/2 sum_j phi_{ti,tj}(r_ij);  fp_i = F'_ti(rho_i)
//   pass 2  s_ij  = fp_i f'_{ti,tj}(r) + fp_j f'_{tj,ti}(r) + phi'_{ti,tj}(r)
//           F_i   = sum_j -s_ij rij / (r + 1e-8)
//           virial = 1/2 sum s r = -tr W ; W = 1/2 sum (-s/(r+1e-8)) rij (x) rij
// The reference computes rho with 2 float atomics per edge and forces with 6
// ("cuda source/eam_cuda_ext_kernel.cu:47-52, 108-113"); here each thread owns
// one atom of the full list, so both passes are register sums. Atoms are
// read as 16-byte fractional fixed-point records (pair_common.h): pass 1 reads
// (s, type) records and writes (s, F'(rho)) records, so that pass 2 gathers a
// neighbour's position and F' in one 16-byte load.
//
// Tables are interleaved so that one 16-byte load yields both grid values a
// linear interpolation needs for two functions: pass1[k] = (f(k), f(k+1),
// phi(k), phi(k+1)), pass2[k] = (f'(k), f'(k+1), phi'(k), phi'(k+1)), with
// the k+1 entry clamped at the last grid point as the reference clamps `next`.
#include <mutex>
#include <type_traits>

#include "pair_common.h"

namespace {

struct Lerp {
    int idx;
    float frac;
};

__device__ __forceinline__ Lerp lerp_r(float r, const sim_eam_params_t& E) {
    const float rc = __builtin_amdgcn_fmed3f(r, 0.0f, E.r_clamp);   // clamp to [0, r_clamp]
    const float xf = rc * E.inv_dr;
    const int idx = (int)xf;
    return Lerp{idx, xf - (float)idx};
}

// a + t (b - a) with one rounding for the product and sum (fused)
__device__ __forceinline__ float lerp2(float a, float b, float t) {
    return __builtin_fmaf(t, b - a, a);
}

// Grid point k of a table pair (both values a linear interpolation needs, of
// two functions): the interleaved float4 table in global memory, or (LT) the
// compact copy in LDS, tl[k] = (first function, second function) at k, one
// extra point past the end repeating the last (the clamp of the interleaved
// table).
template <bool LT>
__device__ __forceinline__ float4 tab_get(const float4* __restrict__ g, const float2* tl,
                                          size_t k) {
    if (LT) {
        const float2 a = tl[k], b = tl[k + 1];
        return make_float4(a.x, b.x, a.y, b.y);
    }
    return g[k];
}

// Fill the LDS copy of an interleaved table of n points (n + 1 entries), 8
// loads per thread in flight.
__device__ __forceinline__ void tab_fill(float2* tl, const float4* __restrict__ g, int n) {
    constexpr int F = 8;
    for (int k0 = threadIdx.x; k0 <= n; k0 += F * blockDim.x) {
        float4 e[F];
#pragma unroll
        for (int u = 0; u < F; ++u) e[u] = g[min(k0 + u * (int)blockDim.x, n - 1)];
#pragma unroll
        for (int u = 0; u < F; ++u) {
            const int k = k0 + u * (int)blockDim.x;
            if (k <= n) tl[k] = make_float2(e[u].x, e[u].z);
        }
    }
    __syncthreads();
}

// Walk of one workgroup over groups of TPW consecutive tiles (256 threads
// per tile; group G holds tiles G * TPW .. G * TPW + TPW - 1, the groups of
// sim_nl_stage when TPW = SIM_STAGE_TILES). Persistent: the groups are dealt
// to the XCDs in chunks of SIM_EAM_XCD_CHUNK consecutive groups, chunk c to
// XCD c % 8 (workgroup b runs on XCD b % 8, grid a multiple of 8), and an
// XCD's workgroups take its groups in turn, one per workgroup and step.
// Neighbouring groups share most of their neighbours, so a chunk keeps them
// in one XCD's L2; dealing chunks round the XCDs spreads replicas of
// different cost (list lengths grow with temperature, and the replicas are
// packed in order) evenly over them. SIM_EAM_XCD_CHUNK 0: one contiguous run
// of groups per XCD. Otherwise one tile per workgroup (xcd_block, TPW = 1).
#ifndef SIM_EAM_XCD_CHUNK
#define SIM_EAM_XCD_CHUNK 4
#endif
struct GroupWalk {
    int G, j, end, step, x, ng;
    bool persistent;
    __device__ __forceinline__ int group_of(int jj) const {
        if (SIM_EAM_XCD_CHUNK == 0) return jj;
        constexpr int C = SIM_EAM_XCD_CHUNK > 0 ? SIM_EAM_XCD_CHUNK : 1;
        return (x + SIM_NXCD * (jj / C)) * C + jj % C;
    }
    __device__ __forceinline__ GroupWalk(bool pers, int ntiles, int tpw) : persistent(pers) {
        if (persistent) {
            ng = (ntiles + tpw - 1) / tpw;
            x = blockIdx.x % SIM_NXCD;
            if (SIM_EAM_XCD_CHUNK == 0) {
                const int q = ng / SIM_NXCD, r = ng % SIM_NXCD;
                const int base = x * q + min(x, r);
                j = base + (int)(blockIdx.x / SIM_NXCD);
                end = base + q + (x < r ? 1 : 0);
            } else {
                // groups of XCD x: chunks x, x + 8, ...; the last may be short
                constexpr int C = SIM_EAM_XCD_CHUNK > 0 ? SIM_EAM_XCD_CHUNK : 1;
                const int nch = (ng + C - 1) / C;
                const int mine = nch > x ? (nch - x + SIM_NXCD - 1) / SIM_NXCD : 0;
                j = (int)(blockIdx.x / SIM_NXCD);
                end = mine * C;   // indices j past ng are skipped in valid()
            }
            step = (int)(gridDim.x / SIM_NXCD);
        } else {
            x = 0;
            ng = 0;
            j = xcd_block(blockIdx.x, gridDim.x);
            end = j + 1;
            step = 1;
        }
        G = persistent ? group_of(j) : j;
        skip();
    }
    // advance past indices whose group lies beyond the last (short last chunk)
    __device__ __forceinline__ void skip() {
        while (persistent && j < end && G >= ng) {
            j += step;
            G = group_of(j);
        }
    }
    __device__ __forceinline__ bool valid() const { return j < end; }
    __device__ __forceinline__ void next() {
        j += step;
        G = persistent ? group_of(j) : j;
        skip();
    }
};

// Entry k of atom i's slot list (sim_stage_t); k is wave-uniform and below
// the list's padded rows.
__device__ __forceinline__ int slot_row(const unsigned short* __restrict__ slot, int ns, int i,
                                        int k) {
    return (int)__builtin_amdgcn_raw_buffer_load_b16(
        buf_rsrc(slot + (size_t)k * ns, (unsigned)ns * 2u), i * 2, 0, 0);
}

// Padded slot rows of a wave whose longest list is cmax (at least 8, which
// the list's row count always reaches).
__device__ __forceinline__ int slot_rows(int cmax) { return max((cmax + 7) & ~7, 8); }

// Copy the records of group G's members into LDS (all threads), 8 members
// per thread in flight: their member indices, then their records, are loaded
// together, so a group costs about two memory round trips, not one per member
// a thread copies.
__device__ __forceinline__ void stage_group(int4* S, const sim_stage_t& ST, int G, int nu,
                                            const int4* __restrict__ src) {
    constexpr int F = 8;
    const int* __restrict__ mem = ST.members + (size_t)G * ST.cap;
    const int nt = (int)blockDim.x;
    for (int q0 = threadIdx.x; q0 < nu; q0 += F * nt) {
        int m[F];
        int4 r[F];
#pragma unroll
        for (int u = 0; u < F; ++u) m[u] = mem[min(q0 + u * nt, nu - 1)];
#pragma unroll
        for (int u = 0; u < F; ++u) r[u] = rec_at(src, m[u]);
#pragma unroll
        for (int u = 0; u < F; ++u)
            if (q0 + u * nt < nu) S[q0 + u * nt] = r[u];
    }
    __syncthreads();
}

// Tile groups per workgroup of the persistent density pass: one table copy
// in LDS serves them all, so more waves per CU share it.
#ifndef SIM_EAM_TPW
#define SIM_EAM_TPW 4
#endif
// List entries in flight per thread in the persistent kernels: at 16 waves
// per CU fewer per thread are faster (measured: density 8, force 4).
#ifndef SIM_EAM_LDS_BATCH
#define SIM_EAM_LDS_BATCH 8
#endif
#ifndef SIM_EAM_FORCE_LDS_BATCH
#define SIM_EAM_FORCE_LDS_BATCH 4
#endif
#ifndef SIM_EAM_FORCE_TPW
#define SIM_EAM_FORCE_TPW 4
#endif
// Workgroups per CU of the persistent density and force kernels (0: 2 / TPW,
// at least 1) and the waves per SIMD their registers are held to (launch
// bound; 1: no bound).
#ifndef SIM_EAM_DENSITY_WG_PER_CU
#define SIM_EAM_DENSITY_WG_PER_CU 0
#endif
#ifndef SIM_EAM_FORCE_WG_PER_CU
#define SIM_EAM_FORCE_WG_PER_CU 0
#endif
#ifndef SIM_EAM_DENSITY_WAVES
#define SIM_EAM_DENSITY_WAVES 1
#endif
#ifndef SIM_EAM_FORCE_WAVES
#define SIM_EAM_FORCE_WAVES 1
#endif
constexpr int eam_wg_per_cu(int set, int tpw) { return set > 0 ? set : (2 / tpw > 0 ? 2 / tpw : 1); }

// ---- pass 1 ---------------------------------------------------------------
// One batch of UB list entries starting at k (indices in j[]). MASKED: the
// batch may run past the end of the list; those entries are self (r = 0) and
// masked out.
// STAGED: j[] are member positions and the records come from the group's
// LDS copy S; otherwise j[] are atoms and the records come from pos.
template <int UB, bool MASKED, int NT, bool LT, bool ORTHO, bool STAGED = false>
__device__ __forceinline__ void density_batch(float& rho, float& ep, const int (&j)[UB], int k,
                                              int cnt, const int4 xi, int ti, const QBox& B,
                                              const int4* __restrict__ pos, const int4* S,
                                              const sim_eam_params_t& E, const float2* tl,
                                              float thr) {
    const float4* __restrict__ T1 = static_cast<const float4*>(E.pass1);
    const int nt = NT == 1 ? 1 : E.ntypes;
    int4 xj[UB];
    float r2[UB];
    Lerp L[UB];
    float4 v[UB];
#pragma unroll
    for (int u = 0; u < UB; ++u) xj[u] = STAGED ? S[j[u]] : rec_at(pos, j[u]);
#pragma unroll
    for (int u = 0; u < UB; ++u) {
        float rx, ry, rz;
        qdisp<ORTHO>(xi, xj[u], B, rx, ry, rz);
        r2[u] = sq3(rx, ry, rz);
        L[u] = lerp_r(pair_sqrt(r2[u]), E);
    }
#pragma unroll
    for (int u = 0; u < UB; ++u) {
        const int tj = NT == 1 ? 0 : xj[u].w;
        v[u] = tab_get<LT>(T1, tl, (size_t)(ti * nt + tj) * E.n_r + L[u].idx);
    }
    // masked (not branched) so that the table loads of the batch stay
    // unconditional and in flight together: a 0/1 weight in one fused
    // multiply-add per sum, fma(1, x, s) = s + x and fma(0, x, s) = s for the
    // finite table values (a select of the term is turned into a branch per
    // entry by the compiler)
#pragma unroll
    for (int u = 0; u < UB; ++u) {
        const bool w = (!MASKED || k + u < cnt) && r2[u] < thr;
        const float wf = w ? 1.0f : 0.0f;
        rho = __builtin_fmaf(wf, lerp2(v[u].x, v[u].y, L[u].frac), rho);
        ep = __builtin_fmaf(wf, lerp2(v[u].z, v[u].w, L[u].frac), ep);
    }
}

// The wave walks its lists in lock step: k is wave-uniform, batches end at
// cmax (the wave's longest list) and are masked once k + U passes cmin (its
// shortest). Masked entries add +0.0f, so the sums are those of each lane's
// own list in order.
template <int NT, bool LT, bool ORTHO>
__device__ __forceinline__ void density_atom(float& rho, float& ep, int i, const int4 xi, int cnt,
                                             int cmin, int cmax, const QBox& B,
                                             const int4* __restrict__ pos,
                                             const int* __restrict__ nbr, int ns,
                                             const sim_eam_params_t& E, const float2* tl,
                                             float thr) {
    constexpr int U = LT ? SIM_EAM_LDS_BATCH : SIM_LIST_BATCH;
    const int ti = NT == 1 ? 0 : xi.w;
    int jn[U];
#pragma unroll
    for (int u = 0; u < U; ++u) jn[u] = nbr_row(nbr, ns, i, u, cnt, cmax);
    for (int k = 0; k < cmax; k += U) {
        int j[U];
#pragma unroll
        for (int u = 0; u < U; ++u) j[u] = jn[u];
        // indices of the next batch load while this batch computes
#pragma unroll
        for (int u = 0; u < U; ++u) jn[u] = nbr_row(nbr, ns, i, k + U + u, cnt, cmax);
        if (k + U <= cmin)
            density_batch<U, false, NT, LT, ORTHO>(rho, ep, j, k, cnt, xi, ti, B, pos, nullptr, E,
                                                   tl, thr);
        else
            density_batch<U, true, NT, LT, ORTHO>(rho, ep, j, k, cnt, xi, ti, B, pos, nullptr, E,
                                                  tl, thr);
    }
}

// The same walk over the slot list, records from the group's LDS copy (one
// element type, LDS table): the same batches and masks, so the same sums.
template <bool ORTHO>
__device__ __forceinline__ void density_atom_staged(float& rho, float& ep, int i, const int4 xi,
                                                    int cnt, int cmin, int cmax, const QBox& B,
                                                    const int4* S,
                                                    const unsigned short* __restrict__ slot,
                                                    int ns, const sim_eam_params_t& E,
                                                    const float2* tl, float thr) {
    constexpr int U = SIM_EAM_LDS_BATCH;
    static_assert(8 % U == 0, "batches must end within the padded slot rows");
    const int rows = slot_rows(cmax);
    int sn[U];
#pragma unroll
    for (int u = 0; u < U; ++u) sn[u] = slot_row(slot, ns, i, min(u, rows - 1));
    for (int k = 0; k < cmax; k += U) {
        int sj[U];
#pragma unroll
        for (int u = 0; u < U; ++u) sj[u] = sn[u];
#pragma unroll
        for (int u = 0; u < U; ++u) sn[u] = slot_row(slot, ns, i, min(k + U + u, rows - 1));
        if (k + U <= cmin)
            density_batch<U, false, 1, true, ORTHO, true>(rho, ep, sj, k, cnt, xi, 0, B, nullptr,
                                                          S, E, tl, thr);
        else
            density_batch<U, true, 1, true, ORTHO, true>(rho, ep, sj, k, cnt, xi, 0, B, nullptr, S,
                                                         E, tl, thr);
    }
}

// STG: a group's member records are staged in LDS (after the table) when
// it has at most kcap members; other groups read the list.
template <int NT, bool LT, int TPW, bool STG>
__global__ void __launch_bounds__(SIM_TILE * TPW, LT ? SIM_EAM_DENSITY_WAVES : 1)
    k_eam_density(const sim_replica_t* __restrict__ reps, const SimTile* __restrict__ tiles,
                  int ntiles, const int4* __restrict__ pos, const int* __restrict__ nbr,
                  const int* __restrict__ nnbr, int n_stride, sim_eam_params_t E,
                  float* __restrict__ rho_out, float* __restrict__ fp_out, int4* __restrict__ rec,
                  double* __restrict__ partial, float thr, sim_stage_t ST, int kcap) {
    extern __shared__ float2 tl[];
    __shared__ double lds[TPW * 4 * 1];
    int4* SR = reinterpret_cast<int4*>(tl + ((E.n_r + 2) & ~1));
    const int g = threadIdx.x / SIM_TILE, lt = threadIdx.x % SIM_TILE;
    if (LT) tab_fill(tl, static_cast<const float4*>(E.pass1), E.n_r);
    for (GroupWalk w(LT, ntiles, TPW); w.valid(); w.next()) {
        const int tb0 = w.G * TPW;
        const bool have = tb0 + g < ntiles;
        const int tb = wave_uniform(have ? tb0 + g : tb0);   // g = threadIdx.x / 256
        bool staged = false;
        if (STG) {
            const int nu = ST.count[w.G];
            staged = nu >= 0 && nu <= kcap;
            if (staged) stage_group(SR, ST, w.G, nu, pos);
        }
        const SimTile t = tiles[tb];
        const sim_replica_t& P = reps[t.rep];
        const int i = t.start + lt;
        const bool active = have && i < P.atom_base + P.natoms;
        double etot = 0.0;
        const int cnt = active ? nnbr[i] : 0;
        const int cmax = wave_max_int(cnt), cmin = wave_min_int(active ? cnt : INT_MAX);
        if (active) {
            const QBox B = load_qbox(P);
            const int4 xi = pos[i];
            const int ti = xi.w;
            float rho = 0.f, ep = 0.f;
            if (STG && staged) {
                if (P.ortho)
                    density_atom_staged<true>(rho, ep, i, xi, cnt, cmin, cmax, B, SR, ST.slot,
                                              n_stride, E, tl, thr);
                else
                    density_atom_staged<false>(rho, ep, i, xi, cnt, cmin, cmax, B, SR, ST.slot,
                                               n_stride, E, tl, thr);
            } else if (P.ortho) {
                density_atom<NT, LT, true>(rho, ep, i, xi, cnt, cmin, cmax, B, pos, nbr, n_stride, E,
                                           tl, thr);
            } else {
                density_atom<NT, LT, false>(rho, ep, i, xi, cnt, cmin, cmax, B, pos, nbr, n_stride,
                                            E, tl, thr);
            }
            // embedding, eam_force_cu.py:448-459
            const float lo = E.rho_lo[ti];
            const float rcl = fminf(fmaxf(rho, lo), E.rho_hi[ti]);
            const float xf = (rcl - lo) * E.inv_drho[ti];
            const int idx = (int)xf;
            const float fr = xf - (float)idx;
            const float2 F = reinterpret_cast<const float2*>(E.emb)[(size_t)ti * E.n_rho + idx];
            const float2 dF = reinterpret_cast<const float2*>(E.demb)[(size_t)ti * E.n_rho + idx];
            const float fpi = lerp2(dF.x, dF.y, fr);
            // rho and F' arrays are optional: the force pass reads F' from rec
            if (rho_out) rho_out[i] = rho;
            if (fp_out) fp_out[i] = fpi;
            if (rec) rec[i] = make_int4(xi.x, xi.y, xi.z, __float_as_int(fpi));
            etot = (double)lerp2(F.x, F.y, fr) + 0.5 * (double)ep;
        }
        double v[1] = {etot};
        group_sum<1>(v, lds, g, lt);
        if (have && lt == 0) partial[(size_t)tb * SIM_NPART + SIM_T_EPOT] = v[0];
        __syncthreads();   // lds (and the staged records) are reused by the next group
    }
}

// ---- pass 2 ---------------------------------------------------------------
struct ForceAcc {
    float fx = 0.f, fy = 0.f, fz = 0.f, vir = 0.f;
    float wxx = 0.f, wyy = 0.f, wzz = 0.f, wxy = 0.f, wxz = 0.f, wyz = 0.f;
};

// s / (r + 1e-8) of the reference is s / r in float32 for every r above
// 0.17 A (1e-8 is below half an ulp of r there); inv_r = 1/r comes from the
// one reciprocal square root that also gives r.
template <bool VIRIAL>
__device__ __forceinline__ void eam_pair_force(ForceAcc& a, float rx, float ry, float rz, float r,
                                               float inv_r, float s) {
    const float g = -s * inv_r;
    a.fx = __builtin_fmaf(g, rx, a.fx);
    a.fy = __builtin_fmaf(g, ry, a.fy);
    a.fz = __builtin_fmaf(g, rz, a.fz);
    if (VIRIAL) {
        // the scalar virial sum s r = -sum g r^2 is minus the trace of the
        // tensor, taken once per atom in the epilogue
        (void)r;
        const float gx = g * rx, gy = g * ry;
        a.wxx = __builtin_fmaf(gx, rx, a.wxx);
        a.wyy = __builtin_fmaf(gy, ry, a.wyy);
        a.wzz = __builtin_fmaf(g * rz, rz, a.wzz);
        a.wxy = __builtin_fmaf(gx, ry, a.wxy);
        a.wxz = __builtin_fmaf(gx, rz, a.wxz);
        a.wyz = __builtin_fmaf(gy, rz, a.wyz);
    }
}

// STAGED as in density_batch.
template <int UB, bool MASKED, int NT, bool LT, bool VIRIAL, bool ORTHO, bool STAGED = false>
__device__ __forceinline__ void force_batch(ForceAcc& a, const int (&j)[UB], int k, int cnt,
                                            const int4 xi, int ti, float fpi, const QBox& B,
                                            const int4* __restrict__ rec, const int4* S,
                                            const int4* __restrict__ typ,
                                            const sim_eam_params_t& E, const float2* tl,
                                            float thr) {
    const float4* __restrict__ T2 = static_cast<const float4*>(E.pass2);
    const int nt = NT == 1 ? 1 : E.ntypes;
    int tj[UB];
    int4 xj[UB];
    float rx[UB], ry[UB], rz[UB], r2[UB], r[UB], ir[UB];
    Lerp L[UB];
    float4 va[UB], vb[UB];
#pragma unroll
    for (int u = 0; u < UB; ++u) xj[u] = STAGED ? S[j[u]] : rec_at(rec, j[u]);
#pragma unroll
    for (int u = 0; u < UB; ++u) tj[u] = NT == 1 ? 0 : rec_at(typ, j[u]).w;
#pragma unroll
    for (int u = 0; u < UB; ++u) {
        qdisp<ORTHO>(xi, xj[u], B, rx[u], ry[u], rz[u]);
        r2[u] = sq3(rx[u], ry[u], rz[u]);
        // masked self entries have r2 = 0: the floor keeps 1/r finite there
        ir[u] = __builtin_amdgcn_rsqf(fmaxf(r2[u], 1e-20f));
        r[u] = r2[u] * ir[u];
        L[u] = lerp_r(r[u], E);
    }
#pragma unroll
    for (int u = 0; u < UB; ++u) {
        if (NT == 1) {
            va[u] = tab_get<LT>(T2, tl, L[u].idx);
        } else {
            va[u] = T2[(size_t)(ti * nt + tj[u]) * E.n_r + L[u].idx];
            vb[u] = T2[(size_t)(tj[u] * nt + ti) * E.n_r + L[u].idx];
        }
    }
#pragma unroll
    for (int u = 0; u < UB; ++u) {
        const bool w = (!MASKED || k + u < cnt) && r2[u] < thr;
        const float dda = lerp2(va[u].x, va[u].y, L[u].frac);
        const float ddb = NT == 1 ? dda : lerp2(vb[u].x, vb[u].y, L[u].frac);
        const float sv = __builtin_fmaf(fpi, dda, __builtin_fmaf(__int_as_float(xj[u].w), ddb,
                                                                 lerp2(va[u].z, va[u].w, L[u].frac)));
        eam_pair_force<VIRIAL>(a, rx[u], ry[u], rz[u], r[u], ir[u], w ? sv : 0.0f);
    }
}

#ifndef SIM_EAM_FORCE_BATCH
#define SIM_EAM_FORCE_BATCH SIM_LIST_BATCH
#endif

// rec: (s, F') records of pass 1; typ: (s, type) records, read for the
// neighbour types only when there are several (NT == 0).
// Lock-step walk as in density_atom.
template <int NT, bool LT, bool VIRIAL, bool ORTHO>
__device__ __forceinline__ void force_atom(ForceAcc& a, int i, const int4 xi, int ti, int cnt,
                                           int cmin, int cmax, const QBox& B,
                                           const int4* __restrict__ rec,
                                           const int4* __restrict__ typ, const int* __restrict__ nbr,
                                           int ns, const sim_eam_params_t& E, const float2* tl,
                                           float thr) {
    constexpr int U = LT ? SIM_EAM_FORCE_LDS_BATCH : SIM_EAM_FORCE_BATCH;
    const float fpi = __int_as_float(xi.w);
    int jn[U];
#pragma unroll
    for (int u = 0; u < U; ++u) jn[u] = nbr_row(nbr, ns, i, u, cnt, cmax);
    for (int k = 0; k < cmax; k += U) {
        int j[U];
#pragma unroll
        for (int u = 0; u < U; ++u) j[u] = jn[u];
#pragma unroll
        for (int u = 0; u < U; ++u) jn[u] = nbr_row(nbr, ns, i, k + U + u, cnt, cmax);
        if (k + U <= cmin)
            force_batch<U, false, NT, LT, VIRIAL, ORTHO>(a, j, k, cnt, xi, ti, fpi, B, rec, nullptr,
                                                         typ, E, tl, thr);
        else
            force_batch<U, true, NT, LT, VIRIAL, ORTHO>(a, j, k, cnt, xi, ti, fpi, B, rec, nullptr,
                                                        typ, E, tl, thr);
    }
}

// The same walk over the slot list, records from the group's LDS copy (one
// element type, LDS table): the same batches and masks, so the same sums.
template <bool VIRIAL, bool ORTHO>
__device__ __forceinline__ void force_atom_staged(ForceAcc& a, int i, const int4 xi, int cnt,
                                                  int cmin, int cmax, const QBox& B,
                                                  const int4* S,
                                                  const unsigned short* __restrict__ slot, int ns,
                                                  const sim_eam_params_t& E, const float2* tl,
                                                  float thr) {
    constexpr int U = SIM_EAM_FORCE_LDS_BATCH;
    static_assert(8 % U == 0, "batches must end within the padded slot rows");
    const float fpi = __int_as_float(xi.w);
    const int rows = slot_rows(cmax);
    int sn[U];
#pragma unroll
    for (int u = 0; u < U; ++u) sn[u] = slot_row(slot, ns, i, min(u, rows - 1));
    for (int k = 0; k < cmax; k += U) {
        int sj[U];
#pragma unroll
        for (int u = 0; u < U; ++u) sj[u] = sn[u];
#pragma unroll
        for (int u = 0; u < U; ++u) sn[u] = slot_row(slot, ns, i, min(k + U + u, rows - 1));
        if (k + U <= cmin)
            force_batch<U, false, 1, true, VIRIAL, ORTHO, true>(a, sj, k, cnt, xi, 0, fpi, B,
                                                                nullptr, S, nullptr, E, tl, thr);
        else
            force_batch<U, true, 1, true, VIRIAL, ORTHO, true>(a, sj, k, cnt, xi, 0, fpi, B,
                                                               nullptr, S, nullptr, E, tl, thr);
    }
}

template <int NT, bool LT, bool VIRIAL, int TPW, bool INTEG, bool STG>
__global__ void __launch_bounds__(SIM_TILE * TPW, LT ? SIM_EAM_FORCE_WAVES : 1)
    k_eam_force(const sim_replica_t* __restrict__ reps, const SimTile* __restrict__ tiles,
                int ntiles, const int4* __restrict__ typ, const int4* __restrict__ rec,
                const int* __restrict__ nbr, const int* __restrict__ nnbr, int n_stride,
                sim_eam_params_t E, float* __restrict__ force, double* __restrict__ partial,
                float thr, SecondHalf S, sim_stage_t ST, int kcap) {
    extern __shared__ float2 tl[];
    constexpr int NV = INTEG ? 14 : 7;
    __shared__ double lds[TPW * 4 * NV];
    __shared__ double wmax[INTEG ? TPW * 4 : 1];
    int4* SR = reinterpret_cast<int4*>(tl + ((E.n_r + 2) & ~1));
    const unsigned long long step = (INTEG && S.step_dev) ? S.step_dev[0] : 0ull;
    const int g = threadIdx.x / SIM_TILE, lt = threadIdx.x % SIM_TILE;
    if (LT) tab_fill(tl, static_cast<const float4*>(E.pass2), E.n_r);
    for (GroupWalk w(LT, ntiles, TPW); w.valid(); w.next()) {
        const int tb0 = w.G * TPW;
        const bool have = tb0 + g < ntiles;
        const int tb = wave_uniform(have ? tb0 + g : tb0);   // g = threadIdx.x / 256
        bool staged = false;
        if (STG) {
            const int nu = ST.count[w.G];
            staged = nu >= 0 && nu <= kcap;
            if (staged) stage_group(SR, ST, w.G, nu, rec);
        }
        const SimTile t = tiles[tb];
        const sim_replica_t& P = reps[t.rep];
        const int i = t.start + lt;
        const bool active = have && i < P.atom_base + P.natoms;
        ForceAcc a;
        const bool next = INTEG && S.pos != nullptr && first_fusable(P);   // tile-uniform
        double4 xn = make_double4(0.0, 0.0, 0.0, 0.0), ln = xn;
        const int cnt = active ? nnbr[i] : 0;
        const int cmax = wave_max_int(cnt), cmin = wave_min_int(active ? cnt : INT_MAX);
        // the second half's O step needs no force: done before the walk, so
        // its v_half load and random draw are off the group's critical path
        float3 w_o = make_float3(0.f, 0.f, 0.f);
        int ty_i = 0;
        if (INTEG && active) {
            ty_i = S.ty.ntypes == 1 ? 0 : typ[i].w;
            w_o = second_half_pre(P, i, ty_i, S.vh, S.noise, S.key, S.seed, step);
            if (next) {   // loaded before the walk, which hides their latency
                xn = S.pos[i];
                ln = S.pos_last[i];
            }
        }
        if (active) {
            const QBox B = load_qbox(P);
            const int4 xi = rec[i];
            const int ti = NT == 1 ? 0 : typ[i].w;
            if (STG && staged) {
                if (P.ortho)
                    force_atom_staged<VIRIAL, true>(a, i, xi, cnt, cmin, cmax, B, SR, ST.slot,
                                                    n_stride, E, tl, thr);
                else
                    force_atom_staged<VIRIAL, false>(a, i, xi, cnt, cmin, cmax, B, SR, ST.slot,
                                                     n_stride, E, tl, thr);
            } else if (P.ortho) {
                force_atom<NT, LT, VIRIAL, true>(a, i, xi, ti, cnt, cmin, cmax, B, rec, typ, nbr,
                                                 n_stride, E, tl, thr);
            } else {
                force_atom<NT, LT, VIRIAL, false>(a, i, xi, ti, cnt, cmin, cmax, B, rec, typ, nbr,
                                                  n_stride, E, tl, thr);
            }
            force[(size_t)i * 3 + 0] = a.fx;
            force[(size_t)i * 3 + 1] = a.fy;
            force[(size_t)i * 3 + 2] = a.fz;
        }
        double v[NV];
        v[0] = -0.5 * (((double)a.wxx + a.wyy) + a.wzz);
        v[1] = 0.5 * a.wxx;
        v[2] = 0.5 * a.wyy;
        v[3] = 0.5 * a.wzz;
        v[4] = 0.5 * a.wxy;
        v[5] = 0.5 * a.wxz;
        v[6] = 0.5 * a.wyz;
        if (INTEG) {
            double kin[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            double d2 = 0.0;
            // one type: a constant kernel-argument offset, so a scalar load
            const float m_i = S.ty.ntypes == 1 ? S.ty.mass[0] : S.ty.mass[ty_i];
            if (active) {
                const float3 vn =
                    second_half_post(P, i, m_i, w_o, a.fx, a.fy, a.fz, S.vel, kin);
                if (next) {   // the next step's first half (k_first's arithmetic)
                    float3 h;
                    d2 = first_half_atom(P, m_i, vn, a.fx, a.fy, a.fz, xn, ln, h);
                    const size_t o = (size_t)i * 3;
                    S.vh[o] = h.x;   // this atom's v_half, read by its own thread above
                    S.vh[o + 1] = h.y;
                    S.vh[o + 2] = h.z;
                    S.pos[i] = xn;
                    S.posq[i] = quant_record(P, xn);
                }
            }
            if (next) {
                d2 = wave_max(d2);
                if ((lt & 63) == 0) wmax[g * 4 + (lt >> 6)] = d2;
            }
#pragma unroll
            for (int k = 0; k < 7; ++k) v[7 + k] = kin[k];
        }
        group_sum<NV>(v, lds, g, lt);   // its barrier also publishes wmax
        if (have && lt == 0) {
            double* p = partial + (size_t)tb * SIM_NPART;
#pragma unroll
            for (int k = 0; k < 7; ++k) p[SIM_T_VIRIAL + k] = v[k];
            if (INTEG) {
#pragma unroll
                for (int k = 0; k < 7; ++k) p[SIM_T_KE + k] = v[7 + k];
            }
            if (next) {
                const double* q = wmax + g * 4;
                const double m = fmax(fmax(q[0], q[1]), fmax(q[2], q[3]));
                // non-negative doubles order like their bit patterns; the plain
                // read first skips most atomics (as k_first)
                const unsigned long long b = (unsigned long long)__double_as_longlong(m);
                if (b > S.maxd2[t.rep]) atomicMax(&S.maxd2[t.rep], b);
            }
        }
        __syncthreads();   // lds (and the staged records) are reused by the next group
    }
}

// LDS-resident tables: one element type, the compact copy within
// SIM_EAM_LDS_TAB_MAX bytes. Grid of the persistent kernels: two workgroups
// per CU (the LDS each holds), a multiple of 8 (one run of tiles per XCD).
#ifndef SIM_EAM_LDS_TAB
#define SIM_EAM_LDS_TAB 1
#endif
#define SIM_EAM_LDS_TAB_MAX (96 * 1024)

size_t lds_tab_bytes(const sim_eam_params_t& E) {
    return (size_t)(E.n_r + 1) * sizeof(float2);
}

// Small launches take the global tables: filling the LDS copy costs each
// workgroup a fixed time that only a long tile walk repays.
#define SIM_EAM_LDS_MIN_TILES 256

bool use_lds_tab(const sim_eam_params_t& E, int ntiles) {
    return SIM_EAM_LDS_TAB && E.ntypes == 1 && lds_tab_bytes(E) <= SIM_EAM_LDS_TAB_MAX &&
           ntiles >= SIM_EAM_LDS_MIN_TILES;
}

// units: work items (tiles, or tile groups); per_cu: workgroups per CU.
int persistent_grid(int units, int per_cu = 2) {
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        ncu <= 0)
        ncu = 256;
    const int g = min(units, per_cu * ncu);
    return (g + SIM_NXCD - 1) / SIM_NXCD * SIM_NXCD;
}

// Raise a kernel's dynamic LDS limit to SIM_EAM_LDS_TAB_MAX (once per kernel).
template <typename K>
bool allow_lds(K kernel) {
    return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               SIM_EAM_LDS_TAB_MAX) == hipSuccess;
}

// Dynamic LDS a staged kernel may take: the device's per-workgroup limit
// less the kernel's static LDS, raised once per kernel; 0 if unavailable.
template <typename K>
int staged_lds(K kernel) {
    int dev = 0, a = 0, b = 0;
    hipFuncAttributes fa{};
    if (hipGetDevice(&dev) != hipSuccess ||
        hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(kernel)) != hipSuccess)
        return 0;
    if (hipDeviceGetAttribute(&a, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        a = 0;
    if (hipDeviceGetAttribute(&b, hipDeviceAttributeSharedMemPerBlockOptin, dev) != hipSuccess)
        b = 0;
    const int dyn = max(a, b) - (int)fa.sharedSizeBytes;
    if (dyn <= 0 || hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                        hipFuncAttributeMaxDynamicSharedMemorySize,
                                        dyn) != hipSuccess)
        return 0;
    return dyn;
}

// Table bytes of the LDS copy rounded up to 16 (the staged records follow).
size_t lds_tab_aligned(const sim_eam_params_t& E) { return (size_t)((E.n_r + 2) & ~1) * 8; }

// Members per group the staged kernel holds in dyn bytes of LDS next to the
// table (at most the stage's capacity); 0: staging off.
int stage_kcap(const sim_stage_t* st, int dyn, const sim_eam_params_t& E) {
    if (!st || dyn <= 0) return 0;
    const long room = ((long)dyn - (long)lds_tab_aligned(E)) / 16;
    return (int)max(0L, min(room, (long)st->cap));
}

bool stage_ok(const sim_stage_t* st) {
    return st == nullptr ||
           (st->members && st->count && st->slot && st->cap >= 1 && st->cap <= SIM_STAGE_CAP_MAX);
}

static_assert(SIM_EAM_TPW == SIM_STAGE_TILES && SIM_EAM_FORCE_TPW == SIM_STAGE_TILES,
              "the persistent kernels walk the stage groups");

}  // namespace

static int density_impl(const void* reps, const int* tiles, int ntiles, const int* posq,
                        const int* nbr, const int* nnbr, int n_stride, sim_eam_params_t eam,
                        float* rho, float* fp, int* rec, double* partial, const sim_stage_t* stg,
                        void* stream) {
    if (!reps || !tiles || ntiles <= 0 || !posq || !nbr || !nnbr || n_stride <= 0 ||
        n_stride >= (1 << 28) || (!fp && !rec) ||
        !partial || !eam.pass1 || !eam.emb || !eam.demb || eam.n_r < 2 || eam.n_rho < 2 ||
        eam.ntypes < 1 || eam.ntypes > SIM_MAX_TYPES || (rec && (const int*)rec == posq) ||
        !stage_ok(stg))
        return SIM_E_ARG;
    const float thr = sqrt_threshold(eam.cutoff);
    auto* P = static_cast<const sim_replica_t*>(reps);
    auto* T = reinterpret_cast<const SimTile*>(tiles);
    auto* X = reinterpret_cast<const int4*>(posq);
    auto* W = reinterpret_cast<int4*>(rec);
    hipStream_t st = as_stream(stream);
    const size_t tb = lds_tab_bytes(eam);
    const sim_stage_t ST = stg ? *stg : sim_stage_t{};
    constexpr int TPW = SIM_EAM_TPW;
    const dim3 pg(persistent_grid((ntiles + TPW - 1) / TPW,
                                  eam_wg_per_cu(SIM_EAM_DENSITY_WG_PER_CU, TPW)));
    static const bool lds_ok = allow_lds(k_eam_density<1, true, TPW, false>);
    static const int dyn_stg = staged_lds(k_eam_density<1, true, TPW, true>);
    const int kcap = stage_kcap(stg, dyn_stg, eam);
    if (use_lds_tab(eam, ntiles) && kcap > 0)
        hipLaunchKernelGGL((k_eam_density<1, true, TPW, true>), pg, dim3(SIM_TILE * TPW),
                           lds_tab_aligned(eam) + (size_t)kcap * 16, st, P, T, ntiles, X, nbr,
                           nnbr, n_stride, eam, rho, fp, W, partial, thr, ST, kcap);
    else if (use_lds_tab(eam, ntiles) && lds_ok)
        hipLaunchKernelGGL((k_eam_density<1, true, TPW, false>), pg, dim3(SIM_TILE * TPW), tb, st,
                           P, T, ntiles, X, nbr, nnbr, n_stride, eam, rho, fp, W, partial, thr, ST,
                           0);
    else if (eam.ntypes == 1)
        hipLaunchKernelGGL((k_eam_density<1, false, 1, false>), dim3(ntiles), dim3(SIM_TILE), 0,
                           st, P, T, ntiles, X, nbr, nnbr, n_stride, eam, rho, fp, W, partial, thr,
                           ST, 0);
    else
        hipLaunchKernelGGL((k_eam_density<0, false, 1, false>), dim3(ntiles), dim3(SIM_TILE), 0,
                           st, P, T, ntiles, X, nbr, nnbr, n_stride, eam, rho, fp, W, partial, thr,
                           ST, 0);
    SIM_CHECK_LAUNCH();
    return SIM_OK;
}

extern "C" int sim_eam_density(int R, const void* reps, const int* tiles, int ntiles,
                               const int* posq, const int* nbr, const int* nnbr, int n_stride,
                               sim_eam_params_t eam, float* rho, float* fp, int* rec,
                               double* partial, void* stream) {
    (void)R;
    return density_impl(reps, tiles, ntiles, posq, nbr, nnbr, n_stride, eam, rho, fp, rec,
                        partial, nullptr, stream);
}

extern "C" int sim_eam_density_staged(int R, const void* reps, const int* tiles, int ntiles,
                                      const int* posq, const int* nbr, const int* nnbr,
                                      int n_stride, sim_eam_params_t eam, float* rho, float* fp,
                                      int* rec, double* partial, const sim_stage_t* stage,
                                      void* stream) {
    (void)R;
    if (!stage) return SIM_E_ARG;
    return density_impl(reps, tiles, ntiles, posq, nbr, nnbr, n_stride, eam, rho, fp, rec,
                        partial, stage, stream);
}

static int force_impl(const void* reps, const int* tiles, int ntiles, const int* posq,
                      const int* rec, const int* nbr, const int* nnbr, int n_stride,
                      sim_eam_params_t eam, float* force, double* partial, int virial,
                      const SecondHalf* S2, const sim_stage_t* stg, void* stream) {
    if (!reps || !tiles || ntiles <= 0 || !posq || !rec || !nbr || !nnbr || n_stride <= 0 ||
        n_stride >= (1 << 28) ||
        !force || !partial || !eam.pass2 || eam.n_r < 2 || eam.ntypes < 1 ||
        eam.ntypes > SIM_MAX_TYPES || !stage_ok(stg))
        return SIM_E_ARG;
    const SecondHalf S = S2 ? *S2 : SecondHalf{};
    const sim_stage_t ST = stg ? *stg : sim_stage_t{};
    const float thr = sqrt_threshold(eam.cutoff);
    auto* P = static_cast<const sim_replica_t*>(reps);
    auto* T = reinterpret_cast<const SimTile*>(tiles);
    auto* X = reinterpret_cast<const int4*>(posq);
    auto* Q = reinterpret_cast<const int4*>(rec);
    hipStream_t st = as_stream(stream);
    const dim3 g(ntiles), b(SIM_TILE);
    const size_t tb = lds_tab_bytes(eam);
    auto go = [&](auto v, auto in) {
        constexpr bool V = decltype(v)::value;
        constexpr bool I = decltype(in)::value;
        constexpr int TPW = SIM_EAM_FORCE_TPW;
        const dim3 pg(persistent_grid((ntiles + TPW - 1) / TPW,
                                      eam_wg_per_cu(SIM_EAM_FORCE_WG_PER_CU, TPW)));
        static const bool lds_ok = allow_lds(k_eam_force<1, true, V, TPW, I, false>);
        static const int dyn_stg = staged_lds(k_eam_force<1, true, V, TPW, I, true>);
        const int kcap = stage_kcap(stg, dyn_stg, eam);
        if (use_lds_tab(eam, ntiles) && kcap > 0)
            hipLaunchKernelGGL((k_eam_force<1, true, V, TPW, I, true>), pg, dim3(SIM_TILE * TPW),
                               lds_tab_aligned(eam) + (size_t)kcap * 16, st, P, T, ntiles, X, Q,
                               nbr, nnbr, n_stride, eam, force, partial, thr, S, ST, kcap);
        else if (use_lds_tab(eam, ntiles) && lds_ok)
            hipLaunchKernelGGL((k_eam_force<1, true, V, TPW, I, false>), pg, dim3(SIM_TILE * TPW),
                               tb, st, P, T, ntiles, X, Q, nbr, nnbr, n_stride, eam, force,
                               partial, thr, S, ST, 0);
        else if (eam.ntypes == 1)
            hipLaunchKernelGGL((k_eam_force<1, false, V, 1, I, false>), g, b, 0, st, P, T, ntiles,
                               X, Q, nbr, nnbr, n_stride, eam, force, partial, thr, S, ST, 0);
        else
            hipLaunchKernelGGL((k_eam_force<0, false, V, 1, I, false>), g, b, 0, st, P, T, ntiles,
                               X, Q, nbr, nnbr, n_stride, eam, force, partial, thr, S, ST, 0);
    };
    if (S2) {
        if (virial) go(std::true_type(), std::true_type());
        else go(std::false_type(), std::true_type());
    } else {
        if (virial) go(std::true_type(), std::false_type());
        else go(std::false_type(), std::false_type());
    }
    SIM_CHECK_LAUNCH();
    return SIM_OK;
}

extern "C" int sim_eam_force(int R, const void* reps, const int* tiles, int ntiles,
                             const int* posq, const int* rec, const int* nbr, const int* nnbr,
                             int n_stride, sim_eam_params_t eam, float* force, double* partial,
                             int virial, void* stream) {
    (void)R;
    return force_impl(reps, tiles, ntiles, posq, rec, nbr, nnbr, n_stride, eam, force, partial,
                      virial, nullptr, nullptr, stream);
}

extern "C" int sim_eam_force_staged(int R, const void* reps, const int* tiles, int ntiles,
                                    const int* posq, const int* rec, const int* nbr,
                                    const int* nnbr, int n_stride, sim_eam_params_t eam,
                                    float* force, double* partial, int virial,
                                    const sim_stage_t* stage, void* stream) {
    (void)R;
    if (!stage) return SIM_E_ARG;
    return force_impl(reps, tiles, ntiles, posq, rec, nbr, nnbr, n_stride, eam, force, partial,
                      virial, nullptr, stage, stream);
}

extern "C" int sim_eam_force_second(int R, const void* reps, const int* tiles, int ntiles,
                                    const int* posq, const int* rec, const int* nbr,
                                    const int* nnbr, int n_stride, sim_eam_params_t eam,
                                    float* force, double* partial, int virial, sim_types_t types,
                                    float* vel_half, float* vel, const float* noise,
                                    const long long* key, unsigned long long seed,
                                    const unsigned long long* step_dev, const sim_stage_t* stage,
                                    double* next_pos4, const double* pos_last, double* maxd2,
                                    int* next_posq, void* stream) {
    (void)R;
    if (!vel_half || !vel || types.ntypes < 1 || types.ntypes > SIM_MAX_TYPES) return SIM_E_ARG;
    // the next step's first half: one element type (the pass then never reads
    // the records it rewrites)
    if (next_pos4 && (!pos_last || !maxd2 || !next_posq || eam.ntypes != 1 || types.ntypes != 1))
        return SIM_E_ARG;
    const SecondHalf S{vel_half,
                       vel,
                       noise,
                       key,
                       seed,
                       step_dev,
                       types,
                       reinterpret_cast<double4*>(next_pos4),
                       reinterpret_cast<const double4*>(pos_last),
                       reinterpret_cast<unsigned long long*>(maxd2),
                       reinterpret_cast<int4*>(next_posq)};
    return force_impl(reps, tiles, ntiles, posq, rec, nbr, nnbr, n_stride, eam, force, partial,
                      virial, &S, stage, stream);
}

// ---- allocation-free EAM evaluation ---------------------------------------
//
// sim_eam_energy_force_noalloc keeps its work arrays (the pass-1 records with
// F', the per-tile partial sums and the reduction chunks) in one file-scope
// device array instead of buffers the caller allocates. Host code never uses
// the array's name as an address: the device address comes from
// hipGetSymbolAddress, once per device, with that device current.
namespace {

constexpr size_t kNoallocRec = (size_t)SIM_NOALLOC_MAX_ATOMS * 4 * sizeof(int);
constexpr size_t kNoallocPartial = (size_t)SIM_NOALLOC_MAX_TILES * SIM_NPART * sizeof(double);
constexpr size_t kNoallocChunks =
    (size_t)SIM_NOALLOC_MAX_REPLICAS * 64 * SIM_NPART * sizeof(double);   // sim_reduce's chunks

static __device__ __align__(256) unsigned char scratch[kNoallocRec + kNoallocPartial +
                                                       kNoallocChunks];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on the first
// call there), or nullptr.
void* scratch_address() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
        g_scratch_dev[dev] = p;
    }
    return g_scratch_dev[dev];
}

}  // namespace

extern "C" int sim_eam_energy_force_noalloc(int R, const void* reps, const int* tiles, int ntiles,
                                            const int* rep_tiles, const int* posq, const int* nbr,
                                            const int* nnbr, int n_stride, sim_eam_params_t eam,
                                            float* force, double* thermo, int virial,
                                            void* stream) {
    if (R <= 0 || !rep_tiles || !thermo) return SIM_E_ARG;
    if (n_stride > SIM_NOALLOC_MAX_ATOMS || ntiles > SIM_NOALLOC_MAX_TILES ||
        R > SIM_NOALLOC_MAX_REPLICAS)
        return SIM_E_SIZE;
    unsigned char* base = static_cast<unsigned char*>(scratch_address());
    if (!base) return SIM_E_LAUNCH;
    int* rec = reinterpret_cast<int*>(base);
    double* partial = reinterpret_cast<double*>(base + kNoallocRec);
    double* chunks = reinterpret_cast<double*>(base + kNoallocRec + kNoallocPartial);
    // pass 1 and pass 2 write every column [0, 8) of every tile's partial
    // row, and the reduction writes every chunk it reads: nothing in the
    // scratch array has to start at zero
    int rc = sim_eam_density(R, reps, tiles, ntiles, posq, nbr, nnbr, n_stride, eam, nullptr,
                             nullptr, rec, partial, stream);
    if (rc != SIM_OK) return rc;
    rc = sim_eam_force(R, reps, tiles, ntiles, posq, rec, nbr, nnbr, n_stride, eam, force, partial,
                       virial, stream);
    if (rc != SIM_OK) return rc;
    return sim_reduce(R, rep_tiles, partial, SIM_T_EPOT, SIM_T_KE, chunks, thermo, stream);
}
