// Lennard-Jones energy / force / virial over the full Verlet list (gfx950).
//
// Physics of core/force/lennard_jones_force.py:15-95 (raw, force-shift and
// polynomial-switch forms; only r < rc counts), restructured for the GPU:
//   * one thread per atom walks its own column of the column-major full list
//     (coalesced index loads), so the per-atom force is a register sum — no
//     atomics, deterministic order (the reference's kernel issues 6 float
//     atomics per edge, "cuda source/lj_energy_force_kernel.cu:109-116");
//   * neighbors are taken SIM_LJ_BATCH at a time with every index and
//     position load of the batch in flight together;
//   * neighbours are read as 16-byte fractional fixed-point records (one
//     gather each); the displacement is formed from them in float64 and
//     rounded once to float32; everything after that is float32;
//   * energy and virial are halved per atom (each pair is seen from both
//     sides), reduced over the workgroup in float64 with wave shuffles, one
//     partial per tile; sim_reduce adds the tiles with compensation.
#include <type_traits>

#include "pair_common.h"

namespace {

struct LJAcc {
    float fx = 0.f, fy = 0.f, fz = 0.f, e = 0.f;
    float wxx = 0.f, wyy = 0.f, wzz = 0.f, wxy = 0.f, wxz = 0.f, wyz = 0.f;
};

template <bool VIRIAL, int MODE>
__device__ __forceinline__ void lj_pair(LJAcc& a, float rx, float ry, float rz, float r2, float thr,
                                        float eps, float sig, const sim_lj_params_t& lj) {
    // raw form: a pair outside the cutoff gets 1/r^2 = 0, which zeroes its
    // energy, force and virial terms without a branch; the smoothed forms
    // add terms that do not vanish with it, so they skip the pair
    if (MODE != 0 && !(r2 < thr)) return;
    const float rc = lj.rc;
    const float inv_r2 = (MODE != 0 || r2 < thr) ? __builtin_amdgcn_rcpf(r2) : 0.0f;   // 1 ulp
    const float s2 = sig * sig * inv_r2;
    const float s6 = s2 * s2 * s2;
    const float s12 = s6 * s6;
    float U = 4.0f * eps * (s12 - s6);
    float fr = 24.0f * eps * inv_r2 * (2.0f * s12 - s6);  // Fmag / r
    if (MODE == 1) {  // force shift, lennard_jones_force.py:65-75
        const float r = sqrtf(r2);
        const float src = sig / rc;
        const float q6 = src * src * src * src * src * src;
        const float q12 = q6 * q6;
        const float Urc = 4.0f * eps * (q12 - q6);
        const float Frc = 24.0f * eps * (1.0f / rc) * (2.0f * q12 - q6);
        U = U - Urc + (r - rc) * Frc;
        fr = fr - Frc / r;
    } else if (MODE == 2) {  // switch, lennard_jones_force.py:48-63
        const float r = sqrtf(r2);
        const float rs = lj.rs;
        if (r > rs) {
            float x = (r - rs) / (rc - rs);
            x = fminf(fmaxf(x, 0.f), 1.f);
            const float x2 = x * x, x3 = x2 * x;
            const float S = 1.0f - 10.0f * x3 + 15.0f * x3 * x - 6.0f * x3 * x2;
            const float dS = (-30.0f * x2 + 60.0f * x3 - 30.0f * x2 * x2) / (rc - rs);
            fr = fr * S - U * dS / r;
            U = U * S;
        }
    }
    a.fx = __builtin_fmaf(fr, rx, a.fx);
    a.fy = __builtin_fmaf(fr, ry, a.fy);
    a.fz = __builtin_fmaf(fr, rz, a.fz);
    a.e += U;
    if (VIRIAL) {
        const float gx = fr * rx, gy = fr * ry;
        a.wxx = __builtin_fmaf(gx, rx, a.wxx);
        a.wyy = __builtin_fmaf(gy, ry, a.wyy);
        a.wzz = __builtin_fmaf(fr * rz, rz, a.wzz);
        a.wxy = __builtin_fmaf(gx, ry, a.wxy);
        a.wxz = __builtin_fmaf(gx, rz, a.wxz);
        a.wyz = __builtin_fmaf(gy, rz, a.wyz);
    }
}

// Pair parameters of atom type ti against every type, held in registers for
// the atom's whole list (NT1: one type, uniform scalars).
template <bool NT1>
struct LJRow {
    float eps[SIM_MAX_TYPES], sig[SIM_MAX_TYPES];
    __device__ __forceinline__ LJRow(const sim_lj_params_t& lj, int ti) {
#pragma unroll
        for (int t = 0; t < SIM_MAX_TYPES; ++t) {
            eps[t] = NT1 ? lj.eps[0] : lj.eps[ti * SIM_MAX_TYPES + t];
            sig[t] = NT1 ? lj.sig[0] : lj.sig[ti * SIM_MAX_TYPES + t];
        }
    }
    __device__ __forceinline__ float e(int tj) const {
        return NT1 ? eps[0] : (tj == 0 ? eps[0] : tj == 1 ? eps[1] : tj == 2 ? eps[2] : eps[3]);
    }
    __device__ __forceinline__ float s(int tj) const {
        return NT1 ? sig[0] : (tj == 0 ? sig[0] : tj == 1 ? sig[1] : tj == 2 ? sig[2] : sig[3]);
    }
};

// One batch of UB list entries from k (indices in j[]); MASKED: entries past
// the end of the list (self) fail the cutoff test.
template <int UB, bool MASKED, bool VIRIAL, int MODE, bool NT1, int GEO>
__device__ __forceinline__ void lj_batch(LJAcc& a, const int (&j)[UB], int k, int cnt,
                                         const int4 xi, const LJRow<NT1>& row, const QBoxD& B,
                                         const int4* __restrict__ pos, const sim_lj_params_t& lj,
                                         float thr) {
    int4 xj[UB];
#pragma unroll
    for (int u = 0; u < UB; ++u) xj[u] = rec_at(pos, j[u]);
#pragma unroll
    for (int u = 0; u < UB; ++u) {
        float rx, ry, rz;
        qdisp_d<GEO>(xi, xj[u], B, rx, ry, rz);
        const float r2 = (!MASKED || k + u < cnt) ? sq3(rx, ry, rz) : thr;
        lj_pair<VIRIAL, MODE>(a, rx, ry, rz, r2, thr, row.e(xj[u].w), row.s(xj[u].w), lj);
    }
}

// LJ walks the list in batches of SIM_LJ_BATCH, unmasked while every lane of
// the wave has entries left, masked after.
#ifndef SIM_LJ_BATCH
#define SIM_LJ_BATCH 8
#endif

template <bool VIRIAL, int MODE, bool NT1, int GEO>
// The wave walks its lists in lock step (k wave-uniform, up to cmax, the
// wave's longest list; batches past cmin, its shortest, are masked).
__device__ __forceinline__ void lj_atom(LJAcc& a, int i, const int4 xi, int cnt, int cmin,
                                        int cmax, const QBoxD& B, const int4* __restrict__ pos,
                                        const int* __restrict__ nbr, int n_stride,
                                        const sim_lj_params_t& lj, float thr) {
    constexpr int U = SIM_LJ_BATCH;
    const LJRow<NT1> row(lj, xi.w);
    int jn[U];
#pragma unroll
    for (int u = 0; u < U; ++u) jn[u] = nbr_row(nbr, n_stride, i, u, cnt, cmax);
    for (int k = 0; k < cmax; k += U) {
        int j[U];
#pragma unroll
        for (int u = 0; u < U; ++u) j[u] = jn[u];
        // indices of the next batch load while this batch computes
#pragma unroll
        for (int u = 0; u < U; ++u) jn[u] = nbr_row(nbr, n_stride, i, k + U + u, cnt, cmax);
        if (k + U <= cmin)   // every lane of the wave has these entries
            lj_batch<U, false, VIRIAL, MODE, NT1, GEO>(a, j, k, cnt, xi, row, B, pos, lj, thr);
        else
            lj_batch<U, true, VIRIAL, MODE, NT1, GEO>(a, j, k, cnt, xi, row, B, pos, lj, thr);
    }
}

// Tiles per workgroup. Adjacent tiles share most neighbours, but four per
// workgroup (one CU's L1 serving them) measured slower than one (1 M atoms:
// 0.184 vs 0.164 ms).
#ifndef SIM_LJ_TPW
#define SIM_LJ_TPW 1
#endif

// INTEG: the step's second half fused in (S; the arithmetic of k_second,
// kinetic terms into columns SIM_T_KE.. of the tile's partial row).
template <bool VIRIAL, int MODE, bool NT1, bool INTEG>
__global__ void __launch_bounds__(SIM_TILE * SIM_LJ_TPW)
    k_lj(const sim_replica_t* __restrict__ reps, const SimTile* __restrict__ tiles, int ntiles,
         const int4* __restrict__ pos, const int* __restrict__ nbr,
         const int* __restrict__ nnbr, int n_stride, sim_lj_params_t lj,
         float* __restrict__ force, double* __restrict__ partial, float thr, SecondHalf S) {
    constexpr int NV = INTEG ? 8 + 7 : 8;
    __shared__ double lds[SIM_LJ_TPW * 4 * NV];
    const int g = threadIdx.x / SIM_TILE, lt = threadIdx.x % SIM_TILE;
    const int tb = xcd_block(blockIdx.x, gridDim.x) * SIM_LJ_TPW + g;
    const bool have = tb < ntiles;
    const SimTile t = tiles[have ? tb : 0];
    const sim_replica_t& P = reps[t.rep];
    const int i = t.start + lt;
    const bool active = have && i < P.atom_base + P.natoms;
    LJAcc a;
    const int cnt = active ? nnbr[i] : 0;
    const int cmax = wave_max_int(cnt), cmin = wave_min_int(active ? cnt : INT_MAX);
    // the second half's O step needs no force: before the walk
    float3 w_o = make_float3(0.f, 0.f, 0.f);
    int ty_i = 0;
    if (INTEG && active) {
        const unsigned long long step = S.step_dev ? S.step_dev[0] : 0ull;
        ty_i = S.ty.ntypes == 1 ? 0 : pos[i].w;
        w_o = second_half_pre(P, i, ty_i, S.vh, S.noise, S.key, S.seed, step);
    }
    if (active) {
        const QBoxD B = load_qbox_d(P);
        const int4 xi = pos[i];
        const int geo = box_geo(P);
        if (geo == 1)
            lj_atom<VIRIAL, MODE, NT1, 1>(a, i, xi, cnt, cmin, cmax, B, pos, nbr, n_stride, lj, thr);
        else if (geo == 2)
            lj_atom<VIRIAL, MODE, NT1, 2>(a, i, xi, cnt, cmin, cmax, B, pos, nbr, n_stride, lj, thr);
        else
            lj_atom<VIRIAL, MODE, NT1, 0>(a, i, xi, cnt, cmin, cmax, B, pos, nbr, n_stride, lj, thr);
        force[(size_t)i * 3 + 0] = a.fx;
        force[(size_t)i * 3 + 1] = a.fy;
        force[(size_t)i * 3 + 2] = a.fz;
    }
    double v[NV];
    v[0] = 0.5 * a.e;
    v[1] = 0.5 * (a.wxx + a.wyy + a.wzz);
    v[2] = 0.5 * a.wxx;
    v[3] = 0.5 * a.wyy;
    v[4] = 0.5 * a.wzz;
    v[5] = 0.5 * a.wxy;
    v[6] = 0.5 * a.wxz;
    v[7] = 0.5 * a.wyz;
    if (INTEG) {
        double kin[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (active) {
            const float m_i = S.ty.ntypes == 1 ? S.ty.mass[0] : S.ty.mass[ty_i];
            (void)second_half_post(P, i, m_i, w_o, a.fx, a.fy, a.fz, S.vel, kin);
        }
#pragma unroll
        for (int k = 0; k < 7; ++k) v[8 + k] = kin[k];
    }
    group_sum<NV>(v, lds, g, lt);
    if (have && lt == 0) {
        double* p = partial + (size_t)tb * SIM_NPART;
#pragma unroll
        for (int k = 0; k < 8; ++k) p[k] = v[k];
        if (INTEG) {
#pragma unroll
            for (int k = 0; k < 7; ++k) p[SIM_T_KE + k] = v[8 + k];
        }
    }
}

template <bool V, bool NT1, bool I>
void launch_mode(int mode, dim3 g, hipStream_t st, const sim_replica_t* reps, const SimTile* tiles,
                 int ntiles, const int4* pos, const int* nbr, const int* nnbr, int ns,
                 sim_lj_params_t lj, float* f, double* part, float thr, const SecondHalf& S) {
    const dim3 b(SIM_TILE * SIM_LJ_TPW);
    if (mode == 1)
        hipLaunchKernelGGL((k_lj<V, 1, NT1, I>), g, b, 0, st, reps, tiles, ntiles, pos, nbr, nnbr,
                           ns, lj, f, part, thr, S);
    else if (mode == 2)
        hipLaunchKernelGGL((k_lj<V, 2, NT1, I>), g, b, 0, st, reps, tiles, ntiles, pos, nbr, nnbr,
                           ns, lj, f, part, thr, S);
    else
        hipLaunchKernelGGL((k_lj<V, 0, NT1, I>), g, b, 0, st, reps, tiles, ntiles, pos, nbr, nnbr,
                           ns, lj, f, part, thr, S);
}

int lj_impl(int R, const void* reps, const int* tiles, int ntiles, const int* posq, const int* nbr,
            const int* nnbr, int n_stride, sim_lj_params_t lj, float* force, double* partial,
            int virial, const SecondHalf* S2, void* stream) {
    if (R <= 0 || !reps || !tiles || ntiles <= 0 || !posq || !nbr || !nnbr || n_stride <= 0 ||
        n_stride >= (1 << 28) ||
        !force || !partial || lj.mode < 0 || lj.mode > 2 || !(lj.rc > 0.f) || lj.ntypes < 1 ||
        lj.ntypes > SIM_MAX_TYPES)
        return SIM_E_ARG;
    const float thr = sqrt_threshold(lj.rc);
    auto* P = static_cast<const sim_replica_t*>(reps);
    auto* T = reinterpret_cast<const SimTile*>(tiles);
    auto* X = reinterpret_cast<const int4*>(posq);
    const dim3 g((ntiles + SIM_LJ_TPW - 1) / SIM_LJ_TPW);
    hipStream_t st = as_stream(stream);
    const SecondHalf S = S2 ? *S2 : SecondHalf{};
    auto go = [&](auto v, auto one, auto in) {
        launch_mode<decltype(v)::value, decltype(one)::value, decltype(in)::value>(
            lj.mode, g, st, P, T, ntiles, X, nbr, nnbr, n_stride, lj, force, partial, thr, S);
    };
    using yes = std::true_type;
    using no = std::false_type;
    auto go2 = [&](auto in) {
        if (lj.ntypes == 1) {
            if (virial) go(yes(), yes(), in);
            else go(no(), yes(), in);
        } else {
            if (virial) go(yes(), no(), in);
            else go(no(), no(), in);
        }
    };
    if (S2) go2(yes());
    else go2(no());
    SIM_CHECK_LAUNCH();
    return SIM_OK;
}

}  // namespace

extern "C" int sim_lj_force(int R, const void* reps, const int* tiles, int ntiles,
                            const int* posq, const int* nbr, const int* nnbr, int n_stride,
                            sim_lj_params_t lj, float* force, double* partial, int virial,
                            void* stream) {
    return lj_impl(R, reps, tiles, ntiles, posq, nbr, nnbr, n_stride, lj, force, partial, virial,
                   nullptr, stream);
}

extern "C" int sim_lj_force_second(int R, const void* reps, const int* tiles, int ntiles,
                                   const int* posq, const int* nbr, const int* nnbr, int n_stride,
                                   sim_lj_params_t lj, float* force, double* partial, int virial,
                                   sim_types_t types, float* vel_half, float* vel,
                                   const float* noise, const long long* key,
                                   unsigned long long seed, const unsigned long long* step_dev,
                                   void* stream) {
    if (!vel_half || !vel || types.ntypes < 1 || types.ntypes > SIM_MAX_TYPES) return SIM_E_ARG;
    const SecondHalf S{vel_half, vel, noise, key, seed, step_dev, types,
                       nullptr, nullptr, nullptr, nullptr};
    return lj_impl(R, reps, tiles, ntiles, posq, nbr, nnbr, n_stride, lj, force, partial, virial,
                   &S, stream);
}

extern "C" int sim_version(char* buf, int len) {
    const char* v = "simulon_hip 0.1 gfx950";
    if (!buf || len <= 0) return SIM_E_ARG;
    int k = 0;
    for (; k < len - 1 && v[k]; ++k) buf[k] = v[k];
    buf[k] = 0;
    return SIM_OK;
}
