// Learned radial-descriptor MLP potential on gfx950.
//
//   sim_mlp_energy   descriptor pass over the full list, then the network and
//                    its input gradient on BF16 MFMA tiles
//   sim_mlp_force    analytic forces and virial from the input gradients
//   sim_mlp_adjoint  dL/d(input gradient) of the force-matching loss
//   sim_mlp_grad     forward-over-reverse pass, weight gradients by MFMA
//   sim_mlp_adam     optimizer step and the bfloat16 weight copies
//
// There is no reference kernel for this network: the reference's learned
// potential is an external model (machine_learning_potentials/
// machine_learning_force.py:85-142 gives the return contract kept here).
//
// Lane layout (every kernel): a wave serves 16 atoms; lane l takes atom
// l % 16 and the descriptor features {4q..4q+3} and {16+4q..16+4q+3},
// q = l / 16.  That is the B-operand layout of v_mfma_f32_16x16x16_bf16 for
// X^T (features x atoms), and the C layout of Z^T = W X^T (weights as the A
// operand) is again the B layout of the next layer, so activations pass from
// layer to layer in registers.  Each activation tile is split x = hi + lo
// into two bfloat16 terms (two MFMAs against the same bfloat16 weights), so
// the products are exact up to float32 accumulation.  The four lanes of an
// atom share its neighbor walk and split the 32 Gaussians 8 ways each.
#include <algorithm>
#include <cmath>

#include "pair_common.h"

namespace {

typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

constexpr int ND = SIM_MLP_NDESC;
constexpr int NH = SIM_MLP_NH;
constexpr int W1S = ND + 4;   // LDS row strides in bf16 elements (padding against bank conflicts)
constexpr int WHS = NH + 4;
constexpr int TS = 20;        // row stride (floats) of the per-wave transpose tiles
constexpr float LOG2E = 1.4426950408889634f;
constexpr int G_W1 = 0, G_B1 = 2048, G_W2 = 2112, G_B2 = 6208, G_W3 = 6272, G_B3 = 6336,
              G_LOSS = 6337;
static_assert(G_LOSS + 1 == SIM_MLP_NGRAD, "gradient layout");

__device__ __forceinline__ unsigned short f2bf(float x) {
    unsigned u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);   // nearest, ties to even
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

__device__ __forceinline__ void split(const f4v x, s4v& hi, s4v& lo) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const unsigned short h = f2bf(x[v]);
        hi[v] = (short)h;
        lo[v] = (short)f2bf(x[v] - bf2f(h));
    }
}

__device__ __forceinline__ f4v mfma(s4v a, s4v b, f4v c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float sigm(float z) {
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-z * LOG2E));
}

// Weights staged once per workgroup, in both orientations.
struct alignas(16) NetLDS {
    unsigned short w1[NH * W1S];    // W1   [64][32]
    unsigned short w1t[ND * WHS];   // W1^T [32][64]
    unsigned short w2[NH * WHS];    // W2   [64][64]
    unsigned short w2t[NH * WHS];   // W2^T [64][64]
};

__device__ __forceinline__ void load_net(NetLDS& L, const sim_mlp_net_t& net) {
    const unsigned short* w1 = static_cast<const unsigned short*>(net.w1);
    const unsigned short* w2 = static_cast<const unsigned short*>(net.w2);
    for (int e = threadIdx.x; e < NH * ND; e += blockDim.x) {
        const int o = e / ND, i = e % ND;
        const unsigned short w = w1[e];
        L.w1[o * W1S + i] = w;
        L.w1t[i * WHS + o] = w;
    }
    for (int e = threadIdx.x; e < NH * NH; e += blockDim.x) {
        const int o = e / NH, i = e % NH;
        const unsigned short w = w2[e];
        L.w2[o * WHS + i] = w;
        L.w2t[i * WHS + o] = w;
    }
    __syncthreads();
}

// out[m] = rows 16m..16m+15 of W X^T for this wave's 16 atoms (C layout);
// W row-major with `stride` elements per row in LDS; X^T as KT hi/lo B tiles.
template <int MT, int KT>
__device__ __forceinline__ void dense(f4v (&out)[MT], const unsigned short* W, int stride,
                                      const s4v (&bh)[KT], const s4v (&bl)[KT]) {
    const int lane = threadIdx.x & 63, i = lane & 15, q = lane >> 4;
#pragma unroll
    for (int m = 0; m < MT; ++m) {
        f4v acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < KT; ++k) {
            const s4v a = *reinterpret_cast<const s4v*>(W + (16 * m + i) * stride + 16 * k + 4 * q);
            acc = mfma(a, bh[k], acc);
            acc = mfma(a, bl[k], acc);
        }
        out[m] = acc;
    }
}

// Radial basis of one lane: Gaussian centres of its 8 features, two runs
// of 4 consecutive centres (mu[0..3] and mu[4..7]).
struct Basis {
    float mu[8];
    float alpha;    // log2(e) / (2 w^2)
    float inv_w2;
    float inv_rc;
    float thr;      // r2 < thr  <=>  sqrtf(r2) < rc
    float dmu;      // centre spacing
    float two_ad;   // 2 alpha dmu
    float ad2;      // alpha dmu^2
    float c2;       // 2^(-2 alpha dmu^2)
};

// The lane's 8 Gaussians exp(-(r - mu_k)^2 / (2 w^2)) and x_k = r - mu_k.
// Within a run of equally spaced centres the ratio of neighbours is
//   G_{k+1} / G_k = 2^(alpha (2 dmu x_k - dmu^2)) = R_k,   R_{k+1} = R_k 2^(-2 alpha dmu^2),
// so a run of 4 costs 2 exponentials and 5 products instead of 4
// exponentials (relative error a few float32 ulps; Gaussians below the
// float32 range come out as 0, as they do when evaluated one by one).
__device__ __forceinline__ void gauss8(float r, const Basis& b, float (&G)[8], float (&x)[8]) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const float x0 = r - b.mu[4 * h];
        float g = __builtin_amdgcn_exp2f(-b.alpha * x0 * x0);
        float R = __builtin_amdgcn_exp2f(__builtin_fmaf(b.two_ad, x0, -b.ad2));
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            G[4 * h + m] = g;
            x[4 * h + m] = x0 - (float)m * b.dmu;
            if (m < 3) {
                g = g * R;
                R = R * b.c2;
            }
        }
    }
}

__device__ __forceinline__ Basis make_basis(const sim_mlp_net_t& net, int q, float thr) {
    Basis b;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        b.mu[v] = net.mu0 + (float)(4 * q + v) * net.dmu;
        b.mu[4 + v] = net.mu0 + (float)(16 + 4 * q + v) * net.dmu;
    }
    b.alpha = 0.5f * net.inv_w2 * LOG2E;
    b.inv_w2 = net.inv_w2;
    b.inv_rc = 1.0f / net.rc;
    b.thr = thr;
    b.dmu = net.dmu;
    b.two_ad = 2.0f * b.alpha * net.dmu;
    b.ad2 = b.alpha * net.dmu * net.dmu;
    b.c2 = __builtin_amdgcn_exp2f(-2.0f * b.ad2);
    return b;
}

// fc(r) and dfc/dr for the cosine cutoff.
__device__ __forceinline__ void cutoff_fn(float r, const Basis& b, float& fc, float& dfc) {
    float s, c;
    sincospif(r * b.inv_rc, &s, &c);
    fc = 0.5f * (c + 1.0f);
    dfc = -0.5f * 3.14159265358979f * b.inv_rc * s;
}

// Walk the full list of atom i in batches of SIM_MLP_BATCH with every index
// and record load of a batch in flight together (branch-free: indices past
// the end of the list are clamped, their entries skipped); body(j, rx, ry,
// rz) sees xi - xj at the minimum image, from the fractional fixed-point
// records (pair_common.h).
#define SIM_MLP_BATCH 8
template <bool ORTHO, typename Body>
__device__ __forceinline__ void walk(int i, const int4 xi, int cnt, int cmax, const QBox& B,
                                     const int4* __restrict__ pos, const int* __restrict__ nbr,
                                     int ns, Body&& body) {
    constexpr int U = SIM_MLP_BATCH;
    int jn[U];
#pragma unroll
    for (int u = 0; u < U; ++u) jn[u] = nbr_row(nbr, ns, i, u, cnt, cmax);
    for (int k = 0; k < cmax; k += U) {   // wave-uniform k, as in the pair kernels
        int j[U];
        int4 xj[U];
#pragma unroll
        for (int u = 0; u < U; ++u) j[u] = jn[u];
#pragma unroll
        for (int u = 0; u < U; ++u) jn[u] = nbr_row(nbr, ns, i, k + U + u, cnt, cmax);
#pragma unroll
        for (int u = 0; u < U; ++u) xj[u] = rec_at(pos, j[u]);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (k + u < cnt) {
                float rx, ry, rz;
                qdisp<ORTHO>(xi, xj[u], B, rx, ry, rz);
                body(j[u], rx, ry, rz);
            }
        }
    }
}

// Atom of this lane for the (tile, quarter) workgroup layout of the list kernels.
__device__ __forceinline__ int lane_atom(const SimTile& t) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    return t.start + (int)(blockIdx.x & 3) * 64 + wid * 16 + (lane & 15);
}

// ---- energy: descriptors, network, input gradient -------------------------

template <bool ORTHO>
__device__ __forceinline__ void descriptors(float (&g)[8], int i, int cmax, const sim_replica_t& P,
                                            const int4* __restrict__ pos,
                                            const int* __restrict__ nbr,
                                            const int* __restrict__ nnbr, int ns, const Basis& b) {
    const QBox B = load_qbox(P);
    const int4 xi = pos[i];
    walk<ORTHO>(i, xi, nnbr[i], cmax, B, pos, nbr, ns, [&](int, float rx, float ry, float rz) {
        const float r2 = sq3(rx, ry, rz);
        if (r2 < b.thr) {
            const float r = sqrtf(r2);
            float fc, dfc;
            cutoff_fn(r, b, fc, dfc);
            float G[8], x[8];
            gauss8(r, b, G, x);
#pragma unroll
            for (int v = 0; v < 8; ++v) g[v] = __builtin_fmaf(G[v], fc, g[v]);
        }
    });
}

__global__ void __launch_bounds__(256)
    k_mlp_energy(const sim_replica_t* __restrict__ reps, const SimTile* __restrict__ tiles,
                 const int4* __restrict__ pos, const int* __restrict__ nbr,
                 const int* __restrict__ nnbr, int ns, sim_mlp_net_t net, float* __restrict__ c,
                 float* __restrict__ gh_out, float* __restrict__ e_atom,
                 double* __restrict__ partial, float thr) {
    __shared__ NetLDS L;
    __shared__ double red[4];
    load_net(L, net);
    const SimTile t = tiles[blockIdx.x >> 2];
    const sim_replica_t& P = reps[t.rep];
    const int lane = threadIdx.x & 63, q = lane >> 4;
    const int i = lane_atom(t);
    const bool active = i < P.atom_base + P.natoms;
    const Basis b = make_basis(net, q, thr);
    const int cmax = wave_max_int(active ? nnbr[i] : 0);
    float g[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (active) {
        if (P.ortho)
            descriptors<true>(g, i, cmax, P, pos, nbr, nnbr, ns, b);
        else
            descriptors<false>(g, i, cmax, P, pos, nbr, nnbr, ns, b);
    }
    const int f0 = 4 * q, f1 = 16 + 4 * q;
    f4v x0, x1;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        x0[v] = (g[v] - net.mean[f0 + v]) * net.inv_std[f0 + v];
        x1[v] = (g[4 + v] - net.mean[f1 + v]) * net.inv_std[f1 + v];
    }
    if (gh_out && active) {
        *reinterpret_cast<f4v*>(gh_out + (size_t)i * ND + f0) = x0;
        *reinterpret_cast<f4v*>(gh_out + (size_t)i * ND + f1) = x1;
    }
    s4v xh[2], xl[2];
    split(x0, xh[0], xl[0]);
    split(x1, xh[1], xl[1]);
    // layer 1
    f4v z[4];
    dense<4, 2>(z, L.w1, W1S, xh, xl);
    f4v s1p[4];
    s4v hh[4], hl[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        f4v h;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const float zz = z[m][v] + net.b1[16 * m + 4 * q + v];
            const float sg = sigm(zz);
            h[v] = zz * sg;
            s1p[m][v] = sg * (1.0f + zz * (1.0f - sg));
        }
        split(h, hh[m], hl[m]);
    }
    // layer 2 and the output unit; dE/dz2 on the way
    dense<4, 4>(z, L.w2, WHS, hh, hl);
    float e = 0.f;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        f4v d2;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int f = 16 * m + 4 * q + v;
            const float zz = z[m][v] + net.b2[f];
            const float sg = sigm(zz);
            e = __builtin_fmaf(net.w3[f], zz * sg, e);
            d2[v] = net.w3[f] * (sg * (1.0f + zz * (1.0f - sg)));
        }
        split(d2, hh[m], hl[m]);
    }
    e += __shfl_xor(e, 16, 64);
    e += __shfl_xor(e, 32, 64);
    e += net.b3[0];
    // back through layer 2 and layer 1
    dense<4, 4>(z, L.w2t, WHS, hh, hl);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        f4v d1;
#pragma unroll
        for (int v = 0; v < 4; ++v) d1[v] = z[m][v] * s1p[m][v];
        split(d1, hh[m], hl[m]);
    }
    f4v dg[2];
    dense<2, 4>(dg, L.w1t, WHS, hh, hl);
    if (active) {
        f4v c0, c1;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            c0[v] = dg[0][v] * net.inv_std[f0 + v];
            c1[v] = dg[1][v] * net.inv_std[f1 + v];
        }
        *reinterpret_cast<f4v*>(c + (size_t)i * ND + f0) = c0;
        *reinterpret_cast<f4v*>(c + (size_t)i * ND + f1) = c1;
        if (e_atom && q == 0) e_atom[i] = e;
    }
    double ev[1] = {(active && q == 0) ? (double)e : 0.0};
    block_sum<1>(ev, red);
    if (threadIdx.x == 0) partial[(size_t)blockIdx.x * SIM_NPART + SIM_T_EPOT] = ev[0];
}

// ---- forces ------------------------------------------------------------------

template <bool VIRIAL, bool ORTHO>
__device__ __forceinline__ void force_atom(float (&acc)[10], int i, int cmax, const sim_replica_t& P,
                                           const int4* __restrict__ pos,
                                           const int* __restrict__ nbr,
                                           const int* __restrict__ nnbr, int ns,
                                           const float* __restrict__ c, const Basis& b, int q) {
    const QBox B = load_qbox(P);
    const int4 xi = pos[i];
    const f4v ci0 = *reinterpret_cast<const f4v*>(c + (size_t)i * ND + 4 * q);
    const f4v ci1 = *reinterpret_cast<const f4v*>(c + (size_t)i * ND + 16 + 4 * q);
    walk<ORTHO>(i, xi, nnbr[i], cmax, B, pos, nbr, ns, [&](int j, float rx, float ry, float rz) {
        const float r2 = sq3(rx, ry, rz);
        if (r2 < b.thr) {
            const f4v cj0 = *reinterpret_cast<const f4v*>(c + (size_t)j * ND + 4 * q);
            const f4v cj1 = *reinterpret_cast<const f4v*>(c + (size_t)j * ND + 16 + 4 * q);
            const float r = sqrtf(r2);
            float fc, dfc;
            cutoff_fn(r, b, fc, dfc);
            const float fcw = fc * b.inv_w2;
            float G[8], x[8];
            gauss8(r, b, G, x);
            float s = 0.f;   // this lane's share of dE/dr
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                const float cc = v < 4 ? ci0[v] + cj0[v] : ci1[v - 4] + cj1[v - 4];
                s = __builtin_fmaf(cc * G[v], dfc - x[v] * fcw, s);
            }
            const float fr = -s * __builtin_amdgcn_rcpf(r);
            acc[0] = __builtin_fmaf(fr, rx, acc[0]);
            acc[1] = __builtin_fmaf(fr, ry, acc[1]);
            acc[2] = __builtin_fmaf(fr, rz, acc[2]);
            if (VIRIAL) {
                acc[3] = __builtin_fmaf(s, r, acc[3]);
                const float gx = fr * rx, gy = fr * ry;
                acc[4] = __builtin_fmaf(gx, rx, acc[4]);
                acc[5] = __builtin_fmaf(gy, ry, acc[5]);
                acc[6] = __builtin_fmaf(fr * rz, rz, acc[6]);
                acc[7] = __builtin_fmaf(gx, ry, acc[7]);
                acc[8] = __builtin_fmaf(gx, rz, acc[8]);
                acc[9] = __builtin_fmaf(gy, rz, acc[9]);
            }
        }
    });
}

template <bool VIRIAL>
__global__ void __launch_bounds__(256)
    k_mlp_force(const sim_replica_t* __restrict__ reps, const SimTile* __restrict__ tiles,
                const int4* __restrict__ pos, const int* __restrict__ nbr,
                const int* __restrict__ nnbr, int ns, sim_mlp_net_t net,
                const float* __restrict__ c, float* __restrict__ force,
                double* __restrict__ partial, float thr) {
    __shared__ double red[4 * 7];
    const SimTile t = tiles[blockIdx.x >> 2];
    const sim_replica_t& P = reps[t.rep];
    const int lane = threadIdx.x & 63, q = lane >> 4;
    const int i = lane_atom(t);
    const bool active = i < P.atom_base + P.natoms;
    const Basis b = make_basis(net, q, thr);
    const int cmax = wave_max_int(active ? nnbr[i] : 0);
    float acc[10] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (active) {
        if (P.ortho)
            force_atom<VIRIAL, true>(acc, i, cmax, P, pos, nbr, nnbr, ns, c, b, q);
        else
            force_atom<VIRIAL, false>(acc, i, cmax, P, pos, nbr, nnbr, ns, c, b, q);
    }
    float fx = acc[0], fy = acc[1], fz = acc[2];
#pragma unroll
    for (int o = 16; o < 64; o <<= 1) {
        fx += __shfl_xor(fx, o, 64);
        fy += __shfl_xor(fy, o, 64);
        fz += __shfl_xor(fz, o, 64);
    }
    if (active && q == 0) {
        force[(size_t)i * 3 + 0] = fx;
        force[(size_t)i * 3 + 1] = fy;
        force[(size_t)i * 3 + 2] = fz;
    }
    if (VIRIAL) {
        double v[7];
#pragma unroll
        for (int k = 0; k < 7; ++k) v[k] = 0.5 * (double)acc[3 + k];
        block_sum<7>(v, red);
        if (threadIdx.x == 0) {
            double* p = partial + (size_t)blockIdx.x * SIM_NPART;
#pragma unroll
            for (int k = 0; k < 7; ++k) p[SIM_T_VIRIAL + k] = v[k];
        }
    }
}

// ---- force-matching adjoint ---------------------------------------------------

template <bool ORTHO>
__device__ __forceinline__ void adjoint_atom(float (&u)[8], int i, int cmax, const sim_replica_t& P,
                                             const int4* __restrict__ pos,
                                             const int* __restrict__ nbr,
                                             const int* __restrict__ nnbr, int ns,
                                             const float* __restrict__ F,
                                             const float* __restrict__ Fr, const Basis& b) {
    const QBox B = load_qbox(P);
    const int4 xi = pos[i];
    const float dix = F[(size_t)i * 3 + 0] - Fr[(size_t)i * 3 + 0];
    const float diy = F[(size_t)i * 3 + 1] - Fr[(size_t)i * 3 + 1];
    const float diz = F[(size_t)i * 3 + 2] - Fr[(size_t)i * 3 + 2];
    walk<ORTHO>(i, xi, nnbr[i], cmax, B, pos, nbr, ns, [&](int j, float rx, float ry, float rz) {
        const float r2 = sq3(rx, ry, rz);
        if (r2 < b.thr) {
            const float djx = F[(size_t)j * 3 + 0] - Fr[(size_t)j * 3 + 0];
            const float djy = F[(size_t)j * 3 + 1] - Fr[(size_t)j * 3 + 1];
            const float djz = F[(size_t)j * 3 + 2] - Fr[(size_t)j * 3 + 2];
            const float r = sqrtf(r2);
            const float proj = ((dix - djx) * rx + (diy - djy) * ry + (diz - djz) * rz) *
                               __builtin_amdgcn_rcpf(r);
            float fc, dfc;
            cutoff_fn(r, b, fc, dfc);
            const float fcw = fc * b.inv_w2;
            float G[8], x[8];
            gauss8(r, b, G, x);
#pragma unroll
            for (int v = 0; v < 8; ++v) u[v] = __builtin_fmaf(-G[v] * (dfc - x[v] * fcw), proj, u[v]);
        }
    });
}

__global__ void __launch_bounds__(256)
    k_mlp_adjoint(const sim_replica_t* __restrict__ reps, const SimTile* __restrict__ tiles,
                  const int4* __restrict__ pos, const int* __restrict__ nbr,
                  const int* __restrict__ nnbr, int ns, sim_mlp_net_t net,
                  const float* __restrict__ F, const float* __restrict__ Fr, float wF,
                  float* __restrict__ vout, float thr) {
    const SimTile t = tiles[blockIdx.x >> 2];
    const sim_replica_t& P = reps[t.rep];
    const int lane = threadIdx.x & 63, q = lane >> 4;
    const int i = lane_atom(t);
    const bool active = i < P.atom_base + P.natoms;
    const int cmax = wave_max_int(active ? nnbr[i] : 0);
    if (!active) return;
    const Basis b = make_basis(net, q, thr);
    float u[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (P.ortho)
        adjoint_atom<true>(u, i, cmax, P, pos, nbr, nnbr, ns, F, Fr, b);
    else
        adjoint_atom<false>(u, i, cmax, P, pos, nbr, nnbr, ns, F, Fr, b);
    const float scale = 2.0f * wF / (3.0f * (float)P.natoms);
    const int f0 = 4 * q, f1 = 16 + 4 * q;
    f4v v0, v1;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        v0[v] = u[v] * scale * net.inv_std[f0 + v];
        v1[v] = u[4 + v] * scale * net.inv_std[f1 + v];
    }
    *reinterpret_cast<f4v*>(vout + (size_t)i * ND + f0) = v0;
    *reinterpret_cast<f4v*>(vout + (size_t)i * ND + f1) = v1;
}

// ---- weight gradient -------------------------------------------------------------

// Wave-local ordering of LDS writes before reads by other lanes of the wave.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Write a C-layout tile set (features 16m+4q+v of atom l%16) into T[f][atom].
template <int MT>
__device__ __forceinline__ void put_c(float* T, const f4v (&x)[MT]) {
    const int lane = threadIdx.x & 63, a = lane & 15, q = lane >> 4;
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int v = 0; v < 4; ++v) T[(16 * m + 4 * q + v) * TS + a] = x[m][v];
}

// acc[mo][ni] += sum over the 16 atoms of L[f=16mo+.][atom] R[g=16ni+.][atom],
// both operands split hi/lo (three products: hi*hi, hi*lo, lo*hi).
template <int MT, int NT>
__device__ __forceinline__ void outer(f4v (&acc)[MT][NT], const float* Lt, const float* Rt) {
    const int lane = threadIdx.x & 63, i = lane & 15, q = lane >> 4;
    s4v bh[NT], bl[NT];
#pragma unroll
    for (int n = 0; n < NT; ++n) {
        const f4v x = *reinterpret_cast<const f4v*>(Rt + (16 * n + i) * TS + 4 * q);
        split(x, bh[n], bl[n]);
    }
#pragma unroll
    for (int m = 0; m < MT; ++m) {
        s4v ah, al;
        const f4v x = *reinterpret_cast<const f4v*>(Lt + (16 * m + i) * TS + 4 * q);
        split(x, ah, al);
#pragma unroll
        for (int n = 0; n < NT; ++n) {
            acc[m][n] = mfma(ah, bh[n], acc[m][n]);
            acc[m][n] = mfma(ah, bl[n], acc[m][n]);
            acc[m][n] = mfma(al, bh[n], acc[m][n]);
        }
    }
}

struct alignas(16) GradLDS {
    NetLDS net;
    float t[4][4][NH * TS];     // per wave: two left and two right transpose tiles
    float g[SIM_MLP_NGRAD];     // workgroup gradient
};

__global__ void __launch_bounds__(256)
    k_mlp_grad(int n, const int* __restrict__ atom_rep, const sim_replica_t* __restrict__ reps,
               const double* __restrict__ thermo, const double* __restrict__ e_ref, float wE,
               float wF, sim_mlp_net_t net, const float* __restrict__ gh,
               const float* __restrict__ vin, const float* __restrict__ F,
               const float* __restrict__ Fr, float* __restrict__ grad_partial) {
    __shared__ GradLDS S;
    load_net(S.net, net);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, q = lane >> 4, al = lane & 15;
    const int f0 = 4 * q, f1 = 16 + 4 * q;
    float* TL0 = S.t[wid][0];
    float* TL1 = S.t[wid][1];
    float* TR0 = S.t[wid][2];
    float* TR1 = S.t[wid][3];
    f4v gW1[4][2], gW2[4][4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
#pragma unroll
        for (int k = 0; k < 2; ++k) gW1[m][k] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < 4; ++k) gW2[m][k] = f4v{0.f, 0.f, 0.f, 0.f};
    }
    float gb1[16], gb2[16], gw3[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) gb1[k] = gb2[k] = gw3[k] = 0.f;
    float gb3 = 0.f, loss = 0.f;

    const int nw = gridDim.x * 4;
    for (int grp = blockIdx.x * 4 + wid; grp * 16 < n; grp += nw) {
        const int a = grp * 16 + al;
        const bool act = a < n;
        f4v x0 = {0.f, 0.f, 0.f, 0.f}, x1 = x0, v0 = x0, v1 = x0;
        float gE = 0.f;
        if (act) {
            x0 = *reinterpret_cast<const f4v*>(gh + (size_t)a * ND + f0);
            x1 = *reinterpret_cast<const f4v*>(gh + (size_t)a * ND + f1);
            v0 = *reinterpret_cast<const f4v*>(vin + (size_t)a * ND + f0);
            v1 = *reinterpret_cast<const f4v*>(vin + (size_t)a * ND + f1);
            const int r = atom_rep[a];
            const double nr = (double)reps[r].natoms;
            const double dE = thermo[(size_t)r * SIM_NPART + SIM_T_EPOT] - e_ref[r];
            gE = (float)(2.0 * (double)wE * dE / (nr * nr));
            if (q == 0) {
                const float dx = F[(size_t)a * 3 + 0] - Fr[(size_t)a * 3 + 0];
                const float dy = F[(size_t)a * 3 + 1] - Fr[(size_t)a * 3 + 1];
                const float dz = F[(size_t)a * 3 + 2] - Fr[(size_t)a * 3 + 2];
                loss += wF * (dx * dx + dy * dy + dz * dz) / (float)(3.0 * nr) +
                        (float)((double)wE * dE * dE / (nr * nr * nr));
                gb3 += gE;
            }
        }
        const float msk = act ? 1.0f : 0.0f;
        s4v xh[2], xl[2], vh[2], vl[2];
        split(x0, xh[0], xl[0]);
        split(x1, xh[1], xl[1]);
        split(v0, vh[0], vl[0]);
        split(v1, vh[1], vl[1]);
        // forward and its tangent along v
        f4v z1[4], zd1[4];
        dense<4, 2>(z1, S.net.w1, W1S, xh, xl);
        dense<4, 2>(zd1, S.net.w1, W1S, vh, vl);
        f4v h1[4], hd1[4], s1p[4], s1pp[4];
        s4v hh[4], hl[4], hdh[4], hdl[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const float zz = z1[m][v] + net.b1[16 * m + 4 * q + v];
                const float sg = sigm(zz);
                h1[m][v] = zz * sg;
                s1p[m][v] = sg * (1.0f + zz * (1.0f - sg));
                s1pp[m][v] = sg * (1.0f - sg) * (2.0f + zz * (1.0f - 2.0f * sg));
                hd1[m][v] = s1p[m][v] * zd1[m][v];
            }
            split(h1[m], hh[m], hl[m]);
            split(hd1[m], hdh[m], hdl[m]);
        }
        f4v z2[4], zd2[4];
        dense<4, 4>(z2, S.net.w2, WHS, hh, hl);
        dense<4, 4>(zd2, S.net.w2, WHS, hdh, hdl);
        f4v d2[4], dd2[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int f = 16 * m + 4 * q + v;
                const float zz = z2[m][v] + net.b2[f];
                const float sg = sigm(zz);
                const float sp = sg * (1.0f + zz * (1.0f - sg));
                const float spp = sg * (1.0f - sg) * (2.0f + zz * (1.0f - 2.0f * sg));
                const float w3 = net.w3[f];
                d2[m][v] = w3 * sp;
                dd2[m][v] = w3 * spp * zd2[m][v];
                gw3[4 * m + v] += msk * (sp * zd2[m][v] + gE * zz * sg);
            }
            split(d2[m], hh[m], hl[m]);
            split(dd2[m], hdh[m], hdl[m]);
        }
        f4v t1[4], td1[4];
        dense<4, 4>(t1, S.net.w2t, WHS, hh, hl);
        dense<4, 4>(td1, S.net.w2t, WHS, hdh, hdl);
        // left factors: P = (d/de of the parameter gradient) + gE * parameter gradient
        f4v p1[4], d1[4], p2[4];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                d1[m][v] = msk * t1[m][v] * s1p[m][v];
                const float dd1 = td1[m][v] * s1p[m][v] + t1[m][v] * s1pp[m][v] * zd1[m][v];
                p1[m][v] = msk * dd1 + gE * d1[m][v];
                p2[m][v] = msk * (dd2[m][v] + gE * d2[m][v]);
                d2[m][v] *= msk;
                h1[m][v] *= msk;
                hd1[m][v] *= msk;
                gb1[4 * m + v] += p1[m][v];
                gb2[4 * m + v] += p2[m][v];
            }
        }
        // W1 gradient: P1 gh^T + D1 v^T
        put_c<4>(TL0, p1);
        put_c<4>(TL1, d1);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            TR0[(f0 + v) * TS + al] = x0[v];
            TR0[(f1 + v) * TS + al] = x1[v];
            TR1[(f0 + v) * TS + al] = v0[v];
            TR1[(f1 + v) * TS + al] = v1[v];
        }
        wave_sync();
        outer<4, 2>(gW1, TL0, TR0);
        outer<4, 2>(gW1, TL1, TR1);
        wave_sync();
        // W2 gradient: P2 h1^T + D2 hdot1^T
        put_c<4>(TL0, p2);
        put_c<4>(TL1, d2);
        put_c<4>(TR0, h1);
        put_c<4>(TR1, hd1);
        wave_sync();
        outer<4, 4>(gW2, TL0, TR0);
        outer<4, 4>(gW2, TL1, TR1);
        wave_sync();
    }
    // bias and output-weight sums over the 16 atoms of the wave
#pragma unroll
    for (int k = 0; k < 16; ++k) {
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
            gb1[k] += __shfl_xor(gb1[k], o, 64);
            gb2[k] += __shfl_xor(gb2[k], o, 64);
            gw3[k] += __shfl_xor(gw3[k], o, 64);
        }
    }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        gb3 += __shfl_xor(gb3, o, 64);
        loss += __shfl_xor(loss, o, 64);
    }
    // workgroup sum, wave by wave in a fixed order
    for (int p = threadIdx.x; p < SIM_MLP_NGRAD; p += blockDim.x) S.g[p] = 0.f;
    for (int w = 0; w < 4; ++w) {
        __syncthreads();
        if (wid != w) continue;
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int f = 16 * m + 4 * q + v;
#pragma unroll
                for (int k = 0; k < 2; ++k) S.g[G_W1 + f * ND + 16 * k + al] += gW1[m][k][v];
#pragma unroll
                for (int k = 0; k < 4; ++k) S.g[G_W2 + f * NH + 16 * k + al] += gW2[m][k][v];
                if (al == 0) {
                    S.g[G_B1 + f] += gb1[4 * m + v];
                    S.g[G_B2 + f] += gb2[4 * m + v];
                    S.g[G_W3 + f] += gw3[4 * m + v];
                }
            }
        if (lane == 0) {
            S.g[G_B3] += gb3;
            S.g[G_LOSS] += loss;
        }
    }
    __syncthreads();
    for (int p = threadIdx.x; p < SIM_MLP_NGRAD; p += blockDim.x)
        grad_partial[(size_t)blockIdx.x * SIM_MLP_NGRAD + p] = S.g[p];
}

__global__ void k_mlp_grad_sum(int nblk, const float* __restrict__ part, float* __restrict__ grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= SIM_MLP_NGRAD) return;
    double s = 0.0;
    for (int b = 0; b < nblk; ++b) s += (double)part[(size_t)b * SIM_MLP_NGRAD + p];
    grad[p] = (float)s;
}

__global__ void k_mlp_adam(float* __restrict__ master, const float* __restrict__ grad,
                           float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                           float b2, float eps, float c1, float c2, sim_mlp_net_t net) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= SIM_MLP_NPARAM) return;
    const float g = grad[p];
    const float mm = b1 * m[p] + (1.0f - b1) * g;
    const float vv = b2 * v[p] + (1.0f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float w = master[p] - lr * (mm * c1) / (sqrtf(vv * c2) + eps);
    master[p] = w;
    if (p < G_B1)
        static_cast<unsigned short*>(net.w1)[p] = f2bf(w);
    else if (p < G_W2)
        net.b1[p - G_B1] = w;
    else if (p < G_B2)
        static_cast<unsigned short*>(net.w2)[p - G_W2] = f2bf(w);
    else if (p < G_W3)
        net.b2[p - G_B2] = w;
    else if (p < G_B3)
        net.w3[p - G_W3] = w;
    else
        net.b3[0] = w;
}

bool net_ok(const sim_mlp_net_t& n) {
    return n.w1 && n.w2 && n.b1 && n.b2 && n.w3 && n.b3 && n.mean && n.inv_std && n.rc > 0.f &&
           n.inv_w2 > 0.f;
}

}  // namespace

extern "C" int sim_mlp_energy(int R, const void* reps, const int* tiles, int ntiles,
                              const int* posq, const int* nbr, const int* nnbr, int n_stride,
                              sim_mlp_net_t net, float* c, float* gh, float* e_atom,
                              double* partial, void* stream) {
    if (R <= 0 || !reps || !tiles || ntiles <= 0 || !posq || !nbr || !nnbr || n_stride <= 0 ||
        !c || !partial || !net_ok(net))
        return SIM_E_ARG;
    hipLaunchKernelGGL(k_mlp_energy, dim3(ntiles * SIM_MLP_ROWS_PER_TILE), dim3(SIM_TILE), 0,
                       as_stream(stream), static_cast<const sim_replica_t*>(reps),
                       reinterpret_cast<const SimTile*>(tiles),
                       reinterpret_cast<const int4*>(posq), nbr, nnbr, n_stride, net, c, gh,
                       e_atom, partial, sqrt_threshold(net.rc));
    SIM_CHECK_LAUNCH();
    return SIM_OK;
}

extern "C" int sim_mlp_force(int R, const void* reps, const int* tiles, int ntiles,
                             const int* posq, const int* nbr, const int* nnbr, int n_stride,
                             sim_mlp_net_t net, const float* c, float* force, double* partial,
                             int virial, void* stream) {
    if (R <= 0 || !reps || !tiles || ntiles <= 0 || !posq || !nbr || !nnbr || n_stride <= 0 ||
        !c || !force || !partial || !net_ok(net))
        return SIM_E_ARG;
    const dim3 g(ntiles * SIM_MLP_ROWS_PER_TILE);
    auto* P = static_cast<const sim_replica_t*>(reps);
    auto* T = reinterpret_cast<const SimTile*>(tiles);
    auto* X = reinterpret_cast<const int4*>(posq);
    const float thr = sqrt_threshold(net.rc);
    if (virial)
        hipLaunchKernelGGL((k_mlp_force<true>), g, dim3(SIM_TILE), 0, as_stream(stream), P, T, X,
                           nbr, nnbr, n_stride, net, c, force, partial, thr);
    else
        hipLaunchKernelGGL((k_mlp_force<false>), g, dim3(SIM_TILE), 0, as_stream(stream), P, T, X,
                           nbr, nnbr, n_stride, net, c, force, partial, thr);
    SIM_CHECK_LAUNCH();
    return SIM_OK;
}

extern "C" int sim_mlp_adjoint(int R, const void* reps, const int* tiles, int ntiles,
                               const int* posq, const int* nbr, const int* nnbr, int n_stride,
                               sim_mlp_net_t net, const float* force, const float* force_ref,
                               float wF, float* v, void* stream) {
    if (R <= 0 || !reps || !tiles || ntiles <= 0 || !posq || !nbr || !nnbr || n_stride <= 0 ||
        !force || !force_ref || !v || !net_ok(net))
        return SIM_E_ARG;
    hipLaunchKernelGGL(k_mlp_adjoint, dim3(ntiles * SIM_MLP_ROWS_PER_TILE), dim3(SIM_TILE), 0,
                       as_stream(stream), static_cast<const sim_replica_t*>(reps),
                       reinterpret_cast<const SimTile*>(tiles),
                       reinterpret_cast<const int4*>(posq), nbr, nnbr, n_stride, net, force,
                       force_ref, wF, v, sqrt_threshold(net.rc));
    SIM_CHECK_LAUNCH();
    return SIM_OK;
}

extern "C" int sim_mlp_grad(int n, const int* atom_rep, const void* reps, const double* thermo,
                            const double* e_ref, float wE, float wF, sim_mlp_net_t net,
                            const float* gh, const float* v, const float* force,
                            const float* force_ref, float* grad_partial, int max_blocks,
                            float* grad, void* stream) {
    if (n <= 0 || !atom_rep || !reps || !thermo || !e_ref || !gh || !v || !force || !force_ref ||
        !grad_partial || max_blocks <= 0 || !grad || !net_ok(net))
        return SIM_E_ARG;
    const int groups = (n + 15) / 16;
    const int nblk = std::min(max_blocks, (groups + 3) / 4);
    hipLaunchKernelGGL(k_mlp_grad, dim3(nblk), dim3(256), 0, as_stream(stream), n, atom_rep,
                       static_cast<const sim_replica_t*>(reps), thermo, e_ref, wE, wF, net, gh, v,
                       force, force_ref, grad_partial);
    SIM_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_mlp_grad_sum, dim3((SIM_MLP_NGRAD + 255) / 256), dim3(256), 0,
                       as_stream(stream), nblk, grad_partial, grad);
    SIM_CHECK_LAUNCH();
    return SIM_OK;
}

extern "C" int sim_mlp_adam(float* master, const float* grad, float* m, float* v, int t, float lr,
                            float beta1, float beta2, float eps, sim_mlp_net_t net,
                            void* stream) {
    if (!master || !grad || !m || !v || t < 1 || !(lr > 0.f) || !(beta1 >= 0.f && beta1 < 1.f) ||
        !(beta2 >= 0.f && beta2 < 1.f) || !net.w1 || !net.w2 || !net.b1 || !net.b2 || !net.w3 ||
        !net.b3)
        return SIM_E_ARG;
    const float c1 = (float)(1.0 / (1.0 - std::pow((double)beta1, t)));
    const float c2 = (float)(1.0 / (1.0 - std::pow((double)beta2, t)));
    hipLaunchKernelGGL(k_mlp_adam, dim3((SIM_MLP_NPARAM + 255) / 256), dim3(256), 0,
                       as_stream(stream), master, grad, m, v, lr, beta1, beta2, eps, c1, c2, net);
    SIM_CHECK_LAUNCH();
    return SIM_OK;
}
