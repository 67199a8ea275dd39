// Per-group neighbour sets for the LDS-staged EAM passes (gfx950).
//
// The global-list EAM kernels read one 16-byte record per list entry; lanes
// of a wave read different cache lines at every entry, so each entry costs
// one L1 tag access and one texture-unit lane slot. A group of
// SIM_STAGE_TILES consecutive tiles (1,024 atoms, one workgroup of the
// persistent EAM kernels) draws its neighbours from a few hundred cells,
// ~5 records per atom against ~70 list entries. This builder lists, per
// group, the atoms of those cells ("members", cell by cell, so a cell's
// records are contiguous) and re-expresses every list entry as a 16-bit
// member position ("slot"); the staged kernels copy the members' records
// into LDS once per group and read the entries from there.
//
// After a rebuild the atoms of every cell are contiguous (sim_nl_build with a
// reorder) and each list names neighbours from its atom's 27-cell stencil.
// Two kernels, for the groups that hold a tile of a replica with flag set:
//   k_stage_cells, a workgroup of 1,024 threads per group (thread t: atom
//   t % 256 of tile 4G + t / 256):
//     1. home cells: the cells of the group's atoms (an LDS hash set);
//     2. union cells: the 27-cell stencils of the home cells, wrapped in
//        their replica's grid (an LDS hash set); their atom counts, prefixed,
//        give each cell's first member position; more members than cap, or a
//        hash set too small, leaves the group unstaged (count -1);
//     3. members[G][m] = cell_start[c] + p for member m = offset(c) + p, and
//        the hash set with each cell's offset and first atom (table) for 4;
//   k_stage_slots, a workgroup per tile, 4 threads per atom (two dependent
//   gathers per entry, the neighbour and its cell, batched 8 at a time):
//     4. slot of entry k of atom i: offset(c) + j - cell_start(c) for its
//        neighbour j in cell c = atom_cell[j]; rows past a list's end, up to
//        the longest list of its wave (64 atoms) rounded up to 8 rows, hold
//        the atom's own slot, so the kernels read valid slots wherever their
//        lock-step walk masks an entry. A neighbour outside the union marks
//        the group unstaged.
// Every hash probe is bounded by the table size. Groups whose tiles all
// belong to replicas with flag 0 keep their sets (their lists did not
// change).
#include "pair_common.h"

namespace {

constexpr int HC = 256;                // home-cell hash slots
constexpr int UC = SIM_STAGE_TABLE;    // union-cell hash slots
constexpr int NT = SIM_TILE * SIM_STAGE_TILES;
constexpr int LPA = NT / SIM_TILE;     // threads per atom of the slot pass
static_assert((UC & (UC - 1)) == 0 && UC % NT == 0, "table size");

template <int S>
__device__ __forceinline__ unsigned hslot(int c) {
    return ((unsigned)c * 2654435761u) >> (32 - __builtin_ctz(S));
}

// Insert key c into an open-addressing table of size S (power of two);
// false when no slot was found within S probes.
template <int S>
__device__ __forceinline__ bool insert(int* keys, int c) {
    unsigned h = hslot<S>(c);
    for (int n = 0; n < S; ++n) {
        const int old = atomicCAS(&keys[h], -1, c);
        if (old == -1 || old == c) return true;
        h = (h + 1) & (S - 1);
    }
    return false;
}

// Table slot of key c, or -1 within S probes.
template <int S>
__device__ __forceinline__ int find(const int* keys, int c) {
    unsigned h = hslot<S>(c);
    for (int n = 0; n < S; ++n) {
        const int k = keys[h];
        if (k == c) return (int)h;
        if (k == -1) return -1;
        h = (h + 1) & (S - 1);
    }
    return -1;
}

// Exclusive prefix of v over the 1,024 threads; *total gets the sum.
__device__ __forceinline__ int block_scan(int v, int* wsum, int* total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(incl, o, 64);
        if (lane >= o) incl += u;
    }
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int w = 0; w < NT / 64; ++w) {
            const int x = wsum[w];
            wsum[w] = run;
            run += x;
        }
        *total = run;
    }
    __syncthreads();
    return wsum[wid] + incl - v;
}

// Persistent walk of a workgroup over work items u = blockIdx.x + k
// gridDim.x (groups, or tiles with GPI = 4 tiles per group), gated by
// whether the item's group holds a tile of a replica with flag set. The
// gates of CH items at a time are read together (their tile and flag loads
// in flight at once) into gate[], so a workgroup skips unchanged items
// without a global round trip each. Every thread calls next() the same way.
template <int GPI>
struct GateWalk {
    static constexpr int CH = 16;
    int* gate;   // [CH] in LDS
    int k, nitems;
    const SimTile* tiles;
    const int* flag;
    int ntiles;
    __device__ GateWalk(int* g, int n, const SimTile* t, const int* f, int nt)
        : gate(g), k(-1), nitems(n), tiles(t), flag(f), ntiles(nt) {}
    // next flagged item of this workgroup, or -1 when done
    __device__ int next() {
        while (true) {
            ++k;
            const int u = (int)blockIdx.x + k * (int)gridDim.x;
            if (u >= nitems) return -1;
            if (k % CH == 0) {   // gates of items k .. k + CH - 1
                __syncthreads();
                if (threadIdx.x < CH) gate[threadIdx.x] = 0;
                __syncthreads();
                if (threadIdx.x < CH * SIM_STAGE_TILES) {
                    const int c = threadIdx.x / SIM_STAGE_TILES, g = threadIdx.x % SIM_STAGE_TILES;
                    const int uc = u + c * (int)gridDim.x;
                    const int G = uc / GPI;
                    const int tb = G * SIM_STAGE_TILES + g;
                    if (uc < nitems && tb < ntiles && (flag == nullptr || flag[tiles[tb].rep] != 0))
                        atomicOr(&gate[c], 1);
                }
                __syncthreads();
            }
            if (gate[k % CH]) return u;
        }
    }
};

__global__ void __launch_bounds__(NT)
    k_stage_cells(const sim_replica_t* __restrict__ reps, const SimTile* __restrict__ tiles,
                  int ntiles, const int* __restrict__ flag, const int* __restrict__ cell_rep,
                  const int* __restrict__ atom_cell, const int* __restrict__ cell_start,
                  const int* __restrict__ cell_count, int* __restrict__ members,
                  int* __restrict__ count, int* __restrict__ table, int cap) {
    __shared__ int hkeys[HC], hlist[HC];
    __shared__ int ukeys[UC], uidx[UC], ucell[UC], uoff[UC], ustart[UC];
    __shared__ int wsum[NT / 64];
    __shared__ int s_bad, s_nh, s_nu, s_total, gate[GateWalk<1>::CH];
    const int ngroups = (ntiles + SIM_STAGE_TILES - 1) / SIM_STAGE_TILES;
    GateWalk<1> walk(gate, ngroups, tiles, flag, ntiles);
    for (int G = walk.next(); G >= 0; G = walk.next()) {
        __syncthreads();   // the previous group's LDS reads are done
        const int t = threadIdx.x, g = t / SIM_TILE, lt = t % SIM_TILE;
        const int tb = G * SIM_STAGE_TILES + g;
        for (int h = t; h < UC; h += NT) ukeys[h] = -1;
        if (t < HC) hkeys[t] = -1;
        if (t == 0) {
            s_bad = 0;
            s_nh = 0;
            s_nu = 0;
        }
        bool active = false;
        int i = 0, base = 0;
        if (tb < ntiles) {
            const SimTile T = tiles[tb];
            const sim_replica_t& P = reps[T.rep];
            i = T.start + lt;
            base = P.atom_base;
            active = i < P.atom_base + P.natoms;
        }
        __syncthreads();
        // 1. home cells (atoms of a cell are contiguous: one insert per run)
        const int hc = active ? atom_cell[i] : -1;
        if (active && (lt == 0 || i == base || atom_cell[i - 1] != hc))
            if (!insert<HC>(hkeys, hc)) s_bad = 1;
        __syncthreads();
        if (t < HC && hkeys[t] >= 0) hlist[atomicAdd(&s_nh, 1)] = hkeys[t];
        __syncthreads();
        // 2. union cells: the wrapped 27-cell stencils of the home cells
        const int nh = s_nh;
        for (int w = t; w < nh * 27; w += NT) {
            const int h = hlist[w / 27], sidx = w % 27;
            const sim_replica_t& P = reps[cell_rep[h]];
            const int nx = P.ncell[0], ny = P.ncell[1], nz = P.ncell[2];
            const int l = h - P.cell_base;
            const int cx = l % nx, cy = (l / nx) % ny, cz = l / (nx * ny);
            int zz = cz + sidx / 9 - 1, yy = cy + (sidx / 3) % 3 - 1, xx = cx + sidx % 3 - 1;
            zz += zz < 0 ? nz : (zz >= nz ? -nz : 0);
            yy += yy < 0 ? ny : (yy >= ny ? -ny : 0);
            xx += xx < 0 ? nx : (xx >= nx ? -nx : 0);
            if (!insert<UC>(ukeys, P.cell_base + (zz * ny + yy) * nx + xx)) s_bad = 1;
        }
        __syncthreads();
        for (int h = t; h < UC; h += NT) {
            if (ukeys[h] >= 0) {
                const int k = atomicAdd(&s_nu, 1);
                ucell[k] = ukeys[h];
                uidx[h] = k;
            }
        }
        __syncthreads();
        // member offsets: prefix of the union cells' atom counts (PT per thread)
        constexpr int PT = UC / NT;
        const int nu = s_nu;
        int cnt_k[PT], sum = 0;
    #pragma unroll
        for (int q = 0; q < PT; ++q) {
            const int k = PT * t + q;
            cnt_k[q] = k < nu ? cell_count[ucell[k]] : 0;
            sum += cnt_k[q];
        }
        int off = block_scan(sum, wsum, &s_total);
        const int total = s_total;
        if (s_bad || total > cap) {
            if (t == 0) count[G] = -1;
            continue;   // uniform
        }
        // 3. members, cell by cell, and the table: (cell, first member
        // position, first atom) at each slot of the hash set
        int* mem = members + (size_t)G * cap;
    #pragma unroll
        for (int q = 0; q < PT; ++q) {
            const int k = PT * t + q;
            if (k < nu) {
                const int s0 = cell_start[ucell[k]];
                for (int p = 0; p < cnt_k[q]; ++p) mem[off + p] = s0 + p;
                uoff[k] = off;
                ustart[k] = s0;
            }
            off += cnt_k[q];
        }
        __syncthreads();
        int* T = table + (size_t)G * 3 * UC;
        for (int h = t; h < UC; h += NT) {
            const int c = ukeys[h];
            T[h] = c;
            T[UC + h] = c >= 0 ? uoff[uidx[h]] : 0;
            T[2 * UC + h] = c >= 0 ? ustart[uidx[h]] : 0;
        }
        if (t == 0) count[G] = total;
    }
}

__global__ void __launch_bounds__(NT)
    k_stage_slots(const sim_replica_t* __restrict__ reps, const SimTile* __restrict__ tiles,
                  int ntiles, const int* __restrict__ flag, const int* __restrict__ nbr,
                  const int* __restrict__ nnbr, int n, int max_nbr,
                  const int* __restrict__ atom_cell, int* __restrict__ count,
                  const int* __restrict__ table, unsigned short* __restrict__ slot) {
    __shared__ int tkey[UC], toff[UC], tstart[UC];
    __shared__ int cntl[SIM_TILE], wmax[SIM_TILE / 64];
    __shared__ int s_bad, s_live, gate[GateWalk<SIM_STAGE_TILES>::CH];
    GateWalk<SIM_STAGE_TILES> walk(gate, ntiles, tiles, flag, ntiles);
    for (int tb0 = walk.next(); tb0 >= 0; tb0 = walk.next()) {
        // tb is the workgroup's: the tile and replica records, the group's
        // count and table pointer come by scalar loads
        const int tb = __builtin_amdgcn_readfirstlane(tb0);
        const int G = tb / SIM_STAGE_TILES;
        const int t = threadIdx.x, a = t / LPA, l = t % LPA;
        const SimTile tile = tiles[tb];
        const sim_replica_t& P = reps[tile.rep];
        const int i = tile.start + a;
        const bool active = i < P.atom_base + P.natoms;
        const int cnt = active ? nnbr[i] : 0;
        const int ci = active ? atom_cell[i] : 0;
        __syncthreads();   // the previous tile's LDS reads are done
        // count[G] may turn -1 under another workgroup of the group: one read,
        // shared, keeps the branch uniform
        if (t == 0) s_live = count[G] >= 0 ? 1 : 0;
        __syncthreads();
        if (!s_live) continue;
        const int* T = table + (size_t)G * 3 * UC;
        for (int h = t; h < UC; h += NT) {
            tkey[h] = T[h];
            toff[h] = T[UC + h];
            tstart[h] = T[2 * UC + h];
        }
        if (l == 0) cntl[a] = cnt;
        if (t == 0) s_bad = 0;
        __syncthreads();
        if (t < SIM_TILE) {   // longest list of each wave (64 atoms) of the tile: waves 0-3
            const int m = wave_max_int(cntl[t]);
            if ((t & 63) == 0) wmax[t >> 6] = m;
        }
        __syncthreads();
        const int pad_end = min((wmax[a / 64] + 7) & ~7, max_nbr);
        if (active) {
            const int hs = find<UC>(tkey, ci);
            if (hs < 0) s_bad = 1;
            const int self = hs < 0 ? 0 : toff[hs] + (i - tstart[hs]);
            constexpr int B = 24;   // rows per thread in flight: one pass for lists up to 96
            for (int k0 = l; k0 < cnt; k0 += LPA * B) {
                int j[B], c[B];
#pragma unroll
                for (int u = 0; u < B; ++u) {
                    const int k = k0 + LPA * u;
                    j[u] = k < cnt ? nbr[(size_t)k * n + i] : i;
                }
#pragma unroll
                for (int u = 0; u < B; ++u) c[u] = atom_cell[j[u]];
#pragma unroll
                for (int u = 0; u < B; ++u) {
                    const int k = k0 + LPA * u;
                    if (k < cnt) {
                        const int h = find<UC>(tkey, c[u]);
                        if (h < 0)
                            s_bad = 1;
                        else
                            slot[(size_t)k * n + i] =
                                (unsigned short)(toff[h] + (j[u] - tstart[h]));
                    }
                }
            }
            for (int k = cnt + l; k < pad_end; k += LPA)
                slot[(size_t)k * n + i] = (unsigned short)self;
        }
        __syncthreads();
        if (t == 0 && s_bad) count[G] = -1;
    }
}

}  // namespace

extern "C" int sim_nl_stage(int R, const void* reps, const int* tiles, int ntiles, const int* flag,
                            const int* nbr, const int* nnbr, int n, int max_nbr,
                            const int* cell_rep, const int* atom_cell, const int* cell_start,
                            const int* cell_count, const sim_stage_t* stage, void* stream) {
    (void)R;
    if (!reps || !tiles || ntiles <= 0 || !nbr || !nnbr || n <= 0 || max_nbr <= 0 ||
        !cell_rep || !atom_cell || !cell_start || !cell_count || !stage || !stage->members ||
        !stage->count || !stage->slot || !stage->table || stage->cap < 1 ||
        stage->cap > SIM_STAGE_CAP_MAX)
        return SIM_E_ARG;
    const int ngroups = (ntiles + SIM_STAGE_TILES - 1) / SIM_STAGE_TILES;
    hipStream_t st = as_stream(stream);
    auto* P = static_cast<const sim_replica_t*>(reps);
    auto* T = reinterpret_cast<const SimTile*>(tiles);
    // persistent grids, two workgroups per CU, over the groups and over the
    // tiles (GateWalk): one replica's groups spread over the CUs
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        ncu <= 0)
        ncu = 256;
    hipLaunchKernelGGL(k_stage_cells, dim3(ngroups < 2 * ncu ? ngroups : 2 * ncu), dim3(NT), 0,
                       st, P, T, ntiles, flag, cell_rep, atom_cell, cell_start, cell_count,
                       stage->members, stage->count, stage->table, stage->cap);
    hipLaunchKernelGGL(k_stage_slots, dim3(ntiles < 2 * ncu ? ntiles : 2 * ncu), dim3(NT), 0, st,
                       P, T, ntiles, flag, nbr, nnbr, n, max_nbr, atom_cell, stage->count,
                       stage->table, stage->slot);
    SIM_CHECK_LAUNCH();
    return SIM_OK;
}
