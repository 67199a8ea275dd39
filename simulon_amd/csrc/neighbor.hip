// Cell-list Verlet neighbor build for packed replicas (gfx950).
//
// Replaces the O(N^2) brute force of core/neighbor_search/gpu_kdtree.py:90-156
// and the cubic-only, 500-capped kernel of
// "cuda source/neighbor_search_kernel.cu:6-138". The pair test is the
// reference's own (minimum image of core/box.py:122-130, squared distance
// (x*x + y*y) + z*z, strict d^2 < r_l^2, gpu_kdtree.py:132-138) evaluated in
// float64 for the MD path, or in float32/float64 exactly as the reference's
// dtype for the find_neighbors API; cells only enumerate candidates, so the
// pair set equals the brute force bit for bit.
//
// Steps (each kernel skips replicas whose rebuild flag is 0):
//   grid     per replica: cells of perpendicular width >= r_l in fractional
//            coordinates (triclinic-safe), capped by the replica's capacity
//   bin      cell id of each atom from its wrapped fractional coordinate,
//            slot within the cell by an integer atomic counter
//   scan     exclusive prefix of the cell counts (one workgroup per replica)
//   scatter  atom -> cell_atoms[start + slot]
//   order    per-cell insertion sort by atom index (deterministic lists)
//   list     thread per atom over the 27 (or fewer, when a dimension has < 3
//            cells) distinct stencil cells; column-major full list, and with
//            a reorder also the 16-bit column slots of the row kernels
//   rows     per-replica prefix of the cell-row counts (ny * nz)
#include <type_traits>

#include "sim_common.h"

namespace {

__device__ __forceinline__ bool rep_active(const int* flag, int r) {
    return flag == nullptr || flag[r] != 0;
}

__device__ void grid_cells(sim_replica_t& P);

// The box of the build: the trigger measures box changes from it.
__device__ __forceinline__ void reset_build_box(sim_replica_t& P) {
    P.mu_acc = 1.0;
    for (int k = 0; k < 9; ++k) P.hb_inv[k] = P.hinv[k];
}

// Cell grid of every rebuilding replica (threads r < R) and the cleared cell
// counts of their reserved cell ranges (threads c < ncells), in one launch.
__global__ void k_grid_clear(int R, sim_replica_t* reps, const int* flag, const int* cell_rep,
                             int ncells, int* cell_count) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < R && rep_active(flag, g)) {
        grid_cells(reps[g]);
        reset_build_box(reps[g]);
    }
    if (g < ncells && rep_active(flag, cell_rep ? cell_rep[g] : 0)) cell_count[g] = 0;
}

// k_grid_clear with the skin trigger (k_trigger) folded in: thread r < R
// decides replica r's flag, counts the build and computes its grid; a cell's
// thread takes the same decision from the same inputs (the displacement
// maximum and the replica's skin, r_l and mu_acc, none of which this launch
// writes: the grid touches only ncell, and mu_acc / hb_inv are reset and the
// maximum cleared by the scan that follows).
__global__ void k_grid_trigger(int R, sim_replica_t* reps, const unsigned long long* maxd2,
                               int* flag, int* nbuild, int force_all, const int* cell_rep,
                               int ncells, int* cell_count) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < R) {
        const int f = skin_trip(reps[g], maxd2[g], force_all) ? 1 : 0;
        flag[g] = f;
        if (nbuild) nbuild[g] += f;
        if (f) grid_cells(reps[g]);
    }
    if (g < ncells) {
        const int r = cell_rep[g];
        if (skin_trip(reps[r], maxd2[r], force_all)) cell_count[g] = 0;
    }
}

__device__ void grid_cells(sim_replica_t& P) {
    const double* h = P.h;
    // perpendicular widths: V / |b x c|, V / |c x a|, V / |a x b|
    double a[3] = {h[0], h[1], h[2]}, b[3] = {h[3], h[4], h[5]}, c[3] = {h[6], h[7], h[8]};
    auto cross_norm = [](const double* u, const double* v) {
        double x = u[1] * v[2] - u[2] * v[1];
        double y = u[2] * v[0] - u[0] * v[2];
        double z = u[0] * v[1] - u[1] * v[0];
        return sqrt(x * x + y * y + z * z);
    };
    double V = fabs(a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
                    a[2] * (b[0] * c[1] - b[1] * c[0]));
    double w[3] = {V / cross_norm(b, c), V / cross_norm(c, a), V / cross_norm(a, b)};
    int nc[3];
    for (int k = 0; k < 3; ++k) {
        double q = floor(w[k] / (P.rl * (1.0 + 1e-6)));
        nc[k] = q < 1.0 ? 1 : (q > 1024.0 ? 1024 : (int)q);
    }
    // shrink the largest dimension until the grid fits the reserved capacity
    while ((long long)nc[0] * nc[1] * nc[2] > P.cell_cap) {
        int m = (nc[0] >= nc[1] && nc[0] >= nc[2]) ? 0 : (nc[1] >= nc[2] ? 1 : 2);
        nc[m] -= 1;
    }
    P.ncell[0] = nc[0];
    P.ncell[1] = nc[1];
    P.ncell[2] = nc[2];
}


template <typename T, int STRIDE>
__device__ __forceinline__ void load_xyz(const T* x, int i, double& a, double& b, double& c) {
    a = (double)x[(size_t)i * STRIDE + 0];
    b = (double)x[(size_t)i * STRIDE + 1];
    c = (double)x[(size_t)i * STRIDE + 2];
}

__device__ __forceinline__ int cell_of(const sim_replica_t& P, double x, double y, double z) {
    const double* hi = P.hinv;
    double s[3];
    s[0] = x * hi[0] + y * hi[3] + z * hi[6];
    s[1] = x * hi[1] + y * hi[4] + z * hi[7];
    s[2] = x * hi[2] + y * hi[5] + z * hi[8];
    int c[3];
    for (int k = 0; k < 3; ++k) {
        double f = s[k] - floor(s[k]);
        int q = (int)(f * P.ncell[k]);
        c[k] = q < 0 ? 0 : (q >= P.ncell[k] ? P.ncell[k] - 1 : q);
    }
    return P.cell_base + (c[2] * P.ncell[1] + c[1]) * P.ncell[0] + c[0];
}

template <typename T, int STRIDE>
__global__ void k_bin(const sim_replica_t* reps, const SimTile* tiles, const T* x, const int* flag,
                      int* cell_count, int* atom_cell, int* atom_slot) {
    const SimTile t = tiles ? tiles[blockIdx.x] : SimTile{0, (int)blockIdx.x * SIM_TILE};
    if (!rep_active(flag, t.rep)) return;   // uniform: the whole tile
    const sim_replica_t& P = reps[t.rep];
    const int i = t.start + threadIdx.x;
    const bool act = i < P.atom_base + P.natoms;
    const int lane = threadIdx.x & 63;
    int cell = -1 - lane;   // inactive lanes: a run of their own, no count
    if (act) {
        double a, b, c;
        load_xyz<T, STRIDE>(x, i, a, b, c);
        cell = cell_of(P, a, b, c);
        atom_cell[i] = cell;
    }
    // Atoms keep the cell order of the last build, so neighbouring lanes
    // mostly share a cell: one atomic per run of equal cells in the wave
    // (its last lane) reserves the run's slots, instead of one per atom on
    // the same counter. rs / re: first and last lane of this lane's run.
    // (both shuffles run on every lane: a lane that skips one would read 0
    // from the lanes it skipped)
    const int up = __shfl_up(cell, 1), down = __shfl_down(cell, 1);
    const bool start = lane == 0 || up != cell;
    const bool end = lane == 63 || down != cell;
    int rs = start ? lane : 0, re = end ? lane : 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(rs, o), d = __shfl_down(re, o);
        if (lane >= o) rs = max(rs, u);
        if (lane + o < 64) re = min(re, d);
    }
    int base = 0;
    if (act && end) base = atomicAdd(&cell_count[cell], lane - rs + 1);
    base = __shfl(base, re);
    if (act) atom_slot[i] = base + (lane - rs);
}

// Exclusive scan of cnt[0..n) into start[0..n], one workgroup of 1024
// threads, each owning a contiguous chunk (half-list counts).
__global__ void __launch_bounds__(1024) k_scan_flat(const int* cnt, int n, int* start) {
    __shared__ int part[1024];
    const int tid = threadIdx.x;
    const int chunk = (n + 1023) / 1024;
    const int b = tid * chunk, e = min(n, b + chunk);
    int s = 0;
    for (int k = b; k < e; ++k) s += cnt[k];
    part[tid] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        int v = tid >= o ? part[tid - o] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    int run = part[tid] - s;
    for (int k = b; k < e; ++k) {
        start[k] = run;
        run += cnt[k];
    }
    if (tid == 1023) start[n] = part[1023];
}

// Exclusive prefix of the cell counts, one workgroup per rebuilding replica
// over its own cell range: replica r's atoms occupy [atom_base, atom_base +
// natoms), so its cell starts begin at atom_base and the other replicas'
// starts stay valid.
__device__ __forceinline__ bool stencil3(const sim_replica_t& P);

// With the trigger folded into the grid launch (trig != nullptr: the
// displacement maxima), workgroup r also clears replica r's maximum and, when
// rebuilding it, resets its build box; workgroup 0 also writes act (the cell
// prefix of the staged list build, act_base's work) when act != nullptr.
__device__ void act_prefix(int R, const sim_replica_t* reps, const int* flag, int* act);

__global__ void __launch_bounds__(1024) k_scan(sim_replica_t* reps, const int* cnt, int* start,
                                               const int* flag, unsigned long long* trig,
                                               int R, int* act) {
    __shared__ int wsum[16];
    __shared__ int carry;
    const int r = blockIdx.x;
    if (trig != nullptr && threadIdx.x == 0) trig[r] = 0ull;
    if (act != nullptr && r == 0) act_prefix(R, reps, flag, act);
    if (!rep_active(flag, r)) return;
    sim_replica_t& P = reps[r];
    if (trig != nullptr && threadIdx.x == 0) reset_build_box(P);
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int cb = P.cell_base, n = P.cell_cap;
    if (tid == 0) carry = P.atom_base;
    __syncthreads();
    for (int b0 = 0; b0 < n; b0 += 1024) {
        const int k = b0 + tid;
        const int v = k < n ? cnt[cb + k] : 0;
        int incl = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(incl, o, 64);
            if (lane >= o) incl += t;
        }
        if (lane == 63) wsum[wid] = incl;
        __syncthreads();
        if (wid == 0) {
            int w = lane < 16 ? wsum[lane] : 0;
#pragma unroll
            for (int o = 1; o < 16; o <<= 1) {
                const int t = __shfl_up(w, o, 64);
                if (lane >= o) w += t;
            }
            if (lane < 16) wsum[lane] = w;   // inclusive over waves
        }
        __syncthreads();
        const int base = carry + (wid > 0 ? wsum[wid - 1] : 0);
        if (k < n) start[cb + k] = base + incl - v;
        __syncthreads();
        if (tid == 0) carry += wsum[15];
        __syncthreads();
    }
    if (tid == 0) start[cb + n] = carry;   // = next replica's first start (or N)
}

// row_base[r] = number of cell rows (ny * nz) of replicas before r; rows of
// cells along x are the work items of the row pair kernels. One workgroup of
// 256 threads (block-wide scan in chunks of 256 replicas).
__global__ void __launch_bounds__(256)
    k_row_base(int R, const sim_replica_t* reps, const int* flag, int* row_base) {
    __shared__ int wsum[4];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    if (flag != nullptr) {
        int any = 0;
        for (int r = tid; r < R; r += 256) any |= flag[r] != 0;
        if (!__syncthreads_or(any)) return;
    }
    int carry = 0;
    for (int base = 0; base < R; base += 256) {
        const int r = base + tid;
        const int v = r < R ? reps[r].ncell[1] * reps[r].ncell[2] : 0;
        int incl = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(incl, o, 64);
            if (lane >= o) incl += t;
        }
        if (lane == 63) wsum[wid] = incl;
        __syncthreads();
        int before = 0;
        for (int w = 0; w < wid; ++w) before += wsum[w];
        const int total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        if (r < R) row_base[r] = carry + before + incl - v;
        carry += total;
        __syncthreads();
    }
    if (tid == 0) row_base[R] = carry;
}

__global__ void k_scatter(const sim_replica_t* reps, const SimTile* tiles, const int* flag,
                          const int* atom_cell, const int* atom_slot, const int* cell_start,
                          int* cell_atoms) {
    const SimTile t = tiles ? tiles[blockIdx.x] : SimTile{0, (int)blockIdx.x * SIM_TILE};
    if (!rep_active(flag, t.rep)) return;
    const sim_replica_t& P = reps[t.rep];
    int i = t.start + threadIdx.x;
    if (i >= P.atom_base + P.natoms) return;
    cell_atoms[cell_start[atom_cell[i]] + atom_slot[i]] = i;
}

// Rank of each atom among the atoms of its cell by index: the binning placed
// them in arrival order (atomics); scattering again by rank gives every cell
// its atoms in ascending index order (deterministic lists).
__global__ void k_cell_rank(const sim_replica_t* reps, const SimTile* tiles, const int* flag,
                            const int* atom_cell, const int* cell_start, const int* cell_count,
                            const int* cell_atoms, int* atom_slot) {
    const SimTile t = tiles ? tiles[blockIdx.x] : SimTile{0, (int)blockIdx.x * SIM_TILE};
    if (!rep_active(flag, t.rep)) return;
    const sim_replica_t& P = reps[t.rep];
    const int i = t.start + threadIdx.x;
    if (i >= P.atom_base + P.natoms) return;
    const int c = atom_cell[i];
    const int* a = cell_atoms + cell_start[c];
    const int n = cell_count[c];
    int rank = 0;
    for (int k = 0; k < n; ++k) rank += a[k] < i ? 1 : 0;
    atom_slot[i] = rank;
}

// Physical reorder of every rebuilding replica into cell order, wrapping each
// position into the primary cell (x -= floor(s) H) and accumulating the
// lattice translation in `image`, so that the atoms of a cell are contiguous
// and the cell-tiled pair kernels can stage whole cells with coalesced loads.
__global__ void k_permute_gather(const sim_replica_t* reps, const SimTile* tiles, const int* flag,
                                 const int* cell_atoms, const int* atom_cell, const double4* pos,
                                 const float* vel, const float* force, const long long* order,
                                 const int4* image, double4* pos_t, float* vel_t, float* force_t,
                                 long long* order_t, int4* image_t, int* cell_t,
                                 const float* vh, float* vh_t) {
    const SimTile t = tiles[blockIdx.x];
    if (!rep_active(flag, t.rep)) return;
    const sim_replica_t& P = reps[t.rep];
    const int s = t.start + threadIdx.x;
    if (s >= P.atom_base + P.natoms) return;
    const int a = cell_atoms[s];
    double4 x = pos[a];
    const double* hi = P.hinv;
    const double* h = P.h;
    const double f0 = floor(x.x * hi[0] + x.y * hi[3] + x.z * hi[6]);
    const double f1 = floor(x.x * hi[1] + x.y * hi[4] + x.z * hi[7]);
    const double f2 = floor(x.x * hi[2] + x.y * hi[5] + x.z * hi[8]);
    x.x -= f0 * h[0] + f1 * h[3] + f2 * h[6];
    x.y -= f0 * h[1] + f1 * h[4] + f2 * h[7];
    x.z -= f0 * h[2] + f1 * h[5] + f2 * h[8];
    pos_t[s] = x;
    int4 im = image[a];
    im.x += (int)f0;
    im.y += (int)f1;
    im.z += (int)f2;
    image_t[s] = im;
    for (int k = 0; k < 3; ++k) {
        vel_t[(size_t)s * 3 + k] = vel[(size_t)a * 3 + k];
        force_t[(size_t)s * 3 + k] = force[(size_t)a * 3 + k];
        vh_t[(size_t)s * 3 + k] = vh[(size_t)a * 3 + k];
    }
    order_t[s] = order[a];
    cell_t[s] = atom_cell[a];
}

// k_cell_rank, the second k_scatter and k_permute_gather in one launch: atom
// i (old slot) takes slot s = cell_start[c] + (its rank by index among the
// atoms of its cell c, from the first scatter) and writes its wrapped
// position, image, velocities, force, caller index and cell to the work
// arrays at s. The gather form reads from cell_atoms[s] = i after the second
// scatter, so both write the same values to the same slots; cell_atoms ends
// as the identity either way (k_permute_store).
__global__ void k_rank_permute(const sim_replica_t* reps, const SimTile* tiles, const int* flag,
                               const int* atom_cell, const int* cell_start, const int* cell_count,
                               const int* cell_atoms, const double4* pos, const float* vel,
                               const float* force, const long long* order, const int4* image,
                               double4* pos_t, float* vel_t, float* force_t, long long* order_t,
                               int4* image_t, int* cell_t, const float* vh, float* vh_t) {
    const SimTile t = tiles[blockIdx.x];
    if (!rep_active(flag, t.rep)) return;
    const sim_replica_t& P = reps[t.rep];
    const int i = t.start + threadIdx.x;
    if (i >= P.atom_base + P.natoms) return;
    const int c = atom_cell[i];
    const int c0 = cell_start[c];
    const int* a = cell_atoms + c0;
    const int n = cell_count[c];
    int rank = 0;
    // the cell's atoms are independent loads: eight in flight at a time
    int k = 0;
    for (; k + 8 <= n; k += 8) {
        int v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = a[k + u];
#pragma unroll
        for (int u = 0; u < 8; ++u) rank += v[u] < i ? 1 : 0;
    }
    for (; k < n; ++k) rank += a[k] < i ? 1 : 0;
    const int s = c0 + rank;
    double4 x = pos[i];
    const double* hi = P.hinv;
    const double* h = P.h;
    const double f0 = floor(x.x * hi[0] + x.y * hi[3] + x.z * hi[6]);
    const double f1 = floor(x.x * hi[1] + x.y * hi[4] + x.z * hi[7]);
    const double f2 = floor(x.x * hi[2] + x.y * hi[5] + x.z * hi[8]);
    x.x -= f0 * h[0] + f1 * h[3] + f2 * h[6];
    x.y -= f0 * h[1] + f1 * h[4] + f2 * h[7];
    x.z -= f0 * h[2] + f1 * h[5] + f2 * h[8];
    pos_t[s] = x;
    int4 im = image[i];
    im.x += (int)f0;
    im.y += (int)f1;
    im.z += (int)f2;
    image_t[s] = im;
    for (int k = 0; k < 3; ++k) {
        vel_t[(size_t)s * 3 + k] = vel[(size_t)i * 3 + k];
        force_t[(size_t)s * 3 + k] = force[(size_t)i * 3 + k];
        vh_t[(size_t)s * 3 + k] = vh[(size_t)i * 3 + k];
    }
    order_t[s] = order[i];
    cell_t[s] = c;
}

__global__ void k_permute_store(const sim_replica_t* reps, const SimTile* tiles, const int* flag,
                                const double4* pos_t, const float* vel_t, const float* force_t,
                                const long long* order_t, const int4* image_t, const int* cell_t,
                                double4* pos, float* vel, float* force, long long* order,
                                int4* image, int* atom_cell, int* cell_atoms, int* slot_of,
                                const float* vh_t, float* vh, int4* posq) {
    const SimTile t = tiles[blockIdx.x];
    if (!rep_active(flag, t.rep)) return;
    const sim_replica_t& P = reps[t.rep];
    const int s = t.start + threadIdx.x;
    if (s >= P.atom_base + P.natoms) return;
    pos[s] = pos_t[s];
    if (posq) posq[s] = quant_record(P, pos_t[s]);   // records of the wrapped positions
    for (int k = 0; k < 3; ++k) {
        vel[(size_t)s * 3 + k] = vel_t[(size_t)s * 3 + k];
        force[(size_t)s * 3 + k] = force_t[(size_t)s * 3 + k];
        vh[(size_t)s * 3 + k] = vh_t[(size_t)s * 3 + k];
    }
    const long long o = order_t[s];
    order[s] = o;
    image[s] = image_t[s];
    atom_cell[s] = cell_t[s];
    cell_atoms[s] = s;
    slot_of[o] = s;
}

// Stencil offsets of one dimension: distinct cells only.
__device__ __forceinline__ int stencil(int nc, int* off) {
    if (nc == 1) {
        off[0] = 0;
        return 1;
    }
    if (nc == 2) {
        off[0] = 0;
        off[1] = 1;
        return 2;
    }
    off[0] = -1;
    off[1] = 0;
    off[2] = 1;
    return 3;
}

__device__ __forceinline__ bool stencil3(const sim_replica_t& P) {
    return P.ncell[0] >= 3 && P.ncell[1] >= 3 && P.ncell[2] >= 3;
}

// WRAPPED: positions were just wrapped into the primary cell and permuted
// into cell order (sim_nl_build with a reorder), so cell_atoms is the identity
// and the image of a candidate is its stencil cell's lattice translation. A
// candidate clearly inside or outside the list radius (relative margin on d^2
// 1e-9 in float64, 1e-5 in float32) is then decided from that displacement;
// only near-ties take the reference's minimum-image arithmetic, so the pair
// set is unchanged.
template <typename TC, typename T, int STRIDE, bool WRAPPED>
__global__ void __launch_bounds__(SIM_TILE)
    k_list(const sim_replica_t* reps, const SimTile* tiles, const T* x, const int* flag,
           const int* atom_cell, const int* cell_count, const int* cell_start,
           const int* cell_atoms, int* nbr, int* nnbr, int n_stride, int max_nbr,
           int* overflow, double* pos_last, unsigned short* nbr16, int* stencil_max,
           int skip_staged) {
    const SimTile t = tiles ? tiles[blockIdx.x] : SimTile{0, (int)blockIdx.x * SIM_TILE};
    if (!rep_active(flag, t.rep)) return;
    const sim_replica_t& P = reps[t.rep];
    if (skip_staged && stencil3(P)) return;   // built by k_list_staged
    const int i = t.start + threadIdx.x;
    if (i >= P.atom_base + P.natoms) return;
    const TC xi = (TC)x[(size_t)i * STRIDE + 0];
    const TC yi = (TC)x[(size_t)i * STRIDE + 1];
    const TC zi = (TC)x[(size_t)i * STRIDE + 2];
    const TC rl2 = (TC)P.rl2;
    const int nx = P.ncell[0], ny = P.ncell[1], nz = P.ncell[2];
    int cl = atom_cell[i] - P.cell_base;
    const int cx = cl % nx, cy = (cl / nx) % ny, cz = cl / (nx * ny);
    int ox[3], oy[3], oz[3];
    const int sx = stencil(nx, ox), sy = stencil(ny, oy), sz = stencil(nz, oz);
    int cnt = 0;
    // Slots for the row kernels: (x column 0..2 of the stencil) << 14 | offset
    // of the atom within that column's 9 cells, the cells taken in (z, y)
    // order (the visiting order below), so a column is one contiguous LDS run.
    int coff[3] = {0, 0, 0};
    const bool fast = WRAPPED && sx == 3 && sy == 3 && sz == 3;
    const double margin = sizeof(TC) == 8 ? 1e-9 : 1e-5;   // far above each type's rounding
    const TC lo2 = rl2 * (TC)(1.0 - margin), hi2 = rl2 * (TC)(1.0 + margin);
    for (int a = 0; a < sz; ++a) {
        int zz = cz + oz[a];
        const int fz = zz < 0 ? -1 : (zz >= nz ? 1 : 0);
        zz -= fz * nz;
        for (int b = 0; b < sy; ++b) {
            int yy = cy + oy[b];
            const int fy = yy < 0 ? -1 : (yy >= ny ? 1 : 0);
            yy -= fy * ny;
            for (int c = 0; c < sx; ++c) {
                int xx = cx + ox[c];
                const int fx = xx < 0 ? -1 : (xx >= nx ? 1 : 0);
                xx -= fx * nx;
                // lattice translation of this stencil cell's atoms
                const TC shx = (TC)(fx * P.h[0] + fy * P.h[3] + fz * P.h[6]);
                const TC shy = (TC)(fx * P.h[1] + fy * P.h[4] + fz * P.h[7]);
                const TC shz = (TC)(fx * P.h[2] + fy * P.h[5] + fz * P.h[8]);
                const int cell = P.cell_base + (zz * ny + yy) * nx + xx;
                const int s0 = cell_start[cell], s1 = s0 + cell_count[cell];
                // candidates in batches of 8: every load of a batch in flight
                // together, the tests into a bit mask, then the appends in order
                constexpr int B = 8;
                for (int sb = s0; sb < s1; sb += B) {
                    TC px[B], py[B], pz[B];
                    int jb[B];
#pragma unroll
                    for (int u = 0; u < B; ++u) {
                        const int ss = min(sb + u, s1 - 1);
                        jb[u] = WRAPPED ? ss : cell_atoms[ss];
                        px[u] = (TC)x[(size_t)jb[u] * STRIDE + 0];
                        py[u] = (TC)x[(size_t)jb[u] * STRIDE + 1];
                        pz[u] = (TC)x[(size_t)jb[u] * STRIDE + 2];
                    }
                    unsigned m = 0;
#pragma unroll
                    for (int u = 0; u < B; ++u) {
                        if (sb + u >= s1 || jb[u] == i) continue;
                        const TC xj = px[u], yj = py[u], zj = pz[u];
                        bool in;
                        const TC ex = xi - (xj + shx), ey = yi - (yj + shy), ez = zi - (zj + shz);
                        const TC e2 = (ex * ex + ey * ey) + ez * ez;
                        if (fast && e2 < lo2) {
                            in = true;
                        } else if (fast && e2 > hi2) {
                            in = false;
                        } else {
                            TC dx = xi - xj, dy = yi - yj, dz = zi - zj;
                            min_image_ref<TC>(dx, dy, dz, P);
                            const TC d2 = (dx * dx + dy * dy) + dz * dz;
                            in = d2 < rl2;
                        }
                        m |= in ? 1u << u : 0u;
                    }
                    while (m != 0) {
                        const int u = __builtin_ctz(m);
                        m &= m - 1;
                        const int ss = sb + u;
                        if (cnt < max_nbr) {
                            nbr[(size_t)cnt * n_stride + i] = WRAPPED ? ss : cell_atoms[ss];
                            if (nbr16 != nullptr)
                                nbr16[(size_t)i * max_nbr + cnt] =
                                    (unsigned short)((c << 14) | (coff[c] + (ss - s0)));
                        }
                        ++cnt;
                    }
                }
                coff[c] += s1 - s0;
            }
        }
    }
    nnbr[i] = cnt < max_nbr ? cnt : max_nbr;
    if (cnt > max_nbr) atomicOr(overflow, 1);
    if (stencil_max != nullptr) {   // read first: rarely raised, the atomics contend
        const int cmax = max(coff[0], max(coff[1], coff[2]));
        const int ctot = coff[0] + coff[1] + coff[2];
        if (cmax > stencil_max[0]) atomicMax(stencil_max, cmax);       // largest column
        if (ctot > stencil_max[1]) atomicMax(stencil_max + 1, ctot);   // largest stencil
        if (cmax > 16383) atomicOr(overflow, 8);   // offsets no longer fit 14 bits
    }
    if (pos_last != nullptr) {
        pos_last[(size_t)i * 4 + 0] = (double)x[(size_t)i * STRIDE + 0];
        pos_last[(size_t)i * 4 + 1] = (double)x[(size_t)i * STRIDE + 1];
        pos_last[(size_t)i * 4 + 2] = (double)x[(size_t)i * STRIDE + 2];
        pos_last[(size_t)i * 4 + 3] = 0.0;
    }
}

// ---- staged list build (after a reorder, replicas with >= 3 cells per
// dimension) -------------------------------------------------------------
// One cell at a time per workgroup: the 27 stencil cells' atoms (contiguous
// per cell after the reorder) are staged in LDS, translated to the image next
// to the cell, and the cell's atoms (8 lanes each, a contiguous run of the
// staged candidates per lane) test them twice: a count pass, a prefix over
// the 8 lanes, and a write pass. The decision per candidate and the order of
// every list (stencil cells in (z, y, x) order, atoms in index order) are
// k_list's (WRAPPED), so the lists are identical. A stencil of more than
// SIM_LIST_CAP atoms reads its candidates from global memory instead (512
// measured far slower on W and Ar: many stencils exceed it).
#ifndef SIM_LIST_CAP
#define SIM_LIST_CAP 1024
#endif
#define SIM_LIST_LANES 8
#ifndef SIM_LIST_GRID
#define SIM_LIST_GRID 2048   // persistent workgroups over the rebuilding cells
#endif
#ifndef SIM_LIST_STAGE_F32
#define SIM_LIST_STAGE_F32 1
#endif
// Replicas whose cell prefix (act) the staged build holds in LDS for its
// search of a cell's replica; more replicas search act in global memory.
#define SIM_LIST_ACT_LDS 512

// act[r] = cells before replica r of the rebuilding replicas built here;
// act[R] = their total. Called by every thread of one workgroup (a multiple
// of 64 threads, at most 1,024).
__device__ void act_prefix(int R, const sim_replica_t* reps, const int* flag, int* act) {
    __shared__ int wsum[16];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int nt = blockDim.x, nw = nt >> 6;
    int carry = 0;
    for (int base = 0; base < R; base += nt) {
        const int r = base + tid;
        int v = 0;
        if (r < R && rep_active(flag, r) && stencil3(reps[r]))
            v = reps[r].ncell[0] * reps[r].ncell[1] * reps[r].ncell[2];
        int incl = v;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(incl, o, 64);
            if (lane >= o) incl += t;
        }
        if (lane == 63) wsum[wid] = incl;
        __syncthreads();
        int before = 0, total = 0;
        for (int w = 0; w < nw; ++w) {
            before += w < wid ? wsum[w] : 0;
            total += wsum[w];
        }
        if (r < R) act[r] = carry + before + incl - v;
        carry += total;
        __syncthreads();
    }
    if (tid == 0) act[R] = carry;
}

__global__ void __launch_bounds__(256)
    k_act_base(int R, const sim_replica_t* reps, const int* flag, int* act) {
    act_prefix(R, reps, flag, act);
}

// Waves per SIMD the staged build's registers are held to (5: 96 VGPRs, no
// spills with float staging; measured faster than 4, 6 spills).
#ifndef SIM_LIST_WAVES
#define SIM_LIST_WAVES 5
#endif
// ACT_LDS: R <= SIM_LIST_ACT_LDS, the replica search reads act from LDS.
template <typename TC, typename T, int STRIDE, bool ACT_LDS>
__global__ void __launch_bounds__(SIM_TILE, SIM_LIST_WAVES)
    k_list_staged(int R, const sim_replica_t* reps, const int* act, const T* x,
                  const int* cell_count, const int* cell_start, int* nbr, int* nnbr, int n_stride,
                  int max_nbr, int* overflow, double* pos_last, unsigned short* nbr16,
                  int* stencil_max) {
    // staged candidates: float offsets from the home cell's first atom (half
    // the LDS of float64, so more workgroups per CU); a candidate they place
    // within 1e-5 of r_l^2 takes the float64 tests below, so the lists stay
    // those of k_list
    using QT = typename std::conditional<SIM_LIST_STAGE_F32 != 0, float, TC>::type;
    __shared__ QT qx[SIM_LIST_CAP], qy[SIM_LIST_CAP], qz[SIM_LIST_CAP];
    __shared__ int qs[SIM_LIST_CAP];
    __shared__ unsigned short q16[SIM_LIST_CAP];
    __shared__ int mstart[27], mcount[27], mpre[28], mslot[27], mcol[3];
    __shared__ TC msh[27][3];
    __shared__ double mref[3];
    __shared__ int sact[ACT_LDS ? SIM_LIST_ACT_LDS + 1 : 1];
    const int tid = threadIdx.x;
    const int total = act[R];
    if ((int)blockIdx.x >= total) return;   // uniform: no cell for this workgroup
    // the replica search of every cell below reads act from LDS (one load
    // round trip per workgroup instead of one per probe)
    if (ACT_LDS) {
        for (int r = tid; r <= R; r += SIM_TILE) sact[r] = act[r];
        __syncthreads();
    }
    const double margin = sizeof(TC) == 8 ? 1e-9 : 1e-5;   // as k_list
    const double smargin = sizeof(QT) == 8 ? 1e-9 : 1e-5;
    for (int ac = blockIdx.x; ac < total; ac += gridDim.x) {
        int lo = 0, hi = R - 1;   // last replica with act[r] <= ac
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if ((ACT_LDS ? sact[mid] : act[mid]) <= ac) lo = mid;
            else hi = mid - 1;
        }
        // equal on every lane (ac is the workgroup's): the replica's fields
        // then come by scalar loads
        lo = __builtin_amdgcn_readfirstlane(lo);
        const sim_replica_t& P = reps[lo];
        const int cl = ac - act[lo];
        const int nx = P.ncell[0], ny = P.ncell[1], nz = P.ncell[2];
        const int cx = cl % nx, cy = (cl / nx) % ny, cz = cl / (nx * ny);
        if (tid < 27) {   // stencil cell k = (dz * 3 + dy) * 3 + dx, offsets -1..1
            const int a = tid / 9, b = (tid / 3) % 3, c = tid % 3;
            int zz = cz + a - 1, yy = cy + b - 1, xx = cx + c - 1;
            const int fz = zz < 0 ? -1 : (zz >= nz ? 1 : 0);
            const int fy = yy < 0 ? -1 : (yy >= ny ? 1 : 0);
            const int fx = xx < 0 ? -1 : (xx >= nx ? 1 : 0);
            zz -= fz * nz;
            yy -= fy * ny;
            xx -= fx * nx;
            const int cell = P.cell_base + (zz * ny + yy) * nx + xx;
            mstart[tid] = cell_start[cell];
            mcount[tid] = cell_count[cell];
            msh[tid][0] = (TC)(fx * P.h[0] + fy * P.h[3] + fz * P.h[6]);
            msh[tid][1] = (TC)(fx * P.h[1] + fy * P.h[4] + fz * P.h[7]);
            msh[tid][2] = (TC)(fx * P.h[2] + fy * P.h[5] + fz * P.h[8]);
        }
        __syncthreads();
        if (tid == 0) {   // prefix in stencil order; column slot bases (k_list's coff)
            int coff[3] = {0, 0, 0};
            mpre[0] = 0;
            for (int k = 0; k < 27; ++k) {
                mpre[k + 1] = mpre[k] + mcount[k];
                mslot[k] = ((k % 3) << 14) | coff[k % 3];
                coff[k % 3] += mcount[k];
            }
            mcol[0] = coff[0];
            mcol[1] = coff[1];
            mcol[2] = coff[2];
            // reference point of the float offsets: the home cell's lower
            // corner (from the box, no memory round trip); a candidate the
            // offsets place near r_l takes the float64 test either way
            const double fx = (double)cx / nx, fy = (double)cy / ny, fz = (double)cz / nz;
            for (int c = 0; c < 3; ++c) mref[c] = fx * P.h[c] + fy * P.h[3 + c] + fz * P.h[6 + c];
        }
        __syncthreads();
        const int nt = mpre[27];
        const bool staged = nt <= SIM_LIST_CAP;
        auto cell_of = [&](int f) {
            int k = 0;
#pragma unroll
            for (int step = 16; step > 0; step >>= 1)
                if (k + step <= 26 && mpre[k + step] <= f) k += step;
            return k;
        };
        if (staged) {
            for (int f = tid; f < nt; f += SIM_TILE) {
                const int k = cell_of(f);
                const int sj = mstart[k] + (f - mpre[k]);
                qx[f] = (QT)((double)((TC)x[(size_t)sj * STRIDE + 0] + msh[k][0]) - mref[0]);
                qy[f] = (QT)((double)((TC)x[(size_t)sj * STRIDE + 1] + msh[k][1]) - mref[1]);
                qz[f] = (QT)((double)((TC)x[(size_t)sj * STRIDE + 2] + msh[k][2]) - mref[2]);
                qs[f] = sj;
                q16[f] = (unsigned short)(mslot[k] + (f - mpre[k]));
            }
        }
        __syncthreads();
        const int h0 = mstart[13], nh = mcount[13];   // the cell itself
        const TC rl2 = (TC)P.rl2;
        const TC lo2 = rl2 * (TC)(1.0 - margin), hi2 = rl2 * (TC)(1.0 + margin);
        const QT slo2 = (QT)(P.rl2 * (1.0 - smargin)), shi2 = (QT)(P.rl2 * (1.0 + smargin));
        constexpr int L = SIM_LIST_LANES;
        for (int a0 = 0; a0 < nh; a0 += SIM_TILE / L) {
            const int a = a0 + tid / L, l = tid % L;
            const bool live = a < nh;
            const int i = h0 + (live ? a : 0);
            const TC xi = (TC)x[(size_t)i * STRIDE + 0];
            const TC yi = (TC)x[(size_t)i * STRIDE + 1];
            const TC zi = (TC)x[(size_t)i * STRIDE + 2];
            const QT xr = (QT)((double)xi - mref[0]), yr = (QT)((double)yi - mref[1]),
                     zr = (QT)((double)zi - mref[2]);
            const int f0 = (int)(((long long)l * nt) / L), f1 = (int)(((long long)(l + 1) * nt) / L);
            // candidate f: index, image position, column slot; then k_list's test
            auto test = [&](int f, int& sj, unsigned short& slot) {
                if (staged) {
                    sj = qs[f];
                    slot = q16[f];
                    if (sj == i) return false;
                    const QT ex = xr - qx[f], ey = yr - qy[f], ez = zr - qz[f];
                    const QT e2 = (ex * ex + ey * ey) + ez * ez;
                    if (e2 < slo2) return true;
                    if (e2 > shi2) return false;
                } else {
                    const int k = cell_of(f);
                    sj = mstart[k] + (f - mpre[k]);
                    const TC px = (TC)x[(size_t)sj * STRIDE + 0] + msh[k][0];
                    const TC py = (TC)x[(size_t)sj * STRIDE + 1] + msh[k][1];
                    const TC pz = (TC)x[(size_t)sj * STRIDE + 2] + msh[k][2];
                    slot = (unsigned short)(mslot[k] + (f - mpre[k]));
                    if (sj == i) return false;
                    const TC ex = xi - px, ey = yi - py, ez = zi - pz;
                    const TC e2 = (ex * ex + ey * ey) + ez * ez;
                    if (e2 < lo2) return true;
                    if (e2 > hi2) return false;
                }
                const TC xj = (TC)x[(size_t)sj * STRIDE + 0];
                const TC yj = (TC)x[(size_t)sj * STRIDE + 1];
                const TC zj = (TC)x[(size_t)sj * STRIDE + 2];
                TC dx = xi - xj, dy = yi - yj, dz = zi - zj;
                min_image_ref<TC>(dx, dy, dz, P);
                return (dx * dx + dy * dy) + dz * dz < rl2;
            };
            // staged: one test pass into a bit mask of the lane's run (at most
            // SIM_LIST_CAP / L = 128 candidates, 8 LDS reads in flight), the
            // writes from the mask; otherwise a count pass and a write pass
            constexpr int W = SIM_LIST_CAP / L / 32;
            unsigned bits[W];
#pragma unroll
            for (int w = 0; w < W; ++w) bits[w] = 0u;
            int mine = 0;
            if (live && staged) {
                // branch-free float test of 8 candidates at a time: inside
                // (below the lower margin) and near-tie bits; the rare near
                // ties then take the float64 test. The atom itself is the
                // home cell's (stencil cell 13's) candidate mpre[13] + a: the
                // stencil cells are distinct, so it appears nowhere else.
                const int fself = mpre[13] + a;
#pragma unroll
                for (int w = 0; w < W; ++w) {
                    for (int u0 = 0; u0 < 32; u0 += 8) {
                        const int fb = f0 + 32 * w + u0;
                        if (fb >= f1) break;
                        unsigned in = 0u, near = 0u;
#pragma unroll
                        for (int u = 0; u < 8; ++u) {
                            const int f = fb + u;
                            const bool ok = f < f1 && f != fself;
                            const int fc = f < f1 ? f : fb;   // a staged index either way
                            const QT ex = xr - qx[fc], ey = yr - qy[fc], ez = zr - qz[fc];
                            const QT e2 = (ex * ex + ey * ey) + ez * ez;
                            in |= (ok && e2 < slo2) ? 1u << u : 0u;
                            near |= (ok && !(e2 < slo2) && !(e2 > shi2)) ? 1u << u : 0u;
                        }
                        while (near != 0u) {
                            const int u = __builtin_ctz(near);
                            near &= near - 1u;
                            int sj;
                            unsigned short slot;
                            if (test(fb + u, sj, slot)) in |= 1u << u;
                        }
                        bits[w] |= in << u0;
                    }
                    mine += __builtin_popcount(bits[w]);
                }
            } else if (live) {
                for (int f = f0; f < f1; ++f) {
                    int sj;
                    unsigned short slot;
                    mine += test(f, sj, slot) ? 1 : 0;
                }
            }
            int incl = mine;
#pragma unroll
            for (int o = 1; o < L; o <<= 1) {
                const int v = __shfl_up(incl, o, L);
                if (l >= o) incl += v;
            }
            const int cnt = __shfl(incl, (int)((threadIdx.x & 63) | (L - 1)), 64);
            if (live) {
                int k = incl - mine;
                auto put = [&](int sj, unsigned short slot) {
                    if (k < max_nbr) {
                        nbr[(size_t)k * n_stride + i] = sj;
                        if (nbr16 != nullptr) nbr16[(size_t)i * max_nbr + k] = slot;
                    }
                    ++k;
                };
                if (staged) {
#pragma unroll
                    for (int w = 0; w < W; ++w) {
                        unsigned m = bits[w];
                        while (m != 0) {
                            const int f = f0 + 32 * w + __builtin_ctz(m);
                            m &= m - 1;
                            put(qs[f], q16[f]);
                        }
                    }
                } else {
                    for (int f = f0; f < f1; ++f) {
                        int sj;
                        unsigned short slot;
                        if (test(f, sj, slot)) put(sj, slot);
                    }
                }
                if (l == 0) {
                    nnbr[i] = cnt < max_nbr ? cnt : max_nbr;
                    if (cnt > max_nbr) atomicOr(overflow, 1);
                    if (pos_last != nullptr) {
                        pos_last[(size_t)i * 4 + 0] = (double)x[(size_t)i * STRIDE + 0];
                        pos_last[(size_t)i * 4 + 1] = (double)x[(size_t)i * STRIDE + 1];
                        pos_last[(size_t)i * 4 + 2] = (double)x[(size_t)i * STRIDE + 2];
                        pos_last[(size_t)i * 4 + 3] = 0.0;
                    }
                }
            }
        }
        if (tid == 0 && stencil_max != nullptr) {
            const int cmax = max(mcol[0], max(mcol[1], mcol[2]));
            if (cmax > stencil_max[0]) atomicMax(stencil_max, cmax);   // read first: rarely raised
            if (nt > stencil_max[1]) atomicMax(stencil_max + 1, nt);
            if (cmax > 16383) atomicOr(overflow, 8);
        }
        __syncthreads();   // LDS is reused by the next cell
    }
}

// The skin trigger folded into sim_nl_update's build (k_grid_trigger).
struct NlTrigger {
    unsigned long long* maxd2;
    int* flag;
    int* nbuild;
    int force_all;
    int* act;   // [R+1] cell prefix of the staged list build
};

template <typename TC, typename T, int STRIDE>
int build_impl(int R, sim_replica_t* reps, const SimTile* tiles, int ntiles, const T* x,
               double* pos_last, const int* flag, const int* cell_rep, int ncells, int* atom_cell,
               int* atom_slot, int* cell_count, int* cell_start, int* cell_atoms, int* nbr,
               int* nnbr, int n_stride, int max_nbr, int* overflow, hipStream_t st,
               const sim_nl_reorder_t* ro = nullptr, const NlTrigger* trig = nullptr) {
    const bool perm = ro != nullptr && tiles != nullptr;
    if (perm && (!ro->vel || !ro->force || !ro->order || !ro->slot_of || !ro->image ||
                 !ro->pos_tmp || !ro->vel_tmp || !ro->force_tmp || !ro->order_tmp ||
                 !ro->image_tmp || !ro->cell_tmp || !ro->vel_half || !ro->vel_half_tmp))
        return SIM_E_ARG;
    // the fused form (trigger in the grid launch, act with the scan, rank and
    // gather in one launch) needs the reorder
    if (trig != nullptr && (!perm || cell_rep == nullptr)) return SIM_E_ARG;
    const dim3 gc((max(ncells, R) + 255) / 256);
    if (trig != nullptr)
        hipLaunchKernelGGL(k_grid_trigger, gc, dim3(256), 0, st, R, reps, trig->maxd2, trig->flag,
                           trig->nbuild, trig->force_all, cell_rep, ncells, cell_count);
    else
        hipLaunchKernelGGL(k_grid_clear, gc, dim3(256), 0, st, R, reps, flag, cell_rep, ncells,
                           cell_count);
    hipLaunchKernelGGL((k_bin<T, STRIDE>), dim3(ntiles), dim3(SIM_TILE), 0, st, reps, tiles, x,
                       flag, cell_count, atom_cell, atom_slot);
    if (trig != nullptr)
        hipLaunchKernelGGL(k_scan, dim3(R), dim3(1024), 0, st, reps, cell_count, cell_start, flag,
                           trig->maxd2, R, trig->act);
    else
        hipLaunchKernelGGL(k_scan, dim3(R), dim3(1024), 0, st, reps, cell_count, cell_start, flag,
                           nullptr, 0, nullptr);
    if (ro != nullptr && ro->row_base != nullptr)
        hipLaunchKernelGGL(k_row_base, dim3(1), dim3(256), 0, st, R, reps, flag, ro->row_base);
    hipLaunchKernelGGL(k_scatter, dim3(ntiles), dim3(SIM_TILE), 0, st, reps, tiles, flag,
                       atom_cell, atom_slot, cell_start, cell_atoms);
    auto* P4 = reinterpret_cast<double4*>(const_cast<T*>(x));
    if (trig != nullptr) {
        hipLaunchKernelGGL(k_rank_permute, dim3(ntiles), dim3(SIM_TILE), 0, st, reps, tiles, flag,
                           atom_cell, cell_start, cell_count, cell_atoms, P4, ro->vel, ro->force,
                           ro->order, reinterpret_cast<const int4*>(ro->image),
                           reinterpret_cast<double4*>(ro->pos_tmp), ro->vel_tmp, ro->force_tmp,
                           ro->order_tmp, reinterpret_cast<int4*>(ro->image_tmp), ro->cell_tmp,
                           ro->vel_half, ro->vel_half_tmp);
    } else {
        hipLaunchKernelGGL(k_cell_rank, dim3(ntiles), dim3(SIM_TILE), 0, st, reps, tiles, flag,
                           atom_cell, cell_start, cell_count, cell_atoms, atom_slot);
        hipLaunchKernelGGL(k_scatter, dim3(ntiles), dim3(SIM_TILE), 0, st, reps, tiles, flag,
                           atom_cell, atom_slot, cell_start, cell_atoms);
        if (perm)
            hipLaunchKernelGGL(k_permute_gather, dim3(ntiles), dim3(SIM_TILE), 0, st, reps, tiles,
                               flag, cell_atoms, atom_cell, P4, ro->vel, ro->force, ro->order,
                               reinterpret_cast<const int4*>(ro->image),
                               reinterpret_cast<double4*>(ro->pos_tmp), ro->vel_tmp,
                               ro->force_tmp, ro->order_tmp,
                               reinterpret_cast<int4*>(ro->image_tmp), ro->cell_tmp,
                               ro->vel_half, ro->vel_half_tmp);
    }
    if (perm) {
        hipLaunchKernelGGL(k_permute_store, dim3(ntiles), dim3(SIM_TILE), 0, st, reps, tiles, flag,
                           reinterpret_cast<const double4*>(ro->pos_tmp), ro->vel_tmp,
                           ro->force_tmp, ro->order_tmp,
                           reinterpret_cast<const int4*>(ro->image_tmp), ro->cell_tmp, P4,
                           ro->vel, ro->force, ro->order, reinterpret_cast<int4*>(ro->image),
                           atom_cell, cell_atoms, ro->slot_of, ro->vel_half_tmp, ro->vel_half,
                           reinterpret_cast<int4*>(ro->posq));
    }
    if (ro != nullptr && tiles != nullptr) {   // positions wrapped and in cell order
        // staged build for replicas with >= 3 cells per dimension (their cell
        // prefix in act, or in atom_slot, free after the scatter); k_list for
        // the rest
        const bool staged = trig != nullptr || n_stride >= R + 1;
        if (staged) {
            int* act = trig != nullptr ? trig->act : atom_slot;
            if (trig == nullptr)
                hipLaunchKernelGGL(k_act_base, dim3(1), dim3(256), 0, st, R, reps, flag, act);
            if (R <= SIM_LIST_ACT_LDS)
                hipLaunchKernelGGL((k_list_staged<TC, T, STRIDE, true>), dim3(SIM_LIST_GRID),
                                   dim3(SIM_TILE), 0, st, R, reps, act, x, cell_count, cell_start,
                                   nbr, nnbr, n_stride, max_nbr, overflow, pos_last, ro->nbr16,
                                   ro->stencil_max);
            else
                hipLaunchKernelGGL((k_list_staged<TC, T, STRIDE, false>), dim3(SIM_LIST_GRID),
                                   dim3(SIM_TILE), 0, st, R, reps, act, x, cell_count, cell_start,
                                   nbr, nnbr, n_stride, max_nbr, overflow, pos_last, ro->nbr16,
                                   ro->stencil_max);
        }
        if (!(staged && ro->stencil3_only))
            hipLaunchKernelGGL((k_list<TC, T, STRIDE, true>), dim3(ntiles), dim3(SIM_TILE), 0,
                               st, reps, tiles, x, flag, atom_cell, cell_count, cell_start,
                               cell_atoms, nbr, nnbr, n_stride, max_nbr, overflow, pos_last,
                               ro->nbr16, ro->stencil_max, staged ? 1 : 0);
    } else {
        hipLaunchKernelGGL((k_list<TC, T, STRIDE, false>), dim3(ntiles), dim3(SIM_TILE), 0, st,
                           reps, tiles, x, flag, atom_cell, cell_count, cell_start, cell_atoms,
                           nbr, nnbr, n_stride, max_nbr, overflow, pos_last,
                           ro ? ro->nbr16 : nullptr, ro ? ro->stencil_max : nullptr, 0);
    }
    SIM_CHECK_LAUNCH();
    return SIM_OK;
}

// ---- half list for the find_neighbors API ---------------------------------

__global__ void k_half_count(const int* nbr, const int* nnbr, int n, int* hc) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int c = 0;
    for (int k = 0; k < nnbr[i]; ++k) c += nbr[(size_t)k * n + i] > i;
    hc[i] = c;
}

template <typename T>
__global__ void k_half_fill(const sim_replica_t* rep, const T* x, int n, const int* nbr,
                            const int* nnbr, const int* hs, long long* ei, long long E, T* dist) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int base = hs[i];
    int m = 0;
    // selection by insertion into the output slice, ascending j
    for (int k = 0; k < nnbr[i]; ++k) {
        const int j = nbr[(size_t)k * n + i];
        if (j <= i) continue;
        int p = m;
        while (p > 0 && ei[E + base + p - 1] > j) {
            ei[E + base + p] = ei[E + base + p - 1];
            --p;
        }
        ei[E + base + p] = j;
        ++m;
    }
    const sim_replica_t& P = *rep;
    for (int k = 0; k < m; ++k) {
        const int j = (int)ei[E + base + k];
        ei[base + k] = i;
        T dx = x[(size_t)i * 3 + 0] - x[(size_t)j * 3 + 0];
        T dy = x[(size_t)i * 3 + 1] - x[(size_t)j * 3 + 1];
        T dz = x[(size_t)i * 3 + 2] - x[(size_t)j * 3 + 2];
        min_image_ref<T>(dx, dy, dz, P);
        dist[base + k] = sqrt((dx * dx + dy * dy) + dz * dz);
    }
}

}  // namespace

extern "C" int sim_nl_build(int R, void* reps, const int* tiles, int ntiles, int n,
                            double* pos4, double* pos_last, const int* flag,
                            const int* cell_rep, int ncells, int* atom_cell, int* atom_slot,
                            int* cell_count, int* cell_start, int* cell_atoms, int* nbr, int* nnbr,
                            int max_nbr, int* overflow, int precise, const sim_nl_reorder_t* ro,
                            void* stream) {
    if (R <= 0 || !reps || !tiles || ntiles <= 0 || n <= 0 || !pos4 || !cell_rep || ncells <= 0 ||
        !atom_cell || !atom_slot || !cell_count || !cell_start || !cell_atoms || !nbr || !nnbr ||
        max_nbr <= 0 || !overflow)
        return SIM_E_ARG;
    auto* P = static_cast<sim_replica_t*>(reps);
    auto* T = reinterpret_cast<const SimTile*>(tiles);
    if (precise)
        return build_impl<double, double, 4>(R, P, T, ntiles, pos4, pos_last, flag, cell_rep, ncells,
                                             atom_cell, atom_slot, cell_count, cell_start,
                                             cell_atoms, nbr, nnbr, n, max_nbr, overflow,
                                             as_stream(stream), ro);
    return build_impl<float, double, 4>(R, P, T, ntiles, pos4, pos_last, flag, cell_rep, ncells,
                                        atom_cell, atom_slot, cell_count, cell_start, cell_atoms,
                                        nbr, nnbr, n, max_nbr, overflow, as_stream(stream), ro);
}

extern "C" int sim_nl_update(int R, void* reps, const int* tiles, int ntiles, int n, double* pos4,
                             double* pos_last, double* maxd2, int* flag, int* nbuild,
                             int force_all, const int* cell_rep, int ncells, int* atom_cell,
                             int* atom_slot, int* cell_count, int* cell_start, int* cell_atoms,
                             int* nbr, int* nnbr, int max_nbr, int* overflow, int precise,
                             int* act, const sim_nl_reorder_t* ro, void* stream) {
    if (R <= 0 || !reps || !tiles || ntiles <= 0 || n <= 0 || !pos4 || !maxd2 || !flag ||
        !cell_rep || ncells <= 0 || !atom_cell || !atom_slot || !cell_count || !cell_start ||
        !cell_atoms || !nbr || !nnbr || max_nbr <= 0 || !overflow || !act || !ro)
        return SIM_E_ARG;
    auto* P = static_cast<sim_replica_t*>(reps);
    auto* T = reinterpret_cast<const SimTile*>(tiles);
    const NlTrigger trig{reinterpret_cast<unsigned long long*>(maxd2), flag, nbuild,
                         force_all ? 1 : 0, act};
    if (precise)
        return build_impl<double, double, 4>(R, P, T, ntiles, pos4, pos_last, flag, cell_rep,
                                             ncells, atom_cell, atom_slot, cell_count, cell_start,
                                             cell_atoms, nbr, nnbr, n, max_nbr, overflow,
                                             as_stream(stream), ro, &trig);
    return build_impl<float, double, 4>(R, P, T, ntiles, pos4, pos_last, flag, cell_rep, ncells,
                                        atom_cell, atom_slot, cell_count, cell_start, cell_atoms,
                                        nbr, nnbr, n, max_nbr, overflow, as_stream(stream), ro,
                                        &trig);
}

extern "C" int sim_nl_build_xyz(void* rep, const void* xyz, int dtype, int n, int* atom_cell,
                                int* atom_slot, int* cell_count, int* cell_start, int* cell_atoms,
                                int ncells, int* nbr, int* nnbr, int max_nbr, int* overflow,
                                void* stream) {
    if (!rep || !xyz || n <= 0 || !atom_cell || !atom_slot || !cell_count || !cell_start ||
        !cell_atoms || ncells <= 0 || !nbr || !nnbr || max_nbr <= 0 || !overflow)
        return SIM_E_ARG;
    auto* P = static_cast<sim_replica_t*>(rep);
    const int ntiles = (n + SIM_TILE - 1) / SIM_TILE;
    hipStream_t st = as_stream(stream);
    if (dtype == 0)
        return build_impl<float, float, 3>(1, P, nullptr, ntiles, static_cast<const float*>(xyz),
                                           nullptr, nullptr, nullptr, ncells, atom_cell, atom_slot,
                                           cell_count, cell_start, cell_atoms, nbr, nnbr, n,
                                           max_nbr, overflow, st);
    if (dtype == 1)
        return build_impl<double, double, 3>(1, P, nullptr, ntiles,
                                             static_cast<const double*>(xyz), nullptr, nullptr,
                                             nullptr, ncells, atom_cell, atom_slot, cell_count,
                                             cell_start, cell_atoms, nbr, nnbr, n, max_nbr,
                                             overflow, st);
    return SIM_E_ARG;
}

extern "C" int sim_nl_half_count(const int* nbr, const int* nnbr, int n, int max_nbr,
                                 int* half_count, int* half_start, void* stream) {
    if (!nbr || !nnbr || n <= 0 || max_nbr <= 0 || !half_count || !half_start) return SIM_E_ARG;
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(k_half_count, dim3((n + 255) / 256), dim3(256), 0, st, nbr, nnbr, n,
                       half_count);
    hipLaunchKernelGGL(k_scan_flat, dim3(1), dim3(1024), 0, st, half_count, n, half_start);
    SIM_CHECK_LAUNCH();
    return SIM_OK;
}

extern "C" int sim_nl_half_fill(void* rep, const void* xyz, int dtype, int n, const int* nbr,
                                const int* nnbr, int max_nbr, const int* half_start,
                                long long* edge_index, long long E, void* dist, void* stream) {
    if (!rep || !xyz || n <= 0 || !nbr || !nnbr || max_nbr <= 0 || !half_start || E < 0)
        return SIM_E_ARG;
    if (E == 0) return SIM_OK;
    if (!edge_index || !dist) return SIM_E_ARG;
    hipStream_t st = as_stream(stream);
    auto* P = static_cast<const sim_replica_t*>(rep);
    if (dtype == 0)
        hipLaunchKernelGGL(k_half_fill<float>, dim3((n + 255) / 256), dim3(256), 0, st, P,
                           static_cast<const float*>(xyz), n, nbr, nnbr, half_start, edge_index,
                           E, static_cast<float*>(dist));
    else if (dtype == 1)
        hipLaunchKernelGGL(k_half_fill<double>, dim3((n + 255) / 256), dim3(256), 0, st, P,
                           static_cast<const double*>(xyz), n, nbr, nnbr, half_start, edge_index,
                           E, static_cast<double*>(dist));
    else
        return SIM_E_ARG;
    SIM_CHECK_LAUNCH();
    return SIM_OK;
}
