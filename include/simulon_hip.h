/*
 * simulon_hip.h — C ABI of the MI355X (gfx950) MD-step library libsimulon_hip.so.
 *
 * Every entry point takes raw device pointers, element counts and a
 * hipStream_t (passed as void*), launches stream-ordered work only (no
 * device-wide synchronisation, no allocation), and returns 0 on success or a
 * negative SIM_E* code.  No entry point keeps global state, except the one
 * allocation-free entry point (sim_eam_energy_force_noalloc, with its fixed
 * per-GPU work array): two calls of the others on different data may run
 * concurrently on different streams.
 *
 * State is packed for R >= 1 independent systems ("replicas"): atoms of
 * replica r occupy [atom_base, atom_base + natoms) of every per-atom array,
 * and a single system is simply R = 1.  Atom-parallel kernels walk a tile
 * table (replica, first atom) with SIM_TILE atoms per tile so that each
 * workgroup reads its replica's box and parameters as uniform scalars.
 *
 * Per-atom arrays:
 *   pos4   double[N][4]  x, y, z (unwrapped, Angstrom) and the atom type as a double
 *   vel    float [N][3]  velocity, Angstrom/ps
 *   force  float [N][3]  eV/Angstrom
 *   nbr    int32 [max_nbr][N]  full Verlet list, column-major (entry k of atom i
 *                              at nbr[k*N + i]), global atom indices
 *   nnbr   int32 [N]
 *   posq   int32 [N][4]  fractional fixed-point records (sim_pos_quantize) read
 *                        by the full-list force kernels
 *
 * The full-list force kernels address records and list rows by 32-bit byte
 * offsets: N < 2^28 and n_stride < 2^30.
 *
 * Reference interfaces replaced (JinZongxiao/Simulon):
 *   sim_nl_build / sim_nl_half_*   simulon_cuda.neighbor_search_cuda
 *        ("cuda source/neighbor_search_kernel.cu:73-138") and the PyTorch
 *        brute force core/neighbor_search/gpu_kdtree.py:90-156
 *   sim_nl_trigger                 skin/2 test of io_utils/reader.py:136-149
 *   sim_nl_update                  sim_nl_trigger + sim_nl_build (with a
 *        reorder) in fewer launches, the same two reference interfaces
 *   sim_lj_force (/ _second)       simulon_cuda.lj_energy_force_cuda
 *        ("cuda source/lj_energy_force.cpp:10-20") + configure_lj_smoothing
 *        ("cuda source/ops.cpp:39-45") and core/force/lennard_jones_force.py:15-95
 *   sim_eam_density / sim_eam_force   simulon_cuda.density_pass / force_pass
 *        ("cuda source/eam_cuda_ext.cpp:30-60") and the rest of
 *        core/force/eam_force_cu.py:461-579; their _staged variants and
 *        sim_eam_force_second (the force pass with the integrator's second
 *        half; with four more arguments also the next step's first half)
 *        the same passes, with sim_nl_stage building the LDS staging (no
 *        counterpart in the reference)
 *   sim_pos_quantize               Box.wrap_positions / minimum_image
 *        (core/box.py:109-142) in fractional fixed point
 *   sim_integrate_first / _second  core/integrator/integrator.py:91-124
 *   sim_reduce                     the torch .sum()/einsum reductions of the
 *        force fields (e.g. core/force/lennard_jones_force.py:92-94),
 *        core/md_model.py:83-86 and integrator.py:118-119
 *   sim_berendsen                  core/barostat.py:36-68
 *   sim_aniso_barostat             AnisotropicNPTBarostat.step, core/barostat.py:131-208
 *   sim_box_deform                 UniaxialTensileLoader.step,
 *        core/mechanics/loading.py:39-78
 *   sim_mlp_*                      the learned-potential backbone slot of
 *        machine_learning_potentials/machine_learning_force.py:85-142 (the
 *        network itself is this library's own, see the MLP section below)
 */
#ifndef SIMULON_HIP_H
#define SIMULON_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define SIM_OK 0
#define SIM_E_ARG (-1)       /* bad argument (null pointer, size out of range) */
#define SIM_E_LAUNCH (-2)    /* kernel launch failed */
#define SIM_E_TYPES (-3)     /* too many atom types */
#define SIM_E_SIZE (-4)      /* problem larger than an entry point's fixed work space */

#define SIM_TILE 256         /* atoms per tile = threads per workgroup */
#define SIM_MAX_TYPES 4
#define SIM_NPART 16         /* doubles per tile in the partial-sum array */

/* Columns of the partial-sum and per-replica thermo arrays (double[16]). */
#define SIM_T_EPOT 0         /* potential energy, eV */
#define SIM_T_VIRIAL 1       /* scalar virial with the force field's own sign */
#define SIM_T_W 2            /* virial tensor sum f (x) r: xx yy zz xy xz yz (2..7) */
#define SIM_T_KE 8           /* kinetic energy 1/2 sum m v^2 */
#define SIM_T_K 9            /* kinetic tensor sum m v_a v_b: xx yy zz xy xz yz (9..14) */
#define SIM_T_MAXD2 15       /* (sim_reduce output) unused */

/* Per-replica state, one struct per replica in device memory. */
typedef struct sim_replica {
    double h[9];          /* lattice vectors as rows, row-major */
    double hinv[9];
    double volume;
    double rl2;           /* squared Verlet list radius (value of fp32 (rc+skin)^2) */
    double trig2;         /* squared rebuild displacement (skin/2)^2 */
    double rl;            /* list radius */
    double skin;
    double mu_acc;        /* lower bound on how much the box changes since the last
                             build shrink any separation (smallest singular value) */
    double dt;            /* time step, ps */
    double mu;            /* last Berendsen scaling factor */
    double press;         /* last instantaneous pressure, bar */
    double target_p;      /* Berendsen target pressure, bar */
    double tau_p;         /* Berendsen time constant, ps */
    double kappa;         /* compressibility, 1/bar */
    double mu_max;
    float damp;           /* Langevin exp(-gamma dt) */
    float rff[SIM_MAX_TYPES]; /* Langevin kick std per type, Angstrom/ps */
    float rcf;            /* force cutoff */
    float dtf;            /* time step as float (reference arithmetic) */
    int atom_base, natoms;
    int cell_base, cell_cap;
    int ncell[3];
    int ortho;
    int langevin;         /* 1: Langevin O step in the second half */
    int barostat;         /* 1: Berendsen, 2: anisotropic, after the second half */
    int pad;
    /* anisotropic barostat (core/barostat.py:78-208) and uniaxial loading
     * (core/mechanics/loading.py:39-78) */
    double eta_dot[3];    /* log strain rates of the lattice rows, 1/ps */
    double p_target3[3];  /* target pressure per lattice axis, bar */
    double p_axis[3];     /* last pressure projected on the lattice axes, bar */
    double gamma_p;       /* strain-rate friction, 1/ps */
    double max_eta_dot;
    double p_tol;         /* deadband around the target, bar */
    double temp_k;        /* temperature of the stochastic term, K */
    double deform[3];     /* loading: row scale factors applied before every step */
    double amap[9];       /* affine map x' = x amap of the last row scaling */
    double hb_inv[9];     /* hinv at the last list build (rebuild trigger bound) */
    float vscale[3];      /* per-axis velocity factors of the last row scaling */
    int axes;             /* anisotropic: bit a set = lattice row a controlled */
    int stochastic;
    int loading;          /* 1: apply deform at the start of every step */
    int pad2;
    unsigned long long noise_key; /* this replica's Philox key; the key of its
                                     Langevin kicks and barostat noise is
                                     (seed argument) XOR noise_key */
    double baro_dt;       /* > 0: time step of the barostat update instead of dt
                             (a standalone barostat call with its own dt) */
} sim_replica_t;

typedef struct sim_types {
    float mass[SIM_MAX_TYPES];   /* internal mass eV ps^2 / A^2, per type */
    int ntypes;
    int pad;
} sim_types_t;

typedef struct sim_lj_params {
    float eps[SIM_MAX_TYPES * SIM_MAX_TYPES];
    float sig[SIM_MAX_TYPES * SIM_MAX_TYPES];
    float rc;        /* cutoff */
    float rs;        /* switch start (mode 2) */
    int mode;        /* 0 raw, 1 force-shift, 2 polynomial switch */
    int ntypes;      /* atom types present, 1..SIM_MAX_TYPES (1: one uniform pair) */
} sim_lj_params_t;

typedef struct sim_eam_params {
    const void* pass1;    /* float4 [nt][nt][n_r]: f(k), f(k+1), phi(k), phi(k+1) */
    const void* pass2;    /* float4 [nt][nt][n_r]: f'(k), f'(k+1), phi'(k), phi'(k+1) */
    const float* emb;     /* float2 [nt][n_rho]: F(k), F(k+1) */
    const float* demb;    /* float2 [nt][n_rho]: F'(k), F'(k+1) */
    int n_r;
    int n_rho;
    int ntypes;
    int pad;
    float inv_dr;         /* float(1/dr) */
    float r_clamp;        /* float(r_max (1 - 1e-7)) */
    float cutoff;         /* r < cutoff counts */
    float rho_lo[SIM_MAX_TYPES];
    float rho_hi[SIM_MAX_TYPES];   /* float(lo + drho (n-1)(1 - 1e-7)) */
    float inv_drho[SIM_MAX_TYPES];
} sim_eam_params_t;

/* Library identification: writes "simulon_hip <version> gfx950". */
int sim_version(char* buf, int len);

/* ---- neighbor list -------------------------------------------------- */

/* Full Verlet list build for every replica whose rebuild flag is set
 * (flag == NULL: all replicas).  Cell grid, counting sort by cell, an
 * exclusive scan over all cells, a per-cell index sort (deterministic order)
 * and the list itself.  Also copies pos4 into pos_last (the reference of the
 * displacement trigger) and resets mu_acc.
 *   cell_rep: int[C] replica of each cell slot (C = total cell capacity)
 *   atom_cell, atom_slot: int[N] work arrays; cell_count/cell_start: int[C+1];
 *   cell_atoms: int[N]; overflow: int[1], set to 1 when a list exceeds max_nbr
 *   n: total atoms (the list's column stride); tiles: int2 (replica, first atom)
 *   precise: 1 = pair test in float64, 0 = in float32 (reference fp32 semantics)
 *   ro: optional physical reorder (NULL: atoms keep their slots). With ro, the
 *       atoms of each rebuilding replica are permuted into cell order and
 *       wrapped into the primary cell before the list is built, and the list is
 *       also written as 16-bit column slots (row kernels; requires >= 3 cells
 *       per dimension) and the cell-row prefix row_base is updated. */
typedef struct sim_nl_reorder {
    float* vel;            /* [N][3], permuted with the atoms */
    float* force;          /* [N][3], permuted with the atoms */
    long long* order;      /* [N] caller index of the atom in each slot */
    int* slot_of;          /* [N] inverse of order */
    int* image;            /* [N][4] lattice translation removed by wrapping */
    double* pos_tmp;       /* work arrays of the same shapes */
    float* vel_tmp;
    float* force_tmp;
    long long* order_tmp;
    int* image_tmp;
    int* cell_tmp;
    unsigned short* nbr16; /* [N][max_nbr] row-major list of column slots, or NULL */
    int* stencil_max;      /* [2] largest stencil-column and 27-cell stencil populations
                              (atomic max), or NULL */
    float* vel_half;       /* [N][3] half-step velocities (a rebuild falls between the halves) */
    float* vel_half_tmp;
    int* row_base;         /* [R+1] prefix of the cell-row counts ny*nz, or NULL */
    int* posq;             /* [N][4] position records rewritten for the reordered
                              replicas (sim_pos_quantize), or NULL */
    int stencil3_only;     /* 1: the caller guarantees >= 3 cells per dimension in every
                              rebuilt replica (the thread-per-atom fallback list kernel,
                              which serves the others, is not launched); 0: launch it */
    int pad_ro;
} sim_nl_reorder_t;

/* LDS staging of the EAM passes' neighbour records: no counterpart in the
 * reference, whose kernels read every neighbour from global memory
 * ("cuda source/eam_cuda_ext_kernel.cu:20-113"). A group is SIM_STAGE_TILES
 * consecutive tiles (1,024 atoms); its members are the atoms of the cells
 * in the 27-cell stencils of its atoms' cells, cell by cell. */
#define SIM_STAGE_TILES 4
#define SIM_STAGE_CAP_MAX 8192
#define SIM_STAGE_TABLE 2048   /* cells of a group's stencil union at most */
typedef struct sim_stage {
    int* members;          /* [ngroups][cap] member atoms of each group */
    int* count;            /* [ngroups] members of each group; -1: unstaged */
    int* table;            /* [ngroups][3 * SIM_STAGE_TABLE] work space of sim_nl_stage */
    unsigned short* slot;  /* [max_nbr][n] column-major like nbr: entry k of atom i as the
                              position of that neighbour in its group's members; rows from
                              the list's end to its wave's longest list (64 atoms of a
                              tile) rounded up to 8 hold the atom's own position */
    int cap;               /* member capacity per group, 1..SIM_STAGE_CAP_MAX */
    int pad;
} sim_stage_t;

/* Members, counts and slots of every group that holds a tile of a replica
 * with flag set (every group when flag is NULL), from the current lists and
 * cells (run right after sim_nl_build with a reorder, same flag and arrays:
 * the atoms of every cell contiguous, atom_cell[i] the cell of atom i);
 * ngroups = ceil(ntiles / SIM_STAGE_TILES). A group whose members exceed
 * cap gets count -1 and is read from the list by the kernels. */
int sim_nl_stage(int R, const void* reps, const int* tiles, int ntiles, const int* flag,
                 const int* nbr, const int* nnbr, int n, int max_nbr, const int* cell_rep,
                 const int* atom_cell, const int* cell_start, const int* cell_count,
                 const sim_stage_t* stage, void* stream);

int sim_nl_build(int R, void* reps, const int* tiles, int ntiles, int n,
                 double* pos4, double* pos_last, const int* flag,
                 const int* cell_rep, int ncells,
                 int* atom_cell, int* atom_slot, int* cell_count, int* cell_start,
                 int* cell_atoms, int* nbr, int* nnbr, int max_nbr, int* overflow,
                 int precise, const sim_nl_reorder_t* ro, void* stream);

/* sim_nl_trigger followed by sim_nl_build with a reorder, in four fewer
 * launches: the skin test (flag, nbuild, the maxima cleared) runs inside the
 * build's grid launch, the cell prefix of the staged list build is written
 * with the cell scan into act (int[R+1] work array, the caller's), and the
 * rank, second scatter and gather of the reorder are one launch. flag,
 * nbuild, the lists, the order and every array end as after the two calls.
 * tiles, cell_rep and ro are required. Same interfaces as those two calls
 * (io_utils/reader.py:136-149 for the trigger). */
int sim_nl_update(int R, void* reps, const int* tiles, int ntiles, int n, double* pos4,
                  double* pos_last, double* maxd2, int* flag, int* nbuild, int force_all,
                  const int* cell_rep, int ncells, int* atom_cell, int* atom_slot,
                  int* cell_count, int* cell_start, int* cell_atoms, int* nbr, int* nnbr,
                  int max_nbr, int* overflow, int precise, int* act,
                  const sim_nl_reorder_t* ro, void* stream);

/* Same build for a single system given as plain [N,3] coordinates, float32
 * (dtype 0) or float64 (dtype 1): the pair test uses exactly the reference's
 * arithmetic in that dtype (core/box.py:122-130, gpu_kdtree.py:130-138).
 * rep: one sim_replica_t in device memory (box, rl2, cell grid). */
int sim_nl_build_xyz(void* rep, const void* xyz, int dtype, int n,
                     int* atom_cell, int* atom_slot, int* cell_count, int* cell_start,
                     int* cell_atoms, int ncells, int* nbr, int* nnbr, int max_nbr,
                     int* overflow, void* stream);

/* Half list (i < j) of a full list: per-atom counts of j > i and their
 * exclusive prefix (int[N+1], last entry = E) ... */
int sim_nl_half_count(const int* nbr, const int* nnbr, int n, int max_nbr,
                      int* half_count, int* half_start, void* stream);
/* ... then edge_index int64 [2][E] in lexicographic order and the distances
 * |minimum_image(x_i - x_j)| in the coordinate dtype. */
int sim_nl_half_fill(void* rep, const void* xyz, int dtype, int n,
                     const int* nbr, const int* nnbr, int max_nbr,
                     const int* half_start, long long* edge_index, long long E,
                     void* dist, void* stream);

/* Skin trigger: flag[r] = force_all || (max displacement over replica r)^2 >
 * threshold, where the threshold accounts for box scaling since the build.
 * maxd2: double[R] running maxima written by sim_integrate_first (reset here);
 * nbuild: optional int[R] rebuild counters (incremented when flag is set). */
int sim_nl_trigger(int R, void* reps, double* maxd2, int* flag, int* nbuild, int force_all,
                   void* stream);

/* ---- force fields ----------------------------------------------------- */

/* Fractional fixed-point position records for the full-list kernels below:
 * posq[i] = (round(s_k 2^32) mod 2^32 for k = 0..2, type), s = x H^-1 of the
 * atom's replica wrapped into [0, 1) (core/box.py:132-142). The wrapped 32-bit
 * difference of two records is the fractional minimum image
 * (core/box.py:109-130); quantisation moves an atom by at most 2^-33 of each
 * box vector. Recompute after positions or boxes change.
 *   posq: int[N][4] */
int sim_pos_quantize(const void* reps, const int* tiles, int ntiles, const double* pos4,
                     int* posq, void* stream);

/* LJ energy/force/virial over the full list (core/force/lennard_jones_force.py
 * :15-95); writes force[N][3] and the SIM_T_EPOT..SIM_T_W columns of
 * partial[ntiles][16]. virial: 0 skips the per-pair virial (columns written
 * as 0). posq: records of sim_pos_quantize. */
int sim_lj_force(int R, const void* reps, const int* tiles, int ntiles,
                 const int* posq, const int* nbr, const int* nnbr, int n_stride,
                 sim_lj_params_t lj, float* force, double* partial, int virial,
                 void* stream);

/* sim_lj_force with the step's second half fused in: after an atom's force
 * the pass does sim_integrate_second's O + B for it (the same bits) and
 * writes the kinetic columns SIM_T_KE..+6 of the tile's partial row, so no
 * second-half launch follows (core/force/lennard_jones_force.py:15-95 and
 * core/integrator/integrator.py:102-124). Arguments after virial as
 * sim_integrate_second's. */
int sim_lj_force_second(int R, const void* reps, const int* tiles, int ntiles,
                        const int* posq, const int* nbr, const int* nnbr, int n_stride,
                        sim_lj_params_t lj, float* force, double* partial, int virial,
                        sim_types_t types, float* vel_half, float* vel, const float* noise,
                        const long long* key, unsigned long long seed,
                        const unsigned long long* step_dev, void* stream);

/* EAM (core/force/eam_force_cu.py:461-579). Pass 1: rho_i (written to rho[N]
 * when rho is not NULL), F(rho_i), F'(rho_i) (written to fp[N] when fp is not
 * NULL) and the energy column; rec (int[N][4], not aliasing posq) receives the
 * records with F'(rho_i) in place of the type, for pass 2. At least one of fp
 * and rec must be given. Pass 2: forces and virial columns from rec
 * (positions and F' of the neighbours) and posq (atom types, read when
 * ntypes > 1). */
int sim_eam_density(int R, const void* reps, const int* tiles, int ntiles,
                    const int* posq, const int* nbr, const int* nnbr, int n_stride,
                    sim_eam_params_t eam, float* rho, float* fp, int* rec, double* partial,
                    void* stream);
int sim_eam_force(int R, const void* reps, const int* tiles, int ntiles,
                  const int* posq, const int* rec, const int* nbr, const int* nnbr,
                  int n_stride, sim_eam_params_t eam, float* force, double* partial,
                  int virial, void* stream);

/* sim_eam_force fused with sim_integrate_second (same arguments as the two,
 * partial rows = tiles): with atom i's force in registers the pass also does
 * the second half of the step for it (O + B, the arithmetic of
 * sim_integrate_second, bit for bit) and writes the kinetic columns
 * SIM_T_KE..SIM_T_KE+6 of partial, so the step needs no separate
 * second-half kernel. Replaces core/integrator/integrator.py:102-124 after
 * the EAM force pass of core/force/eam_force_cu.py:555-570. */
/* next_pos4 (optional, one element type; no counterpart in the reference):
 * for the replicas with neither a barostat nor loading, the pass then also
 * does the next step's first half (sim_integrate_first_q's arithmetic, bit
 * for bit): vel_half, next_pos4 and the records next_posq (the posq of this
 * call) of the new positions, and the displacement maxima since the build
 * (pos_last) into maxd2; that step then calls sim_integrate_first_q with
 * unfused_only = 1. */
int sim_eam_force_second(int R, const void* reps, const int* tiles, int ntiles, const int* posq,
                         const int* rec, const int* nbr, const int* nnbr, int n_stride,
                         sim_eam_params_t eam, float* force, double* partial, int virial,
                         sim_types_t types, float* vel_half, float* vel, const float* noise,
                         const long long* key, unsigned long long seed,
                         const unsigned long long* step_dev, const sim_stage_t* stage,
                         double* next_pos4, const double* pos_last, double* maxd2,
                         int* next_posq, void* stream);

/* sim_eam_density and sim_eam_force with each group's member records staged
 * in LDS (stage: sim_nl_stage of the current lists; sim_eam_force_second
 * takes the same stage, or NULL). Every result equals the unstaged entry
 * point's bit for bit: the same entries in the same order and masks, read
 * from LDS instead of global memory. Used by the persistent one-type
 * kernels (>= 256 tiles, table within the LDS limit), per group where its
 * members fit in the LDS beside the table; the rest read the list. Replace
 * the passes of core/force/eam_force_cu.py:461-579 like the unstaged ones. */
int sim_eam_density_staged(int R, const void* reps, const int* tiles, int ntiles, const int* posq,
                           const int* nbr, const int* nnbr, int n_stride, sim_eam_params_t eam,
                           float* rho, float* fp, int* rec, double* partial,
                           const sim_stage_t* stage, void* stream);
int sim_eam_force_staged(int R, const void* reps, const int* tiles, int ntiles, const int* posq,
                         const int* rec, const int* nbr, const int* nnbr, int n_stride,
                         sim_eam_params_t eam, float* force, double* partial, int virial,
                         const sim_stage_t* stage, void* stream);

/* Allocation-free EAM evaluation: no counterpart in the reference. For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): sim_eam_density +
 * sim_eam_force + sim_reduce of their columns [SIM_T_EPOT, SIM_T_KE) in one
 * call, with the records, per-tile partials and reduction chunks in a
 * fixed work array inside the library (one per GPU, held until the process
 * ends) instead of caller buffers. The caller passes inputs and outputs:
 * force[N][3] and thermo[R][16] (columns 0..7 written: energy, virial, W).
 * rep_tiles: int[R+1] first tile of each replica (tiles grouped by replica).
 * Limits: n_stride (atoms) <= SIM_NOALLOC_MAX_ATOMS, ntiles <=
 * SIM_NOALLOC_MAX_TILES, R <= SIM_NOALLOC_MAX_REPLICAS, else SIM_E_SIZE; the
 * work array is 36 MiB. Every call on one GPU uses that GPU's array: calls
 * that overlap on one GPU (concurrent streams, or threads) race. Results
 * equal those of the three entry points with caller buffers, bit for bit. */
#define SIM_NOALLOC_MAX_ATOMS (1 << 21)
#define SIM_NOALLOC_MAX_TILES 16384
#define SIM_NOALLOC_MAX_REPLICAS 256
int sim_eam_energy_force_noalloc(int R, const void* reps, const int* tiles, int ntiles,
                                 const int* rep_tiles, const int* posq, const int* nbr,
                                 const int* nnbr, int n_stride, sim_eam_params_t eam,
                                 float* force, double* thermo, int virial, void* stream);

/* Cell variants (one workgroup per cell slot, ncells = total cell
 * capacity): the 27-cell neighbourhood of the cell is staged in LDS column by
 * column (at most cap atoms, cap <= 2048; a larger one is read from global
 * memory instead and sets overflow bit 2) and the lists are read as the
 * 16-bit column slots of nbr16. Partials are written per cell slot:
 * partial[ncells][16]. The EAM variants take one element type. Suited to
 * small systems; the row variants below to large ones. */
int sim_lj_force_cell(int ncells, const void* reps, const int* cell_rep, const int* cell_start,
                      const int* cell_count, const double* pos4, const unsigned short* nbr16,
                      const int* nnbr, int max_nbr, int cap, sim_lj_params_t lj, float* force,
                      double* partial, int virial, int* overflow, void* stream);
int sim_eam_density_cell(int ncells, const void* reps, const int* cell_rep,
                         const int* cell_start, const int* cell_count, const double* pos4,
                         const unsigned short* nbr16, const int* nnbr, int max_nbr, int cap,
                         sim_eam_params_t eam, float* rho, float* fp, double* partial,
                         int* overflow, void* stream);
int sim_eam_force_cell(int ncells, const void* reps, const int* cell_rep, const int* cell_start,
                       const int* cell_count, const double* pos4, const unsigned short* nbr16,
                       const int* nnbr, int max_nbr, int cap, sim_eam_params_t eam,
                       const float* fp, float* force, double* partial, int virial, int* overflow,
                       void* stream);

/* Row variants: one workgroup per row of cells along x (ncells = total
 * cell capacity; row_base[R+1] from sim_nl_build with a reorder). A ring of
 * four stencil columns (9 cells each) is staged in LDS, at most cap atoms per
 * column (cap <= 512; a larger column is read from global memory instead and
 * sets overflow bit 2), and the lists are read as the 16-bit column slots of
 * nbr16. Partials are written per cell slot: partial[ncells][16] (zeros for
 * slots outside a replica's current grid). The EAM variants take one element
 * type; ntypes selects the LJ type handling. */
int sim_lj_force_rows(int ncells, const void* reps, const int* cell_rep, const int* row_base,
                      int R, const int* cell_start, const int* cell_count, const double* pos4,
                      const unsigned short* nbr16, const int* nnbr, int max_nbr, int cap,
                      int ntypes, sim_lj_params_t lj, float* force, double* partial, int virial,
                      int* overflow, void* stream);
int sim_eam_density_rows(int ncells, const void* reps, const int* cell_rep, const int* row_base,
                         int R, const int* cell_start, const int* cell_count, const double* pos4,
                         const unsigned short* nbr16, const int* nnbr, int max_nbr, int cap,
                         sim_eam_params_t eam, float* rho, float* fp, double* partial,
                         int* overflow, void* stream);
int sim_eam_force_rows(int ncells, const void* reps, const int* cell_rep, const int* row_base,
                       int R, const int* cell_start, const int* cell_count, const double* pos4,
                       const unsigned short* nbr16, const int* nnbr, int max_nbr, int cap,
                       sim_eam_params_t eam, const float* fp, float* force, double* partial,
                       int virial, int* overflow, void* stream);

/* ---- integrator --------------------------------------------------------- */

/* First half (B + A): vh = v + 1/2 F/m dt (float32), x += vh dt (float64),
 * per-replica max |x - pos_last|^2 into maxd2 (double bits, atomic max). */
int sim_integrate_first(const void* reps, const int* tiles, int ntiles,
                        sim_types_t types, double* pos4, const float* vel,
                        const float* force, float* vel_half, const double* pos_last,
                        double* maxd2, void* stream);

/* sim_integrate_first that also writes the position records of the new
 * positions (as sim_pos_quantize) into posq [N][4]: a step whose list build
 * passes the same posq in sim_nl_reorder_t needs no sim_pos_quantize.
 * flags: SIM_FIRST_UNFUSED_ONLY: only the replicas with a barostat or
 * loading (the others' first half was done by the previous step's
 * sim_eam_force_second); SIM_FIRST_BERENDSEN_SCALE: for the Berendsen
 * replicas, first the scaling of x, pos_last and vel by mu that the previous
 * step's sim_berendsen_box left undone (sim_berendsen's arithmetic; vel and
 * pos_last are then written). */
#define SIM_FIRST_UNFUSED_ONLY 1
#define SIM_FIRST_BERENDSEN_SCALE 2
int sim_integrate_first_q(const void* reps, const int* tiles, int ntiles,
                          sim_types_t types, double* pos4, float* vel,
                          const float* force, float* vel_half, double* pos_last,
                          double* maxd2, int* posq, int flags, void* stream);

/* Second half (O + B): v = vh damp + xi rff (Langevin replicas), then
 * v += 1/2 F/m dt; KE and kinetic-tensor columns of partial.
 * noise: optional float[N][3] of N(0,1) draws (NULL = Philox4x32-10 keyed by
 * seed, counter = (atom id, step)). key: optional int64[N] caller index of the
 * atom in each slot; the atom id is then key - atom_base of its replica (the
 * same draws for an atom whatever its storage slot, and in a packed engine as
 * in a single-replica one); NULL: the slot index. */
int sim_integrate_second(const void* reps, const int* tiles, int ntiles,
                         sim_types_t types, const double* pos4, const float* vel_half,
                         const float* force, float* vel, const float* noise, const long long* key,
                         unsigned long long seed, unsigned long long step,
                         const unsigned long long* step_dev, double* partial, void* stream);

/* step_dev[0] += 1 (device step counter read by sim_integrate_second, so that
 * a captured step graph draws fresh Langevin noise on every replay). */
int sim_step_advance(unsigned long long* step_dev, void* stream);

/* Per-replica compensated FP64 sums of partial columns [col0, col1) over the
 * replica's rows of `partial` (rep_rows: int[R+1], tiles or cell slots) into
 * thermo[R][16]; chunks: double[R][64][16] work array. */
int sim_reduce(int R, const int* rep_rows, const double* partial, int col0, int col1,
               double* chunks, double* thermo, void* stream);

/* sim_reduce that also advances the device step counter (as sim_step_advance)
 * once the sums are taken: the last call of a step, after the second half has
 * read the counter, saves the step a launch of its own. */
int sim_reduce_step(int R, const int* rep_rows, const double* partial, int col0, int col1,
                    double* chunks, double* thermo, unsigned long long* step_dev, void* stream);

/* The box half of sim_berendsen (pressure, mu, H, volume, the trigger's box
 * bound) without scaling the atoms: a caller that runs several steps back to
 * back can leave that scaling to the next step's sim_integrate_first_q with
 * SIM_FIRST_BERENDSEN_SCALE (the same bits); the last step of a run calls
 * sim_berendsen. core/barostat.py:36-68. */
int sim_berendsen_box(int R, void* reps, const double* thermo, void* stream);

/* Berendsen (barostat.py:36-68) for replicas with barostat = 1: pressure from
 * thermo K + W and the volume, mu, box update in reps; then x, v and
 * pos_last scaled by mu. */
int sim_berendsen(int R, void* reps, const int* tiles, int ntiles, const double* thermo,
                  double* pos4, float* vel, double* pos_last, void* stream);

/* Anisotropic barostat (barostat.py:131-208) for replicas with barostat = 2:
 * the pressure tensor (K + W)/V projected on the unit lattice rows, the
 * deadband/friction/clamp update of eta_dot (with a Philox N(0,1) term keyed
 * by (seed, step_dev, replica) when stochastic), lattice rows scaled by
 * exp(dt eta_dot), x and pos_last mapped affinely with the box, v scaled by
 * exp(-dt eta_dot) per Cartesian axis; p_axis holds the projected pressure. */
int sim_aniso_barostat(int R, void* reps, const int* tiles, int ntiles, const double* thermo,
                       double* pos4, float* vel, double* pos_last, unsigned long long seed,
                       const unsigned long long* step_dev, void* stream);

/* Uniaxial loading (loading.py:39-78) for replicas with loading = 1: lattice
 * rows scaled by deform[3], x and pos_last mapped affinely with the box
 * (fractional coordinates unchanged); velocities are not scaled. */
int sim_box_deform(int R, void* reps, const int* tiles, int ntiles, double* pos4,
                   double* pos_last, void* stream);

/* Copies positions into a float32 [N,3] (dtype 0) or float64 [N,3] (dtype 1)
 * array, and back (sim_pos_import). */
int sim_pos_export(const double* pos4, int n, void* xyz, int dtype, void* stream);
int sim_pos_import(double* pos4, int n, const void* xyz, int dtype, void* stream);

/* ---- learned radial-descriptor MLP potential ----------------------------
 *
 * The reference's learned potential is an external model behind the return
 * contract of machine_learning_potentials/machine_learning_force.py:85-142
 * ({'energy', 'forces', 'virial'}); these entry points implement this
 * library's own network under the same contract:
 *   g_ik = sum_{j, r_ij < rc} exp(-(r_ij - mu_k)^2 / (2 w^2)) fc(r_ij),
 *          mu_k = mu0 + k dmu (k < 32), fc(r) = (cos(pi r / rc) + 1) / 2
 *   E_i  = w3 . silu(W2 silu(W1 (g_i - mean) inv_std + b1) + b2) + b3
 * W1 [64][32] and W2 [64][64] are bfloat16 (MFMA operands, float32
 * accumulation); biases, w3, mean and inv_std are float32.
 * Partial rows: SIM_MLP_ROWS_PER_TILE rows per tile (partial[ntiles*4][16]);
 * sim_reduce over rep_rows = 4 * tile offsets gives the replica totals. */
#define SIM_MLP_NDESC 32
#define SIM_MLP_NH 64
#define SIM_MLP_ROWS_PER_TILE 4
/* flat parameter / gradient vector: W1 (2048) b1 (64) W2 (4096) b2 (64) w3 (64) b3 (1) */
#define SIM_MLP_NPARAM 6337
#define SIM_MLP_NGRAD (SIM_MLP_NPARAM + 1) /* gradient, then the loss value */

typedef struct sim_mlp_net {
    void* w1;             /* bf16 [64][32] row-major (hidden unit, descriptor) */
    void* w2;             /* bf16 [64][64] */
    float* b1;            /* [64] */
    float* b2;            /* [64] */
    float* w3;            /* [64] */
    float* b3;            /* [1] */
    float* mean;          /* [32] descriptor shift */
    float* inv_std;       /* [32] descriptor scale */
    float rc;             /* cutoff, Angstrom */
    float mu0;            /* first Gaussian centre */
    float dmu;            /* centre spacing */
    float inv_w2;         /* 1 / w^2 */
} sim_mlp_net_t;

/* Descriptors and the network for every atom: per-atom dE_i/dg_i (c, float
 * [N][32], already divided by the descriptor scale), optional per-atom
 * energies e_atom [N], optional normalised descriptors gh [N][32], and the
 * energy column of partial. posq (all MLP entry points): the fractional
 * fixed-point records of sim_pos_quantize. */
int sim_mlp_energy(int R, const void* reps, const int* tiles, int ntiles, const int* posq,
                   const int* nbr, const int* nnbr, int n_stride, sim_mlp_net_t net, float* c,
                   float* gh, float* e_atom, double* partial, void* stream);

/* Analytic forces F_i = -sum_j (c_i + c_j) . dphi(r_ij)/dr r_ij/|r_ij| and the
 * virial columns of partial (same convention as sim_eam_force). */
int sim_mlp_force(int R, const void* reps, const int* tiles, int ntiles, const int* posq,
                  const int* nbr, const int* nnbr, int n_stride, sim_mlp_net_t net, const float* c,
                  float* force, double* partial, int virial, void* stream);

/* Force-matching adjoint: with R_i = 2 wF (F_i - Fref_i) / (3 N_replica),
 * v_i = inv_std * dL/dc_i = -inv_std * sum_j dphi(r_ij)/dr ((R_i - R_j) . r_ij/|r_ij|). */
int sim_mlp_adjoint(int R, const void* reps, const int* tiles, int ntiles, const int* posq,
                    const int* nbr, const int* nnbr, int n_stride, sim_mlp_net_t net,
                    const float* force, const float* force_ref, float wF, float* v, void* stream);

/* Weight gradient of the loss
 *   L = sum_r [ wE (E_r - Eref_r)^2 / N_r^2 + wF sum_{i in r} |F_i - Fref_i|^2 / (3 N_r) ]
 * from gh (sim_mlp_energy), v (sim_mlp_adjoint), the replica energies
 * thermo[r][0] (sim_reduce of the energy column) and the forces: a
 * forward-over-reverse pass through the network per atom, BF16-MFMA outer
 * products over 16-atom groups, per-workgroup partials (grad_partial, float
 * [max_blocks][SIM_MLP_NGRAD]) and a fixed-order float64 sum into grad
 * [SIM_MLP_NGRAD] (the last entry is L).  atom_rep: int[n] replica of each atom. */
int sim_mlp_grad(int n, const int* atom_rep, const void* reps, const double* thermo,
                 const double* e_ref, float wE, float wF, sim_mlp_net_t net, const float* gh,
                 const float* v, const float* force, const float* force_ref, float* grad_partial,
                 int max_blocks, float* grad, void* stream);

/* Adam step on the float32 master parameters (flat layout above), then the
 * network arrays of net are rewritten from them (W1, W2 rounded to bfloat16,
 * nearest-even).  t: 1-based step number. */
int sim_mlp_adam(float* master, const float* grad, float* m, float* v, int t, float lr,
                 float beta1, float beta2, float eps, sim_mlp_net_t net, void* stream);

#ifdef __cplusplus
}
#endif
#endif
