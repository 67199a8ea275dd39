/*
 * Oracle (test infrastructure only): float64 cell-list restatement of the
 * reference's half-list Lennard-Jones and two-pass EAM for large systems.
 *
 * The numpy oracle (oracle/neighbor.py, lj.py, eam.py) is exact but takes
 * minutes at a million atoms; this file computes the same quantities in C so
 * that the full-size configurations can be checked atom by atom.
 *
 * Restates, in float64 with the numpy oracle's operation order:
 *   core/box.py:109-130                 minimum image (orthogonal: r - L rint(r/L);
 *                                       triclinic: s = r H^-1, s -= rint(s), r = s H,
 *                                       each product summed (a0 + a1) + a2)
 *   core/neighbor_search/gpu_kdtree.py:130-138, io_utils/reader.py:477-484
 *                                       pair (i < j) kept when (x^2 + y^2) + z^2 < r_l^2
 *   core/force/lennard_jones_force.py:32-94  raw / force-shift / switch LJ
 *   core/force/eam_force_cu.py:437-579  clamped table lerp, rho, F(rho), F'(rho),
 *                                       s = F'_i f'_{ti,tj} + F'_j f'_{tj,ti} + phi',
 *                                       f = -s rij/(r + 1e-8)
 *
 * Candidates come from a cell grid whose perpendicular widths are at least
 * r_l (27 stencil cells, fewer when a dimension has 1 or 2 cells); every
 * candidate is accepted or rejected with the brute-force criterion, so the
 * pair set is the brute force's. Each atom's row (all j != i in the list
 * radius) is summed by one thread (OpenMP over atoms): the force on i in
 * full, half of each pair's energy and virial. Totals therefore agree with
 * the numpy oracle's half-list sums to float64 rounding, not bit for bit.
 *
 * pinned by: equality with the numpy oracle (pair counts exactly; forces,
 * energy, virial and density to 1e-12 relative) on small orthogonal and
 * triclinic systems, tests/test_oracle_cell.py; the numpy oracle is pinned by
 * the reference's golden vectors and the closed forms listed in its modules.
 *
 * Build: gcc -O2 -fopenmp -ffp-contract=off -fPIC -shared oracle/cell_md.c -lm
 * (-ffp-contract=off keeps every product and sum a separate rounding, as numpy
 * computes them).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    double H[9], Hi[9];
    int ortho;
    int nc[3];
    int noff[3];
    int off[3][3];
    int ncell;
    int* cell_start;   /* [ncell + 1] */
    int* cell_atoms;   /* [n] atoms sorted by cell */
    int* cell_of;      /* [n] */
    int* cc;           /* [n][3] cell coordinates */
} grid_t;

static void min_image(const grid_t* g, double* r) {
    if (g->ortho) {
        for (int k = 0; k < 3; ++k) {
            const double L = g->H[4 * k];
            r[k] = r[k] - L * rint(r[k] / L);
        }
        return;
    }
    double s[3];
    for (int k = 0; k < 3; ++k) {
        s[k] = (r[0] * g->Hi[k] + r[1] * g->Hi[3 + k]) + r[2] * g->Hi[6 + k];
        s[k] = s[k] - rint(s[k]);
    }
    for (int m = 0; m < 3; ++m) r[m] = (s[0] * g->H[m] + s[1] * g->H[3 + m]) + s[2] * g->H[6 + m];
}

static double perp_width(const double* H, int k) {
    const double* a = H + 3 * ((k + 1) % 3);
    const double* b = H + 3 * ((k + 2) % 3);
    const double c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                         a[0] * b[1] - a[1] * b[0]};
    const double det = H[0] * (H[4] * H[8] - H[5] * H[7]) - H[1] * (H[3] * H[8] - H[5] * H[6]) +
                       H[2] * (H[3] * H[7] - H[4] * H[6]);
    return fabs(det) / sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
}

static int grid_build(grid_t* g, int n, const double* pos, const double* H, const double* Hi,
                      int ortho, double rl) {
    memcpy(g->H, H, sizeof(g->H));
    memcpy(g->Hi, Hi, sizeof(g->Hi));
    g->ortho = ortho;
    for (int k = 0; k < 3; ++k) {
        int c = (int)floor(perp_width(H, k) / (rl * (1.0 + 1e-6)));
        g->nc[k] = c < 1 ? 1 : c;
        if (g->nc[k] == 1) {
            g->noff[k] = 1;
            g->off[k][0] = 0;
        } else if (g->nc[k] == 2) {
            g->noff[k] = 2;
            g->off[k][0] = 0;
            g->off[k][1] = 1;
        } else {
            g->noff[k] = 3;
            g->off[k][0] = -1;
            g->off[k][1] = 0;
            g->off[k][2] = 1;
        }
    }
    g->ncell = g->nc[0] * g->nc[1] * g->nc[2];
    g->cell_start = calloc((size_t)g->ncell + 1, sizeof(int));
    g->cell_atoms = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    g->cell_of = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    g->cc = malloc(sizeof(int) * 3 * (size_t)(n > 0 ? n : 1));
    if (!g->cell_start || !g->cell_atoms || !g->cell_of || !g->cc) return -1;
    for (int i = 0; i < n; ++i) {
        const double* x = pos + 3 * (size_t)i;
        for (int k = 0; k < 3; ++k) {
            double s = x[0] * Hi[k] + x[1] * Hi[3 + k] + x[2] * Hi[6 + k];
            s -= floor(s);
            int c = (int)(s * g->nc[k]);
            if (c > g->nc[k] - 1) c = g->nc[k] - 1;
            if (c < 0) c = 0;
            g->cc[3 * i + k] = c;
        }
        g->cell_of[i] = (g->cc[3 * i + 2] * g->nc[1] + g->cc[3 * i + 1]) * g->nc[0] + g->cc[3 * i];
        g->cell_start[g->cell_of[i] + 1]++;
    }
    for (int c = 0; c < g->ncell; ++c) g->cell_start[c + 1] += g->cell_start[c];
    int* fill = malloc(sizeof(int) * (size_t)g->ncell);
    if (!fill) return -1;
    memcpy(fill, g->cell_start, sizeof(int) * (size_t)g->ncell);
    for (int i = 0; i < n; ++i) g->cell_atoms[fill[g->cell_of[i]]++] = i;   /* ascending i per cell */
    free(fill);
    return 0;
}

static void grid_free(grid_t* g) {
    free(g->cell_start);
    free(g->cell_atoms);
    free(g->cell_of);
    free(g->cc);
}

/* Row of atom i: every j != i with |mi(x_i - x_j)|^2 < rl2, in cell-visit
 * order; fn(i, j, rij, d2, ctx) with rij = mi(x_i - x_j). The minimum image is
 * odd (mi(-r) = -mi(r) exactly), so each pair of the half list is seen from
 * both ends with opposite rij. */
typedef void (*pair_fn)(int i, int j, const double* rij, double d2, void* ctx);

static long long visit_row(const grid_t* g, const double* pos, double rl2, int i, pair_fn fn,
                           void* ctx) {
    long long np = 0;
    const int* ci = g->cc + 3 * i;
    for (int a = 0; a < g->noff[2]; ++a)
        for (int b = 0; b < g->noff[1]; ++b)
            for (int c = 0; c < g->noff[0]; ++c) {
                int z = (ci[2] + g->off[2][a] + g->nc[2]) % g->nc[2];
                int y = (ci[1] + g->off[1][b] + g->nc[1]) % g->nc[1];
                int x = (ci[0] + g->off[0][c] + g->nc[0]) % g->nc[0];
                const int cell = (z * g->nc[1] + y) * g->nc[0] + x;
                for (int k = g->cell_start[cell]; k < g->cell_start[cell + 1]; ++k) {
                    const int j = g->cell_atoms[k];
                    if (j == i) continue;
                    double r[3] = {pos[3 * (size_t)i] - pos[3 * (size_t)j],
                                   pos[3 * (size_t)i + 1] - pos[3 * (size_t)j + 1],
                                   pos[3 * (size_t)i + 2] - pos[3 * (size_t)j + 2]};
                    min_image(g, r);
                    const double d2 = (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2];
                    if (d2 < rl2) {
                        ++np;
                        fn(i, j, r, d2, ctx);
                    }
                }
            }
    return np;
}

/* Per-thread sums of a row visitor. */
typedef struct {
    double e, vir, W[9];
    long long ncut;
} acc_t;

/* ---- Lennard-Jones (core/force/lennard_jones_force.py:32-94) ------------- */

typedef struct {
    const int* types;
    int nt;
    const double* eps;
    const double* sig;
    double rc, rs;
    int mode;           /* 0 raw, 1 shift, 2 switch */
    double* forces;
    acc_t* acc;         /* this thread's sums */
    int* tie;           /* optional [n]: pairs of i with |r - rc| < tie_tol */
    double tie_tol;
} lj_ctx;

/* Half of the pair's energy and virial (each pair is visited from both
 * ends); the whole force on i. */
static void lj_pair(int i, int j, const double* rij, double d2, void* vctx) {
    lj_ctx* c = vctx;
    const double r2 = d2 > 1e-24 ? d2 : 1e-24;
    const double r = sqrt(r2);
    /* a pair this close to the cutoff may fall on the other side of it in
     * float32 arithmetic; raw LJ jumps by Fmag(rc) there */
    if (c->tie && fabs(r - c->rc) < c->tie_tol) c->tie[i]++;
    if (!(r < c->rc)) return;
    const int ti = c->types ? c->types[i] : 0, tj = c->types ? c->types[j] : 0;
    const double eps = c->eps[ti * c->nt + tj], sig = c->sig[ti * c->nt + tj];
    const double inv_r = 1.0 / r;
    const double sr = sig * inv_r;
    const double sr6 = pow(sr, 6.0);
    const double sr12 = sr6 * sr6;
    double U = (4.0 * eps) * (sr12 - sr6);
    double F = ((24.0 * eps) * inv_r) * (2.0 * sr12 - sr6);
    if (c->mode == 1) {
        const double src = sig / c->rc;
        const double s6 = pow(src, 6.0), s12 = s6 * s6;
        const double Urc = (4.0 * eps) * (s12 - s6);
        const double Frc = ((24.0 * eps) * (1.0 / c->rc)) * (2.0 * s12 - s6);
        U = (U - Urc) + (r - c->rc) * Frc;
        F = F - Frc;
    } else if (c->mode == 2) {
        double S = 1.0, dS = 0.0;
        if (r > c->rs) {
            double x = (r - c->rs) / (c->rc - c->rs);
            x = x < 0.0 ? 0.0 : (x > 1.0 ? 1.0 : x);
            const double x2 = x * x, x3 = pow(x, 3.0), x4 = pow(x, 4.0), x5 = pow(x, 5.0);
            S = ((1.0 - 10.0 * x3) + 15.0 * x4) - 6.0 * x5;
            dS = ((-30.0 * x2 + 60.0 * x3) - 30.0 * x4) / (c->rc - c->rs);
        }
        const double Ueff = U * S;
        F = F * S - U * dS;
        U = Ueff;
    }
    acc_t* a = c->acc;
    a->ncut++;
    a->e += 0.5 * U;
    a->vir += 0.5 * (F * r);
    double f[3];
    for (int k = 0; k < 3; ++k) f[k] = F * (rij[k] * (1.0 / r));
    for (int k = 0; k < 3; ++k) c->forces[3 * (size_t)i + k] += f[k];
    for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) a->W[3 * p + q] += 0.5 * (f[p] * rij[q]);
}

static void acc_add(acc_t* dst, const acc_t* src) {
    dst->e += src->e;
    dst->vir += src->vir;
    for (int k = 0; k < 9; ++k) dst->W[k] += src->W[k];
    dst->ncut += src->ncut;
}

/* out: [0] energy, [1] virial, [2..10] virial tensor (row-major), [11] pairs
 * within r_l, [12] pairs within rc. tie (optional, int[n]): per atom, the
 * pairs with |r - rc| < tie_tol. Returns 0, or -1 when memory runs out. */
int orc_lj(int n, const double* pos, const int* types, int nt, const double* eps,
           const double* sig, double rc, int mode, double switch_ratio, const double* H,
           const double* Hinv, int ortho, double rl2, double* forces, double* out, int* tie,
           double tie_tol) {
    grid_t g;
    memset(&g, 0, sizeof(g));
    if (grid_build(&g, n, pos, H, Hinv, ortho, sqrt(rl2)) != 0) {
        grid_free(&g);
        return -1;
    }
    memset(forces, 0, sizeof(double) * 3 * (size_t)n);
    if (tie) memset(tie, 0, sizeof(int) * (size_t)n);
    acc_t tot;
    memset(&tot, 0, sizeof(tot));
    long long np = 0;
#pragma omp parallel reduction(+ : np)
    {
        acc_t a;
        memset(&a, 0, sizeof(a));
        lj_ctx c = {types, nt, eps, sig, rc, switch_ratio * rc, mode, forces, &a, tie, tie_tol};
#pragma omp for schedule(static)
        for (int i = 0; i < n; ++i) np += visit_row(&g, pos, rl2, i, lj_pair, &c);
#pragma omp critical
        acc_add(&tot, &a);
    }
    out[0] = tot.e;
    out[1] = tot.vir;
    for (int k = 0; k < 9; ++k) out[2 + k] = tot.W[k];
    out[11] = (double)(np / 2);
    out[12] = (double)(tot.ncut / 2);
    grid_free(&g);
    return 0;
}

/* ---- EAM (core/force/eam_force_cu.py:437-579) ----------------------------- */

typedef struct {
    const int* types;
    int ne, n_r;
    double cutoff, r_max, inv_dr;
    const double *dens, *ddens, *phi, *dphi;   /* [ne][ne][n_r] */
    const double *F, *dF;                      /* [ne][n_rho] */
    const double *rho_lo, *drho, *inv_drho;    /* [ne] */
    int n_rho;
    double* rho;
    double* dFi;
    double* forces;
    acc_t* acc;
} eam_ctx;

static double interp_r(const eam_ctx* c, double r, const double* tab, int a, int b) {
    const double hi = c->r_max * (1 - 1e-7);
    const double rcl = r < 0.0 ? 0.0 : (r > hi ? hi : r);
    const double idx_f = rcl * c->inv_dr;
    const long long idx = (long long)idx_f;
    const double frac = idx_f - (double)idx;
    const long long nxt = idx + 1 < c->n_r - 1 ? idx + 1 : c->n_r - 1;
    const double* t = tab + ((size_t)a * c->ne + b) * c->n_r;
    return t[idx] + frac * (t[nxt] - t[idx]);
}

static double interp_rho(const eam_ctx* c, double rho, int a, const double* tab) {
    const double lo = c->rho_lo[a], d = c->drho[a], inv = c->inv_drho[a];
    const double hi = lo + d * (c->n_rho - 1) * (1 - 1e-7);
    const double rcl = rho < lo ? lo : (rho > hi ? hi : rho);
    const double idx_f = (rcl - lo) * inv;
    const long long idx = (long long)idx_f;
    const double frac = idx_f - (double)idx;
    const long long nxt = idx + 1 < c->n_rho - 1 ? idx + 1 : c->n_rho - 1;
    const double* t = tab + (size_t)a * c->n_rho;
    return t[idx] + frac * (t[nxt] - t[idx]);
}

/* rho_i = sum_j f_{ti,tj}(r_ij) over the row of i */
static void eam_rho_pair(int i, int j, const double* rij, double d2, void* vctx) {
    eam_ctx* c = vctx;
    const double r = sqrt(d2);
    if (!(r < c->cutoff)) return;
    const int ti = c->types ? c->types[i] : 0, tj = c->types ? c->types[j] : 0;
    c->rho[i] += interp_r(c, r, c->dens, ti, tj);
    (void)rij;
}

/* Force on i of the pair (i, j): f = -s rij/(r + 1e-8) with
 * s = F'_i f'_{tj,ti} + F'_j f'_{ti,tj} + phi'_{ti,tj} (the half list's
 * s of the edge (i, j), symmetric in i and j); half the pair's energy and
 * virial. */
static void eam_force_pair(int i, int j, const double* rij, double d2, void* vctx) {
    eam_ctx* c = vctx;
    const double r = sqrt(d2);
    if (!(r < c->cutoff)) return;
    const int ti = c->types ? c->types[i] : 0, tj = c->types ? c->types[j] : 0;
    acc_t* a = c->acc;
    a->ncut++;
    a->e += 0.5 * interp_r(c, r, c->phi, ti, tj);
    const double dd_i = interp_r(c, r, c->ddens, tj, ti);   /* f'_{tj,ti}: density i feels */
    const double dd_j = interp_r(c, r, c->ddens, ti, tj);
    const double dphi = interp_r(c, r, c->dphi, ti, tj);
    const double s = (c->dFi[i] * dd_j + c->dFi[j] * dd_i) + dphi;
    double f[3];
    for (int k = 0; k < 3; ++k) f[k] = -s * (rij[k] / (r + 1e-8));
    for (int k = 0; k < 3; ++k) c->forces[3 * (size_t)i + k] += f[k];
    a->vir += 0.5 * (s * r);
    for (int p = 0; p < 3; ++p)
        for (int q = 0; q < 3; ++q) a->W[3 * p + q] += 0.5 * (f[p] * rij[q]);
}

/* out: [0] energy, [1] virial, [2..10] virial tensor, [11] pairs within r_l,
 * [12] pairs within the cutoff, [13] embedding energy, [14] pair energy.
 * rho, dF: per-atom outputs [n]. */
int orc_eam(int n, const double* pos, const int* types, int ne, int n_r, double cutoff,
            double r_max, double inv_dr, const double* dens, const double* ddens,
            const double* phi, const double* dphi, int n_rho, const double* F, const double* dF,
            const double* rho_lo, const double* drho, const double* inv_drho, const double* H,
            const double* Hinv, int ortho, double rl2, double* forces, double* rho,
            double* dFi, double* out) {
    grid_t g;
    memset(&g, 0, sizeof(g));
    if (grid_build(&g, n, pos, H, Hinv, ortho, sqrt(rl2)) != 0) {
        grid_free(&g);
        return -1;
    }
    memset(rho, 0, sizeof(double) * (size_t)n);
    memset(forces, 0, sizeof(double) * 3 * (size_t)n);
    acc_t tot;
    memset(&tot, 0, sizeof(tot));
    long long np = 0;
    double emb = 0.0;
#pragma omp parallel reduction(+ : np, emb)
    {
        acc_t a;
        memset(&a, 0, sizeof(a));
        eam_ctx c = {types, ne, n_r, cutoff, r_max, inv_dr, dens, ddens, phi, dphi, F, dF,
                     rho_lo, drho, inv_drho, n_rho, rho, dFi, forces, &a};
#pragma omp for schedule(static)
        for (int i = 0; i < n; ++i) np += visit_row(&g, pos, rl2, i, eam_rho_pair, &c);
#pragma omp for schedule(static)
        for (int i = 0; i < n; ++i) {
            const int t = types ? types[i] : 0;
            emb += interp_rho(&c, rho[i], t, F);
            dFi[i] = interp_rho(&c, rho[i], t, dF);
        }
#pragma omp for schedule(static)
        for (int i = 0; i < n; ++i) visit_row(&g, pos, rl2, i, eam_force_pair, &c);
#pragma omp critical
        acc_add(&tot, &a);
    }
    out[0] = emb + tot.e;
    out[1] = tot.vir;
    for (int k = 0; k < 9; ++k) out[2 + k] = tot.W[k];
    out[11] = (double)(np / 2);
    out[12] = (double)(tot.ncut / 2);
    out[13] = emb;
    out[14] = tot.e;
    grid_free(&g);
    return 0;
}

/* Full neighbor rows of the atoms idx[0..m): for each, the j != i with
 * |mi(x_i - x_j)|^2 < rl2, written as (row k, j) pairs in ascending j.
 * Returns the number of pairs (at most cap are written), -1 on error. */
long long orc_rows(int n, const double* pos, const double* H, const double* Hinv, int ortho,
                   double rl2, int m, const int* idx, long long cap, int* prow, int* pcol) {
    grid_t g;
    memset(&g, 0, sizeof(g));
    if (grid_build(&g, n, pos, H, Hinv, ortho, sqrt(rl2)) != 0) {
        grid_free(&g);
        return -1;
    }
    long long np = 0;
    int* buf = malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    if (!buf) {
        grid_free(&g);
        return -1;
    }
    for (int k = 0; k < m; ++k) {
        const int i = idx[k];
        const int* ci = g.cc + 3 * i;
        int cnt = 0;
        for (int a = 0; a < g.noff[2]; ++a)
            for (int b = 0; b < g.noff[1]; ++b)
                for (int c = 0; c < g.noff[0]; ++c) {
                    int z = (ci[2] + g.off[2][a] + g.nc[2]) % g.nc[2];
                    int y = (ci[1] + g.off[1][b] + g.nc[1]) % g.nc[1];
                    int x = (ci[0] + g.off[0][c] + g.nc[0]) % g.nc[0];
                    const int cell = (z * g.nc[1] + y) * g.nc[0] + x;
                    for (int q = g.cell_start[cell]; q < g.cell_start[cell + 1]; ++q) {
                        const int j = g.cell_atoms[q];
                        if (j == i) continue;
                        double r[3] = {pos[3 * (size_t)i] - pos[3 * (size_t)j],
                                       pos[3 * (size_t)i + 1] - pos[3 * (size_t)j + 1],
                                       pos[3 * (size_t)i + 2] - pos[3 * (size_t)j + 2]};
                        min_image(&g, r);
                        if ((r[0] * r[0] + r[1] * r[1]) + r[2] * r[2] < rl2) buf[cnt++] = j;
                    }
                }
        /* ascending j (insertion sort; rows are short) */
        for (int p = 1; p < cnt; ++p) {
            const int v = buf[p];
            int q = p - 1;
            while (q >= 0 && buf[q] > v) {
                buf[q + 1] = buf[q];
                --q;
            }
            buf[q + 1] = v;
        }
        for (int p = 0; p < cnt; ++p, ++np)
            if (np < cap) {
                prow[np] = k;
                pcol[np] = buf[p];
            }
    }
    free(buf);
    grid_free(&g);
    return np;
}
