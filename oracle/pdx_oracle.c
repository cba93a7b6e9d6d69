/*
 * CPU oracle for the force-density hot path of Peridynamics.jl.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C restatement of the reference
 * algorithm, used as the checker by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  Nothing in the product path links or calls it.
 *
 * Every function follows the reference source op for op (Julia evaluation order,
 * no FMA contraction: build with -ffp-contract=off) and cites the file:line it
 * restates.  Indices are 0-based; point vectors are column-major 3 x n (x, y, z of a
 * point contiguous), as Julia's Matrix{Float64}(3, n).
 *
 * Parity pins (see tests/test_oracle_*.py): bond lists of the reference unit tests,
 * the closed-form BB force, the OSB and CPD (NeoHookePenalty) regression vectors,
 * the force-density invariants (zero at reference, momentum, mirror symmetry,
 * pulled-corner damage), the stable time step closed form, brute force vs cell list;
 * the constitutive models by P = dPsi/dF and their closed forms, the ZEMWan C_1 by the
 * isotropic closed form and a brute-force Hooke contraction, its force by a bond sum.
 *
 * Parity unpinned (no reference test fixes them, and the reference cannot run here):
 *  - the order of bonds across grid cells of PointNeighbors' foreach_neighbor; this
 *    oracle lists the neighbours of a point in ascending index order, which is the
 *    reference order whenever all neighbours share one grid cell (pinned by
 *    test_bond_system.jl:105-157);
 *  - StaticArrays' exact op order of norm/inv/det beyond the ~1e-8 the reference
 *    tests check; this file uses left-to-right sums without FMA.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_OK 0
#define ORC_ERR_ARG -1
#define ORC_ERR_OOM -3
#define ORC_ERR_DUPLICATE -6

static int set_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
    return omp_get_max_threads();
#else
    (void)nthreads;
    return 1;
#endif
}

int orc_max_threads(void) { return set_threads(0); }

/* ------------------------------------------------------------------------------ */
/* Neighbour search: find_bonds / find_bonds! / check_point_duplicates            */
/* bond_system.jl:96-132 with PointNeighbors' GridNeighborhoodSearch (cell size =  */
/* maximum horizon, bond_system.jl:97-99).  Distance of a pair as PointNeighbors   */
/* forms it: pos_diff = x_i - x_j, d2 = dot(pos_diff, pos_diff), accepted if       */
/* d2 <= delta_i^2, length L = sqrt(d2).                                          */
/* ------------------------------------------------------------------------------ */

typedef struct {
    int64_t nx, ny, nz;
    double x0, y0, z0, cs;
    int64_t *cell_start; /* ncell + 1 */
    int64_t *cell_pts;   /* n */
} grid_t;

static int64_t cell_coord(double v, double v0, double cs, int64_t nmax)
{
    int64_t c = (int64_t)floor((v - v0) / cs);
    if (c < 0) c = 0;
    if (c >= nmax) c = nmax - 1;
    return c;
}

static int grid_build(grid_t *g, int64_t n, const double *X, double cs)
{
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) {
            double v = X[3 * i + d];
            if (v < mn[d]) mn[d] = v;
            if (v > mx[d]) mx[d] = v;
        }
    if (n == 0) { mn[0] = mn[1] = mn[2] = 0; mx[0] = mx[1] = mx[2] = 0; }
    if (!(cs > 0)) cs = 1.0;
    /* grow the cell if the grid would be huge compared to the point count */
    for (;;) {
        double ncell = 1.0;
        for (int d = 0; d < 3; d++) ncell *= floor((mx[d] - mn[d]) / cs) + 1.0;
        if (ncell <= 8.0 * (double)n + 4096.0) break;
        cs *= 2.0;
    }
    g->cs = cs;
    g->x0 = mn[0]; g->y0 = mn[1]; g->z0 = mn[2];
    g->nx = (int64_t)floor((mx[0] - mn[0]) / cs) + 1;
    g->ny = (int64_t)floor((mx[1] - mn[1]) / cs) + 1;
    g->nz = (int64_t)floor((mx[2] - mn[2]) / cs) + 1;
    int64_t ncell = g->nx * g->ny * g->nz;
    g->cell_start = (int64_t *)calloc((size_t)ncell + 1, sizeof(int64_t));
    g->cell_pts = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    int64_t *cid = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    if (!g->cell_start || !g->cell_pts || !cid) { free(cid); return ORC_ERR_OOM; }
    for (int64_t i = 0; i < n; i++) {
        int64_t cx = cell_coord(X[3 * i], g->x0, cs, g->nx);
        int64_t cy = cell_coord(X[3 * i + 1], g->y0, cs, g->ny);
        int64_t cz = cell_coord(X[3 * i + 2], g->z0, cs, g->nz);
        cid[i] = cx + g->nx * (cy + g->ny * cz);
        g->cell_start[cid[i] + 1]++;
    }
    for (int64_t c = 0; c < ncell; c++) g->cell_start[c + 1] += g->cell_start[c];
    int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)(ncell > 0 ? ncell : 1));
    if (!fill) { free(cid); return ORC_ERR_OOM; }
    memcpy(fill, g->cell_start, sizeof(int64_t) * (size_t)ncell);
    for (int64_t i = 0; i < n; i++) g->cell_pts[fill[cid[i]]++] = i;
    free(fill);
    free(cid);
    return ORC_OK;
}

static void grid_free(grid_t *g)
{
    free(g->cell_start);
    free(g->cell_pts);
}

static int cmp_i64(const void *a, const void *b)
{
    int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return (x > y) - (x < y);
}

static inline double pair_d2(const double *X, int64_t i, int64_t j)
{
    double p0 = X[3 * i] - X[3 * j];
    double p1 = X[3 * i + 1] - X[3 * j + 1];
    double p2 = X[3 * i + 2] - X[3 * j + 2];
    return (p0 * p0 + p1 * p1) + p2 * p2;
}

/* candidates of point i within delta_i, sorted ascending; returns count */
static int64_t point_neighbors(const grid_t *g, const double *X, const double *delta,
                               int64_t i, int64_t *buf, int64_t cap, int *dup, int64_t *dup_j)
{
    double di = delta[i];
    double r2 = di * di;
    int64_t cx = cell_coord(X[3 * i], g->x0, g->cs, g->nx);
    int64_t cy = cell_coord(X[3 * i + 1], g->y0, g->cs, g->ny);
    int64_t cz = cell_coord(X[3 * i + 2], g->z0, g->cs, g->nz);
    int64_t cnt = 0;
    for (int64_t z = cz - 1; z <= cz + 1; z++) {
        if (z < 0 || z >= g->nz) continue;
        for (int64_t y = cy - 1; y <= cy + 1; y++) {
            if (y < 0 || y >= g->ny) continue;
            for (int64_t x = cx - 1; x <= cx + 1; x++) {
                if (x < 0 || x >= g->nx) continue;
                int64_t c = x + g->nx * (y + g->ny * z);
                for (int64_t k = g->cell_start[c]; k < g->cell_start[c + 1]; k++) {
                    int64_t j = g->cell_pts[k];
                    if (j == i) continue;
                    double d2 = pair_d2(X, i, j);
                    if (d2 <= r2) {
                        if (sqrt(d2) < DBL_EPSILON && !*dup) { *dup = 1; *dup_j = j; }
                        if (cnt < cap) buf[cnt] = j;
                        cnt++;
                    }
                }
            }
        }
    }
    if (cnt <= cap) qsort(buf, (size_t)cnt, sizeof(int64_t), cmp_i64);
    return cnt;
}

/* Number of bonds of every point (counts[n]); returns total bonds, or a negative
 * status.  On a duplicate point (L < eps, bond_system.jl:125-132) returns
 * ORC_ERR_DUPLICATE and stores the pair in dup_pair[2]. */
int64_t orc_find_bonds_count(int64_t n, const double *X, const double *delta,
                             int64_t *counts, int64_t *dup_pair, int nthreads)
{
    if (n < 0 || !X || !delta || !counts) return ORC_ERR_ARG;
    set_threads(nthreads);
    double dmax = 0;
    for (int64_t i = 0; i < n; i++) if (delta[i] > dmax) dmax = delta[i];
    grid_t g;
    if (grid_build(&g, n, X, dmax) != ORC_OK) return ORC_ERR_OOM;
    int64_t total = 0;
    int any_dup = 0;
    int64_t di = -1, dj = -1;
#pragma omp parallel
    {
        int64_t cap = 4096;
        int64_t *buf = (int64_t *)malloc(sizeof(int64_t) * (size_t)cap);
#pragma omp for schedule(dynamic, 256) reduction(+ : total)
        for (int64_t i = 0; i < n; i++) {
            int dup = 0;
            int64_t j = -1;
            int64_t c = point_neighbors(&g, X, delta, i, buf, 0, &dup, &j);
            counts[i] = c;
            total += c;
            if (dup) {
#pragma omp critical
                {
                    if (!any_dup || i < di) { any_dup = 1; di = i; dj = j; }
                }
            }
        }
        free(buf);
    }
    grid_free(&g);
    if (any_dup) {
        if (dup_pair) { dup_pair[0] = di; dup_pair[1] = dj; }
        return ORC_ERR_DUPLICATE;
    }
    return total;
}

/* Fill the CSR bond list: row_ptr[n+1] from the counts, neighbours ascending,
 * L and fail_permit[i] & fail_permit[j] per bond (Bond, bond_system.jl:14-18,118). */
int orc_find_bonds_fill(int64_t n, const double *X, const double *delta,
                        const uint8_t *fail_permit, const int64_t *row_ptr, int32_t *nbr,
                        double *L, uint8_t *bond_fp, int nthreads)
{
    if (n < 0 || !X || !delta || !row_ptr || !nbr || !L || !bond_fp) return ORC_ERR_ARG;
    set_threads(nthreads);
    double dmax = 0;
    for (int64_t i = 0; i < n; i++) if (delta[i] > dmax) dmax = delta[i];
    grid_t g;
    if (grid_build(&g, n, X, dmax) != ORC_OK) return ORC_ERR_OOM;
    int bad = 0;
#pragma omp parallel
    {
        int64_t cap = 4096;
        int64_t *buf = (int64_t *)malloc(sizeof(int64_t) * (size_t)cap);
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < n; i++) {
            int dup = 0;
            int64_t j = -1;
            int64_t want = row_ptr[i + 1] - row_ptr[i];
            if (want > cap) {
                cap = want;
                buf = (int64_t *)realloc(buf, sizeof(int64_t) * (size_t)cap);
            }
            int64_t c = point_neighbors(&g, X, delta, i, buf, cap, &dup, &j);
            if (c != want) { bad = 1; continue; }
            for (int64_t k = 0; k < c; k++) {
                int64_t jj = buf[k];
                int64_t b = row_ptr[i] + k;
                nbr[b] = (int32_t)jj;
                L[b] = sqrt(pair_d2(X, i, jj));
                bond_fp[b] = (uint8_t)((fail_permit ? fail_permit[i] : 1) &
                                       (fail_permit ? fail_permit[jj] : 1));
            }
        }
        free(buf);
    }
    grid_free(&g);
    return bad ? ORC_ERR_ARG : ORC_OK;
}

/* O(n^2) brute-force search, an independent second oracle for small inputs. */
int64_t orc_find_bonds_brute(int64_t n, const double *X, const double *delta,
                             int64_t *counts, int64_t *row_ptr, int32_t *nbr, double *L)
{
    int64_t total = 0;
    for (int64_t i = 0; i < n; i++) {
        double r2 = delta[i] * delta[i];
        int64_t c = 0;
        for (int64_t j = 0; j < n; j++) {
            if (j == i) continue;
            double d2 = pair_d2(X, i, j);
            if (d2 <= r2) {
                if (row_ptr && nbr) {
                    int64_t b = row_ptr[i] + c;
                    nbr[b] = (int32_t)j;
                    if (L) L[b] = sqrt(d2);
                }
                c++;
            }
        }
        if (counts) counts[i] = c;
        total += c;
    }
    return total;
}

/* Bonds that join set A and set B (filter_bonds_by_crack! / break_bonds!,
 * bond_system.jl:142-165, 249-267): hit[b] = 1 for (i in A and j in B) or
 * (i in B and j in A). */
void orc_precrack_hits(int64_t n, const int64_t *row_ptr, const int32_t *nbr,
                       const uint8_t *in_a, const uint8_t *in_b, uint8_t *hit)
{
    for (int64_t i = 0; i < n; i++)
        for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
            int64_t j = nbr[b];
            int pa = in_a[i], pb = in_b[i], na = in_a[j], nb = in_b[j];
            hit[b] = (uint8_t)((pa && nb) || (pb && na));
        }
}

/* ------------------------------------------------------------------------------ */
/* Bond-based force density                                                        */
/* calc_force_density!(storage, ...)  bond_system.jl:285-297                       */
/* calc_failure! (BB, stores l)       bond_based.jl:125-144                        */
/* calc_damage!                       bond_system.jl:338-343                       */
/* force_density_point! (BB)          bond_based.jl:146-161, NoCorrection = 1      */
/* The per-point loop has no scatter, so any split of points over threads gives     */
/* the serial result bit for bit (the threads data handler, one chunk per thread). */
/* ------------------------------------------------------------------------------ */
void orc_bb_force(int64_t n, const double *x, const double *vol, const int64_t *row_ptr,
                  const int32_t *nbr, const double *L, const uint8_t *bond_fp,
                  uint8_t *active, double *bond_length, int64_t *n_active, double *damage,
                  double bc, double epsc, double *b_int, double *b_abs, int nthreads,
                  const double *scf)
{
    set_threads(nthreads);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        /* b_int .= 0, n_active_bonds .= 0 */
        b_int[3 * i] = 0.0; b_int[3 * i + 1] = 0.0; b_int[3 * i + 2] = 0.0;
        int64_t na = 0;
        /* calc_failure! */
        for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
            int64_t j = nbr[b];
            double dx = x[3 * j] - x[3 * i];
            double dy = x[3 * j + 1] - x[3 * i + 1];
            double dz = x[3 * j + 2] - x[3 * i + 2];
            double l = sqrt((dx * dx + dy * dy) + dz * dz);
            double eps = (l - L[b]) / L[b];
            bond_length[b] = l;
            if (eps > epsc && bond_fp[b]) active[b] = 0;
            na += active[b];
        }
        n_active[i] = na;
        /* calc_damage! */
        damage[i] = 1.0 - (double)na / (double)(row_ptr[i + 1] - row_ptr[i]);
        /* force_density_point! */
        double bx = 0.0, by = 0.0, bz = 0.0, ax = 0.0, ay = 0.0, az = 0.0;
        for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
            int64_t j = nbr[b];
            double dx = x[3 * j] - x[3 * i];
            double dy = x[3 * j + 1] - x[3 * i + 1];
            double dz = x[3 * j + 2] - x[3 * i + 2];
            double l = bond_length[b];
            double eps = (l - L[b]) / L[b];
            /* Bool * surface_correction_factor (1, or scfactor[b]) */
            double w = active[b] ? (scf ? scf[b] : 1.0) : 0.0;
            double s = ((w * bc) * eps) * vol[j];
            double tx = (s * dx) / l, ty = (s * dy) / l, tz = (s * dz) / l;
            bx += tx; by += ty; bz += tz;
            ax += fabs(tx); ay += fabs(ty); az += fabs(tz);
        }
        b_int[3 * i] += bx; b_int[3 * i + 1] += by; b_int[3 * i + 2] += bz;
        if (b_abs) { b_abs[3 * i] = ax; b_abs[3 * i + 1] = ay; b_abs[3 * i + 2] = az; }
    }
}

/* The same loop with the reference's exact accumulation order: update_add_vector!
 * adds every bond's vector straight into b_int (static_arrays.jl:191-197).  The
 * two agree bit for bit because b_int starts at zero: 0 + t1 + t2 ... */

/* ------------------------------------------------------------------------------ */
/* Scatter loops run in two phases.  The reference's OSB and CPD loops add each bond's */
/* force to both ends (b_int[:, i] += t V_j, b_int[:, j] -= t V_i), so the serial    */
/* order of the additions into one point's b_int is: the bonds of points i < k that  */
/* list k, then k's own bonds, then those of points i > k.  Phase 1 (points in       */
/* parallel) stores every bond's vector t; phase 2 (points in parallel) gathers      */
/* b_int[k] in exactly that order over the reverse bond index, so the result equals  */
/* the serial loop bit for bit for any thread count.                                */
/* ------------------------------------------------------------------------------ */

/* Reverse bond index: the bonds grouped by their neighbour k = nbr[b], each group in
 * ascending owner order (a stable counting sort by nbr).  in_ptr: n + 1, in_bond and
 * in_src (the owner): one per bond. */
int orc_reverse_index(int64_t n, const int64_t *row_ptr, const int32_t *nbr, int64_t *in_ptr,
                      int64_t *in_bond, int32_t *in_src)
{
    const int64_t nb = row_ptr[n];
    int64_t *pos = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    if (!pos) return ORC_ERR_OOM;
    for (int64_t k = 0; k <= n; k++) in_ptr[k] = 0;
    for (int64_t b = 0; b < nb; b++) in_ptr[nbr[b] + 1]++;
    for (int64_t k = 0; k < n; k++) in_ptr[k + 1] += in_ptr[k];
    memcpy(pos, in_ptr, sizeof(int64_t) * (size_t)(n + 1));
    for (int64_t i = 0; i < n; i++)
        for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
            const int64_t q = pos[nbr[b]]++;
            in_bond[q] = b;
            in_src[q] = (int32_t)i;
        }
    free(pos);
    return ORC_OK;
}

/* per-bond vectors of phase 1 (grow-only; the oracle is called from one host thread) */
static double *g_tb = NULL, *g_tabs = NULL;
static size_t g_tb_cap = 0, g_tabs_cap = 0;
static uint8_t *g_skip = NULL;
static size_t g_skip_cap = 0;

static int grow(void **p, size_t *cap, size_t bytes)
{
    if (bytes <= *cap) return ORC_OK;
    free(*p);
    *p = malloc(bytes);
    *cap = *p ? bytes : 0;
    return *p ? ORC_OK : ORC_ERR_OOM;
}

typedef struct {
    const int64_t *in_ptr, *in_bond;
    const int32_t *in_src;
    int64_t *own_ptr, *own_bond; /* built here when the caller passes no index */
    int32_t *own_src;
} rev_t;

static int rev_get(rev_t *r, int64_t n, const int64_t *row_ptr, const int32_t *nbr,
                   const int64_t *in_ptr, const int64_t *in_bond, const int32_t *in_src)
{
    memset(r, 0, sizeof *r);
    if (in_ptr && in_bond && in_src) {
        r->in_ptr = in_ptr; r->in_bond = in_bond; r->in_src = in_src;
        return ORC_OK;
    }
    const int64_t nb = row_ptr[n];
    r->own_ptr = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    r->own_bond = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nb > 0 ? nb : 1));
    r->own_src = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nb > 0 ? nb : 1));
    if (!r->own_ptr || !r->own_bond || !r->own_src) return ORC_ERR_OOM;
    if (orc_reverse_index(n, row_ptr, nbr, r->own_ptr, r->own_bond, r->own_src) != ORC_OK)
        return ORC_ERR_OOM;
    r->in_ptr = r->own_ptr; r->in_bond = r->own_bond; r->in_src = r->own_src;
    return ORC_OK;
}

static void rev_free(rev_t *r)
{
    free(r->own_ptr); free(r->own_bond); free(r->own_src);
}

/* Phase 2.  abs_mode 0: b_abs[i] += |t V_j|, b_abs[j] += |t V_i| (OSB); abs_mode 1:
 * b_abs[i] += tabs V_j, b_abs[j] += tabs V_i (CPD's rounding scale per bond). */
static void gather(int64_t n, const int64_t *row_ptr, const int32_t *nbr, const rev_t *r,
                   const uint8_t *skip, const double *tb, const double *tabs, int abs_mode,
                   const double *vol, double *b_int, double *b_abs)
{
#pragma omp parallel for schedule(dynamic, 512)
    for (int64_t k = 0; k < n; k++) {
        double acc[3] = {0.0, 0.0, 0.0}, aab[3] = {0.0, 0.0, 0.0};
        int64_t q = r->in_ptr[k];
        const int64_t e = r->in_ptr[k + 1];
        for (int pass = 0; pass < 2; pass++) {
            /* pass 0: owners i < k; pass 1: owners i > k */
            for (; q < e && (pass == 1 || r->in_src[q] < k); q++) {
                const int64_t i = r->in_src[q], b = r->in_bond[q];
                if (skip[i]) continue;
                for (int d = 0; d < 3; d++) {
                    acc[d] += -tb[3 * b + d] * vol[i];
                    if (b_abs) aab[d] += abs_mode ? tabs[3 * b + d] * vol[i] : fabs(tb[3 * b + d] * vol[i]);
                }
            }
            if (pass == 0 && !skip[k])
                for (int64_t b = row_ptr[k]; b < row_ptr[k + 1]; b++) {
                    const int64_t j = nbr[b];
                    for (int d = 0; d < 3; d++) {
                        acc[d] += tb[3 * b + d] * vol[j];
                        if (b_abs) aab[d] += abs_mode ? tabs[3 * b + d] * vol[j] : fabs(tb[3 * b + d] * vol[j]);
                    }
                }
        }
        for (int d = 0; d < 3; d++) {
            b_int[3 * k + d] = acc[d];
            if (b_abs) b_abs[3 * k + d] = aab[d];
        }
    }
}

/* ------------------------------------------------------------------------------ */
/* Ordinary state-based force density (scatter form, as the reference)            */
/* calc_failure! (OSB)         ordinary_state_based.jl:107-126                     */
/* force_density_point!        ordinary_state_based.jl:128-149 (one set),          */
/*                             :151-174 (several sets: per bond                    */
/*                             c1 = 15 (G_i + G_j) / (2 wvol_i),                   */
/*                             c2 = dil_i (3 (K_i + K_j) / (2 wvol_i) - c1 / 3))   */
/* calc_weighted_volume        ordinary_state_based.jl:176-189                     */
/* calc_dilatation             ordinary_state_based.jl:191-204                     */
/* omega[b] = kernel(delta_i, L) (find_kernels, bond_system.jl:184-196)            */
/* Points with wvol = 0 add nothing.  multi = 0: G_p, K_p, epsc_p hold one value;   */
/* multi = 1: one value per point.                                                 */
/* ------------------------------------------------------------------------------ */
static int osb_core(int64_t n, const double *X, const double *x, const double *vol,
                    const int64_t *row_ptr, const int32_t *nbr, const double *L,
                    const uint8_t *bond_fp, uint8_t *active, double *bond_length,
                    const double *omega, int64_t *n_active, double *damage, const double *G_p,
                    const double *K_p, const double *epsc_p, int multi, double *b_int,
                    double *wvol_out, double *dil_out, double *b_abs, const double *scf,
                    const int64_t *in_ptr, const int64_t *in_bond, const int32_t *in_src,
                    int nthreads)
{
    set_threads(nthreads);
    const int64_t nb = row_ptr[n];
    const int64_t ps = multi ? 1 : 0;
    if (grow((void **)&g_tb, &g_tb_cap, sizeof(double) * 3 * (size_t)(nb > 0 ? nb : 1)) ||
        grow((void **)&g_skip, &g_skip_cap, (size_t)(n > 0 ? n : 1)))
        return ORC_ERR_OOM;
    rev_t r;
    if (rev_get(&r, n, row_ptr, nbr, in_ptr, in_bond, in_src) != ORC_OK) {
        rev_free(&r);
        return ORC_ERR_OOM;
    }
    double *tb = g_tb;
    uint8_t *skip = g_skip;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; i++) {
        const double epsc = epsc_p[ps * i];
        int64_t na = 0;
        for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
            int64_t j = nbr[b];
            double dx = x[3 * j] - x[3 * i];
            double dy = x[3 * j + 1] - x[3 * i + 1];
            double dz = x[3 * j + 2] - x[3 * i + 2];
            double l = sqrt((dx * dx + dy * dy) + dz * dz);
            bond_length[b] = l;
            double eps = (l - L[b]) / L[b];
            if (eps > epsc && bond_fp[b]) active[b] = 0;
            na += active[b];
        }
        n_active[i] = na;
        damage[i] = 1.0 - (double)na / (double)(row_ptr[i + 1] - row_ptr[i]);
        /* calc_weighted_volume */
        double wvol = 0.0;
        for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
            int64_t j = nbr[b];
            double Dx = X[3 * j] - X[3 * i];
            double Dy = X[3 * j + 1] - X[3 * i + 1];
            double Dz = X[3 * j + 2] - X[3 * i + 2];
            double dsq = (Dx * Dx + Dy * Dy) + Dz * Dz;
            double w = active[b] ? omega[b] : 0.0;
            wvol += ((w * (scf ? scf[b] : 1.0)) * dsq) * vol[j];
        }
        if (wvol_out) wvol_out[i] = wvol;
        if (dil_out) dil_out[i] = 0.0;
        skip[i] = wvol == 0.0;
        if (skip[i]) continue;
        /* calc_dilatation */
        double dil = 0.0;
        double c1d = 3.0 / wvol;
        for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
            int64_t j = nbr[b];
            double l = bond_length[b];
            double w = active[b] ? omega[b] : 0.0;
            dil += ((((w * (scf ? scf[b] : 1.0)) * c1d) * L[b]) * (l - L[b])) * vol[j];
        }
        if (dil_out) dil_out[i] = dil;
        for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
            int64_t j = nbr[b];
            double dx = x[3 * j] - x[3 * i];
            double dy = x[3 * j + 1] - x[3 * i + 1];
            double dz = x[3 * j + 2] - x[3 * i + 2];
            double l = bond_length[b];
            double c1, c2;
            if (multi) {
                c1 = 15.0 * (G_p[i] + G_p[j]) / (2.0 * wvol);
                c2 = dil * (3.0 * (K_p[i] + K_p[j]) / (2.0 * wvol) - c1 / 3.0);
            } else {
                c1 = (15.0 * G_p[0]) / wvol;
                c2 = dil * ((3.0 * K_p[0]) / wvol - c1 / 3.0);
            }
            double w = active[b] ? omega[b] : 0.0;
            double s = ((w * (scf ? scf[b] : 1.0)) * (c2 * L[b] + c1 * (l - L[b]))) / l;
            tb[3 * b] = s * dx;
            tb[3 * b + 1] = s * dy;
            tb[3 * b + 2] = s * dz;
        }
    }
    gather(n, row_ptr, nbr, &r, skip, tb, NULL, 0, vol, b_int, b_abs);
    rev_free(&r);
    return ORC_OK;
}

int orc_osb_force(int64_t n, const double *X, const double *x, const double *vol,
                  const int64_t *row_ptr, const int32_t *nbr, const double *L,
                  const uint8_t *bond_fp, uint8_t *active, double *bond_length,
                  const double *omega, int64_t *n_active, double *damage, double G,
                  double K, double epsc, double *b_int, double *wvol_out, double *dil_out,
                  double *b_abs, const double *scf, const int64_t *in_ptr,
                  const int64_t *in_bond, const int32_t *in_src, int nthreads)
{
    return osb_core(n, X, x, vol, row_ptr, nbr, L, bond_fp, active, bond_length, omega,
                    n_active, damage, &G, &K, &epsc, 0, b_int, wvol_out, dil_out, b_abs, scf,
                    in_ptr, in_bond, in_src, nthreads);
}

/* ------------------------------------------------------------------------------ */
/* Several parameter sets on one body (ParameterHandler, src/core/parameter_handler.jl:36-81):
 * per-point arrays hold each point's set.
 * BB: calc_failure! with the owner's eps_c (bond_system.jl:309-326), force_density_point!
 * with (bc_i + bc_j) / 2 (bond_based.jl:163-180):
 *   b = ((((w * (bc_i + bc_j)) / 2) * eps) * V_j) * dx / l                                  */
void orc_bb_force_sets(int64_t n, const double *x, const double *vol, const int64_t *row_ptr,
                       const int32_t *nbr, const double *L, const uint8_t *bond_fp,
                       uint8_t *active, double *bond_length, int64_t *n_active, double *damage,
                       const double *bc_pt, const double *epsc_pt, double *b_int, double *b_abs,
                       const double *scf)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; i++) {
        int64_t na = 0;
        for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
            int64_t j = nbr[b];
            double dx = x[3 * j] - x[3 * i];
            double dy = x[3 * j + 1] - x[3 * i + 1];
            double dz = x[3 * j + 2] - x[3 * i + 2];
            double l = sqrt((dx * dx + dy * dy) + dz * dz);
            double eps = (l - L[b]) / L[b];
            bond_length[b] = l;
            if (eps > epsc_pt[i] && bond_fp[b]) active[b] = 0;
            na += active[b];
        }
        n_active[i] = na;
        damage[i] = 1.0 - (double)na / (double)(row_ptr[i + 1] - row_ptr[i]);
        double bx = 0.0, by = 0.0, bz = 0.0, ax = 0.0, ay = 0.0, az = 0.0;
        for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
            int64_t j = nbr[b];
            double dx = x[3 * j] - x[3 * i];
            double dy = x[3 * j + 1] - x[3 * i + 1];
            double dz = x[3 * j + 2] - x[3 * i + 2];
            double l = bond_length[b];
            double eps = (l - L[b]) / L[b];
            double w = active[b] ? (scf ? scf[b] : 1.0) : 0.0;
            double s = (((w * (bc_pt[i] + bc_pt[j])) / 2.0) * eps) * vol[j];
            double tx = (s * dx) / l, ty = (s * dy) / l, tz = (s * dz) / l;
            bx += tx; by += ty; bz += tz;
            ax += fabs(tx); ay += fabs(ty); az += fabs(tz);
        }
        b_int[3 * i] = bx; b_int[3 * i + 1] = by; b_int[3 * i + 2] = bz;
        if (b_abs) { b_abs[3 * i] = ax; b_abs[3 * i + 1] = ay; b_abs[3 * i + 2] = az; }
    }
}

int orc_osb_force_sets(int64_t n, const double *X, const double *x, const double *vol,
                       const int64_t *row_ptr, const int32_t *nbr, const double *L,
                       const uint8_t *bond_fp, uint8_t *active, double *bond_length,
                       const double *omega, int64_t *n_active, double *damage,
                       const double *G_pt, const double *K_pt, const double *epsc_pt,
                       double *b_int, double *b_abs, const double *scf,
                       const int64_t *in_ptr, const int64_t *in_bond, const int32_t *in_src,
                       int nthreads)
{
    return osb_core(n, X, x, vol, row_ptr, nbr, L, bond_fp, active, bond_length, omega,
                    n_active, damage, G_pt, K_pt, epsc_pt, 1, b_int, NULL, NULL, b_abs, scf,
                    in_ptr, in_bond, in_src, nthreads);
}

/* EnergySurfaceCorrection factors (bond_system_corrections.jl:98-204).  Per-point parameter
 * arrays; multi = 0 takes each point's own (lambda, mu) (get_averaged_lame_parameters with
 * one parameter set, :124-127), multi = 1 averages them over the active bonds (:129-145).
 * kind 0 = BB (strain_energy_density_point!, bond_based.jl:183-201), 1 = OSB
 * (ordinary_state_based.jl:207-235), both at scfactor = 1 and the reference configuration
 * stretched by 1.001 along one axis at a time.  mfactor[3 i + d], scfactor[b].  omega: the
 * OSB influence function per bond (kernel(system, bond_id), ordinary_state_based.jl:225), or
 * null for linear_kernel delta / L. */
static double orc_norm3(double a, double b, double c) { return sqrt((a * a + b * b) + c * c); }
static double sq(double a) { return a * a; } /* Julia's literal x^2 */
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

void orc_esc_factors(int64_t n, int kind, int multi, const double *X, const double *vol,
                     const int64_t *row_ptr, const int32_t *nbr, const double *L,
                     const uint8_t *active, double delta, const double *lam_pt,
                     const double *mu_pt, const double *bc_pt, const double *G_pt,
                     const double *K_pt, double *mfactor, double *scfactor,
                     const double *omega)
{
    const double st = 1.001;
    double *x = (double *)malloc(sizeof(double) * 3 * n);
    for (int d = 0; d < 3; d++) {
        for (int64_t k = 0; k < 3 * n; k++) x[k] = X[k];
        for (int64_t i = 0; i < n; i++) x[3 * i + d] = st * X[3 * i + d];
        for (int64_t i = 0; i < n; i++) {
            double la = lam_pt[i], mu = mu_pt[i];
            if (multi) {
                double sl = 0.0, sm = 0.0;
                int64_t na = 0;
                for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
                    if (!active[b]) continue;
                    int64_t j = nbr[b];
                    sl += (lam_pt[i] + lam_pt[j]) / 2.0;
                    sm += (mu_pt[i] + mu_pt[j]) / 2.0;
                    na++;
                }
                la = sl / (double)na;
                mu = sm / (double)na;
            }
            double psi_t = (0.5 * la + mu) * ((st - 1.0) * (st - 1.0));
            double psi = 0.0;
            if (kind == 0) {
                for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
                    int64_t j = nbr[b];
                    double l = orc_norm3(x[3 * j] - x[3 * i], x[3 * j + 1] - x[3 * i + 1],
                                         x[3 * j + 2] - x[3 * i + 2]);
                    double eps = (l - L[b]) / L[b];
                    double w = active[b] ? 1.0 : 0.0;
                    double bc = (bc_pt[i] + bc_pt[j]) / 2.0;
                    psi += (((((0.25 * w) * bc) * eps) * eps) * L[b]) * vol[j];
                }
            } else {
                double wvol = 0.0;
                for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
                    int64_t j = nbr[b];
                    double Dx = X[3 * j] - X[3 * i], Dy = X[3 * j + 1] - X[3 * i + 1],
                           Dz = X[3 * j + 2] - X[3 * i + 2];
                    double w = active[b] ? (omega ? omega[b] : delta / L[b]) : 0.0;
                    wvol += ((w * 1.0) * ((Dx * Dx + Dy * Dy) + Dz * Dz)) * vol[j];
                }
                if (wvol != 0.0) {
                    double c1 = 3.0 / wvol, dil = 0.0;
                    for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
                        int64_t j = nbr[b];
                        double l = orc_norm3(x[3 * j] - x[3 * i], x[3 * j + 1] - x[3 * i + 1],
                                             x[3 * j + 2] - x[3 * i + 2]);
                        double w = active[b] ? (omega ? omega[b] : delta / L[b]) : 0.0;
                        dil += ((((w * 1.0) * c1) * L[b]) * (l - L[b])) * vol[j];
                    }
                    double psi_vol = (0.5 * K_pt[i]) * (dil * dil), psi_dev = 0.0;
                    for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
                        int64_t j = nbr[b];
                        double l = orc_norm3(x[3 * j] - x[3 * i], x[3 * j + 1] - x[3 * i + 1],
                                             x[3 * j + 2] - x[3 * i + 2]);
                        double e = l - L[b];
                        double edev = e - ((1.0 / 3.0) * dil) * L[b];
                        double w = active[b] ? (omega ? omega[b] : delta / L[b]) : 0.0;
                        double G = (G_pt[i] + G_pt[j]) / 2.0;
                        double cdev = (15.0 * G) / (2.0 * wvol);
                        psi_dev += ((((cdev * w) * 1.0) * edev) * edev) * vol[j];
                    }
                    psi = psi_vol + psi_dev;
                }
            }
            mfactor[3 * i + d] = psi_t / psi;
        }
    }
    free(x);
    const double right = (90.0 * M_PI) / 180.0;
    for (int64_t i = 0; i < n; i++) {
        for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
            int64_t j = nbr[b];
            double dx = X[3 * j] - X[3 * i], dy = X[3 * j + 1] - X[3 * i + 1],
                   dz = X[3 * j + 2] - X[3 * i + 2];
            double scx = (mfactor[3 * i] + mfactor[3 * j]) / 2.0;
            double scy = (mfactor[3 * i + 1] + mfactor[3 * j + 1]) / 2.0;
            double scz = (mfactor[3 * i + 2] + mfactor[3 * j + 2]) / 2.0;
            double th, ph, scr;
            if (fabs(dz) <= 1e-10) {
                if (fabs(dy) <= 1e-10) th = 0.0;
                else if (fabs(dx) <= 1e-10) th = right;
                else th = atan(fabs(dy) / fabs(dx));
                ph = right;
                scr = sqrt(1.0 / ((sq(cos(th) * sin(ph)) / sq(scx) +
                                   sq(sin(th) * sin(ph)) / sq(scy)) +
                                  sq(cos(ph)) / sq(scz)));
            } else if (fabs(dx) <= 1e-10 && fabs(dy) <= 1e-10) {
                scr = scz;
            } else {
                th = atan(fabs(dy) / fabs(dx));
                ph = acos(fabs(dz) / L[b]);
                scr = sqrt(1.0 / ((sq(cos(th) * sin(ph)) / sq(scx) +
                                   sq(sin(th) * sin(ph)) / sq(scy)) +
                                  sq(cos(ph)) / sq(scz)));
            }
            scfactor[b] = scr;
        }
    }
}

/* calc_timestep with per-point parameters (velocity_verlet.jl:155-163, bond_system.jl:269-277) */
double orc_min_timestep_sets(int64_t n, const double *vol, const int64_t *row_ptr,
                             const int32_t *nbr, const double *L, const double *bc_pt,
                             const double *rho_pt)
{
    double mn = INFINITY;
    for (int64_t i = 0; i < n; i++) {
        double dtsum = 0.0;
        for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) dtsum += vol[nbr[b]] * bc_pt[i] / L[b];
        double dt = sqrt(2.0 * rho_pt[i] / dtsum);
        if (dt < mn) mn = dt;
    }
    return mn;
}

/* update_acc_and_vel! with per-point rho (velocity_verlet.jl:262-275) */
void orc_update_acc_and_vel_sets(int64_t n, const double *b_int, const double *b_ext,
                                 const double *rho_pt, const double *vh, double dth, double *a,
                                 double *v)
{
    for (int64_t k = 0; k < 3 * n; k++) {
        a[k] = (b_int[k] + b_ext[k]) / rho_pt[k / 3];
        v[k] = vh[k] + a[k] * dth;
    }
}

/* ------------------------------------------------------------------------------ */
/* 3x3 helpers, column-major like Julia (A[r + 3c])                                */
/* ------------------------------------------------------------------------------ */
static void m3_mul(const double *A, const double *B, double *C)
{
    double T[9];
    for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++)
            T[r + 3 * c] = (A[r] * B[3 * c] + A[r + 3] * B[1 + 3 * c]) + A[r + 6] * B[2 + 3 * c];
    memcpy(C, T, sizeof T);
}

/* A' * B */
static void m3_tmul(const double *A, const double *B, double *C)
{
    double T[9];
    for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++)
            T[r + 3 * c] = (A[3 * r] * B[3 * c] + A[1 + 3 * r] * B[1 + 3 * c]) +
                           A[2 + 3 * r] * B[2 + 3 * c];
    memcpy(C, T, sizeof T);
}

/* A * B' */
static void m3_mult(const double *A, const double *B, double *C)
{
    double T[9];
    for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++)
            T[r + 3 * c] = (A[r] * B[c] + A[r + 3] * B[c + 3]) + A[r + 6] * B[c + 6];
    memcpy(C, T, sizeof T);
}

static void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

static double dot3(const double *a, const double *b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

/* det of a 3x3 as the column triple product x0 . (x1 x x2) */
static double m3_det(const double *A)
{
    double y[3];
    cross3(A + 3, A + 6, y);
    return dot3(A, y);
}

/* inverse via the adjugate rows x1 x x2, x2 x x0, x0 x x1 over the determinant */
static void m3_inv(const double *A, double *Ainv)
{
    double x0[3] = {A[0], A[1], A[2]}, x1[3] = {A[3], A[4], A[5]}, x2[3] = {A[6], A[7], A[8]};
    double y0[3], y1[3], y2[3];
    cross3(x1, x2, y0);
    double d = dot3(x0, y0);
    for (int k = 0; k < 3; k++) { x0[k] = x0[k] / d; y0[k] = y0[k] / d; }
    cross3(x2, x0, y1);
    cross3(x0, x1, y2);
    Ainv[0] = y0[0]; Ainv[1] = y1[0]; Ainv[2] = y2[0];
    Ainv[3] = y0[1]; Ainv[4] = y1[1]; Ainv[5] = y2[1];
    Ainv[6] = y0[2]; Ainv[7] = y1[2]; Ainv[8] = y2[2];
}

static int m3_hasnan(const double *A)
{
    for (int k = 0; k < 9; k++) if (isnan(A[k])) return 1;
    return 0;
}

/* first_piola_kirchhoff, SaintVenantKirchhoff: constitutive_models.jl:35-41 */
static void pk1_svk(const double *F, double lambda, double mu, double *P)
{
    double C[9], E[9], S[9];
    m3_tmul(F, F, C);
    for (int k = 0; k < 9; k++) E[k] = 0.5 * (C[k] - ((k % 4 == 0) ? 1.0 : 0.0));
    double trE = (E[0] + E[4]) + E[8];
    double lt = lambda * trE;
    double tm = 2.0 * mu;
    for (int k = 0; k < 9; k++) S[k] = tm * E[k];
    S[0] += lt; S[4] += lt; S[8] += lt;
    m3_mul(F, S, P);
}

/* first_piola_kirchhoff, NeoHookePenalty: constitutive_models.jl:231-243 */
static void pk1_nhp(const double *F, double G, double K, double *P)
{
    double J = m3_det(F);
    if (J < DBL_EPSILON || isnan(J)) { for (int k = 0; k < 9; k++) P[k] = 0.0; return; }
    double C[9], Ci[9], S[9];
    m3_tmul(F, F, C);
    m3_inv(C, Ci);
    double trC = (C[0] + C[4]) + C[8];
    double a = (1.0 / 3.0) * trC;
    double Jp = pow(J, -2.0 / 3.0);
    double iJ = 1.0 / J;
    double vol = (K / 4.0) * (J * J - iJ * iJ);
    for (int k = 0; k < 9; k++) {
        double Iab = (k % 4 == 0) ? 1.0 : 0.0;
        double X = Iab - a * Ci[k];
        S[k] = (G * X) * Jp + vol * Ci[k];
    }
    m3_mul(F, S, P);
}

/* first_piola_kirchhoff, LinearElastic: constitutive_models.jl:83-97 (S = C_voigt E_voigt,
 * E_voigt = (E11, E22, E33, 2 E23, 2 E31, 2 E12), get_hooke_matrix_voigt :113-124) */
static void pk1_le(const double *F, double lambda, double mu, double nu, double *P)
{
    double C[9], E[9], S[9];
    m3_tmul(F, F, C);
    for (int k = 0; k < 9; k++) E[k] = 0.5 * (C[k] - ((k % 4 == 0) ? 1.0 : 0.0));
    double a = ((1.0 - nu) * lambda) / nu;
    double ev[6] = {E[0], E[4], E[8], 2.0 * E[7], 2.0 * E[2], 2.0 * E[3]};
    double cv[36] = {a, lambda, lambda, 0, 0, 0, lambda, a, lambda, 0, 0, 0,
                     lambda, lambda, a, 0, 0, 0, 0, 0, 0, mu, 0, 0,
                     0, 0, 0, 0, mu, 0, 0, 0, 0, 0, 0, mu};
    double sv[6];
    for (int r = 0; r < 6; r++) {
        double acc = 0.0;
        for (int k = 0; k < 6; k++) acc += cv[r + 6 * k] * ev[k];
        sv[r] = acc;
    }
    /* S = [s1 s6 s5; s6 s2 s4; s5 s4 s3] */
    S[0] = sv[0]; S[1] = sv[5]; S[2] = sv[4];
    S[3] = sv[5]; S[4] = sv[1]; S[5] = sv[3];
    S[6] = sv[4]; S[7] = sv[3]; S[8] = sv[2];
    m3_mul(F, S, P);
}

/* first_piola_kirchhoff, NeoHooke: constitutive_models.jl:169-175,
 * S = mu (I - C^-1) + lambda log(J) C^-1 */
static void pk1_nh(const double *F, double lambda, double mu, double *P)
{
    double J = m3_det(F);
    double C[9], Ci[9], S[9];
    m3_tmul(F, F, C);
    m3_inv(C, Ci);
    for (int k = 0; k < 9; k++) {
        double Iab = (k % 4 == 0) ? 1.0 : 0.0;
        S[k] = mu * (Iab - Ci[k]) + (lambda * log(J)) * Ci[k];
    }
    m3_mul(F, S, P);
}

/* the constitutive models by code: 0 SVK, 1 NeoHookePenalty, 2 LinearElastic, 3 NeoHooke */
void orc_pk1(int model, const double *F, double lambda, double mu, double G, double K, double nu,
             double *P)
{
    if (model == 1) pk1_nhp(F, G, K, P);
    else if (model == 2) pk1_le(F, lambda, mu, nu, P);
    else if (model == 3) pk1_nh(F, lambda, mu, P);
    else pk1_svk(F, lambda, mu, P);
}

/* cauchy_stress: stress.jl:112-116, sigma = (1/J) P F' */
static void cauchy(const double *P, const double *F, double *sig)
{
    double J = m3_det(F);
    double s = 1.0 / J;
    double Ps[9];
    for (int k = 0; k < 9; k++) Ps[k] = s * P[k];
    m3_mult(Ps, F, sig);
}

/* ------------------------------------------------------------------------------ */
/* Correspondence force density (scatter form, as the reference)                   */
/* calc_failure! (generic)      bond_system.jl:309-327                             */
/* force_density_point!         correspondence.jl:183-191                          */
/* calc_deformation_gradient!   correspondence.jl:193-214                          */
/* too_much_damage!             correspondence.jl:288-298                          */
/* calc_first_piola_kirchhoff!  correspondence.jl:216-224                          */
/* c_force_density! ZEMSilling  correspondence.jl:226-251                          */
/* c_force_density! ZEMWan     correspondence.jl:253-285                          */
/* model & 0xff: 0 SaintVenantKirchhoff, 1 NeoHookePenalty, 2 LinearElastic,      */
/* 3 NeoHooke; model >> 8: 0 ZEMSilling, 1 ZEMWan                                 */
/* ------------------------------------------------------------------------------ */
/* calc_zem_stiffness_tensor (zem_stabilization.jl:32-42): C_1[i,j] = sum_kl C[i,j,k,l]
 * Kinv[k,l] of the Hooke tensor get_hooke_matrix (constitutive_models.jl:107-133): Voigt
 * a = (1 - nu) lambda / nu on the normal diagonal, lambda off it, mu for the three shears,
 * each shear entry at the four index orders of its pair.  Column-major 3x3. */
void orc_zem_c1(const double *Kinv, double lambda, double mu, double nu, double *C1)
{
    const double a = ((1.0 - nu) * lambda) / nu;
    for (int c = 0; c < 3; c++)
        for (int r = 0; r < 3; r++) {
            if (r == c) {
                double s = 0.0;
                for (int k = 0; k < 3; k++) s += (k == r ? a : lambda) * Kinv[k + 3 * k];
                C1[r + 3 * c] = s;
            } else {
                C1[r + 3 * c] = mu * Kinv[r + 3 * c] + mu * Kinv[c + 3 * r];
            }
        }
}

/* Parameters per point (`ps` = 1: arrays of n, one value per point's parameter set, the
 * params of the point whose loop runs, as force_density_point! receives them with a
 * ParameterHandler) or one value for all (`ps` = 0). */
static int cpd_force_core(int64_t n, const double *X, const double *x, const double *vol,
                           const int64_t *row_ptr, const int32_t *nbr, const double *L,
                           const uint8_t *bond_fp, uint8_t *active, const double *omega,
                           int64_t *n_active, double *damage, int model, const double *lambda_p,
                           const double *mu_p, const double *G_p, const double *K_p,
                           const double *nu_p, const double *bc_p, const double *epsc_p,
                           int64_t ps, double delta,
                           double Cs, double maxdmg, double *defgrad, double *cauchy_out,
                           double *b_int, double *b_abs, const int64_t *in_ptr,
                           const int64_t *in_bond, const int32_t *in_src, int nthreads)
{
    const int wan = (model >> 8) == 1;
    model &= 0xff;
    set_threads(nthreads);
    const int64_t nb = row_ptr[n];
    const size_t bvec = sizeof(double) * 3 * (size_t)(nb > 0 ? nb : 1);
    if (grow((void **)&g_tb, &g_tb_cap, bvec) || grow((void **)&g_tabs, &g_tabs_cap, bvec) ||
        grow((void **)&g_skip, &g_skip_cap, (size_t)(n > 0 ? n : 1)))
        return ORC_ERR_OOM;
    rev_t rv;
    if (rev_get(&rv, n, row_ptr, nbr, in_ptr, in_bond, in_src) != ORC_OK) {
        rev_free(&rv);
        return ORC_ERR_OOM;
    }
    double *tb = g_tb, *tba = g_tabs;
    uint8_t *skip = g_skip;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; i++) {
        const double lambda = lambda_p[ps * i], mu = mu_p[ps * i], G = G_p[ps * i], K = K_p[ps * i];
        const double nu = nu_p[ps * i];
        const double epsc = epsc_p[ps * i];
        const double beta = (Cs * bc_p[ps * i]) / delta;
        int64_t na = 0;
        for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
            int64_t j = nbr[b];
            double dx = x[3 * j] - x[3 * i];
            double dy = x[3 * j + 1] - x[3 * i + 1];
            double dz = x[3 * j + 2] - x[3 * i + 2];
            double l = sqrt((dx * dx + dy * dy) + dz * dz);
            double eps = (l - L[b]) / L[b];
            if (eps > epsc && bond_fp[b]) active[b] = 0;
            na += active[b];
        }
        n_active[i] = na;
        damage[i] = 1.0 - (double)na / (double)(row_ptr[i + 1] - row_ptr[i]);
        /* calc_deformation_gradient! */
        double Km[9] = {0}, Ft[9] = {0};
        for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
            int64_t j = nbr[b];
            double DX[3], dxv[3];
            for (int d = 0; d < 3; d++) {
                DX[d] = X[3 * j + d] - X[3 * i + d];
                dxv[d] = x[3 * j + d] - x[3 * i + d];
            }
            double w = active[b] ? omega[b] : 0.0;
            double temp = w * vol[j];
            for (int c = 0; c < 3; c++)
                for (int r = 0; r < 3; r++) {
                    Km[r + 3 * c] += temp * (DX[r] * DX[c]);
                    Ft[r + 3 * c] += temp * (dxv[r] * DX[c]);
                }
        }
        double Kinv[9], F[9];
        m3_inv(Km, Kinv);
        m3_mul(Ft, Kinv, F);
        if (defgrad) memcpy(defgrad + 9 * i, F, sizeof F);
        /* too_much_damage! */
        skip[i] = damage[i] > maxdmg || m3_hasnan(F);
        if (skip[i]) {
            for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) active[b] = 0;
            n_active[i] = 0;
            continue;
        }
        /* calc_first_piola_kirchhoff! */
        double P[9], PKinv[9], sig[9];
        orc_pk1(model, F, lambda, mu, G, K, nu, P);
        m3_mul(P, Kinv, PKinv);
        cauchy(P, F, sig);
        if (cauchy_out) memcpy(cauchy_out + 9 * i, sig, sizeof sig);
        double C1[9] = {0};
        if (wan) orc_zem_c1(Kinv, lambda, mu, nu, C1);
        /* c_force_density! (ZEMSilling: tzem = w beta z; ZEMWan: tzem = (w C_1) z) */
        for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) {
            int64_t j = nbr[b];
            double DX[3], dxv[3];
            for (int d = 0; d < 3; d++) {
                DX[d] = X[3 * j + d] - X[3 * i + d];
                dxv[d] = x[3 * j + d] - x[3 * i + d];
            }
            double w = active[b] ? omega[b] : 0.0;
            double FX[3], z[3], t[3];
            for (int r = 0; r < 3; r++) FX[r] = (F[r] * DX[0] + F[r + 3] * DX[1]) + F[r + 6] * DX[2];
            for (int r = 0; r < 3; r++) z[r] = dxv[r] - FX[r];
            double wb = w * beta;
            double tabs[3];
            for (int r = 0; r < 3; r++) {
                double tz, zabs;
                if (wan) {
                    tz = ((w * C1[r]) * z[0] + (w * C1[r + 3]) * z[1]) + (w * C1[r + 6]) * z[2];
                    zabs = 0.0;
                    for (int c = 0; c < 3; c++)
                        zabs += fabs(w * C1[r + 3 * c]) *
                                (fabs(dxv[c]) + fabs(F[c] * DX[0]) + fabs(F[c + 3] * DX[1]) +
                                 fabs(F[c + 6] * DX[2]));
                } else {
                    tz = wb * z[r];
                    zabs = fabs(wb) * (fabs(dxv[r]) + fabs(F[r] * DX[0]) + fabs(F[r + 3] * DX[1]) +
                                       fabs(F[r + 6] * DX[2]));
                }
                double m0 = w * PKinv[r], m1 = w * PKinv[r + 3], m2 = w * PKinv[r + 6];
                t[r] = ((m0 * DX[0] + m1 * DX[1]) + m2 * DX[2]) + tz;
                /* rounding scale of t: its parts before the cancellations inside
                 * z = dx - F dX and P K^-1 dX */
                tabs[r] = (fabs(m0 * DX[0]) + fabs(m1 * DX[1]) + fabs(m2 * DX[2])) + zabs;
            }
            for (int r = 0; r < 3; r++) {
                tb[3 * b + r] = t[r];
                tba[3 * b + r] = tabs[r];
            }
        }
    }
    gather(n, row_ptr, nbr, &rv, skip, tb, tba, 1, vol, b_int, b_abs);
    rev_free(&rv);
    return ORC_OK;
}

int orc_cpd_force(int64_t n, const double *X, const double *x, const double *vol,
                  const int64_t *row_ptr, const int32_t *nbr, const double *L,
                  const uint8_t *bond_fp, uint8_t *active, const double *omega,
                  int64_t *n_active, double *damage, int model, double lambda, double mu,
                  double G, double K, double nu, double bc, double delta, double Cs,
                  double maxdmg, double epsc, double *defgrad, double *cauchy_out,
                  double *b_int, double *b_abs, const int64_t *in_ptr, const int64_t *in_bond,
                  const int32_t *in_src, int nthreads)
{
    return cpd_force_core(n, X, x, vol, row_ptr, nbr, L, bond_fp, active, omega, n_active,
                          damage, model, &lambda, &mu, &G, &K, &nu, &bc, &epsc, 0, delta, Cs,
                          maxdmg, defgrad, cauchy_out, b_int, b_abs, in_ptr, in_bond, in_src,
                          nthreads);
}

/* several parameter sets: lambda, mu, G, K, bc, eps_c per point (the sets share delta, the
 * material's model, Cs and maxdmg) */
int orc_cpd_force_sets(int64_t n, const double *X, const double *x, const double *vol,
                       const int64_t *row_ptr, const int32_t *nbr, const double *L,
                       const uint8_t *bond_fp, uint8_t *active, const double *omega,
                       int64_t *n_active, double *damage, int model, const double *lambda_pt,
                       const double *mu_pt, const double *G_pt, const double *K_pt,
                       const double *nu_pt, const double *bc_pt, const double *epsc_pt,
                       double delta, double Cs, double maxdmg, double *defgrad,
                       double *cauchy_out, double *b_int, double *b_abs,
                       const int64_t *in_ptr, const int64_t *in_bond, const int32_t *in_src,
                       int nthreads)
{
    return cpd_force_core(n, X, x, vol, row_ptr, nbr, L, bond_fp, active, omega, n_active,
                          damage, model, lambda_pt, mu_pt, G_pt, K_pt, nu_pt, bc_pt, epsc_pt, 1,
                          delta, Cs, maxdmg, defgrad, cauchy_out, b_int, b_abs, in_ptr, in_bond,
                          in_src, nthreads);
}

/* ------------------------------------------------------------------------------ */
/* Time integration pieces                                                         */
/* ------------------------------------------------------------------------------ */

/* calc_timestep_point: bond_system.jl:269-277; calc_timestep / calc_stable_timestep:
 * velocity_verlet.jl:136-163 (min over points, times the safety factor outside). */
double orc_min_timestep(int64_t n, const double *vol, const int64_t *row_ptr,
                        const int32_t *nbr, const double *L, double bc, double rho)
{
    double mn = INFINITY;
    for (int64_t i = 0; i < n; i++) {
        double dtsum = 0.0;
        for (int64_t b = row_ptr[i]; b < row_ptr[i + 1]; b++) dtsum += vol[nbr[b]] * bc / L[b];
        double dt = sqrt(2.0 * rho / dtsum);
        if (dt < mn) mn = dt;
    }
    return mn;
}

/* update_vel_half!: velocity_verlet.jl:239-245 */
void orc_update_vel_half(int64_t ndof, const double *v, const double *a, double dth, double *vh)
{
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < ndof; k++) vh[k] = v[k] + a[k] * dth;
}

/* update_disp_and_pos!: velocity_verlet.jl:247-255 */
void orc_update_disp_and_pos(int64_t ndof, const double *vh, double dt, double *u, double *x)
{
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < ndof; k++) {
        double du = vh[k] * dt;
        u[k] += du;
        x[k] += du;
    }
}

/* _update_acc_and_vel! (one parameter set): velocity_verlet.jl:280-295 */
void orc_update_acc_and_vel(int64_t ndof, const double *b_int, const double *b_ext, double rho,
                            const double *vh, double dth, double *a, double *v)
{
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < ndof; k++) {
        a[k] = (b_int[k] + b_ext[k]) / rho;
        v[k] = vh[k] + a[k] * dth;
    }
}

/* check_for_nans: nans.jl:10-18 */
int orc_has_nan(int64_t m, const double *a)
{
    for (int64_t k = 0; k < m; k++) if (isnan(a[k])) return 1;
    return 0;
}
