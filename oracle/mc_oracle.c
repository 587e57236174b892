/*
 * mc_oracle.c -- CPU oracle for the miniCOMBUST flow-solver SIMPLE step.
 *
 * TEST INFRASTRUCTURE ONLY (see mc_oracle.h).  Single-threaded C restatement
 * of the reference CPU path at one flow rank (no halos: nhalos = 0, the
 * boundary map is empty, local_cells_disp = shmem_cell_disp = 0).  Each
 * function cites the reference lines it restates.  Build with
 * -ffp-contract=off so every expression rounds as written.
 *
 * Allocation note: the reference mallocs face_fields uninitialised and its
 * constructor assembles from them (FlowSolver.hpp:201,538-539); the oracle
 * zero-initialises every array, which is what fresh pages hold.
 */
#include "mc_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MESH_BOUNDARY UINT64_MAX

/* std::min / std::max semantics: min(a,b) = (b < a) ? b : a,
   max(a,b) = (a < b) ? b : a. */
static inline double dmin(double a, double b) { return (b < a) ? b : a; }
static inline double dmax(double a, double b) { return (a < b) ? b : a; }

/* vec helpers, include/utils/utils.hpp:184-354 (operand order preserved) */
static inline mco_vec vadd(mco_vec a, mco_vec b) { mco_vec r = {a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
static inline mco_vec vsub(mco_vec a, mco_vec b) { mco_vec r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static inline mco_vec vmul(mco_vec a, double b) { mco_vec r = {a.x * b, a.y * b, a.z * b}; return r; }
static inline mco_vec vdiv(mco_vec a, double b) { mco_vec r = {a.x / b, a.y / b, a.z / b}; return r; }
static inline double  vdot(mco_vec a, mco_vec b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline double  vmag(mco_vec v) { return sqrt(v.x * v.x + v.y * v.y + v.z * v.z); }
static inline mco_vec vcross(mco_vec a, mco_vec b)
{
    mco_vec r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
    return r;
}
static inline mco_vec vnormalise(mco_vec a) { return vdiv(a, vmag(a)); }
static inline double  vcosangle(mco_vec a, mco_vec b) { return vdot(a, b) / (vmag(a) * vmag(b)); }

/* Mesh.hpp:114-120 */
static const double GAS_VEL_X = 50.0, GAS_VEL_Y = 0.0, GAS_VEL_Z = 0.0;
static const double GAS_PRE = 4000.0, GAS_TEM = 2000.0, GAS_TE = 0.0001;
static const double GAS_ED = 0.0, GAS_FUL = 0.1, GAS_PRO = 0.001;

/* Mesh.hpp:20-28 */
static const int CUBE_FACE_VERTEX_MAP[6][4] = {
    {0, 1, 2, 3}, {4, 5, 6, 7}, {0, 2, 4, 6}, {1, 3, 5, 7}, {0, 1, 4, 5}, {2, 3, 6, 7}};

/* ======================================================================
 * A1: mesh generator at one flow rank (src/examples/mesh_examples.cpp)
 * ====================================================================== */
int mco_mesh_build(int64_t m, mco_mesh *out)
{
    memset(out, 0, sizeof(*out));
    /* src/minicombust.cpp:123-124 */
    const double mdx = 0.05 * (double)m * 2, mdy = 0.05 * (double)m, mdz = 0.05 * (double)m;
    const uint64_t ex = 2 * m, ey = m, ez = m;
    /* mesh_examples.cpp:125-126 */
    const uint64_t px = ex + 1, py = ey + 1, pz = ez + 1;
    const double edx = mdx / (double)ex, edy = mdy / (double)ey, edz = mdz / (double)ez;
    const uint64_t N = ex * ey * ez, NP = px * py * pz;
    out->m = m; out->nx = ex; out->ny = ey; out->nz = ez;
    out->n_cells = N; out->n_points = NP;

    out->points = calloc(NP, sizeof(mco_vec));
    out->cells_per_point = calloc(NP, 1);
    /* mesh_examples.cpp:286-302 */
    for (uint64_t index = 0; index < NP; index++) {
        const uint64_t x = index % px, y = (index / px) % py, z = index / (px * py);
        out->points[index].x = (double)x * edx;
        out->points[index].y = (double)y * edy;
        out->points[index].z = (double)z * edz;
        const int first_dims = (x == 0) + (y == 0) + (z == 0);
        const int last_dims = (x == px - 1) + (y == py - 1) + (z == pz - 1);
        out->cells_per_point[index] = (uint8_t)pow(2, 3 - first_dims - last_dims);
    }

    out->cells = calloc(8 * N, sizeof(uint64_t));
    out->cell_neighbours = calloc(6 * N, sizeof(uint64_t));
    /* mesh_examples.cpp:305-384 with one block; fill_neighbours :13-118 */
    for (uint64_t z = 0; z < ez; z++)
        for (uint64_t y = 0; y < ey; y++)
            for (uint64_t x = 0; x < ex; x++) {
                const uint64_t c = z * ex * ey + y * ex + x;
                const uint64_t gp = x + y * px + z * px * py;
                uint64_t *nb = &out->cell_neighbours[6 * c];
                nb[0] = (z == 0) ? MESH_BOUNDARY : c - ex * ey;      /* FRONT */
                nb[1] = (z == ez - 1) ? MESH_BOUNDARY : c + ex * ey; /* BACK  */
                nb[2] = (x == 0) ? MESH_BOUNDARY : c - 1;            /* LEFT  */
                nb[3] = (x == ex - 1) ? MESH_BOUNDARY : c + 1;       /* RIGHT */
                nb[4] = (y == 0) ? MESH_BOUNDARY : c - ex;           /* DOWN  */
                nb[5] = (y == ey - 1) ? MESH_BOUNDARY : c + ex;      /* UP    */
                uint64_t *cv = &out->cells[8 * c];
                cv[0] = gp;                       /* A */
                cv[7] = gp + px * py + px + 1;    /* H */
                cv[1] = gp + 1;                   /* B */
                cv[2] = gp + px;                  /* C */
                cv[3] = gp + px + 1;              /* D */
                cv[4] = gp + px * py;             /* E */
                cv[5] = gp + px * py + 1;         /* F */
                cv[6] = gp + px * py + px;        /* G */
            }

    /* mesh_examples.cpp:469-474 */
    out->boundary_types[0] = MCO_WALL;  out->boundary_types[1] = MCO_WALL;
    out->boundary_types[2] = MCO_INLET; out->boundary_types[3] = MCO_OUTLET;
    out->boundary_types[4] = MCO_WALL;  out->boundary_types[5] = MCO_WALL;

    /* mesh_examples.cpp:486-564 */
    const uint64_t F = (ex + 1) * ey * ez + (ey + 1) * ex * ez + (ez + 1) * ex * ey;
    out->n_faces = F;
    out->faces = calloc(2 * F, sizeof(uint64_t));
    out->cell_faces = calloc(6 * N, sizeof(uint64_t));
    uint64_t *face_indexes = malloc(6 * N * sizeof(uint64_t));
    for (uint64_t i = 0; i < 6 * N; i++) face_indexes[i] = UINT64_MAX;
    uint64_t face_count = 0;
    for (uint64_t cell = 0; cell < N; cell++) {
        for (uint64_t f_id = 0; f_id < 6; f_id++) {
            uint64_t nbc = out->cell_neighbours[cell * 6 + f_id];
            const uint64_t nb_block = nbc; /* cell_block_disp = 0 */
            if (face_indexes[cell * 6 + f_id] == UINT64_MAX) {
                if (nbc == MESH_BOUNDARY) nbc = N + f_id;
                out->faces[2 * face_count + 0] = cell < nbc ? cell : nbc;
                out->faces[2 * face_count + 1] = cell < nbc ? nbc : cell;
                if (nb_block < N) face_indexes[nb_block * 6 + (f_id ^ 1)] = face_count;
                out->cell_faces[cell * 6 + f_id] = face_count;
                face_count++;
            } else {
                out->cell_faces[cell * 6 + f_id] = face_indexes[cell * 6 + f_id];
            }
        }
    }
    free(face_indexes);
    if (face_count != F) { fprintf(stderr, "mco_mesh_build: face count %lu != %lu\n", (unsigned long)face_count, (unsigned long)F); return 1; }

    /* Mesh.hpp:54-67 */
    out->cell_centers = calloc(N, sizeof(mco_vec));
    for (uint64_t c = 0; c < N; c++) {
        mco_vec s = {0.0, 0.0, 0.0};
        for (int i = 0; i < 8; i++) s = vadd(s, out->points[out->cells[c * 8 + i]]);
        out->cell_centers[c] = vdiv(s, 8.0);
    }
    return 0;
}

void mco_mesh_free(mco_mesh *mesh)
{
    free(mesh->points); free(mesh->cells); free(mesh->cell_neighbours);
    free(mesh->faces); free(mesh->cell_faces); free(mesh->cells_per_point);
    free(mesh->cell_centers);
    memset(mesh, 0, sizeof(*mesh));
}

/* ======================================================================
 * Linear algebra (stand-ins for PETSc KSP; see header)
 * ====================================================================== */
void mco_spmv(int64_t n, const int64_t *rp, const int64_t *col, const double *v, const double *x, double *y)
{
    for (int64_t i = 0; i < n; i++) {
        double s = 0.0;
        for (int64_t k = rp[i]; k < rp[i + 1]; k++) s += v[k] * x[col[k]];
        y[i] = s;
    }
}

static double dot(int64_t n, const double *a, const double *b)
{
    double s = 0.0;
    for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/* Unpreconditioned BiCGStab (van der Vorst 1992), x0 = 0 as set by
   VecZeroEntries(u) before KSPSolve (FlowSolver.inl:1348). Converged when
   ||r|| <= rtol ||b||, PETSc's default test with x0 = 0.
   warm != 0 (not in the reference; the library's simple-mode warm start):
   x holds x0, r0 = b - A x0, and a guess whose residual is not below ||b||
   is dropped for x0 = 0; the shadow residual is b and the test
   ||r|| <= rtol ||b|| either way. */
int mco_bicgstab_x0(int64_t n, const int64_t *rp, const int64_t *col, const double *A,
                    const double *b, double *x, double rtol, int maxit, double *relres, int warm)
{
    double *r = malloc(n * sizeof(double)), *rh = malloc(n * sizeof(double));
    double *p = calloc(n, sizeof(double)), *v = calloc(n, sizeof(double));
    double *s = malloc(n * sizeof(double)), *t = malloc(n * sizeof(double));
    const double bnorm = sqrt(dot(n, b, b));
    const double tol = rtol * bnorm;
    double rn = bnorm;
    int zero_start = 1;
    if (warm && bnorm > 0.0) {
        mco_spmv(n, rp, col, A, x, t);
        for (int64_t i = 0; i < n; i++) r[i] = b[i] - t[i];
        rn = sqrt(dot(n, r, r));
        zero_start = !(rn < bnorm);
        if (zero_start) rn = bnorm;
    }
    if (zero_start)
        for (int64_t i = 0; i < n; i++) { x[i] = 0.0; r[i] = b[i]; }
    for (int64_t i = 0; i < n; i++) rh[i] = b[i];
    double rho_old = 1.0, alpha = 1.0, omega = 1.0;
    int it = 0;
    if (bnorm == 0.0) { if (relres) *relres = 0.0; goto done; }
    if (rn <= tol) { if (relres) *relres = rn / bnorm; goto done; }
    for (it = 1; it <= maxit; it++) {
        const double rho = dot(n, rh, r);
        if (rho == 0.0) break;
        const double beta = (rho / rho_old) * (alpha / omega);
        for (int64_t i = 0; i < n; i++) p[i] = r[i] + beta * (p[i] - omega * v[i]);
        mco_spmv(n, rp, col, A, p, v);
        alpha = rho / dot(n, rh, v);
        for (int64_t i = 0; i < n; i++) s[i] = r[i] - alpha * v[i];
        const double sn = sqrt(dot(n, s, s));
        if (sn <= tol) {
            for (int64_t i = 0; i < n; i++) x[i] += alpha * p[i];
            rn = sn;
            break;
        }
        mco_spmv(n, rp, col, A, s, t);
        omega = dot(n, t, s) / dot(n, t, t);
        for (int64_t i = 0; i < n; i++) { x[i] += alpha * p[i] + omega * s[i]; r[i] = s[i] - omega * t[i]; }
        rn = sqrt(dot(n, r, r));
        rho_old = rho;
        if (rn <= tol) break;
    }
    if (relres) *relres = rn / bnorm;
done:
    free(r); free(rh); free(p); free(v); free(s); free(t);
    return it;
}

int mco_bicgstab(int64_t n, const int64_t *rp, const int64_t *col, const double *A,
                 const double *b, double *x, double rtol, int maxit, double *relres)
{
    return mco_bicgstab_x0(n, rp, col, A, b, x, rtol, maxit, relres, 0);
}

/* Jacobi-preconditioned CG for the singular, consistent pressure-correction
   system (pure Neumann Laplacian, constant null space).  The RHS is projected
   to mean zero and the returned solution has mean zero.  Converged when
   ||b' - A x|| <= rtol ||b'||. */
int mco_pcg_singular(int64_t n, const int64_t *rp, const int64_t *col, const double *A,
                     const double *bin, double *x, double rtol, double atol, int maxit, double *relres,
                     double *bnorm_out)
{
    double *b = malloc(n * sizeof(double)), *r = malloc(n * sizeof(double));
    double *z = malloc(n * sizeof(double)), *p = malloc(n * sizeof(double));
    double *q = malloc(n * sizeof(double)), *dinv = malloc(n * sizeof(double));
    double mean = 0.0;
    for (int64_t i = 0; i < n; i++) mean += bin[i];
    mean /= (double)n;
    for (int64_t i = 0; i < n; i++) b[i] = bin[i] - mean;
    for (int64_t i = 0; i < n; i++) {
        double d = 0.0;
        for (int64_t k = rp[i]; k < rp[i + 1]; k++) if (col[k] == i) d = A[k];
        dinv[i] = (d != 0.0) ? 1.0 / d : 0.0;
    }
    const double bnorm = sqrt(dot(n, b, b));
    const double tol = (rtol * bnorm > atol) ? rtol * bnorm : atol;
    if (bnorm_out) *bnorm_out = bnorm;
    for (int64_t i = 0; i < n; i++) { x[i] = 0.0; r[i] = b[i]; z[i] = dinv[i] * r[i]; p[i] = z[i]; }
    double rz = dot(n, r, z), rn = bnorm;
    int it = 0;
    if (bnorm > tol) {
        for (it = 1; it <= maxit; it++) {
            mco_spmv(n, rp, col, A, p, q);
            const double alpha = rz / dot(n, p, q);
            for (int64_t i = 0; i < n; i++) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
            rn = sqrt(dot(n, r, r));
            if (rn <= tol) break;
            for (int64_t i = 0; i < n; i++) z[i] = dinv[i] * r[i];
            const double rz_new = dot(n, r, z);
            const double beta = rz_new / rz;
            rz = rz_new;
            for (int64_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
        }
    }
    double xm = 0.0;
    for (int64_t i = 0; i < n; i++) xm += x[i];
    xm /= (double)n;
    for (int64_t i = 0; i < n; i++) x[i] -= xm;
    if (relres) *relres = (bnorm > 0.0) ? rn / bnorm : 0.0;
    free(b); free(r); free(z); free(p); free(q); free(dinv);
    return it;
}

/* ----------------------------------------------------------------------
 * Aggregation-AMG preconditioned CG (CPU).  The reference preconditions its
 * pressure solve with PETSc GAMG (FlowSolver.hpp:521-527); this plain-C
 * unsmoothed-aggregation V-cycle gives the CPU path a multigrid solve of the
 * same class: three passes of greedy pairwise matching per level (strongest
 * unmatched neighbour), damped Jacobi (2 sweeps before and after), scaled
 * coarse correction, dense Cholesky of the regularised coarsest operator.
 * ---------------------------------------------------------------------- */
typedef struct {
    int64_t n, nnz;
    int64_t *rp, *col;
    double *val, *dinv;
    int64_t nc;
    int64_t *agg;           /* aggregate of each row */
    double *x, *b, *t, *r;
} amg_level;

typedef struct {
    int nlev;
    amg_level lev[32];
    double *chol;           /* coarsest: Cholesky factor (n x n, lower) */
} amg_h;

static amg_h *g_amg = NULL;
static const int64_t *g_amg_rp = NULL;

static void csr_free(amg_level *L)
{
    free(L->rp); free(L->col); free(L->val); free(L->dinv); free(L->agg);
    free(L->x); free(L->b); free(L->t); free(L->r);
}

static void pairwise(int64_t n, const int64_t *rp, const int64_t *col, const double *val, int64_t *agg, int64_t *nc)
{
    for (int64_t i = 0; i < n; i++) agg[i] = -1;
    int64_t k = 0;
    for (int64_t i = 0; i < n; i++) {
        if (agg[i] >= 0) continue;
        int64_t best = -1;
        double bw = 0.0;
        for (int64_t e = rp[i]; e < rp[i + 1]; e++) {
            const int64_t j = col[e];
            if (j == i || agg[j] >= 0) continue;
            const double w = -val[e];
            if (w > bw * (1.0 + 1e-12)) { best = j; bw = w; }
        }
        agg[i] = k;
        if (best >= 0) agg[best] = k;
        k++;
    }
    *nc = k;
}

/* coarse CSR of P^T A P for piecewise-constant P (columns sorted) */
static void galerkin_c(int64_t n, const int64_t *rp, const int64_t *col, const double *val, const int64_t *agg,
                       int64_t nc, int64_t **rpc, int64_t **colc, double **valc, int64_t *nnzc)
{
    int64_t *cnt = calloc(nc + 1, sizeof(int64_t)), *mem = malloc(n * sizeof(int64_t));
    for (int64_t i = 0; i < n; i++) cnt[agg[i] + 1]++;
    for (int64_t I = 0; I < nc; I++) cnt[I + 1] += cnt[I];
    int64_t *pos = malloc(nc * sizeof(int64_t));
    for (int64_t I = 0; I < nc; I++) pos[I] = cnt[I];
    for (int64_t i = 0; i < n; i++) mem[pos[agg[i]]++] = i;
    int64_t cap = rp[n] + 16, nz = 0;
    int64_t *rc = malloc((nc + 1) * sizeof(int64_t)), *cc = malloc(cap * sizeof(int64_t));
    double *vc = malloc(cap * sizeof(double));
    int64_t *mark = malloc(nc * sizeof(int64_t));
    for (int64_t I = 0; I < nc; I++) mark[I] = -1;
    rc[0] = 0;
    for (int64_t I = 0; I < nc; I++) {
        const int64_t start = nz;
        for (int64_t t = cnt[I]; t < cnt[I + 1]; t++) {
            const int64_t i = mem[t];
            for (int64_t e = rp[i]; e < rp[i + 1]; e++) {
                const int64_t J = agg[col[e]];
                if (mark[J] < start) { mark[J] = nz; cc[nz] = J; vc[nz] = 0.0; nz++; }
                vc[mark[J]] += val[e];
            }
        }
        /* sort the row by column */
        for (int64_t a = start + 1; a < nz; a++) {
            const int64_t cj = cc[a]; const double cv = vc[a];
            int64_t b = a - 1;
            while (b >= start && cc[b] > cj) { cc[b + 1] = cc[b]; vc[b + 1] = vc[b]; b--; }
            cc[b + 1] = cj; vc[b + 1] = cv;
        }
        for (int64_t a = start; a < nz; a++) mark[cc[a]] = a;
        rc[I + 1] = nz;
    }
    free(cnt); free(mem); free(pos); free(mark);
    *rpc = rc; *colc = cc; *valc = vc; *nnzc = nz;
}

static void amg_level_vectors(amg_level *L)
{
    L->dinv = malloc(L->n * sizeof(double));
    for (int64_t i = 0; i < L->n; i++) {
        double d = 0.0;
        for (int64_t e = L->rp[i]; e < L->rp[i + 1]; e++) if (L->col[e] == i) d = L->val[e];
        L->dinv[i] = (d != 0.0) ? 1.0 / d : 0.0;
    }
    L->x = calloc(L->n, sizeof(double)); L->b = calloc(L->n, sizeof(double));
    L->t = calloc(L->n, sizeof(double)); L->r = calloc(L->n, sizeof(double));
}

static amg_h *amg_build(int64_t n, const int64_t *rp, const int64_t *col, const double *val)
{
    amg_h *h = calloc(1, sizeof(amg_h));
    amg_level *L = &h->lev[0];
    L->n = n; L->nnz = rp[n];
    L->rp = malloc((n + 1) * sizeof(int64_t)); memcpy(L->rp, rp, (n + 1) * sizeof(int64_t));
    L->col = malloc(L->nnz * sizeof(int64_t)); memcpy(L->col, col, L->nnz * sizeof(int64_t));
    L->val = malloc(L->nnz * sizeof(double)); memcpy(L->val, val, L->nnz * sizeof(double));
    amg_level_vectors(L);
    int l = 0;
    while (h->lev[l].n > 256 && l < 30) {
        amg_level *F = &h->lev[l];
        int64_t *agg = malloc(F->n * sizeof(int64_t));
        for (int64_t i = 0; i < F->n; i++) agg[i] = i;
        int64_t gn = F->n, *grp = F->rp, *gcol = F->col, gnnz = F->nnz, nc = F->n;
        double *gval = F->val;
        int own = 0;
        for (int pass = 0; pass < 3; pass++) {
            int64_t *a1 = malloc(gn * sizeof(int64_t)), ncn = 0;
            pairwise(gn, grp, gcol, gval, a1, &ncn);
            for (int64_t i = 0; i < F->n; i++) agg[i] = a1[agg[i]];
            int64_t *r2, *c2; double *v2;
            galerkin_c(gn, grp, gcol, gval, a1, ncn, &r2, &c2, &v2, &gnnz);
            if (own) { free(grp); free(gcol); free(gval); }
            grp = r2; gcol = c2; gval = v2; gn = ncn; own = 1; nc = ncn;
            free(a1);
            if (nc <= 256) break;
        }
        F->agg = agg; F->nc = nc;
        amg_level *C = &h->lev[l + 1];
        C->n = nc; C->nnz = gnnz; C->rp = grp; C->col = gcol; C->val = gval;
        amg_level_vectors(C);
        l++;
    }
    h->nlev = l + 1;
    /* coarsest: Cholesky of A + alpha 1 1^T, alpha = mean diagonal / n */
    amg_level *C = &h->lev[l];
    const int64_t nn = C->n;
    double *M = calloc(nn * nn, sizeof(double));
    double dsum = 0.0;
    for (int64_t i = 0; i < nn; i++)
        for (int64_t e = C->rp[i]; e < C->rp[i + 1]; e++) if (C->col[e] == i) dsum += C->val[e];
    const double alpha = (dsum / nn) / nn;
    for (int64_t i = 0; i < nn * nn; i++) M[i] = alpha;
    for (int64_t i = 0; i < nn; i++)
        for (int64_t e = C->rp[i]; e < C->rp[i + 1]; e++) M[i * nn + C->col[e]] += C->val[e];
    for (int64_t j = 0; j < nn; j++) {
        double s = M[j * nn + j];
        for (int64_t k = 0; k < j; k++) s -= M[j * nn + k] * M[j * nn + k];
        const double d = sqrt(s);
        M[j * nn + j] = d;
        for (int64_t i = j + 1; i < nn; i++) {
            double t = M[i * nn + j];
            for (int64_t k = 0; k < j; k++) t -= M[i * nn + k] * M[j * nn + k];
            M[i * nn + j] = t / d;
        }
    }
    h->chol = M;
    return h;
}

static void amg_free(amg_h *h)
{
    if (!h) return;
    for (int l = 0; l < h->nlev; l++) csr_free(&h->lev[l]);
    free(h->chol);
    free(h);
}

/* values of every level from the fine values (pattern unchanged) */
static void amg_refresh(amg_h *h, const double *val)
{
    memcpy(h->lev[0].val, val, h->lev[0].nnz * sizeof(double));
    for (int l = 0; l + 1 < h->nlev; l++) {
        amg_level *F = &h->lev[l], *C = &h->lev[l + 1];
        int64_t *r2, *c2, nz; double *v2;
        galerkin_c(F->n, F->rp, F->col, F->val, F->agg, F->nc, &r2, &c2, &v2, &nz);
        memcpy(C->val, v2, nz * sizeof(double));
        free(r2); free(c2); free(v2);
    }
    for (int l = 0; l < h->nlev; l++) {
        amg_level *L = &h->lev[l];
        for (int64_t i = 0; i < L->n; i++) {
            double d = 0.0;
            for (int64_t e = L->rp[i]; e < L->rp[i + 1]; e++) if (L->col[e] == i) d = L->val[e];
            L->dinv[i] = (d != 0.0) ? 1.0 / d : 0.0;
        }
    }
    /* refactor the coarsest */
    amg_level *C = &h->lev[h->nlev - 1];
    const int64_t nn = C->n;
    double *M = h->chol;
    double dsum = 0.0;
    for (int64_t i = 0; i < nn; i++)
        for (int64_t e = C->rp[i]; e < C->rp[i + 1]; e++) if (C->col[e] == i) dsum += C->val[e];
    const double alpha = (dsum / nn) / nn;
    for (int64_t i = 0; i < nn * nn; i++) M[i] = alpha;
    for (int64_t i = 0; i < nn; i++)
        for (int64_t e = C->rp[i]; e < C->rp[i + 1]; e++) M[i * nn + C->col[e]] += C->val[e];
    for (int64_t j = 0; j < nn; j++) {
        double s = M[j * nn + j];
        for (int64_t k = 0; k < j; k++) s -= M[j * nn + k] * M[j * nn + k];
        const double d = sqrt(s);
        M[j * nn + j] = d;
        for (int64_t i = j + 1; i < nn; i++) {
            double t = M[i * nn + j];
            for (int64_t k = 0; k < j; k++) t -= M[i * nn + k] * M[j * nn + k];
            M[i * nn + j] = t / d;
        }
    }
}

static void amg_vcycle(amg_h *h, int l, const double *b, double *x)
{
    amg_level *L = &h->lev[l];
    const int64_t n = L->n;
    const double om = 0.8, cs = 1.8;
    if (l == h->nlev - 1) {
        const double *M = h->chol;
        for (int64_t i = 0; i < n; i++) {
            double s = b[i];
            for (int64_t k = 0; k < i; k++) s -= M[i * n + k] * x[k];
            x[i] = s / M[i * n + i];
        }
        for (int64_t i = n - 1; i >= 0; i--) {
            double s = x[i];
            for (int64_t k = i + 1; k < n; k++) s -= M[k * n + i] * x[k];
            x[i] = s / M[i * n + i];
        }
        return;
    }
    for (int64_t i = 0; i < n; i++) x[i] = om * L->dinv[i] * b[i];
    for (int sweep = 1; sweep < 2; sweep++) {
        mco_spmv(n, L->rp, L->col, L->val, x, L->t);
        for (int64_t i = 0; i < n; i++) x[i] += om * L->dinv[i] * (b[i] - L->t[i]);
    }
    mco_spmv(n, L->rp, L->col, L->val, x, L->t);
    amg_level *C = &h->lev[l + 1];
    for (int64_t I = 0; I < C->n; I++) C->b[I] = 0.0;
    for (int64_t i = 0; i < n; i++) C->b[L->agg[i]] += b[i] - L->t[i];
    amg_vcycle(h, l + 1, C->b, C->x);
    for (int64_t i = 0; i < n; i++) x[i] += cs * C->x[L->agg[i]];
    for (int sweep = 0; sweep < 2; sweep++) {
        mco_spmv(n, L->rp, L->col, L->val, x, L->t);
        for (int64_t i = 0; i < n; i++) x[i] += om * L->dinv[i] * (b[i] - L->t[i]);
    }
}

int mco_pcg_amg(int64_t n, const int64_t *rp, const int64_t *col, const double *A, const double *bin, double *x,
                double rtol, double atol, int maxit, double *relres, double *bnorm_out)
{
    if (!g_amg || g_amg_rp != rp || g_amg->lev[0].n != n) {
        amg_free(g_amg);
        g_amg = amg_build(n, rp, col, A);
        g_amg_rp = rp;
    } else {
        amg_refresh(g_amg, A);
    }
    double *b = malloc(n * sizeof(double)), *r = malloc(n * sizeof(double));
    double *z = malloc(n * sizeof(double)), *p = malloc(n * sizeof(double)), *q = malloc(n * sizeof(double));
    double mean = 0.0;
    for (int64_t i = 0; i < n; i++) mean += bin[i];
    mean /= (double)n;
    for (int64_t i = 0; i < n; i++) b[i] = bin[i] - mean;
    const double bnorm = sqrt(dot(n, b, b));
    const double tol = (rtol * bnorm > atol) ? rtol * bnorm : atol;
    if (bnorm_out) *bnorm_out = bnorm;
    for (int64_t i = 0; i < n; i++) { x[i] = 0.0; r[i] = b[i]; p[i] = 0.0; }
    double rn = bnorm, rz_old = 0.0;
    int it = 0;
    if (bnorm > tol) {
        for (it = 1; it <= maxit; it++) {
            amg_vcycle(g_amg, 0, r, z);
            const double rz = dot(n, r, z);
            const double beta = (it == 1) ? 0.0 : rz / rz_old;
            rz_old = rz;
            for (int64_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
            mco_spmv(n, rp, col, A, p, q);
            const double alpha = rz / dot(n, p, q);
            for (int64_t i = 0; i < n; i++) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
            rn = sqrt(dot(n, r, r));
            if (rn <= tol) break;
        }
    }
    double xm = 0.0;
    for (int64_t i = 0; i < n; i++) xm += x[i];
    xm /= (double)n;
    for (int64_t i = 0; i < n; i++) x[i] -= xm;
    if (relres) *relres = (bnorm > 0.0) ? rn / bnorm : 0.0;
    free(b); free(r); free(z); free(p); free(q);
    return it;
}

/* ======================================================================
 * Constructor (include/flow/FlowSolver.hpp:147-539)
 * ====================================================================== */
void mco_default_opts(mco_opts *o)
{
    o->mode = MCO_MODE_COMPAT;
    o->delta = 1.0e-5;
    o->mom_rtol = 1.0e-5;   /* PETSc default rtol */
    o->mom_maxit = 10000;   /* PETSc default max_it */
    o->p_rtol = 1.0e-8;
    o->p_maxit = 100000;
    o->solve_scalars = 1;
    o->p_solver = 0;
    o->limit_grad = 0;
    o->warm_start = 0;
}

static void setup_sparse_matrix(mco_state *s, double URFactor, double *A_c, double *phi_c, double *S_c);

mco_state *mco_create(int64_t m, const mco_opts *opts)
{
    mco_state *s = calloc(1, sizeof(mco_state));
    if (mco_mesh_build(m, &s->mesh)) { free(s); return NULL; }
    if (opts) s->opts = *opts; else mco_default_opts(&s->opts);
    mco_mesh *M = &s->mesh;
    const int64_t N = M->n_cells, F = M->n_faces, NS = N + 6;
    s->N = N; s->F = F;
    s->face_areas = calloc(F, sizeof(double));
    s->face_lambdas = calloc(F, sizeof(double));
    s->face_rlencos = calloc(F, sizeof(double));
    s->face_normals = calloc(F, sizeof(mco_vec));
    s->face_centers = calloc(F, sizeof(mco_vec));
    s->cell_volumes = calloc(N, sizeof(double));
    s->cell_densities = calloc(N, sizeof(double));
    for (int f = 0; f < MCO_NFIELDS; f++) {
        s->phi[f] = calloc(NS, sizeof(double));
        s->grad[f] = calloc(NS, sizeof(mco_vec));
    }
    s->A_U = calloc(NS, sizeof(double)); s->A_V = calloc(NS, sizeof(double)); s->A_W = calloc(NS, sizeof(double));
    s->S_U = calloc(NS, sizeof(double)); s->S_V = calloc(NS, sizeof(double)); s->S_W = calloc(NS, sizeof(double));
    s->ff0 = calloc(F, sizeof(double)); s->ff1 = calloc(F, sizeof(double));
    s->mflux = calloc(F, sizeof(double));

    /* interior face geometry, FlowSolver.hpp:252-294 */
    for (int64_t face = 0; face < F; face++) {
        const uint64_t c0 = M->faces[2 * face], c1 = M->faces[2 * face + 1];
        if (c1 >= (uint64_t)N) continue;
        uint64_t ids[8];
        int nc = 0;
        for (int n0 = 0; n0 < 8; n0++)
            for (int n1 = 0; n1 < 8; n1++)
                if (M->cells[c0 * 8 + n0] == M->cells[c1 * 8 + n1]) ids[nc++] = M->cells[c0 * 8 + n0];
        const mco_vec fn0 = M->points[ids[0]], fn1 = M->points[ids[1]], fn2 = M->points[ids[2]], fn3 = M->points[ids[3]];
        const mco_vec c01 = vsub(M->cell_centers[c1], M->cell_centers[c0]);
        s->face_areas[face] = vmag(vsub(fn2, fn0)) * vmag(vsub(fn1, fn0));
        s->face_centers[face] = vadd(vadd(vadd(fn0, fn1), fn2), fn3);
        s->face_centers[face] = vdiv(s->face_centers[face], 4.0);
        s->face_lambdas[face] = vmag(vsub(s->face_centers[face], M->cell_centers[c0])) / vmag(c01);
        s->face_normals[face] = vcross(vsub(fn2, fn0), vsub(fn1, fn0));
        if (vdot(s->face_normals[face], vsub(M->cell_centers[c1], M->cell_centers[c0])) < 0)
            s->face_normals[face] = vmul(s->face_normals[face], -1.);
        s->face_rlencos[face] = s->face_areas[face] / vmag(c01) / vcosangle(s->face_normals[face], c01);
        s->mflux[face] = 0.0;
    }
    s->effective_viscosity = 0.000014; /* FlowSolver.hpp:296-297 */

    /* cell data and boundary face geometry, FlowSolver.hpp:301-398 */
    for (int64_t c = 0; c < N; c++) {
        const uint64_t *cn = &M->cells[c * 8];
        s->phi[MCO_U][c] = GAS_VEL_X; s->phi[MCO_V][c] = GAS_VEL_Y; s->phi[MCO_W][c] = GAS_VEL_Z;
        s->phi[MCO_P][c] = GAS_PRE; s->phi[MCO_PP][c] = 0.0; s->phi[MCO_TE][c] = GAS_TE;
        s->phi[MCO_ED][c] = GAS_ED; s->phi[MCO_TP][c] = 0.0; s->phi[MCO_TEM][c] = GAS_TEM;
        s->phi[MCO_FUL][c] = GAS_FUL; s->phi[MCO_PRO][c] = GAS_PRO; s->phi[MCO_VARF][c] = GAS_FUL;
        s->phi[MCO_VARP][c] = GAS_PRO;
        s->cell_densities[c] = 1.2;
        const double width = vmag(vsub(M->points[cn[1]], M->points[cn[0]]));
        const double height = vmag(vsub(M->points[cn[2]], M->points[cn[0]]));
        const double length = vmag(vsub(M->points[cn[4]], M->points[cn[0]]));
        s->cell_volumes[c] = width * height * length;
        for (int f = 0; f < 6; f++) {
            if (M->cell_neighbours[c * 6 + f] < (uint64_t)N) continue;
            const uint64_t face = M->cell_faces[c * 6 + f];
            const mco_vec fn0 = M->points[cn[CUBE_FACE_VERTEX_MAP[f][0]]], fn1 = M->points[cn[CUBE_FACE_VERTEX_MAP[f][1]]];
            const mco_vec fn2 = M->points[cn[CUBE_FACE_VERTEX_MAP[f][2]]], fn3 = M->points[cn[CUBE_FACE_VERTEX_MAP[f][3]]];
            s->face_lambdas[face] = 1.0;
            s->face_areas[face] = vmag(vsub(fn2, fn0)) * vmag(vsub(fn1, fn0));
            s->face_centers[face] = vdiv(vadd(vadd(vadd(fn0, fn1), fn2), fn3), 4.0);
            s->face_normals[face] = vcross(vsub(fn2, fn0), vsub(fn1, fn0));
            const uint64_t bc = M->faces[2 * face + 1] - N;
            const uint64_t bt = M->boundary_types[bc];
            if (bt == MCO_INLET || bt == MCO_OUTLET || bt == MCO_WALL) {
                s->mflux[face] = 0.0;
                if (vdot(s->face_normals[face], vsub(M->cell_centers[c], s->face_centers[face])) > 0)
                    s->face_normals[face] = vmul(s->face_normals[face], -1.);
            }
            const mco_vec cf = vsub(s->face_centers[face], M->cell_centers[c]);
            s->face_rlencos[face] = s->face_areas[face] / vmag(cf) / vcosangle(s->face_normals[face], cf);
        }
    }

    /* boundary slots, FlowSolver.hpp:402-485 */
    for (int bc = 0; bc < 6; bc++) {
        const int64_t i = N + bc;
        const double velmag2 = pow(GAS_VEL_X, 2) + pow(GAS_VEL_Y, 2) + pow(GAS_VEL_Z, 2);
        s->inlet_effective_viscosity = s->effective_viscosity + 1.2 * 0.09 * (3.0 / 2.0 * ((0.1 * 0.1) * velmag2)) /
            ((pow(0.09, 0.75) * pow((3.0 / 2.0 * ((0.1 * 0.1) * velmag2)), 1.5)) + 0.00000000000000000001);
        double *p = NULL;
        (void)p;
        const uint64_t bt = M->boundary_types[bc];
        s->phi[MCO_U][i] = 0.0; s->phi[MCO_V][i] = 0.0; s->phi[MCO_W][i] = 0.0; s->phi[MCO_P][i] = GAS_PRE;
        s->phi[MCO_PP][i] = 0.0; s->phi[MCO_TE][i] = GAS_TE; s->phi[MCO_ED][i] = GAS_ED; s->phi[MCO_TP][i] = 0.0;
        s->phi[MCO_TEM][i] = GAS_TEM; s->phi[MCO_FUL][i] = GAS_FUL; s->phi[MCO_PRO][i] = GAS_PRO;
        s->phi[MCO_VARF][i] = GAS_FUL; s->phi[MCO_VARP][i] = GAS_PRO;
        if (bt == MCO_INLET) {
            s->phi[MCO_U][i] = GAS_VEL_X; s->phi[MCO_V][i] = 0.0; s->phi[MCO_W][i] = 0.0; s->phi[MCO_P][i] = GAS_PRE;
            s->phi[MCO_PP][i] = 0.0;
            s->phi[MCO_TE][i] = 3.0 / 2.0 * ((0.1 * 0.1) * velmag2);
            s->phi[MCO_ED][i] = pow(0.09, 0.75) * pow(s->phi[MCO_TE][i], 1.5);
            s->phi[MCO_TP][i] = 0.0; s->phi[MCO_TEM][i] = GAS_TEM; s->phi[MCO_FUL][i] = GAS_FUL;
            s->phi[MCO_PRO][i] = 0.0; s->phi[MCO_VARF][i] = GAS_FUL; s->phi[MCO_VARP][i] = 0.0;
        } else if (bt == MCO_OUTLET) {
            s->phi[MCO_U][i] = GAS_VEL_X; s->phi[MCO_V][i] = 0.0; s->phi[MCO_W][i] = 0.0; s->phi[MCO_P][i] = GAS_PRE;
            s->phi[MCO_PP][i] = 0.0; s->phi[MCO_TE][i] = GAS_TE; s->phi[MCO_ED][i] = GAS_ED; s->phi[MCO_TP][i] = 0.0;
            s->phi[MCO_TEM][i] = GAS_TEM; s->phi[MCO_FUL][i] = GAS_FUL; s->phi[MCO_PRO][i] = GAS_PRO;
            s->phi[MCO_VARF][i] = GAS_FUL; s->phi[MCO_VARP][i] = GAS_PRO;
        } else if (bt == MCO_WALL) {
            s->phi[MCO_U][i] = 0.0; s->phi[MCO_V][i] = 0.0; s->phi[MCO_W][i] = 0.0; s->phi[MCO_P][i] = GAS_PRE;
            s->phi[MCO_PP][i] = 0.0; s->phi[MCO_TE][i] = GAS_TE; s->phi[MCO_ED][i] = GAS_ED; s->phi[MCO_TP][i] = 0.0;
            s->phi[MCO_TEM][i] = GAS_TEM; s->phi[MCO_FUL][i] = 0.0; s->phi[MCO_PRO][i] = 0.0;
            s->phi[MCO_VARF][i] = 0.0; s->phi[MCO_VARP][i] = 0.0;
        }
    }

    /* CSR pattern of MatSeqAIJ (FlowSolver.hpp:491-502): diag + interior
       neighbours with sorted columns. */
    s->row_ptr = calloc(N + 1, sizeof(int64_t));
    for (int64_t c = 0; c < N; c++) {
        int cnt = 1;
        for (int f = 0; f < 6; f++) if (M->cell_neighbours[c * 6 + f] < (uint64_t)N) cnt++;
        s->row_ptr[c + 1] = s->row_ptr[c] + cnt;
    }
    s->nnz = s->row_ptr[N];
    s->col = calloc(s->nnz, sizeof(int64_t));
    s->diag_idx = calloc(N, sizeof(int64_t));
    for (int64_t c = 0; c < N; c++) {
        int64_t cols[7]; int cnt = 0;
        cols[cnt++] = c;
        for (int f = 0; f < 6; f++) if (M->cell_neighbours[c * 6 + f] < (uint64_t)N) cols[cnt++] = (int64_t)M->cell_neighbours[c * 6 + f];
        for (int a = 1; a < cnt; a++) { int64_t v = cols[a]; int b = a - 1; while (b >= 0 && cols[b] > v) { cols[b + 1] = cols[b]; b--; } cols[b + 1] = v; }
        for (int k = 0; k < cnt; k++) {
            s->col[s->row_ptr[c] + k] = cols[k];
            if (cols[k] == c) s->diag_idx[c] = s->row_ptr[c] + k;
        }
    }
    s->face_idx01 = malloc(F * sizeof(int64_t));
    s->face_idx10 = malloc(F * sizeof(int64_t));
    for (int64_t face = 0; face < F; face++) {
        const uint64_t c0 = M->faces[2 * face], c1 = M->faces[2 * face + 1];
        s->face_idx01[face] = -1; s->face_idx10[face] = -1;
        if (c1 >= (uint64_t)N) continue;
        for (int64_t k = s->row_ptr[c0]; k < s->row_ptr[c0 + 1]; k++) if (s->col[k] == (int64_t)c1) s->face_idx01[face] = k;
        for (int64_t k = s->row_ptr[c1]; k < s->row_ptr[c1 + 1]; k++) if (s->col[k] == (int64_t)c0) s->face_idx10[face] = k;
    }
    s->A_vals = calloc(s->nnz, sizeof(double));
    s->PA_vals = calloc(s->nnz, sizeof(double));
    s->b = calloc(N, sizeof(double));
    s->Pb = calloc(N, sizeof(double));
    s->u = calloc(N, sizeof(double));
    s->first_pressure = 1;
    s->timestep_count = 0;

    /* FlowSolver.hpp:538-539: first assemblies (face_fields are zero here) */
    setup_sparse_matrix(s, 1.0, s->A_U, s->phi[MCO_U], s->S_U);
    mco_setup_pressure_matrix(s);
    return s;
}

void mco_destroy(mco_state *s)
{
    if (!s) return;
    mco_mesh_free(&s->mesh);
    free(s->face_areas); free(s->face_lambdas); free(s->face_rlencos); free(s->face_normals); free(s->face_centers);
    free(s->cell_volumes); free(s->cell_densities);
    for (int f = 0; f < MCO_NFIELDS; f++) { free(s->phi[f]); free(s->grad[f]); }
    free(s->A_U); free(s->A_V); free(s->A_W); free(s->S_U); free(s->S_V); free(s->S_W);
    free(s->ff0); free(s->ff1); free(s->mflux);
    free(s->row_ptr); free(s->col); free(s->diag_idx); free(s->face_idx01); free(s->face_idx10);
    free(s->A_vals); free(s->PA_vals); free(s->b); free(s->Pb); free(s->u);
    free(s);
}

double *mco_phi(mco_state *s, int field) { return s->phi[field]; }
mco_vec *mco_grad(mco_state *s, int field) { return s->grad[field]; }
int64_t mco_n_cells(mco_state *s) { return s->N; }
int64_t mco_n_faces(mco_state *s) { return s->F; }

/* ======================================================================
 * A4/A5: least-squares gradients (FlowSolver.inl:738-803, 895-1063)
 * ====================================================================== */
static void solve3(const double A[3][3], const double *b, double *out)
{
    /* FlowSolver.inl:895-915 */
    const double det = (A[0][0] * (A[1][1] * A[2][2] - A[2][1] * A[1][2])) -
                       (A[0][1] * (A[1][0] * A[2][2] - A[2][0] * A[1][2])) +
                       (A[0][2] * (A[1][0] * A[2][1] - A[2][0] * A[1][1]));
    double invdet = 0.0;
    if (det != 0.0) invdet = 1 / det;
    out[0] = (b[0] * invdet * (A[1][1] * A[2][2] - A[2][1] * A[1][2])) +
             (b[1] * invdet * (A[2][1] * A[0][2] - A[0][1] * A[2][2])) +
             (b[2] * invdet * (A[0][1] * A[1][2] - A[1][1] * A[0][2]));
    out[1] = (b[0] * invdet * (A[1][2] * A[2][0] - A[1][0] * A[2][2])) +
             (b[1] * invdet * (A[0][0] * A[2][2] - A[2][0] * A[0][2])) +
             (b[2] * invdet * (A[1][0] * A[0][2] - A[0][0] * A[1][2]));
    out[2] = (b[0] * invdet * (A[1][0] * A[2][1] - A[2][0] * A[1][1])) +
             (b[1] * invdet * (A[2][0] * A[0][1] - A[0][0] * A[2][1])) +
             (b[2] * invdet * (A[0][0] * A[1][1] - A[1][0] * A[0][1]));
}

/* Generic LSQ gradient over a list of fields.  zero_boundary[k] selects the
   pressure rule (boundary dphi = 0, FlowSolver.inl:773-776, 980). */
static void lsq_gradients(mco_state *s, int nf, const int *fields, const int *zero_boundary)
{
    const mco_mesh *M = &s->mesh;
    const int64_t N = s->N;
    for (int64_t c = 0; c < N; c++) {
        double A[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        double b[MCO_NFIELDS][3];
        memset(b, 0, sizeof(b));
        for (int f = 0; f < 6; f++) {
            const uint64_t face = M->cell_faces[c * 6 + f];
            const uint64_t c0 = M->faces[2 * face], c1 = M->faces[2 * face + 1];
            double dphi[MCO_NFIELDS];
            mco_vec dX;
            if (c1 < (uint64_t)N) {
                const double mask = (c0 == (uint64_t)c) ? 1. : -1.;
                for (int k = 0; k < nf; k++) dphi[k] = mask * (s->phi[fields[k]][c1] - s->phi[fields[k]][c0]);
                dX = vmul(vsub(M->cell_centers[c1], M->cell_centers[c0]), mask);
            } else {
                const uint64_t bc = c1 - N;
                for (int k = 0; k < nf; k++)
                    dphi[k] = zero_boundary[k] ? 0.0 : s->phi[fields[k]][N + bc] - s->phi[fields[k]][c0];
                dX = vsub(s->face_centers[face], M->cell_centers[c0]);
            }
            A[0][0] += (dX.x * dX.x); A[1][0] += (dX.x * dX.y); A[2][0] += (dX.x * dX.z);
            A[0][1] += (dX.y * dX.x); A[1][1] += (dX.y * dX.y); A[2][1] += (dX.y * dX.z);
            A[0][2] += (dX.z * dX.x); A[1][2] += (dX.z * dX.y); A[2][2] += (dX.z * dX.z);
            for (int k = 0; k < nf; k++) {
                b[k][0] += (dX.x * dphi[k]); b[k][1] += (dX.y * dphi[k]); b[k][2] += (dX.z * dphi[k]);
            }
        }
        for (int k = 0; k < nf; k++) {
            double out[3];
            solve3((const double(*)[3])A, b[k], out);
            s->grad[fields[k]][c].x = out[0]; s->grad[fields[k]][c].y = out[1]; s->grad[fields[k]][c].z = out[2];
        }
    }
}

void mco_get_phi_gradients(mco_state *s)
{
    /* FlowSolver.inl:917-1063: U V W P TE ED TEM FUL PRO VARF VARP, dP = 0 on boundaries */
    static const int fields[11] = {MCO_U, MCO_V, MCO_W, MCO_P, MCO_TE, MCO_ED, MCO_TEM, MCO_FUL, MCO_PRO, MCO_VARF, MCO_VARP};
    static const int zb[11] = {0, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0};
    lsq_gradients(s, 11, fields, zb);
}

/* limit_phi_gradient (FlowSolver.inl:804-873), Venkatakrishnan's limiter:
   phi's min and max over the cell and its six face neighbours (boundary
   slots at the boundary), then the gradient scaled by the least
   (r^2 + 2r) / (r^2 + r + 2) over the faces, r = (max or min - phi_c) over
   grad . (x_face - x_c), r = 1000 where |grad . (x_face - x_c)| < 6e-6.
   The reference's pow(r, 2) is evaluated as r * r, the correctly rounded
   square (glibc's pow is within 0.52 ulp of it). */
static void limit_phi_gradient(mco_state *s, int field)
{
    const mco_mesh *M = &s->mesh;
    const int64_t N = s->N;
    const double *phi = s->phi[field];
    mco_vec *g = s->grad[field];
    for (int64_t c = 0; c < N; c++) {
        double pmax = phi[c], pmin = phi[c];
        for (int f = 0; f < 6; f++) {
            const uint64_t face = M->cell_faces[c * 6 + f];
            const uint64_t c0 = M->faces[2 * face], c1 = M->faces[2 * face + 1];
            double v;
            if (c1 < (uint64_t)N) v = (c0 == (uint64_t)c) ? phi[c1] : phi[c0];
            else v = phi[N + (c1 - N)];
            pmax = dmax(pmax, v);
            pmin = dmin(pmin, v);
        }
        const double deltamax = pmax - phi[c], deltamin = pmin - phi[c];
        double alpha = 1.0;
        for (int f = 0; f < 6; f++) {
            const uint64_t face = M->cell_faces[c * 6 + f];
            const mco_vec ds = vsub(s->face_centers[face], M->cell_centers[c]);
            const double delta_face = g[c].x * ds.x + g[c].y * ds.y + g[c].z * ds.z;
            double r;
            if (fabs(delta_face) < 0.000006) r = 1000.0;
            else if (delta_face > 0.0) r = deltamax / delta_face;
            else r = deltamin / delta_face;
            alpha = dmin(alpha, (r * r + 2.0 * r) / (r * r + r + 2.0));
        }
        g[c].x = g[c].x * alpha; g[c].y = g[c].y * alpha; g[c].z = g[c].z * alpha;
    }
}

void mco_limit_phi_gradients(mco_state *s)
{
    /* FlowSolver.inl:875-893, the same 11 fields in the same order */
    static const int fields[11] = {MCO_U, MCO_V, MCO_W, MCO_P, MCO_TE, MCO_ED, MCO_TEM, MCO_FUL, MCO_PRO, MCO_VARF, MCO_VARP};
    for (int k = 0; k < 11; k++) limit_phi_gradient(s, fields[k]);
}

void mco_get_phi_gradient(mco_state *s, int field, int pressure)
{
    /* FlowSolver.inl:738-803 */
    int zb = pressure ? 1 : 0;
    lsq_gradients(s, 1, &field, &zb);
}

/* ======================================================================
 * A6: first-iteration momentum diagonal (FlowSolver.inl:1065-1113)
 * ====================================================================== */
void mco_precomp_AU(mco_state *s)
{
    const mco_mesh *M = &s->mesh;
    const int64_t N = s->N;
    for (int64_t i = 0; i < N; i++) s->A_U[i] = 0.0;
    for (int64_t face = 0; face < s->F; face++) {
        const uint64_t c0 = M->faces[2 * face], c1 = M->faces[2 * face + 1];
        if (c1 < (uint64_t)N) continue;
        const uint64_t bt = M->boundary_types[c1 - N];
        if (bt == MCO_INLET || bt == MCO_OUTLET) {
            const double Visac = s->effective_viscosity;
            const double VisFace = Visac * s->face_rlencos[face];
            const double f = -VisFace + dmin(s->mflux[face], 0.0);
            s->A_U[c0] = s->A_U[c0] - f;
        }
    }
    const double rdelta = 1.0 / s->opts.delta;
    for (int64_t i = 0; i < N; i++) {
        const double f = s->cell_densities[i] * s->cell_volumes[i] * rdelta;
        s->A_U[i] += f;
    }
}

void mco_set_up_field(mco_state *s)
{
    mco_precomp_AU(s);
    mco_calculate_mass_flux(s);
}

/* ======================================================================
 * A7: momentum face fluxes (FlowSolver.inl:1377-1653)
 * ====================================================================== */
void mco_calculate_flux_UVW(mco_state *s)
{
    const mco_mesh *M = &s->mesh;
    const int64_t N = s->N;
    double *U = s->phi[MCO_U], *V = s->phi[MCO_V], *W = s->phi[MCO_W];
    const mco_vec *gU = s->grad[MCO_U], *gV = s->grad[MCO_V], *gW = s->grad[MCO_W];
    const double GammaBlend = 0.0;
    for (int64_t face = 0; face < s->F; face++) {
        const uint64_t c0 = M->faces[2 * face], c1 = M->faces[2 * face + 1];
        const double mf = s->mflux[face];
        if (c1 < (uint64_t)N) {
            const double lambda0 = s->face_lambdas[face], lambda1 = 1.0 - lambda0;
            const mco_vec dUdXac = vadd(vmul(gU[c0], lambda0), vmul(gU[c1], lambda1));
            const mco_vec dVdXac = vadd(vmul(gV[c0], lambda0), vmul(gV[c1], lambda1));
            const mco_vec dWdXac = vadd(vmul(gW[c0], lambda0), vmul(gW[c1], lambda1));
            const double Visac = s->effective_viscosity * lambda0 + s->effective_viscosity * lambda1;
            const double VisFace = Visac * s->face_rlencos[face];
            const mco_vec Xpn = vsub(M->cell_centers[c1], M->cell_centers[c0]);
            double UFace, VFace, WFace;
            if (mf >= 0.0) { UFace = U[c0]; VFace = V[c0]; WFace = W[c0]; }
            else { UFace = U[c1]; VFace = V[c1]; WFace = W[c1]; }
            const double fuce = mf * UFace, fvce = mf * VFace, fwce = mf * WFace;
            const double sx = s->face_normals[face].x, sy = s->face_normals[face].y, sz = s->face_normals[face].z;
            const double fude = Visac * ((dUdXac.x + dUdXac.x) * sx + (dUdXac.y + dVdXac.x) * sy + (dUdXac.z + dWdXac.x) * sz);
            const double fvde = Visac * ((dUdXac.y + dVdXac.x) * sx + (dVdXac.y + dVdXac.y) * sy + (dVdXac.z + dWdXac.y) * sz);
            const double fwde = Visac * ((dUdXac.z + dWdXac.x) * sx + (dWdXac.y + dVdXac.z) * sy + (dWdXac.z + dWdXac.z) * sz);
            const double fmin = dmin(mf, 0.0), fmax = dmax(mf, 0.0);
            const double fuci = fmin * U[c0] + fmax * U[c1];
            const double fvci = fmin * V[c0] + fmax * V[c1];
            const double fwci = fmin * W[c0] + fmax * W[c1];
            const double fudi = VisFace * vdot(dUdXac, Xpn);
            const double fvdi = VisFace * vdot(dVdXac, Xpn);
            const double fwdi = VisFace * vdot(dWdXac, Xpn);
            s->ff0[face] = -VisFace - dmax(mf, 0.0);
            s->ff1[face] = -VisFace + dmin(mf, 0.0);
            const double blend_u = GammaBlend * (fuce - fuci);
            const double blend_v = GammaBlend * (fvce - fvci);
            const double blend_w = GammaBlend * (fwce - fwci);
            s->S_U[c0] = s->S_U[c0] - blend_u + fude - fudi;
            s->S_V[c0] = s->S_V[c0] - blend_v + fvde - fvdi;
            s->S_W[c0] = s->S_W[c0] - blend_w + fwde - fwdi;
            s->S_U[c1] = s->S_U[c1] + blend_u - fude + fudi;
            s->S_V[c1] = s->S_V[c1] + blend_v - fvde + fvdi;
            s->S_W[c1] = s->S_W[c1] + blend_w - fwde + fwdi;
        } else {
            const uint64_t bc = c1 - N;
            const uint64_t bt = M->boundary_types[bc];
            const mco_vec Xpn = vsub(s->face_centers[face], M->cell_centers[c0]);
            if (bt == MCO_INLET || bt == MCO_OUTLET) {
                const mco_vec dUdXac = gU[c0], dVdXac = gV[c0], dWdXac = gW[c0];
                double UFace, VFace, WFace, Visac;
                if (bt == MCO_INLET) {
                    UFace = GAS_VEL_X; VFace = GAS_VEL_Y; WFace = GAS_VEL_Z;
                    Visac = s->inlet_effective_viscosity;
                } else {
                    Visac = s->effective_viscosity;
                    UFace = U[c0]; VFace = V[c0]; WFace = W[c0];
                }
                const double VisFace = Visac * s->face_rlencos[face];
                const double sx = s->face_normals[face].x, sy = s->face_normals[face].y, sz = s->face_normals[face].z;
                const double fude = Visac * ((dUdXac.x + dUdXac.x) * sx + (dUdXac.y + dVdXac.x) * sy + (dUdXac.z + dWdXac.x) * sz);
                const double fvde = Visac * ((dUdXac.y + dVdXac.x) * sx + (dVdXac.y + dVdXac.y) * sy + (dVdXac.z + dWdXac.y) * sz);
                const double fwde = Visac * ((dUdXac.z + dWdXac.x) * sx + (dWdXac.y + dVdXac.z) * sy + (dWdXac.z + dWdXac.z) * sz);
                const double fudi = VisFace * vdot(dUdXac, Xpn);
                const double fvdi = VisFace * vdot(dVdXac, Xpn);
                const double fwdi = VisFace * vdot(dWdXac, Xpn);
                if (bt == MCO_OUTLET && s->mflux[face] < 0.0) {
                    /* FlowSolver.inl:1576-1580 */
                    fprintf(stderr, "MAIN COMP UVW NEGATIVE OUTFLOW %3.18f\n", s->mflux[face]);
                    s->mflux[face] = 1e-15;
                }
                const double f = -VisFace + dmin(s->mflux[face], 0.0);
                s->A_U[c0] = s->A_U[c0] - f;
                s->S_U[c0] = s->S_U[c0] - f * UFace + fude - fudi;
                U[N + bc] = UFace;
                s->A_V[c0] = s->A_V[c0] - f;
                s->S_V[c0] = s->S_V[c0] - f * VFace + fvde - fvdi;
                V[N + bc] = VFace;
                s->A_W[c0] = s->A_W[c0] - f;
                s->S_W[c0] = s->S_W[c0] - f * WFace + fwde - fwdi;
                W[N + bc] = WFace;
            } else if (bt == MCO_WALL) {
                const double UFace = 0., VFace = 0., WFace = 0.;
                const double Visac = s->effective_viscosity;
                const double coef = Visac * s->face_rlencos[face];
                mco_vec Up = {U[c0] - UFace, V[c0] - VFace, W[c0] - WFace};
                const mco_vec nh = vnormalise(s->face_normals[face]);
                const double dp = vdot(Up, nh);
                const mco_vec Ut = vsub(Up, vmul(nh, dp));
                const double Uvel = fabs(Ut.x) + fabs(Ut.y) + fabs(Ut.z);
                mco_vec force;
                if (Uvel > 0.0) {
                    const double distance_to_face = vmag(Xpn);
                    force = vdiv(vmul(Ut, s->face_areas[face] * Visac), distance_to_face);
                } else {
                    force.x = 0.0; force.y = 0.0; force.z = 0.0;
                }
                s->A_U[c0] = s->A_U[c0] + coef;
                s->A_V[c0] = s->A_V[c0] + coef;
                s->A_W[c0] = s->A_W[c0] + coef;
                s->S_U[c0] = s->S_U[c0] + coef * U[c0] - force.x;
                s->S_V[c0] = s->S_V[c0] + coef * V[c0] - force.y;
                s->S_W[c0] = s->S_W[c0] + coef * W[c0] - force.z;
                U[N + bc] = UFace; V[N + bc] = VFace; W[N + bc] = WFace;
            }
        }
    }
}

/* ======================================================================
 * A9: CSR assembly (FlowSolver.inl:1186-1339)
 * ====================================================================== */
static void setup_sparse_matrix(mco_state *s, double URFactor, double *A_c, double *phi_c, double *S_c)
{
    const mco_mesh *M = &s->mesh;
    const int64_t N = s->N;
    const double RURF = 1. / URFactor;
    memset(s->A_vals, 0, s->nnz * sizeof(double)); /* MatZeroEntries */
    memset(s->b, 0, N * sizeof(double));           /* VecZeroEntries */
    for (int64_t face = 0; face < s->F; face++) {
        const uint64_t c0 = M->faces[2 * face], c1 = M->faces[2 * face + 1];
        if (c1 >= (uint64_t)N) continue;
        s->A_vals[s->face_idx01[face]] = s->ff1[face];
        s->A_vals[s->face_idx10[face]] = s->ff0[face];
        A_c[c0] -= s->ff1[face];
        A_c[c1] -= s->ff0[face];
    }
    for (int64_t i = 0; i < N; i++) {
        A_c[i] *= RURF;
        S_c[i] = S_c[i] + (1.0 - URFactor) * A_c[i] * phi_c[i];
        s->A_vals[s->diag_idx[i]] = A_c[i];
        s->b[i] = S_c[i];
    }
}

static void update_sparse_matrix(mco_state *s, double URFactor, double *A_c, double *phi_c, double *S_c)
{
    const mco_mesh *M = &s->mesh;
    const int64_t N = s->N;
    const double RURF = 1. / URFactor;
    for (int64_t face = 0; face < s->F; face++) {
        const uint64_t c0 = M->faces[2 * face], c1 = M->faces[2 * face + 1];
        if (c1 >= (uint64_t)N) continue;
        A_c[c0] -= s->ff1[face];
        A_c[c1] -= s->ff0[face];
    }
    for (int64_t i = 0; i < N; i++) {
        A_c[i] *= RURF;
        S_c[i] = S_c[i] + (1.0 - URFactor) * A_c[i] * phi_c[i];
        s->A_vals[s->diag_idx[i]] = A_c[i];
        s->b[i] = S_c[i];
    }
}

/* A10: FlowSolver.inl:1341-1375 */
static int solve_sparse_matrix(mco_state *s, double *phi_c)
{
    double rr;
    if (s->opts.warm_start) memcpy(s->u, phi_c, s->N * sizeof(double));
    const int it = mco_bicgstab_x0(s->N, s->row_ptr, s->col, s->A_vals, s->b, s->u, s->opts.mom_rtol,
                                   s->opts.mom_maxit, &rr, s->opts.warm_start);
    memcpy(phi_c, s->u, s->N * sizeof(double));
    return it;
}

/* ======================================================================
 * A8 + A9 + A10: calculate_UVW (FlowSolver.inl:1655-1810)
 * ====================================================================== */
void mco_momentum_assemble(mco_state *s)
{
    const int64_t N = s->N;
    for (int64_t i = 0; i < N; i++) {
        s->A_U[i] = 0.0; s->A_V[i] = 0.0; s->A_W[i] = 0.0;
        s->S_U[i] = 0.0; s->S_V[i] = 0.0; s->S_W[i] = 0.0;
    }
    mco_calculate_flux_UVW(s);
    for (int64_t i = 0; i < N; i++) {
        /* FlowSolver.inl:1689-1697 */
        const double BodyForce = -0.001 * s->cell_densities[i] * s->cell_volumes[i] * (s->phi[MCO_TEM][i] - GAS_TEM);
        const double gravity[3] = {0.0, -9.81, 0.0};
        s->S_U[i] += gravity[0] * BodyForce;
        s->S_V[i] += gravity[1] * BodyForce;
        s->S_W[i] += gravity[2] * BodyForce;
    }
    for (int64_t i = 0; i < N; i++) {
        /* FlowSolver.inl:1700-1705 */
        s->S_U[i] -= s->grad[MCO_P][i].x * s->cell_volumes[i];
        s->S_V[i] -= s->grad[MCO_P][i].y * s->cell_volumes[i];
        s->S_W[i] -= s->grad[MCO_P][i].z * s->cell_volumes[i];
    }
    const double rdelta = 1.0 / s->opts.delta;
    for (int64_t i = 0; i < N; i++) {
        /* FlowSolver.inl:1708-1721 */
        const double f = s->cell_densities[i] * s->cell_volumes[i] * rdelta;
        s->S_U[i] += f * s->phi[MCO_U][i];
        s->S_V[i] += f * s->phi[MCO_V][i];
        s->S_W[i] += f * s->phi[MCO_W][i];
        s->A_U[i] += f; s->A_V[i] += f; s->A_W[i] += f;
    }
    /* setup_sparse_matrix for U, update_sparse_matrix for V and W
       (FlowSolver.inl:1731-1758); the V/W updates read only V/W data, so
       computing them before the U solve changes nothing. */
    setup_sparse_matrix(s, 0.5, s->A_U, s->phi[MCO_U], s->S_U);
}

void mco_momentum_solve(mco_state *s, int comp)
{
    if (comp == 0) {
        s->stats.mom_iters += solve_sparse_matrix(s, s->phi[MCO_U]);
    } else if (comp == 1) {
        update_sparse_matrix(s, 0.5, s->A_V, s->phi[MCO_V], s->S_V);
        s->stats.mom_iters += solve_sparse_matrix(s, s->phi[MCO_V]);
    } else {
        update_sparse_matrix(s, 0.5, s->A_W, s->phi[MCO_W], s->S_W);
        s->stats.mom_iters += solve_sparse_matrix(s, s->phi[MCO_W]);
    }
}

void mco_calculate_UVW(mco_state *s)
{
    mco_momentum_assemble(s);
    mco_momentum_solve(s, 0);
    mco_momentum_solve(s, 1);
    mco_momentum_solve(s, 2);
}

/* ======================================================================
 * A11: Rhie-Chow mass flux (FlowSolver.inl:1880-2104)
 * ====================================================================== */
void mco_calculate_mass_flux(mco_state *s)
{
    const mco_mesh *M = &s->mesh;
    const int64_t N = s->N;
    const double *U = s->phi[MCO_U], *V = s->phi[MCO_V], *W = s->phi[MCO_W], *P = s->phi[MCO_P];
    const mco_vec *gU = s->grad[MCO_U], *gV = s->grad[MCO_V], *gW = s->grad[MCO_W], *gP = s->grad[MCO_P];
    const double *rho = s->cell_densities, *vol = s->cell_volumes;
    for (int64_t face = 0; face < s->F; face++) {
        const uint64_t c0 = M->faces[2 * face], c1 = M->faces[2 * face + 1];
        if (c1 < (uint64_t)N) {
            const double lambda0 = s->face_lambdas[face], lambda1 = 1.0 - lambda0;
            const mco_vec dUdXac = vadd(vmul(gU[c0], lambda0), vmul(gU[c1], lambda1));
            const mco_vec dVdXac = vadd(vmul(gV[c0], lambda0), vmul(gV[c1], lambda1));
            const mco_vec dWdXac = vadd(vmul(gW[c0], lambda0), vmul(gW[c1], lambda1));
            const mco_vec cc0 = M->cell_centers[c0], cc1 = M->cell_centers[c1];
            const mco_vec Xac = vadd(vmul(cc1, lambda1), vmul(cc0, lambda0));
            const mco_vec fc = s->face_centers[face];
            const mco_vec delta = vsub(fc, Xac);
            const double UFace = U[c1] * lambda1 + U[c0] * lambda0 + vdot(dUdXac, delta);
            const double VFace = V[c1] * lambda1 + V[c0] * lambda0 + vdot(dVdXac, delta);
            const double WFace = W[c1] * lambda1 + W[c0] * lambda0 + vdot(dWdXac, delta);
            const double densityf = rho[c0] * lambda0 + rho[c1] * lambda1;
            const mco_vec n = s->face_normals[face];
            s->mflux[face] = densityf * (UFace * n.x + VFace * n.y + WFace * n.z);
            const mco_vec nh = vnormalise(n);
            const mco_vec Xpac = vsub(fc, vmul(nh, vdot(vsub(fc, cc0), nh)));
            const mco_vec Xnac = vsub(fc, vmul(nh, vdot(vsub(fc, cc1), nh)));
            const mco_vec delp = vsub(Xpac, cc0);
            const mco_vec deln = vsub(Xnac, cc1);
            const double cell0_P = P[c0] + vdot(gP[c0], delp);
            const double cell1_P = P[c1] + vdot(gP[c1], deln);
            const mco_vec Xpn = vsub(Xnac, Xpac);
            const mco_vec Xpn2 = vsub(cc1, cc0);
            const double ApV0 = (s->A_U[c0] != 0.0) ? 1.0 / s->A_U[c0] : 0.0;
            const double ApV1 = (s->A_U[c1] != 0.0) ? 1.0 / s->A_U[c1] : 0.0;
            double ApV = rho[c0] * ApV0 * lambda0 + rho[c1] * ApV1 * lambda1;
            const double volume_avg = vol[c0] * lambda0 + vol[c1] * lambda1;
            ApV = ApV * s->face_areas[face] * volume_avg / vdot(Xpn2, nh);
            const double dpx = (gP[c1].x * lambda1 + gP[c0].x * lambda0) * Xpn.x;
            const double dpy = (gP[c1].y * lambda1 + gP[c0].y * lambda0) * Xpn.y;
            const double dpz = (gP[c1].z * lambda1 + gP[c0].z * lambda0) * Xpn.z;
            s->ff0[face] = -ApV;
            s->ff1[face] = -ApV;
            s->mflux[face] -= ApV * ((cell1_P - cell0_P) - dpx - dpy - dpz);
        } else {
            const uint64_t bc = c1 - N;
            const uint64_t bt = M->boundary_types[bc];
            const mco_vec vel = {GAS_VEL_X, GAS_VEL_Y, GAS_VEL_Z};
            if (bt == MCO_INLET) {
                const double Din = 1.2;
                s->mflux[face] = Din * vdot(vel, s->face_normals[face]);
                s->S_U[c0] = s->S_U[c0] - s->mflux[face];
            } else if (bt == MCO_OUTLET) {
                const double Din = 1.2;
                s->mflux[face] = Din * vdot(vel, s->face_normals[face]);
                if (s->mflux[face] < 0.0) {
                    fprintf(stderr, "MAIN COMP PRES NEGATIVE OUTFLOW %3.18f\n", s->mflux[face]);
                    s->mflux[face] = 1e-15;
                    s->phi[MCO_TE][N + bc] = s->phi[MCO_TE][c0];
                    s->phi[MCO_ED][N + bc] = s->phi[MCO_ED][c0];
                    s->phi[MCO_TEM][N + bc] = s->phi[MCO_TEM][c0];
                }
            } else if (bt == MCO_WALL) {
                s->mflux[face] = 0.0;
            }
        }
    }
    /* outlet balance, FlowSolver.inl:2010-2103 */
    double FlowOut = 0.0, FlowIn = 0.0, areaout = 0.0;
    int count_out = 0;
    for (int64_t face = 0; face < s->F; face++) {
        const uint64_t c1 = M->faces[2 * face + 1];
        if (c1 < (uint64_t)N) continue;
        const uint64_t bt = M->boundary_types[c1 - N];
        if (bt == MCO_INLET) FlowIn += s->mflux[face];
        else if (bt == MCO_OUTLET) { FlowOut += s->mflux[face]; count_out++; areaout += s->face_areas[face]; }
    }
    (void)count_out;
    if (FlowOut < 0.0000000001) {
        const double ratearea = -FlowIn / areaout;
        FlowOut = 0.0;
        for (int64_t face = 0; face < s->F; face++) {
            const uint64_t c1 = M->faces[2 * face + 1];
            if (c1 < (uint64_t)N) continue;
            const uint64_t bc = c1 - N;
            if (M->boundary_types[bc] == MCO_OUTLET) {
                s->mflux[face] = ratearea * s->face_areas[face];
                s->phi[MCO_U][N + bc] = GAS_VEL_X;
                s->phi[MCO_V][N + bc] = GAS_VEL_Y;
                s->phi[MCO_W][N + bc] = GAS_VEL_Z;
                FlowOut += s->mflux[face];
            }
        }
    }
    for (int64_t face = 0; face < s->F; face++) {
        const uint64_t c0 = M->faces[2 * face], c1 = M->faces[2 * face + 1];
        if (c1 < (uint64_t)N) continue;
        if (M->boundary_types[c1 - N] == MCO_OUTLET) s->S_U[c0] -= s->mflux[face];
    }
}

/* ======================================================================
 * A12: pressure-correction matrix (FlowSolver.inl:2106-2201)
 * Compat: Pressure_A while first_pressure, else the momentum matrix "A"
 * (which is never solved with).  Simple: always the pressure matrix.
 * ====================================================================== */
void mco_setup_pressure_matrix(mco_state *s)
{
    const mco_mesh *M = &s->mesh;
    const int64_t N = s->N;
    const int into_pressure = s->first_pressure || s->opts.mode == MCO_MODE_SIMPLE;
    double *vals = into_pressure ? s->PA_vals : s->A_vals;
    memset(vals, 0, s->nnz * sizeof(double));
    for (int64_t face = 0; face < s->F; face++) {
        const uint64_t c0 = M->faces[2 * face], c1 = M->faces[2 * face + 1];
        if (c1 >= (uint64_t)N) continue;
        vals[s->face_idx01[face]] = s->ff1[face];
        vals[s->face_idx10[face]] = s->ff0[face];
        s->S_U[c0] -= s->mflux[face];
        s->S_U[c1] += s->mflux[face];
        s->A_V[c0] -= s->ff1[face];
        s->A_V[c1] -= s->ff0[face];
    }
    for (int64_t i = 0; i < N; i++) vals[s->diag_idx[i]] = s->A_V[i];
}

/* A13: FlowSolver.inl:2203-2273.  KSPSolve always uses Pressure_A and
   Pressure_b; in compat mode Pressure_b is only written while
   first_pressure holds. */
void mco_solve_pressure_matrix(mco_state *s)
{
    const int64_t N = s->N;
    if (s->first_pressure || s->opts.mode == MCO_MODE_SIMPLE) memcpy(s->Pb, s->S_U, N * sizeof(double));
    else memcpy(s->b, s->S_U, N * sizeof(double));
    double rr = 0.0;
    const int it = (s->opts.p_solver == 1 ? mco_pcg_amg : mco_pcg_singular)(
                                    N, s->row_ptr, s->col, s->PA_vals, s->Pb, s->u, s->opts.p_rtol, s->p_atol,
                                    s->opts.p_maxit, &rr, &s->p_bnorm_last);
    s->stats.p_iters += it;
    s->stats.p_solves += 1;
    s->stats.last_p_relres = rr;
    memcpy(s->phi[MCO_PP], s->u, N * sizeof(double));
}

/* FlowSolver.inl:2275-2290 */
void mco_update_P_at_boundaries(mco_state *s, int field)
{
    const mco_mesh *M = &s->mesh;
    const int64_t N = s->N;
    double *phi = s->phi[field];
    for (int64_t face = 0; face < s->F; face++) {
        const uint64_t c0 = M->faces[2 * face], c1 = M->faces[2 * face + 1];
        if (c1 < (uint64_t)N) continue;
        phi[N + (c1 - N)] = phi[c0];
    }
}

/* A15: FlowSolver.inl:2292-2315 */
void mco_update_P(mco_state *s, int field)
{
    const mco_mesh *M = &s->mesh;
    const int64_t N = s->N;
    double *phi = s->phi[field];
    const mco_vec *g = s->grad[field];
    for (int64_t face = 0; face < s->F; face++) {
        const uint64_t c0 = M->faces[2 * face], c1 = M->faces[2 * face + 1];
        if (c1 < (uint64_t)N) continue;
        const uint64_t bc = c1 - N;
        if (M->boundary_types[bc] == MCO_OUTLET) phi[N + bc] = phi[c0];
        else {
            const mco_vec ds = vsub(s->face_centers[face], M->cell_centers[c0]);
            phi[N + bc] = phi[c0] + vdot(g[c0], ds);
        }
    }
}

/* FlowSolver.inl:1812-1878 */
void mco_update_mass_flux(mco_state *s)
{
    const mco_mesh *M = &s->mesh;
    const int64_t N = s->N;
    const mco_vec *gPP = s->grad[MCO_PP];
    for (int64_t face = 0; face < s->F; face++) {
        const uint64_t c0 = M->faces[2 * face], c1 = M->faces[2 * face + 1];
        if (c1 >= (uint64_t)N) continue;
        const mco_vec fc = s->face_centers[face], cc0 = M->cell_centers[c0], cc1 = M->cell_centers[c1];
        const mco_vec nh = vnormalise(s->face_normals[face]);
        const mco_vec Xpac = vsub(fc, vmul(nh, vdot(vsub(fc, cc0), nh)));
        const mco_vec Xnac = vsub(fc, vmul(nh, vdot(vsub(fc, cc1), nh)));
        mco_vec Xn = vsub(Xnac, cc1), Xp = vsub(Xpac, cc0);
        if (fabs(Xn.x) < 0.0000000000000003) Xn.x = 0.0;
        if (fabs(Xn.y) < 0.0000000000000003) Xn.y = 0.0;
        if (fabs(Xn.z) < 0.0000000000000003) Xn.z = 0.0;
        if (fabs(Xp.x) < 0.0000000000000003) Xp.x = 0.0;
        if (fabs(Xp.y) < 0.0000000000000003) Xp.y = 0.0;
        if (fabs(Xp.z) < 0.0000000000000003) Xp.z = 0.0;
        const double fact = s->ff0[face];
        const double dpx = gPP[c1].x * Xn.x - gPP[c0].x * Xp.x;
        const double dpy = gPP[c1].y * Xn.y - gPP[c0].y * Xp.y;
        const double dpz = gPP[c1].z * Xn.z - gPP[c0].z * Xp.z;
        const double fc_ = fact * (dpx + dpy + dpz) * 0.8;
        s->mflux[face] += fc_;
        s->S_U[c0] -= fc_;
        s->S_U[c1] += fc_;
    }
}

/* ======================================================================
 * A14: pressure-correction loop (FlowSolver.inl:2317-2463)
 * ====================================================================== */
/* One loop's corrections from the solved PP (FlowSolver.inl:2391-2433):
   PP boundary slots and gradient, face fluxes += ff0 (PP1 - PP0), and the
   cell corrections P += 0.2 PP, U -= grad PP V / A_U -- in compat mode only
   while first_pressure holds, which the first corrected cell clears (so the
   reference corrects cell 0 of the first loop only). */
void mco_apply_pressure_correction(mco_state *s)
{
    const mco_mesh *M = &s->mesh;
    const int64_t N = s->N;
    double *PP = s->phi[MCO_PP], *P = s->phi[MCO_P];
    mco_update_P_at_boundaries(s, MCO_PP);
    mco_get_phi_gradient(s, MCO_PP, 1);
    for (int64_t face = 0; face < s->F; face++) {
        const uint64_t c0 = M->faces[2 * face], c1 = M->faces[2 * face + 1];
        if (c1 >= (uint64_t)N) continue;
        s->mflux[face] += (s->ff0[face] * (PP[c1] - PP[c0]));
    }
    for (int64_t cell = 0; cell < N; cell++) {
        const double Ar = (s->A_U[cell] != 0.0) ? 1.0 / s->A_U[cell] : 0.0;
        const double fact = s->cell_volumes[cell] * Ar;
        if (s->first_pressure || s->opts.mode == MCO_MODE_SIMPLE) {
            P[cell] += 0.2 * PP[cell];
            s->phi[MCO_U][cell] -= s->grad[MCO_PP][cell].x * fact;
            s->phi[MCO_V][cell] -= s->grad[MCO_PP][cell].y * fact;
            s->phi[MCO_W][cell] -= s->grad[MCO_PP][cell].z * fact;
            s->first_pressure = 0;
        }
    }
}

void mco_calculate_pressure(mco_state *s)
{
    const int64_t N = s->N;
    double *PP = s->phi[MCO_PP];
    int Loop_num = 0, Loop_continue = 1;
    double Pressure_correction_max = 0, Pressure_correction_ref = 0;
    double first_bnorm = 0.0;
    for (int64_t i = 0; i < N; i++) { s->A_V[i] = 0.0; s->S_U[i] = 0.0; }
    for (int64_t face = 0; face < s->F; face++) { s->mflux[face] = 0.0; s->ff0[face] = 0.0; s->ff1[face] = 0.0; }
    mco_calculate_mass_flux(s);
    mco_setup_pressure_matrix(s);
    while (Loop_continue) {
        Loop_num++;
        /* Loops after the first solve only the O(1e-16) non-orthogonal
           remainder of update_mass_flux; their tolerance is taken relative to
           the first loop's RHS (an absolute floor), not to their own RHS. */
        s->p_atol = (Loop_num == 1) ? 0.0 : s->opts.p_rtol * first_bnorm;
        mco_solve_pressure_matrix(s);
        if (Loop_num == 1) first_bnorm = s->p_bnorm_last;
        Pressure_correction_max = PP[0];
        for (int64_t i = 1; i < N; i++)
            if (fabs(PP[i]) > Pressure_correction_max) Pressure_correction_max = fabs(PP[i]);
        if (Loop_num == 1) Pressure_correction_ref = Pressure_correction_max;
        mco_apply_pressure_correction(s);
        for (int64_t i = 0; i < N; i++) s->S_U[i] = 0.0;
        mco_update_mass_flux(s);
        for (int64_t i = 0; i < N + 6; i++) PP[i] = 0.0;
        if (Loop_num >= 4 || Pressure_correction_max <= 0.25 * Pressure_correction_ref) Loop_continue = 0;
    }
    s->stats.p_loops += Loop_num;
    s->pressure_loops_last = Loop_num;
    s->pressure_correction_max_last = Pressure_correction_ref;
    mco_update_P_at_boundaries(s, MCO_P);
    mco_get_phi_gradient(s, MCO_P, 1);
    mco_update_P(s, MCO_P);
}

/* ======================================================================
 * A16: scalar transport (FlowSolver.inl:2516-3076)
 * ====================================================================== */
static double scalar_visac(int type, double ev, double Visac)
{
    if (type == MCO_TEMP) return (ev + Visac / 0.9) / 0.6905;
    if (type == MCO_TERBTE) return ev + Visac;
    if (type == MCO_TERBED) return ev + Visac / 1.219;
    return (ev + Visac / 0.9) / 0.9;
}

void mco_flux_scalar(mco_state *s, int type, int field)
{
    const mco_mesh *M = &s->mesh;
    const int64_t N = s->N;
    double *phi = s->phi[field];
    const mco_vec *g = s->grad[field];
    const double ev = s->effective_viscosity;
    const double GammaBlend = 0.0;
    for (int64_t face = 0; face < s->F; face++) {
        const uint64_t c0 = M->faces[2 * face], c1 = M->faces[2 * face + 1];
        const double mf = s->mflux[face];
        if (c1 < (uint64_t)N) {
            const double lambda0 = s->face_lambdas[face], lambda1 = 1.0 - lambda0;
            double Visac = ev * lambda0 + ev * lambda1;
            Visac -= ev;
            Visac = scalar_visac(type, ev, Visac);
            const mco_vec dPhiXac = vadd(vmul(g[c0], lambda0), vmul(g[c1], lambda1));
            const mco_vec Xpn = vsub(M->cell_centers[c1], M->cell_centers[c0]);
            const double VisFace = Visac * s->face_rlencos[face];
            const double PhiFace = (mf >= 0.0) ? phi[c0] : phi[c1];
            const double fce = mf * PhiFace;
            const double fde1 = Visac * vdot(dPhiXac, s->face_normals[face]);
            const double fci = dmin(mf, 0.0) * phi[c0] + dmax(mf, 0.0) * phi[c1];
            const double fdi = VisFace * vdot(dPhiXac, Xpn);
            s->ff0[face] = -VisFace - dmax(mf, 0.0);
            s->ff1[face] = -VisFace + dmin(mf, 0.0);
            const double blend = GammaBlend * (fce - fci);
            s->S_U[c0] = s->S_U[c0] - blend + fde1 - fdi;
            s->S_U[c1] = s->S_U[c1] + blend - fde1 + fdi;
        } else {
            const uint64_t bc = c1 - N;
            const uint64_t bt = M->boundary_types[bc];
            const mco_vec Xpn = vsub(s->face_centers[face], M->cell_centers[c0]);
            if (bt == MCO_INLET) {
                const mco_vec dPhidXac = g[c0];
                const double velmag2 = pow(GAS_VEL_X, 2) + pow(GAS_VEL_Y, 2) + pow(GAS_VEL_Z, 2);
                double PhiFace;
                if (type == MCO_TEMP) PhiFace = GAS_TEM;
                else if (type == MCO_TERBTE) PhiFace = 3.0 / 2.0 * ((0.1 * 0.1) * velmag2);
                else if (type == MCO_TERBED) PhiFace = pow(0.09, 0.75) * pow((3.0 / 2.0 * ((0.1 * 0.1) * velmag2)), 1.5);
                else if (type == MCO_FUEL) PhiFace = GAS_FUL;
                else if (type == MCO_PROG) PhiFace = 0.0;
                else if (type == MCO_VARFU) PhiFace = GAS_FUL;
                else PhiFace = 0.0;
                double Visac = s->inlet_effective_viscosity;
                Visac -= ev;
                Visac = scalar_visac(type, ev, Visac);
                const double VisFace = Visac * s->face_rlencos[face];
                const double fde = Visac * vdot(dPhidXac, s->face_normals[face]);
                const double fdi = VisFace * vdot(dPhidXac, Xpn);
                const double f = -VisFace + dmin(mf, 0.0);
                s->A_V[c0] = s->A_V[c0] - f;
                s->S_U[c0] = s->S_U[c0] - f * PhiFace + fde - fdi;
                phi[N + bc] = PhiFace;
            } else if (bt == MCO_OUTLET) {
                const mco_vec dPhidXac = g[c0];
                double Visac = ev;
                Visac -= ev;
                Visac = scalar_visac(type, ev, Visac);
                const double PhiFace = phi[c0] + vdot(dPhidXac, Xpn);
                const double VisFace = Visac * s->face_rlencos[face];
                const double fde = Visac * vdot(dPhidXac, s->face_normals[face]);
                const double fdi = VisFace * vdot(dPhidXac, Xpn);
                s->S_U[c0] = s->S_U[c0] + fde - fdi;
                phi[N + bc] = PhiFace;
            } else if (bt == MCO_WALL) {
                /* `type != TERBTE or type != TERBED` always holds */
                phi[N + bc] = phi[c0];
            }
        }
    }
}

void mco_solve_turbulence_models(mco_state *s, int type)
{
    const mco_mesh *M = &s->mesh;
    const int64_t N = s->N;
    const double ev = s->effective_viscosity;
    double *TE = s->phi[MCO_TE], *ED = s->phi[MCO_ED], *TP = s->phi[MCO_TP];
    const double *rho = s->cell_densities, *vol = s->cell_volumes;
    const double Cmu = 0.09, Cmu75 = pow(Cmu, 0.75);
    if (type == MCO_TERBTE) {
        for (int64_t i = 0; i < N; i++) {
            const mco_vec dUdXp = s->grad[MCO_U][i], dVdXp = s->grad[MCO_V][i], dWdXp = s->grad[MCO_W][i];
            const double s1 = (dUdXp.x + dUdXp.x) * dUdXp.x + (dUdXp.y + dVdXp.x) * dUdXp.y + (dUdXp.z + dWdXp.x) * dUdXp.z;
            const double s2 = (dVdXp.x + dUdXp.y) * dVdXp.x + (dVdXp.y + dVdXp.y) * dVdXp.y + (dVdXp.z + dWdXp.y) * dVdXp.z;
            const double s3 = (dWdXp.x + dUdXp.z) * dWdXp.x + (dWdXp.y + dVdXp.z) * dWdXp.y + (dWdXp.z + dWdXp.z) * dWdXp.z;
            const double VisT = ev - ev;
            const double Pk = VisT * (s1 + s2 + s3);
            TP[i] = Pk;
            const double Dis = rho[i] * ED[i];
            s->S_U[i] = s->S_U[i] + TP[i] * vol[i];
            s->A_V[i] = s->A_V[i] + Dis / (TE[i] + 0.000000000000000001) * vol[i];
        }
        for (int64_t face = 0; face < s->F; face++) {
            const uint64_t c0 = M->faces[2 * face], c1 = M->faces[2 * face + 1];
            if (c1 < (uint64_t)N) continue;
            const uint64_t bc = c1 - N;
            if (M->boundary_types[bc] != MCO_WALL) continue;
            s->S_U[c0] = s->S_U[c0] - TP[c0] * vol[c0];
            const double UFace = 0., VFace = 0., WFace = 0.;
            const double Visc = ev;
            const mco_vec Xpn = vsub(s->face_centers[face], M->cell_centers[c0]);
            const mco_vec Up = {s->phi[MCO_U][c0] - UFace, s->phi[MCO_V][c0] - VFace, s->phi[MCO_W][c0] - WFace};
            const mco_vec nh = vnormalise(s->face_normals[face]);
            const double dp = vdot(Up, nh);
            const mco_vec Ut = vsub(Up, vmul(nh, dp));
            const double Uvel = sqrt(vdot(Ut, Ut));
            const double distance_to_face = vmag(Xpn);
            const double rkapdn = 1.0 / (0.419 * distance_to_face);
            const double Tau_w = Visc * Uvel / distance_to_face;
            const double Utau = sqrt(Tau_w / rho[c0]);
            TP[c0] = Tau_w * Utau * rkapdn;
            TE[N + bc] = TE[c0];
            s->S_U[c0] = s->S_U[c0] + TP[c0] * vol[c0];
            const double DisP = Cmu75 * sqrt(TE[c0]) * rkapdn;
            s->A_V[c0] = s->A_V[c0] + rho[c0] * DisP * vol[c0];
        }
    } else if (type == MCO_TERBED) {
        for (int64_t i = 0; i < N; i++) {
            const double fact = ED[i] / (TE[i] + 0.000000000000000001) * vol[i];
            s->S_U[i] = s->S_U[i] + 1.44 * fact * TP[i];
            s->A_V[i] = s->A_V[i] + 1.92 * fact * rho[i];
        }
        for (int64_t face = 0; face < s->F; face++) {
            const uint64_t c0 = M->faces[2 * face], c1 = M->faces[2 * face + 1];
            if (c1 < (uint64_t)N) continue;
            const uint64_t bc = c1 - N;
            if (M->boundary_types[bc] != MCO_WALL) continue;
            const double turb = TE[c0];
            const double distance = vmag(vsub(s->face_centers[face], M->cell_centers[c0]));
            const double Dis = Cmu75 * pow(turb, 1.5) / (distance * 0.419);
            for (int j = 0; j < 6; j++) {
                const uint64_t nf = M->cell_faces[c0 * 6 + j];
                const uint64_t n0 = M->faces[2 * nf], n1 = M->faces[2 * nf + 1];
                if (n0 < (uint64_t)N && n1 < (uint64_t)N) {
                    if (n0 == c0) s->ff1[nf] = 0.0;
                    else if (n1 == c0) s->ff0[nf] = 0.0;
                }
            }
            ED[c0] = Dis;
            s->S_U[c0] = Dis;
            s->A_V[c0] = 1;
            ED[N + bc] = ED[c0];
        }
    }
}

void mco_scalar_assemble(mco_state *s, int type, int field)
{
    const int64_t N = s->N;
    double *phi = s->phi[field];
    for (int64_t i = 0; i < N; i++) { s->A_V[i] = 0.0; s->S_U[i] = 0.0; }
    mco_flux_scalar(s, type, field);
    const double rdelta = 1.0 / s->opts.delta;
    for (int64_t i = 0; i < N; i++) {
        const double f = s->cell_densities[i] * s->cell_volumes[i] * rdelta;
        s->S_U[i] += f * phi[i];
        s->A_V[i] += f;
    }
    if (type == MCO_TERBTE || type == MCO_TERBED) mco_solve_turbulence_models(s, type);
    double underrelax;
    if (type == MCO_TERBTE || type == MCO_TERBED) underrelax = 0.5;
    else underrelax = 0.95;
    setup_sparse_matrix(s, underrelax, s->A_V, phi, s->S_U);
}

/* FlowSolver.inl:2956-2985: non-negativity repair, in place, ascending cell
   order (a negative cell takes the mean of its interior neighbours as they
   stand, at least 1e-12) */
void mco_repair_nonneg(mco_state *s, int field)
{
    const mco_mesh *M = &s->mesh;
    const int64_t N = s->N;
    double *phi = s->phi[field];
    for (int64_t i = 0; i < N; i++) {
        int count = 0;
        double phisum = 0.0;
        if (phi[i] < 0.0) {
            for (int j = 0; j < 6; j++) {
                const uint64_t nb = M->cell_neighbours[i * 6 + j];
                if (nb < (uint64_t)N) { count++; phisum += phi[nb]; }
            }
            phisum = (phisum / count);
            phi[i] = dmax(phisum, 0.000000000001);
        }
    }
}

void mco_scalar_solve(mco_state *s, int type, int field)
{
    double *phi = s->phi[field];
    mco_scalar_assemble(s, type, field);
    s->stats.scalar_iters += solve_sparse_matrix(s, phi);
    if (type == MCO_TERBTE || type == MCO_TERBED) mco_repair_nonneg(s, field);
}

/* ======================================================================
 * Particle -> flow coupling, the flow side of update_flow_field without its
 * MPI transport (FlowSolver.inl:496-700)
 * ====================================================================== */
/* FlowSolver.inl:675-683: each received (cell, particle_aos) entry is added
   to the cell's particle terms in receive order.  pt: [5 N] momentum x y z,
   energy, fuel per cell (particle_aos<double>). */
void mco_particle_terms_add(mco_state *s, int64_t n, const uint64_t *cells, const double *terms, double *pt)
{
    (void)s;
    for (int64_t i = 0; i < n; i++) {
        double *d = pt + 5 * cells[i];
        const double *t = terms + 5 * i;
        d[0] = d[0] + t[0]; d[1] = d[1] + t[1]; d[2] = d[2] + t[2];
        d[3] += t[3];
        d[4] += t[4];
    }
}

/* Node values for the particle side: each node of the interpolation cells
   starts from the mesh's dummy gas state weighted by its missing cells
   (FlowSolver.inl:297-335, Mesh.hpp:114-116), then every cell of the set adds
   (phi_c + grad phi_c . (x_node - x_c)) / 8 to each of its eight nodes
   (interpolate_to_nodes, :428-493), cells in the order given (the reference
   iterates an unordered_set; duplicates count once).  out: [5 nnodes]
   flow_aos<double> (vel x y z, pressure, temp) of the requested nodes. */
void mco_interpolate_to_nodes(mco_state *s, int64_t ncells, const uint64_t *cells, int64_t nnodes,
                              const uint64_t *nodes, double *out)
{
    const mco_mesh *M = &s->mesh;
    const double node_neighbours = 8;
    uint8_t *want = calloc((size_t)M->n_points, 1);
    uint8_t *seen = calloc((size_t)s->N, 1);
    double *val = malloc(sizeof(double) * 5 * (size_t)M->n_points);
    for (int64_t k = 0; k < nnodes; k++) {
        const uint64_t nd = nodes[k];
        want[nd] = 1;
        const double b = node_neighbours - M->cells_per_point[nd];
        double *o = val + 5 * nd;
        o[0] = 50. * (b / node_neighbours);
        o[1] = 0.00 * (b / node_neighbours);
        o[2] = 0.00 * (b / node_neighbours);
        o[3] = 4000.0 * (b / node_neighbours);
        o[4] = 2000.0 * (b / node_neighbours);
    }
    for (int64_t i = 0; i < ncells; i++) {
        const uint64_t c = cells[i];
        if (seen[c]) continue;
        seen[c] = 1;
        const double T = s->phi[MCO_TEM][c], P = s->phi[MCO_P][c];
        const double U = s->phi[MCO_U][c], V = s->phi[MCO_V][c], W = s->phi[MCO_W][c];
        const mco_vec gT = s->grad[MCO_TEM][c], gP = s->grad[MCO_P][c];
        const mco_vec gU = s->grad[MCO_U][c], gV = s->grad[MCO_V][c], gW = s->grad[MCO_W][c];
        const mco_vec cc = M->cell_centers[c];
        for (int n = 0; n < 8; n++) {
            const uint64_t nd = M->cells[8 * c + n];
            if (!want[nd]) continue;
            const mco_vec d = vsub(M->points[nd], cc);
            double *o = val + 5 * nd;
            o[4] += (T + (gT.x * d.x + gT.y * d.y + gT.z * d.z)) / node_neighbours;
            o[3] += (P + (gP.x * d.x + gP.y * d.y + gP.z * d.z)) / node_neighbours;
            o[0] += (U + (gU.x * d.x + gU.y * d.y + gU.z * d.z)) / node_neighbours;
            o[1] += (V + (gV.x * d.x + gV.y * d.y + gV.z * d.z)) / node_neighbours;
            o[2] += (W + (gW.x * d.x + gW.y * d.y + gW.z * d.z)) / node_neighbours;
        }
    }
    for (int64_t k = 0; k < nnodes; k++)
        for (int q = 0; q < 5; q++) out[5 * k + q] = val[5 * nodes[k] + q];
    free(want);
    free(seen);
    free(val);
}

/* ======================================================================
 * A17: FlowSolver::timestep (FlowSolver.inl:3148-3356)
 * ====================================================================== */
void mco_simple_iteration(mco_state *s, int inner)
{
    mco_get_phi_gradients(s);
    if (s->opts.limit_grad) mco_limit_phi_gradients(s);   /* FlowSolver.inl:3197-3198 */
    if (s->timestep_count == 0 && inner == 0) mco_set_up_field(s);
    mco_calculate_UVW(s);
    mco_calculate_pressure(s);
    if (s->opts.solve_scalars) {
        mco_scalar_solve(s, MCO_TERBTE, MCO_TE);
        mco_scalar_solve(s, MCO_TERBED, MCO_ED);
        mco_scalar_solve(s, MCO_TEMP, MCO_TEM);
        mco_scalar_solve(s, MCO_FUEL, MCO_FUL);
        mco_scalar_solve(s, MCO_PROG, MCO_PRO);
        mco_scalar_solve(s, MCO_VARFU, MCO_VARF);
        mco_scalar_solve(s, MCO_VARPR, MCO_VARP);
    }
    /* FGM_loop_up (FlowSolver.inl:3265) writes S_phi.U, which the next
       calculate_UVW zeroes before use: no effect on any field. */
}

void mco_timestep(mco_state *s)
{
    for (int i = 0; i < 3; i++) mco_simple_iteration(s, i);
    s->timestep_count++;
}
