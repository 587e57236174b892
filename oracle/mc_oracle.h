/*
 * mc_oracle.h -- CPU oracle for the miniCOMBUST flow-solver SIMPLE step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, single-threaded restatement of
 * the reference CPU FlowSolver (include/flow/FlowSolver.{hpp,inl}) and of the
 * reference mesh generator (src/examples/mesh_examples.cpp) at one flow rank.
 * It exists to check the HIP implementation; nothing in the product links it.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Every function cites the reference lines it follows.  Compiled with
 * -ffp-contract=off so that floating-point expressions are evaluated exactly
 * as written, in the reference's operand order.
 *
 * Linear solves: the reference calls PETSc (KSPBCGS+PCNONE for momentum and
 * scalars, KSPFGMRES+PCGAMG for pressure; FlowSolver.hpp:513-527).  PETSc is
 * not part of this image; the oracle uses its own textbook BiCGStab and a
 * Jacobi-preconditioned CG on the mean-projected (singular, consistent)
 * pressure-correction system.  Their results agree with any converged solver
 * to the solver tolerance, not bit for bit.
 */
#ifndef MC_ORACLE_H
#define MC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* phi_vector field order (include/utils/utils.hpp:158-174). */
enum { MCO_U = 0, MCO_V, MCO_W, MCO_P, MCO_PP, MCO_TE, MCO_ED, MCO_TP,
       MCO_TEM, MCO_FUL, MCO_PRO, MCO_VARF, MCO_VARP, MCO_NFIELDS };

/* Scalar equation types (include/utils/utils.hpp:31-37). */
enum { MCO_TERBTE = 0, MCO_TERBED = 1, MCO_TEMP = 2, MCO_FUEL = 3,
       MCO_PROG = 4, MCO_VARFU = 5, MCO_VARPR = 6 };

/* Boundary types (include/geometry/Mesh.hpp:17). */
enum { MCO_NOT_BOUNDARY = 0, MCO_WALL = 1, MCO_INLET = 2, MCO_OUTLET = 3 };

/* Semantics of the pressure-correction loop.
 *  COMPAT: the reference CPU path as written (frozen pressure system after the
 *          first solve, cell correction on cell 0 of the first loop only).
 *  SIMPLE: the pressure-correction system is assembled from the current mass
 *          fluxes and its current RHS is solved every loop; every cell is
 *          corrected every loop. */
enum { MCO_MODE_COMPAT = 0, MCO_MODE_SIMPLE = 1 };

typedef struct { double x, y, z; } mco_vec;

typedef struct mco_mesh {
    int64_t m;                 /* cells modifier                      */
    int64_t nx, ny, nz;        /* 2m, m, m                            */
    int64_t n_cells, n_faces, n_points;
    mco_vec  *points;          /* [n_points]                          */
    uint64_t *cells;           /* [8 n_cells] vertex ids A..H         */
    uint64_t *cell_neighbours; /* [6 n_cells], UINT64_MAX = boundary  */
    uint64_t *faces;           /* [2 n_faces] (cell0, cell1)          */
    uint64_t *cell_faces;      /* [6 n_cells]                         */
    uint8_t  *cells_per_point; /* [n_points]                          */
    mco_vec  *cell_centers;    /* [n_cells]                           */
    uint64_t boundary_types[6];
} mco_mesh;

typedef struct mco_opts {
    int    mode;             /* MCO_MODE_COMPAT / MCO_MODE_SIMPLE            */
    double delta;            /* time step (src/minicombust.cpp:120: 1e-5)     */
    double mom_rtol;         /* BiCGStab relative tolerance (momentum,scalars)*/
    int    mom_maxit;
    double p_rtol;           /* pressure-correction relative tolerance        */
    int    p_maxit;
    int    solve_scalars;    /* 1: run the 7 scalar transport solves          */
    int    p_solver;         /* 0: Jacobi-PCG (checker), 1: aggregation-AMG PCG */
    int    limit_grad;       /* 1: Venkatakrishnan limiter after the gradients
                                (FLOW_SOLVER_LIMIT_GRAD, utils.hpp:26; off there) */
    int    warm_start;       /* 1: momentum and scalar BiCGStab start from the
                                current field (the library's simple-mode option;
                                the reference starts from zero,
                                FlowSolver.inl:1348)                          */
} mco_opts;

typedef struct mco_stats {
    int64_t mom_iters;       /* accumulated BiCGStab iterations */
    int64_t scalar_iters;
    int64_t p_iters;         /* accumulated pressure CG iterations */
    int64_t p_solves;
    int64_t p_loops;         /* correction-loop passes */
    double  last_p_relres;
} mco_stats;

typedef struct mco_state {
    mco_mesh mesh;
    mco_opts opts;
    mco_stats stats;
    int64_t N, F;
    /* geometry (FlowSolver.hpp:252-398) */
    double  *face_areas, *face_lambdas, *face_rlencos;
    mco_vec *face_normals, *face_centers;
    double  *cell_volumes, *cell_densities;
    /* fields, length N + 6 (6 boundary slots after the cells) */
    double  *phi[MCO_NFIELDS];
    mco_vec *grad[MCO_NFIELDS];       /* grad[MCO_TP] unused */
    double  *A_U, *A_V, *A_W, *S_U, *S_V, *S_W;
    double  *ff0, *ff1;               /* face_fields (cell0, cell1) */
    double  *mflux;                   /* face_mass_fluxes */
    /* CSR pattern of the PETSc SeqAIJ matrix: diag + interior neighbours,
       columns sorted (MatSetValue into MatSeqAIJSetPreallocation(A,7)). */
    int64_t  nnz;
    int64_t *row_ptr;                 /* [N+1] */
    int64_t *col;                     /* [nnz] */
    int64_t *diag_idx;                /* [N] position of (i,i) */
    int64_t *face_idx01;              /* [F] position of (cell0,cell1), -1 for boundary */
    int64_t *face_idx10;              /* [F] position of (cell1,cell0) */
    double  *A_vals, *b;              /* momentum / scalar matrix "A", vector "b" */
    double  *PA_vals, *Pb;            /* "Pressure_A", "Pressure_b" */
    double  *u;                       /* solution vector "u" */
    int      first_pressure;
    int64_t  timestep_count;
    double   effective_viscosity, inlet_effective_viscosity;
    double   pressure_correction_max_last;
    int      pressure_loops_last;
    double   p_atol;                  /* absolute floor of the current solve */
    double   p_bnorm_last;            /* ||b - mean b|| of the last solve     */
} mco_state;

/* ---- mesh (src/examples/mesh_examples.cpp:13-571 at one flow rank) ---- */
int  mco_mesh_build(int64_t m, mco_mesh *out);
void mco_mesh_free(mco_mesh *mesh);

/* ---- solver state ---- */
void mco_default_opts(mco_opts *o);
mco_state *mco_create(int64_t m, const mco_opts *opts);
void mco_destroy(mco_state *s);

/* ---- stages (names follow FlowSolver.inl) ---- */
void mco_get_phi_gradients(mco_state *s);
void mco_limit_phi_gradients(mco_state *s);
void mco_repair_nonneg(mco_state *s, int field);
void mco_particle_terms_add(mco_state *s, int64_t n, const uint64_t *cells, const double *terms, double *pt);
void mco_interpolate_to_nodes(mco_state *s, int64_t ncells, const uint64_t *cells, int64_t nnodes,
                              const uint64_t *nodes, double *out);
void mco_get_phi_gradient(mco_state *s, int field, int pressure);
void mco_precomp_AU(mco_state *s);
void mco_set_up_field(mco_state *s);
void mco_calculate_flux_UVW(mco_state *s);
void mco_momentum_assemble(mco_state *s);   /* calculate_UVW minus the solves */
void mco_momentum_solve(mco_state *s, int comp); /* comp 0,1,2 */
void mco_calculate_UVW(mco_state *s);
void mco_calculate_mass_flux(mco_state *s);
void mco_setup_pressure_matrix(mco_state *s);
void mco_solve_pressure_matrix(mco_state *s);
void mco_update_P_at_boundaries(mco_state *s, int field);
void mco_update_P(mco_state *s, int field);
void mco_update_mass_flux(mco_state *s);
void mco_apply_pressure_correction(mco_state *s);   /* one loop's corrections from phi[PP] */
void mco_calculate_pressure(mco_state *s);
void mco_flux_scalar(mco_state *s, int type, int field);
void mco_solve_turbulence_models(mco_state *s, int type);
void mco_scalar_assemble(mco_state *s, int type, int field); /* Scalar_solve minus solve */
void mco_scalar_solve(mco_state *s, int type, int field);
void mco_simple_iteration(mco_state *s, int inner);   /* one pass of timestep() */
void mco_timestep(mco_state *s);                       /* 3 passes */

/* linear algebra used by the stages (exported for tests) */
int  mco_bicgstab(int64_t n, const int64_t *row_ptr, const int64_t *col,
                  const double *vals, const double *b, double *x,
                  double rtol, int maxit, double *relres);
/* the same with warm != 0: x holds the starting guess on entry */
int  mco_bicgstab_x0(int64_t n, const int64_t *row_ptr, const int64_t *col,
                     const double *vals, const double *b, double *x,
                     double rtol, int maxit, double *relres, int warm);
int  mco_pcg_singular(int64_t n, const int64_t *row_ptr, const int64_t *col,
                      const double *vals, const double *b, double *x,
                      double rtol, double atol, int maxit, double *relres, double *bnorm);
int  mco_pcg_amg(int64_t n, const int64_t *row_ptr, const int64_t *col,
                 const double *vals, const double *b, double *x,
                 double rtol, double atol, int maxit, double *relres, double *bnorm);
void mco_spmv(int64_t n, const int64_t *row_ptr, const int64_t *col,
              const double *vals, const double *x, double *y);

/* accessors for ctypes */
double  *mco_phi(mco_state *s, int field);
mco_vec *mco_grad(mco_state *s, int field);
int64_t  mco_n_cells(mco_state *s);
int64_t  mco_n_faces(mco_state *s);

#ifdef __cplusplus
}
#endif
#endif
