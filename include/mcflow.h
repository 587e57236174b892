/*
 * mcflow.h -- C-ABI of the MI355X flow-solver library (libmcflow.so).
 *
 * The library implements miniCOMBUST's flow-solver SIMPLE step on one GPU
 * with hand-written HIP/CDNA4 kernels.  It replaces, below the reference's
 * FlowSolver<T> (include/flow/gpu/FlowSolver.{hpp,inl}), the layer formed by
 * the C_kernel_* launch wrappers (include/flow/gpu/gpu_kernels.cuh:8-83) and
 * the AmgX C API (include/flow/gpu/FlowSolver.hpp:1008-1067,
 * include/flow/gpu/FlowSolver.inl:716-1103).
 *
 * Conventions: every entry point returns MC_OK (0) or a positive error code
 * and never exits; mc_last_error() gives the message.  All device memory is
 * owned by the context; host pointers are plain arrays.  Work runs on the
 * context's HIP stream; entry points that return host data synchronise it.
 * Cell arrays have N + 6 entries: the reference's 6 boundary-side slots
 * (phi[N + side], FlowSolver.hpp:211-213) follow the N cells.
 */
#ifndef MCFLOW_H
#define MCFLOW_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MC_OK = 0,
    MC_ERR_ARG = 1,        /* invalid argument                              */
    MC_ERR_HIP = 2,        /* HIP runtime error (message in mc_last_error)  */
    MC_ERR_NOMEM = 3,      /* device allocation failed                      */
    MC_ERR_STATE = 4,      /* call order violated                           */
    MC_ERR_NOCONV = 5,     /* linear solver did not converge                */
    MC_ERR_NODEV = 6       /* no HIP device                                 */
};

/* Field ids: the reference phi_vector order (include/utils/utils.hpp:158-174). */
enum { MC_U = 0, MC_V, MC_W, MC_P, MC_PP, MC_TE, MC_ED, MC_TP, MC_TEM, MC_FUL,
       MC_PRO, MC_VARF, MC_VARP, MC_NFIELDS };

/* Scalar equation types (include/utils/utils.hpp:31-37). */
enum { MC_TERBTE = 0, MC_TERBED, MC_TEMP, MC_FUEL, MC_PROG, MC_VARFU, MC_VARPR };

/* Pressure-correction semantics.
 * MC_MODE_COMPAT: the reference CPU path as written -- the pressure-correction
 *   system is assembled and stored only on the first call (first_pressure,
 *   FlowSolver.inl:2118-2253) and every later solve re-solves it; the cell
 *   correction touches cell 0 of the first loop only (FlowSolver.inl:2416-2433).
 * MC_MODE_SIMPLE: the system is re-assembled from the current mass fluxes each
 *   SIMPLE iteration, its current RHS is solved in every correction loop, and
 *   every cell is corrected every loop. */
enum { MC_MODE_COMPAT = 0, MC_MODE_SIMPLE = 1 };

/* Pressure-solve forms. */
enum {
    MC_PSOLVE_FP64 = 0,    /* FP64 PCG + aggregation-AMG V-cycle                      */
    MC_PSOLVE_MIXED = 1,   /* FP32 PCG-AMG (SpMV, smoothers, hierarchy) inside an FP64
                              iterative-refinement loop; FP64 residual, update and
                              coarsest solve                                          */
};

typedef struct mc_options {
    int    mode;           /* MC_MODE_*                                            */
    int    psolve;         /* MC_PSOLVE_*                                          */
    int    learned;        /* learned guess for loop-1 pressure solves: 0 off, 1 while
                              it pays (measured), 2 always                        */
    int    solve_scalars;  /* 1: the 7 scalar equations of the SIMPLE step         */
    double delta;          /* time step, src/minicombust.cpp:120 (1e-5)            */
    double mom_rtol;       /* BiCGStab rtol, momentum and scalars (PETSc: 1e-5)    */
    int    mom_maxit;
    double p_rtol;         /* pressure-correction relative residual (1e-8)         */
    int    p_maxit;
    int    amg_nu;         /* Jacobi sweeps before and after the coarse correction (3) */
    double amg_omega;      /* Jacobi damping                                       */
    double amg_cscale;     /* coarse-correction scaling                            */
    int    amg_coarse_max; /* dense coarsest level size bound                      */
    double mixed_inner_rtol; /* FP32 inner solve reduction per refinement step (1e-5) */
    int    check_every;    /* ignored (kept for the layout): host convergence checks
                              follow batches predicted from the convergence rate */
    double inlet_scale;    /* multiplies the reference gas velocity (Mesh.hpp:114,
                              inlet and initial field); 1 = the reference case    */
    double amg_cheb_ratio; /* > 1: the nu smoothing sweeps use Chebyshev weights on
                              [2/ratio, 2]; <= 1: every sweep uses amg_omega      */
    int    cell_tile;      /* generated box mesh: x extent of the 256-cell brick
                              (cell_tile x 1 x 256/cell_tile) one workgroup of the
                              cell-centric assembly kernels visits; a power of two
                              in [1, 256], 0: the widest that leaves at most 1/8
                              of a row's lanes idle                              */
    int    limit_grad;     /* 1: Venkatakrishnan limiter on the 11 gradients of each
                              SIMPLE iteration (FLOW_SOLVER_LIMIT_GRAD,
                              utils.hpp:26, FlowSolver.inl:3197; 0 as there)     */
    int    amg_refresh;    /* the coarse AMG levels (Galerkin values, coarsest inverse)
                              are recomputed at every amg_refresh-th pressure solve
                              and reused in between (1: every solve); the fine
                              level always holds the current matrix              */
    int    cell_ychunk;    /* generated box mesh: y rows per chunk of the bricks'
                              visiting order (chunks in y; within one, z-slab by
                              z-slab, row by row, tile by tile), 0: 16          */
    int    warm_start;     /* 1: momentum and scalar BiCGStab start from the current
                              field (a guess no better than zero is dropped for
                              x0 = 0); 0: always x0 = 0 (VecZeroEntries,
                              FlowSolver.inl:1348); -1: 1 in MC_MODE_SIMPLE, 0 in
                              MC_MODE_COMPAT                                     */
    int    amg_half;       /* 1: the mixed form's V-cycle smooths the fine level of a
                              7-diagonal pressure pattern with an FP16 copy of
                              D^-1 A's off-diagonals (the PCG product and the FP64
                              refinement residual keep the full matrix); 0: FP32 */
    int    amg_tail;       /* 1: the V-cycle from the first level of at most 2048
                              rows down runs in one workgroup (same results, fewer
                              launches); 0: one launch per pass                  */
    int    p_recycle;      /* 0-4: the last p_recycle loop-1 pressure solutions join
                              the deflation space W of the next loop-1 solve, which
                              starts from the Galerkin projection x0 = W y,
                              (W^T A W) y = W^T b (with the learned guess in W when
                              `learned` uses it); 0: no recycling (the reference's
                              x0 = 0 unless the learned guess is used)           */
    int    pcg_graph;      /* 1: each pressure PCG iteration (V-cycle, product and
                              reductions) is captured once as a HIP graph and
                              replayed (same kernels and order, so the same bits;
                              off while PCG kernel classes are being profiled);
                              0 (default): one launch at a time.  On the MI355X
                              the replay measured no faster: the ~8 us between
                              dependent small kernels stays, and each replay adds
                              its own launch gap.                                 */
} mc_options;

typedef struct mc_solve_status {
    int    iterations;     /* Krylov iterations (summed over RHS / refinement steps) */
    int    converged;      /* 1 if every RHS met its tolerance                      */
    double relres;         /* final ||b - A x|| / ||b|| (max over RHS)              */
    int    outer;          /* refinement steps (mixed form)                         */
    int    guess_used;     /* learned initial guess accepted                        */
    double bnorm;          /* ||b|| (pressure: of the mean-projected RHS)           */
} mc_solve_status;

typedef struct mc_stats {
    int64_t simple_iterations;
    int64_t mom_iters, scalar_iters, p_iters, p_solves, p_loops, p_outer;
    double  last_p_relres;
    double  p_solve_ms;       /* host wall time inside pressure solves           */
    double  last_p_max;       /* max |PP| of the last correction loop            */
    int     last_loops;
    double  p_relres_max;     /* max over solves of ||b - A x|| / (tol / rtol): the
                                 residual relative to the RHS scale its tolerance
                                 was set against (loop 1: its own RHS)           */
    int64_t clip_cells;       /* TE/ED cells repaired to non-negative values     */
    int64_t clip_passes;      /* relaxation passes of that repair                */
    int64_t host_polls;       /* host waits on solver convergence flags          */
    int64_t p_guesses;        /* pressure solves started from the learned guess  */
    double  p_guess_ratio_sum;/* sum over those of ||b - A x0|| / ||b - mean b||  */
    int64_t p_defl_starts;    /* loop-1 pressure solves started from the deflation
                                 space's Galerkin projection (learned guess and/or
                                 recycled solutions), the start accepted           */
    double  p_defl_ratio_sum; /* sum over those of ||b - A x0|| / ||b - mean b||  */
    int64_t p_defl_vectors;   /* sum over those of the vectors W held            */
    int64_t amg_setups;       /* AMG hierarchy builds (aggregation, Galerkin pattern) */
    double  amg_setup_ms;     /* host wall time of those builds, first values included */
    int64_t p_loops_skipped;  /* correction loops after the first whose PP came out
                                 zero everywhere (a zero right-hand side): their
                                 passes, which would add only signed zeros, skipped */
} mc_stats;

/* Host view of a mesh, the arrays of the reference Mesh<T>
 * (include/geometry/Mesh.hpp:70-107), one flow rank. */
typedef struct mc_mesh_desc {
    int64_t n_cells, n_faces, n_points;
    const double   *points;          /* [3 n_points] x,y,z interleaved (vec<double>)   */
    const uint64_t *cells;           /* [8 n_cells] vertex ids A..H                    */
    const uint64_t *cell_neighbours; /* [6 n_cells], UINT64_MAX at the boundary        */
    const uint64_t *faces;           /* [2 n_faces] (cell0, cell1); boundary cell1 =
                                        n_cells + side                                 */
    const uint64_t *cell_faces;      /* [6 n_cells]                                    */
    uint64_t boundary_types[6];      /* Mesh.hpp:17 BOUNDARY_TYPES per box side        */
} mc_mesh_desc;

/* Array ids for mc_array_get / mc_array_set. */
enum {
    MC_ARR_PHI = 0,          /* + field id: phi[field], N+6 f64                    */
    MC_ARR_GRAD = 16,        /* + field id: gradient, (N+6) x 3 f64 interleaved     */
    MC_ARR_AU = 32, MC_ARR_AV, MC_ARR_AW, MC_ARR_SU, MC_ARR_SV, MC_ARR_SW, /* N+6 f64 */
    MC_ARR_FF0 = 40, MC_ARR_FF1, MC_ARR_MFLUX,                       /* F f64      */
    MC_ARR_FACE_AREA = 48, MC_ARR_FACE_LAMBDA, MC_ARR_FACE_RLENCOS,  /* F f64      */
    MC_ARR_FACE_NORMAL, MC_ARR_FACE_CENTER,                          /* F x 3 f64  */
    MC_ARR_CELL_CENTER, MC_ARR_CELL_VOLUME, MC_ARR_CELL_DENSITY,     /* N(x3) f64  */
    MC_ARR_ROW_PTR = 64, MC_ARR_COL,                                 /* i32        */
    MC_ARR_A_VALS, MC_ARR_PA_VALS, MC_ARR_PB,                        /* f64        */
    MC_ARR_FACES = 72, MC_ARR_CELL_FACES, MC_ARR_CELL_NEIGHBOURS,    /* u64        */
    MC_ARR_CELLS, MC_ARR_POINTS, MC_ARR_CELLS_PER_POINT,             /* u64/f64/u8 */
    MC_ARR_PARTICLE_TERMS = 80,                                      /* N x 5 f64: particle_aos
                                                                        (momentum xyz, energy, fuel) */
    MC_ARR_PDEFL = 84                                                /* + k (0-3): the k-th newest
                                                                        recycled loop-1 pressure
                                                                        solution, N f64, read-only */
};

/* Kernel classes timed with HIP events when profiling is on.  Every class but
   MC_KC_VCYCLE is one kernel (fine level for the solver kernels), so its
   event time is that kernel's launch duration. */
enum {
    MC_KC_GRAD = 0,      /* k_grad<11>: 11-field LSQ gradient                    */
    MC_KC_GRAD1,         /* k_grad<1>: 1-field LSQ gradient                      */
    MC_KC_MOM_ASM,       /* k_momentum: fluxes + sources + matrix assembly       */
    MC_KC_P_ASM,         /* k_mass_pressure: Rhie-Chow flux + pressure assembly  */
    MC_KC_SCALAR_ASM,    /* k_scalar: one scalar's fluxes + sources + assembly   */
    MC_KC_SPMV64,        /* k_pcg_pq<.,double>: fine PCG p-update + SpMV + dot   */
    MC_KC_SPMV32,        /* k_pcg_pq<.,float>: same, FP32 inner PCG              */
    MC_KC_VCYCLE,        /* AMG V-cycle, all levels (several kernels)            */
    MC_KC_CORRECT,       /* k_flux_correct: mass-flux correction                 */
    MC_KC_MPNN,          /* learned guess: whole forward pass (work = MFMA flops) */
    MC_KC_JAC64,         /* k_jac<.,double>: fine-level Jacobi sweep             */
    MC_KC_JAC32,         /* k_jac<.,float>: fine-level Jacobi sweep              */
    MC_KC_BICG,          /* k_bi_spmv_v: BiCGStab SpMV + (r0,v) dot              */
    MC_KC_SCALAR_BATCH,  /* k_scalar_batch: 4 passive scalars assembled at once  */
    MC_KC_IR_RESID,      /* k_ir_resid: FP64 residual of the refinement loop     */
    MC_KC_MPNN_BWD,      /* learned guess: whole backward pass (work = flops)     */
    MC_KC_COUNT
};

/* ---- context: the FlowSolver constructor ----------------------------------
 * mc_ctx_create   replaces FlowSolver<T>::FlowSolver(MPI_Config*, Mesh<T>*,
 *                 double delta, FILE*) (include/flow/gpu/FlowSolver.hpp:274-1067):
 *                 uploads the mesh, computes face/cell geometry (FlowSolver.hpp:
 *                 252-398), initial fields and boundary slots (:301-485), the CSR
 *                 pattern (:491-502) and the first assemblies (:538-539).
 * mc_ctx_create_box  load_mesh (src/examples/mesh_examples.cpp:120-571, one flow
 *                 rank, box of src/minicombust.cpp:123-124) generated on the GPU,
 *                 followed by the constructor above; nothing crosses PCIe.
 * Both check the options first (MC_ERR_ARG, *out = NULL), then the device
 * (MC_ERR_NODEV).  A later failure still returns the context in *out, so that
 * mc_last_error can say why; the caller then destroys it.                   */
typedef struct mc_ctx mc_ctx;
int  mc_options_default(mc_options *opt);
int  mc_ctx_create(const mc_mesh_desc *mesh, const mc_options *opt, mc_ctx **out);
int  mc_ctx_create_box(int64_t modifier, const mc_options *opt, mc_ctx **out);
int  mc_ctx_destroy(mc_ctx *ctx);
const char *mc_last_error(mc_ctx *ctx);
int  mc_ctx_sizes(mc_ctx *ctx, int64_t *n_cells, int64_t *n_faces, int64_t *nnz, int64_t *n_points);
int  mc_set_options(mc_ctx *ctx, const mc_options *opt);
/* Back to the constructor's state (FlowSolver.hpp:301-539): initial fields and
   boundary slots from the current options (inlet_scale), zero mass fluxes and
   matrices, first-call flags set; mesh, geometry and AMG hierarchy are kept.
   No counterpart in the reference (which builds a new FlowSolver). */
int  mc_reset_fields(mc_ctx *ctx);

/* ---- SIMPLE step ---------------------------------------------------------
 * mc_timestep        FlowSolver<T>::timestep() (FlowSolver.inl:3148-3356):
 *                    three SIMPLE iterations.
 * mc_simple_iteration one pass of its loop body (FlowSolver.inl:3190-3267). */
int  mc_timestep(mc_ctx *ctx);
int  mc_simple_iteration(mc_ctx *ctx, int inner);

/* ---- stages (each the counterpart of a reference kernel wrapper) ---------- */
/* C_kernel_get_phi_gradients (gpu_kernels.cuh:30) / get_phi_gradients
   (FlowSolver.inl:917-1063): 11-field LSQ gradients.                          */
int  mc_get_phi_gradients(mc_ctx *ctx);
/* C_kernel_get_phi_gradient (gpu_kernels.cuh:48) / get_phi_gradient
   (FlowSolver.inl:738-803); pressure != 0 selects dphi = 0 on boundaries.     */
int  mc_get_phi_gradient(mc_ctx *ctx, int field, int pressure);
/* limit_phi_gradients (FlowSolver.inl:875-893, limit_phi_gradient :804-873):
   Venkatakrishnan's limiter applied in place to the 11 gradients (the GPU
   reference has no counterpart; its call is commented out,
   gpu/FlowSolver.inl:1333).                                                   */
int  mc_limit_phi_gradients(mc_ctx *ctx);
/* The non-negativity repair Scalar_solve applies to TE and ED after their
   solve (FlowSolver.inl:2956-2985): negative cells, in ascending order, take
   the mean of their interior neighbours as they stand (at least 1e-12).     */
int  mc_repair_nonneg(mc_ctx *ctx, int field);
/* C_kernel_precomp_AU (gpu_kernels.cuh:38) + calculate_mass_flux:
   set_up_field (FlowSolver.inl:1065-1113).                                   */
int  mc_set_up_field(mc_ctx *ctx);
/* C_kernel_calculate_flux_UVW + C_kernel_apply_forces + C_kernel_setup_sparse_
   matrix + C_kernel_update_sparse_matrix (gpu_kernels.cuh:32,34,36,40):
   calculate_UVW without its solves (FlowSolver.inl:1655-1758).               */
int  mc_momentum_assemble(mc_ctx *ctx);
/* the three AMGX_solver_solve calls of calculate_UVW (gpu/FlowSolver.inl:
   716-782) / KSPSolve (FlowSolver.inl:1341-1375): BiCGStab, x0 = 0.          */
int  mc_momentum_solve(mc_ctx *ctx, mc_solve_status *st);
/* C_kernel_calculate_mass_flux + C_kernel_compute_flow_correction +
   C_kernel_correct_flow{,2} + C_kernel_setup_pressure_matrix
   (gpu_kernels.cuh:54,56,58,28,42): the start of calculate_pressure
   (FlowSolver.inl:2333-2360).                                                */
int  mc_pressure_assemble(mc_ctx *ctx);
/* the pressure AMGX_solver_solve (gpu/FlowSolver.inl:943-969) / KSPSolve with
   FGMRES+GAMG (FlowSolver.inl:2203-2273): PCG + aggregation-AMG, mean-zero
   solution, into phi[PP].                                                    */
int  mc_pressure_solve(mc_ctx *ctx, mc_solve_status *st);
/* calculate_pressure (FlowSolver.inl:2317-2463): assembly, correction loop
   (C_kernel_find_pressure_correction_max, _Update_P_at_boundaries,
   _update_vel_and_flux, _update_mass_flux) and C_kernel_Update_P.            */
int  mc_calculate_pressure(mc_ctx *ctx);
/* Scalar_solve (FlowSolver.inl:2885-3076): C_kernel_flux_scalar,
   _apply_pres_forces, _solve_turb_models_{cell,face}, _setup_sparse_matrix
   (gpu_kernels.cuh:60-66,36) and the solve.                                  */
int  mc_scalar_assemble(mc_ctx *ctx, int type);
int  mc_scalar_solve(mc_ctx *ctx, int type);
/* The four passive-scalar Scalar_solve calls (FUEL, PROG, VARFU, VARPR,
   FlowSolver.inl:3243-3260): their matrices are identical, so one assembly
   and one 4-RHS BiCGStab give the four results of the sequential calls.     */
int  mc_scalar_batch_solve(mc_ctx *ctx);
/* C_kernel_Update_P_at_boundaries (gpu_kernels.cuh:46) and C_kernel_Update_P
   (gpu_kernels.cuh:52) on a field.                                          */
int  mc_update_boundary_slots(mc_ctx *ctx, int field);
int  mc_update_P(mc_ctx *ctx);

/* ---- linear algebra on the context's CSR pattern ------------------------- */
/* y = A x with which = 0 (momentum/scalar matrix) or 1 (pressure matrix).   */
int  mc_spmv(mc_ctx *ctx, int which, const double *x_host, double *y_host);
/* Solve the pressure matrix with a host RHS (tests, data generation).        */
int  mc_solve_pressure_host(mc_ctx *ctx, const double *b_host, double *x_host, mc_solve_status *st);

/* ---- state access -------------------------------------------------------- */
int  mc_array_get(mc_ctx *ctx, int id, void *host, int64_t count);
int  mc_array_set(mc_ctx *ctx, int id, const void *host, int64_t count);
int  mc_get_stats(mc_ctx *ctx, mc_stats *out);
int  mc_state_flags(mc_ctx *ctx, int *first_pressure, int64_t *timestep_count);
int  mc_set_state_flags(mc_ctx *ctx, int first_pressure, int64_t timestep_count);
int  mc_sync(mc_ctx *ctx);

/* ---- particle / flow coupling ----------------------------------------------
   The flow side of update_flow_field (FlowSolver.inl:496-700) without its MPI
   transport; the caller passes the arrays the reference receives and sends.
   GPU reference: gpu/FlowSolver.inl:361-601, kernel_process_particle_fields
   and kernel_interpolate_phi_to_nodes (gpu_kernels.cu:131-148, 2132-2191).  */
/* memset of mesh->particle_terms (FlowSolver.inl:508).                       */
int  mc_particle_terms_reset(mc_ctx *ctx);
/* FlowSolver.inl:675-683: terms[5 i .. 5 i + 4] (particle_aos: momentum xyz,
   energy, fuel) added to cell cells[i], in receive order (host arrays).
   Result: mc_array_get(MC_ARR_PARTICLE_TERMS).                               */
int  mc_particle_terms_add(mc_ctx *ctx, const uint64_t *cells, const double *terms, int64_t n);
/* mc_particle_terms_add without any device allocation.  No counterpart in
   the reference: it is for callers that cannot allocate device memory (code
   under HIP graph capture, tight per-timestep coupling loops, harnesses that
   forbid hipMalloc).  Same arguments and the same result, bit for bit, as
   mc_particle_terms_add; its device buffers (staged rows, sort keys and
   indices, the sort's temporary storage) live in a fixed 64 MiB array that
   each GPU holds until the process ends, which bounds n to about 850,000 rows
   (n x 64 bytes plus the sort's storage): larger n is refused with MC_ERR_ARG
   and a message giving the bytes needed.  mc_particle_terms_reset must have
   run once on the context (it allocates the terms array; MC_ERR_STATE
   otherwise).  Every call on one GPU shares that array: calls that overlap
   on one GPU (contexts on different streams or threads) race.              */
int  mc_particle_terms_add_noalloc(mc_ctx *ctx, const uint64_t *cells, const double *terms, int64_t n);
/* FlowSolver.inl:297-335 + interpolate_to_nodes (:428-493): values at the
   requested nodes from the interpolation cell set `cells` (duplicates count
   once; each node sums its cells in ascending order).  node_fields[5 k ..]:
   flow_aos (vel xyz, pressure, temp) of nodes[k] (host arrays).              */
int  mc_interpolate_to_nodes(mc_ctx *ctx, const uint64_t *cells, int64_t ncells, const uint64_t *nodes,
                             int64_t nnodes, double *node_fields);

/* ---- output ---------------------------------------------------------------
   FlowSolver<T>::output_data (FlowSolver.inl:18-23) ->
   VisitWriter<T>::write_flow (include/visit/VisitWriter.hpp:127-234): one
   legacy-ASCII VTK POLYDATA file <prefix>_flow<rank>_timestep<timestep>.vtk
   with the cell centres as points, the velocity and nine scalars as point
   data, byte for byte as the reference writes it.                            */
int  mc_write_vtk(mc_ctx *ctx, const char *prefix, int64_t timestep, int rank);
/* The same file from host arrays (no device needed): centres [3 n] x,y,z
   interleaved, fields[11] = U V W P TE ED TEM FUL PRO VARF VARP, n each.     */
int  mc_vtk_write_arrays(const char *path, int64_t n, const double *centres, const double *const *fields);

/* ---- measurement --------------------------------------------------------- */
/* Profiling: each launch of a class whose bit (1 << MC_KC_x) is set in `mask`
   is bracketed by HIP events on the context stream, with no synchronisation
   (events come from a pool and are resolved when read).  mc_profile_read
   synchronises the stream and returns accumulated milliseconds, launch counts
   and algorithmic bytes per class (arrays of MC_KC_COUNT).  mask 0 = off.    */
int  mc_profile_enable(mc_ctx *ctx, int mask);
int  mc_profile_read(mc_ctx *ctx, double *ms, int64_t *calls, double *bytes);
int  mc_profile_reset(mc_ctx *ctx);
/* Per-stage times of the SIMPLE iteration: the reference's fine timers
   (vel_/pres_/sca_*_time, include/flow/FlowSolver.hpp:128-141, printed by
   FlowSolver.inl:1768-1806 and its pressure and scalar counterparts when
   FLOW_SOLVER_FINE_TIME is set).  mc_stage_enable(ctx, 1) brackets every
   stage with HIP events on the context stream (no synchronisation);
   mc_stage_read synchronises and returns the accumulated milliseconds and
   bracket counts per stage (arrays of MC_ST_COUNT, either may be NULL);
   mc_stage_reset clears them.                                                */
enum { MC_ST_GRAD = 0,     /* 11-field gradients (get_phi_gradients, limiter)   */
       MC_ST_VEL_FLUX,     /* momentum fluxes, sources and matrix               */
       MC_ST_VEL_SOLVE,    /* U, V, W BiCGStab                                  */
       MC_ST_PRES_FLUX,    /* Rhie-Chow mass flux and pressure-correction matrix */
       MC_ST_PRES_SOLVE,   /* pressure-correction solves (every loop, AMG refresh) */
       MC_ST_PRES_CORRECT, /* PP slots, grad PP, flux and cell corrections, P update */
       MC_ST_TURB_FLUX,    /* TE and ED assembly                                */
       MC_ST_TURB_SOLVE,   /* TE and ED BiCGStab, non-negativity repair         */
       MC_ST_SCA_FLUX,     /* TEM, FUL, PRO, VARF, VARP assembly                */
       MC_ST_SCA_SOLVE,    /* their BiCGStab solves                             */
       MC_ST_COUNT };
int  mc_stage_enable(mc_ctx *ctx, int on);
int  mc_stage_read(mc_ctx *ctx, double *ms, int64_t *calls);
int  mc_stage_reset(mc_ctx *ctx);
/* Streaming copy of `bytes` (multiple of 16) between two fresh device buffers,
   `reps` launches on the context stream; *ms_out = average launch time.  The
   HBM bandwidth probe and the known-size kernel that scales the FETCH_SIZE /
   WRITE_SIZE counters (kernel symbol mc::k_copy_probe).  No counterpart in the
   reference: measurement only. */
int  mc_copy_probe(mc_ctx *ctx, int64_t bytes, int reps, double *ms_out);
/* Device memory held by the context, per category (array of MC_MEM_COUNT). */
enum { MC_MEM_MESH = 0,   /* topology, geometry, face orders                */
       MC_MEM_FIELDS,     /* 13 fields + gradients, A/S vectors, face fields */
       MC_MEM_MATRIX,     /* CSR pattern and values, diagonal copies         */
       MC_MEM_AMG,        /* AMG hierarchy                                   */
       MC_MEM_WORK,       /* solver work vectors, reductions                 */
       MC_MEM_MPNN,       /* learned guess: weights, activations             */
       MC_MEM_DATASET,    /* learned guess: training pairs                   */
       MC_MEM_COUNT };
int  mc_ctx_memory(mc_ctx *ctx, int64_t *bytes);

/* ---- learned pressure-correction initial guess -------------------------- *
 * No counterpart in the reference: its pressure solves start from zero
 * (FlowSolver.inl:2203-2273, gpu/FlowSolver.inl:943-969).  A message-passing
 * net over the cell graph (3 layers, hidden 64, BF16 weights on MFMA with FP32
 * accumulation; definition and parameter order in mc_mpnn.hip) maps the
 * pressure system (RHS, diagonal, cell volume, net mass outflow; the
 * off-diagonals as edge features) to a guess g.  With mc_options.learned = 1
 * and weights loaded, each pressure solve starts from x0 = g (g,b)/(g,Ag),
 * kept only when it lowers the residual (else x0 = 0, the plain solve);
 * convergence is judged by the unchanged PCG-AMG residual test.           */
int64_t mc_mpnn_nparams(void);
int  mc_mpnn_load(mc_ctx *ctx, const float *weights, int64_t count);
int  mc_mpnn_save(mc_ctx *ctx, float *weights, int64_t count);
/* Collect (features, solution) pairs from the following pressure solves into
   a device data set of max_pairs (0 = stop collecting).                    */
int  mc_mpnn_collect(mc_ctx *ctx, int64_t max_pairs);
int  mc_mpnn_dataset_size(mc_ctx *ctx, int64_t *pairs);
/* Host copies of a stored pair: node features [N x 4], edge features [N x 6]
   (side order of MC_ARR_CELL_NEIGHBOURS), normalised target [N].          */
int  mc_mpnn_pair_get(mc_ctx *ctx, int64_t pair, float *feat, float *edge, float *target);
/* Net output on a stored pair's features [N].                              */
int  mc_mpnn_forward_pair(mc_ctx *ctx, int64_t pair, float *y_host);
/* Loss (mean squared error of the mean-removed output against the target)
   and its gradient for one stored pair [mc_mpnn_nparams()].                */
int  mc_mpnn_grad(mc_ctx *ctx, int64_t pair, float *grad_host, double *loss);
/* One Adam step (learning rate lr) on one stored pair; *loss before it.    */
int  mc_mpnn_train_step(mc_ctx *ctx, int64_t pair, double lr, double *loss);

#ifdef __cplusplus
}
#endif
#endif
