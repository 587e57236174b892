// mc_internal.hpp -- shared internals of the HIP flow-solver library.
//
// Device layout (one context = one GPU, one stream):
//   * cells are numbered as the reference mesh generator numbers them
//     (src/examples/mesh_examples.cpp:305-384: x fastest, one block);
//   * every per-cell and per-face quantity is a separate SoA array so that a
//     wavefront of 64 consecutive cells/faces issues fully coalesced loads;
//   * vector quantities (centres, normals, gradients) are three SoA arrays;
//   * cell arrays have N + 6 entries: the 6 boundary-side slots of the
//     reference's phi_vector (FlowSolver.hpp:211-213) sit after the cells;
//   * indices are int32 (N, F < 2^31 for every mesh that fits 288 GB).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/mcflow.h"

namespace mc {

constexpr int NFIELDS = 13;
enum Field { U = 0, V, W, P, PP, TE, ED, TP, TEM, FUL, PRO, VARF, VARP };
enum ScalarType { TERBTE = 0, TERBED, TEMP, FUEL, PROG, VARFU, VARPR };
enum BType { NOT_BOUNDARY = 0, WALL = 1, INLET = 2, OUTLET = 3 };

// Physical constants fixed by the reference (Mesh.hpp:114-120,
// FlowSolver.hpp:296-297,342).
struct Consts {
    double gas_vel_x, gas_vel_y, gas_vel_z, gas_pre, gas_tem, gas_te, gas_ed, gas_ful, gas_pro;
    double ev;            // effective_viscosity
    double inlet_ev;      // inlet_effective_viscosity (FlowSolver.hpp:421-422)
    double rdelta;        // 1 / delta
    double inlet_te;      // 3/2 (0.1^2 |U|^2)
    double inlet_ed;      // 0.09^0.75 inlet_te^1.5
    double cmu75;         // 0.09^0.75
};

constexpr int DIA_MAX = 32;     // most diagonals a diagonal-storage matrix may have
constexpr int RED_PARTIALS = 16 * 8192;   // doubles in the reduction partials buffer
constexpr int DENSE_MAX = 128;  // largest coarsest AMG level (inverted in LDS)
constexpr int DEFL_MAX = 4;     // recycled pressure solutions in the deflation space

// Per-level storage of the aggregation-AMG hierarchy.
struct AmgLevel {
    int n = 0, nnz = 0;
    int *row_ptr = nullptr, *col = nullptr;   // CSR pattern of this level
    int nd = 0;                               // diagonals of the diagonal copy (0: none)
    int off[DIA_MAX] = {};
    int sym = 0;                              // symmetric pattern: the FP64 copy stores only the
                                              // diagonals with offset >= 0 (nd / 2 + 1 arrays)
    int *dia_map = nullptr;                   // CSR index -> d*n + i
    double *dia = nullptr;
    float *dia32 = nullptr;
    double *vals = nullptr;                   // FP64 values
    float *vals32 = nullptr;                  // FP32 copy (mixed form)
    double *dinv = nullptr;                   // 1/diag
    float *dinv32 = nullptr;
    // fine level of a 7-diagonal pattern: the off-diagonals of D^-1 A in FP16
    // (6 arrays, main diagonal left out) and the FP32 diagonal, for the FP32
    // V-cycle's smoothing sweeps and residual
    _Float16 *hoff = nullptr;
    float *diag32 = nullptr;
    // aggregation to the next level
    int nc = 0;
    int *agg = nullptr;                       // [n] aggregate id
    int *agg_ptr = nullptr, *agg_members = nullptr; // coarse row -> fine members
    // Galerkin gather map for the next level's values: coarse nnz k sums
    // this level's values gmap[gptr[k] .. gptr[k+1])
    int *gptr = nullptr, *gmap = nullptr;
    // vectors
    double *x = nullptr, *b = nullptr, *t = nullptr;
    float *x32 = nullptr, *b32 = nullptr, *t32 = nullptr;
};

struct Ctx {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int device = -1;               // the HIP device current at creation
    // device address, on this context's device, of the fixed scratch array
    // of the allocation-free coupling entry point (mc_couple.hip)
    void *noalloc_scratch = nullptr;
    // the coarsest AMG level's inverse is formed on `side` (one workgroup)
    // while `stream` runs the first V-cycle's descent; ev_dense marks it done
    hipStream_t side = nullptr;
    hipEvent_t ev_coarse = nullptr, ev_dense = nullptr;
    bool dense_pending = false;
    // one PCG iteration captured as a HIP graph, per precision (0 FP32, 1 FP64)
    // and parity of the p buffers; valid while `gen` equals graph_gen (bumped
    // by a hierarchy build and by mc_set_options) and the vectors are the same
    struct PcgGraph {
        hipGraphExec_t exec = nullptr;
        uint64_t gen = 0;
        const void *key[8] = {};
    };
    PcgGraph pcg_graphs[2][2];
    uint64_t graph_gen = 1;
    mc_options opt{};
    Consts k{};
    std::string err;

    // ---- mesh ----
    int64_t m = 0;              // box modifier (0 when the mesh came from the host)
    int N = 0, F = 0, NP = 0, Fb = 0;
    int *fc0 = nullptr, *fc1 = nullptr;     // [F] face cells; boundary: fc1 = N + side
    // Face arrays are stored in an order of their own: on the generated box
    // side-major (x faces, y faces, z faces by creating cell, then the three
    // low boundary sides), so the k-th face of 64 consecutive cells is 64
    // consecutive entries; elsewhere in reference order.  Face indices held in
    // device arrays (cell_faces, forder, bfaces, last_face) are storage
    // positions; sums over faces still run in reference face-id order.
    int *face_ref = nullptr;                 // [F] reference face id of each storage position
    int *face_pos = nullptr;                 // [F] storage position of each reference face id
    std::vector<int> h_face_pos, h_face_ref; // host copies for array get / set
    int *cell_faces = nullptr;               // [6N] side order (FRONT..UP)
    int *forder = nullptr;                   // [6N] the cell's faces in ascending face id
    int *fother = nullptr;                   // [6N] other cell of each forder face (boundary: N + side),
                                             //      bit 31 set when the cell is the face's cell1
    int *nbr = nullptr;                      // [6N] neighbour per side, -1 at the boundary
    int *bfaces = nullptr;                   // [Fb] boundary faces, ascending
    int btype[6] = {0, 0, 0, 0, 0, 0};
    int last_face[6] = {-1, -1, -1, -1, -1, -1};   // highest face id per box side
    int last_cell[6] = {-1, -1, -1, -1, -1, -1};
    int wall_cells = 0;
    uint8_t *is_wall_cell = nullptr;         // [N]
    // kept for mesh download / geometry
    double *px = nullptr, *py = nullptr, *pz = nullptr;   // [NP] points
    uint64_t *cells = nullptr;               // [8N] vertex ids (freed after setup unless kept)
    uint8_t *cells_per_point = nullptr;

    // ---- geometry ----
    double *cx = nullptr, *cy = nullptr, *cz = nullptr;        // [N] cell centres
    double *fcx = nullptr, *fcy = nullptr, *fcz = nullptr;     // [F] face centres
    double *fnx = nullptr, *fny = nullptr, *fnz = nullptr;     // [F] area normals
    double *farea = nullptr, *flam = nullptr, *frlc = nullptr; // [F]
    double *vol = nullptr, *rho = nullptr;                     // [N]

    // ---- fields ----
    double *phi[NFIELDS] = {};
    double **field_ptrs = nullptr;   // device table: phi, gx, gy, gz of every field
    double *gx[NFIELDS] = {}, *gy[NFIELDS] = {}, *gz[NFIELDS] = {};
    double *AU = nullptr, *AV = nullptr, *AW = nullptr;
    double *SU = nullptr, *SV = nullptr, *SW = nullptr;
    double *ff0 = nullptr, *ff1 = nullptr, *mflux = nullptr, *fcorr = nullptr;

    // ---- CSR (PETSc SeqAIJ pattern) ----
    int nnz = 0;
    int *row_ptr = nullptr, *col = nullptr, *diag_idx = nullptr;
    double *A_vals = nullptr;      // momentum / scalar matrix
    double *A_vals2 = nullptr;     // temperature matrix, assembled beside the passive scalars'
    double *PA_vals = nullptr;     // pressure-correction matrix
    double *Pb = nullptr;          // pressure RHS (compat: frozen copy)
    double *u = nullptr;           // solution scratch
    int fine_nd = 0;               // diagonal copy of the fine pattern (0: CSR kernels)
    int fine_off[DIA_MAX] = {};
    int *fine_map = nullptr;
    double *A_dia = nullptr;       // momentum / scalar matrix in diagonal storage; with a diagonal
                                   // pattern the assembly kernels write it directly (A_vals unused)
    double *A_dia2 = nullptr;      // the same for the temperature matrix (A_vals2)
    int64_t amg_updates = 0;       // pressure-matrix value updates since the hierarchy was built
    bool pa_dia_fresh = false;     // the pressure assembly also wrote the AMG fine level's diagonal copies
    bool first_pressure = true;
    bool p_zero = false;          // the last pressure solve met its tolerance at x = 0 (x zeroed, no iteration)
    bool field_set_up = false;   // set_up_field has run (first pass of timestep 0)
    int64_t timestep_count = 0;
    int64_t global_iter = 0;

    // ---- solver workspace ----
    double *srhs[5] = {};          // right-hand sides of the batched scalar solve (4: TEM)
    // the batched scalar pass's matrix, diagonal and face coefficients when it
    // runs on `side` beside the ED solve (swapped with A_dia, AV, ff0, ff1 at
    // the join); ev_fork / ev_join order it between the two streams
    double *A_dia3 = nullptr, *AV2 = nullptr, *ff0b = nullptr, *ff1b = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    double *r0buf[3] = {};         // b - A x0 formed by an assembly for the warm-started solve after it
    int *neg_list = nullptr;       // compacted negative cells (TE/ED repair)
    double *wk[24] = {};           // N-length work vectors
    float *wk32[8] = {};
    double *red = nullptr;         // reduction partials
    double *dscal = nullptr;       // device scalars for the solvers
    double *hscal = nullptr;       // pinned host mirror
    int clip_batch[2] = {8, 8};    // TE / ED repair: passes to launch before the first check
    // BiCGStab: iterations the last solve for the same unknowns took (its first
    // batch before a host poll), keyed by the solution vector
    static constexpr int BI_HIST = 8;
    const void *bi_hist_key[BI_HIST] = {};
    int bi_hist_its[BI_HIST] = {};
    // particle/flow coupling (allocated on first use)
    double *pterms = nullptr;      // [5N] particle_aos per cell: momentum xyz, energy, fuel
    int *p2c_ptr = nullptr, *p2c_cell = nullptr;   // point -> cells, ascending
    uint8_t *cell_flag = nullptr;  // [N] interpolation cell set
    void *cpl_buf = nullptr;       // grow-only scratch of the coupling calls
    size_t cpl_bytes = 0;
    int red_blocks = 0;

    // ---- AMG ----
    std::vector<AmgLevel> amg;
    bool amg_ready = false;
    int amg_coarse_n = 0;
    double *amg_dense = nullptr;   // [n*n] coarsest (regularised) inverse
    double *amg_work = nullptr;

    // ---- deflation space of loop-1 pressure solves (mc_linalg.hip) ----
    // the last p_recycle loop-1 solutions, mean-zero (a ring of defl_cap
    // slots, defl_x[defl_head] the newest), and A W of the current system
    // (allocated on first use)
    double *defl_x[DEFL_MAX] = {};
    double *defl_aw[DEFL_MAX + 1] = {};
    int defl_count = 0, defl_head = 0, defl_cap = 0;

    // ---- learned initial guess ----
    void *mpnn = nullptr;
    // measured use policy (loop-1 pressure solves): running means of the
    // iterations with / without the guess, the guess's cost and the cost of
    // one Krylov iteration; the guess stays on while it pays for itself
    double lg_it_on = -1.0, lg_it_off = -1.0, lg_cost_ms = -1.0, lg_iter_ms = -1.0, lg_base_ms = -1.0;
    int64_t lg_count = 0;
    int64_t lg_last = 0;       // solve count of the last re-check of the other choice
    int lg_period = 16;        // solves between re-checks (doubles while they agree)
    bool lg_on = true;

    // ---- device allocations owned by the context ----
    std::vector<void *> allocs;
    std::vector<int> alloc_cat;            // category of each, as charged
    std::vector<int64_t> alloc_bytes;
    int mem_cat = MC_MEM_MESH;             // category charged by alloc()
    int64_t mem_bytes[MC_MEM_COUNT] = {};

    // ---- stats / timing ----
    mc_stats st{};
    unsigned profile = 0;                // class mask of timed kernels
    std::vector<hipEvent_t> ev_pool;     // 2 events per pending launch
    std::vector<int> ev_kc;              // class of each pending pair
    std::vector<double> ev_bytes;
    size_t ev_used = 0;                  // pairs recorded since the last fold
    // accumulated time, brackets and algorithmic bytes per kernel class,
    // then per stage (class MC_KC_COUNT + MC_ST_x, bit of the same number in
    // `profile`)
    double kt_ms[MC_KC_COUNT + MC_ST_COUNT] = {};
    int64_t kt_calls[MC_KC_COUNT + MC_ST_COUNT] = {};
    double kt_bytes[MC_KC_COUNT + MC_ST_COUNT] = {};
};

// ---------------------------------------------------------------- errors --
#define MC_HIP(ctx, call)                                                            \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess) {                                                      \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);          \
            return MC_ERR_HIP;                                                       \
        }                                                                            \
    } while (0)

#define MC_TRY(expr)                    \
    do {                                        \
        int r_ = (expr);                        \
        if (r_ != MC_OK) return r_;             \
    } while (0)

#define MC_LAUNCH_CHECK(ctx) MC_HIP(ctx, hipGetLastError())

// ---------------------------------------------------------- device utils --
// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X
// (workgroup i -> XCD i % 8), and each XCD has its own L2.  Remap so that
// every XCD processes one contiguous range of cells: neighbours at +-1 and
// +-nx then sit in the same L2 as the cell that gathers them.
__device__ __forceinline__ int xcd_block(int bid, int nblocks)
{
    constexpr int NXCD = 8;
    const int per = nblocks / NXCD, rem = nblocks % NXCD;
    const int x = bid % NXCD, local = bid / NXCD;
    return x * per + (x < rem ? x : rem) + local;
}

// Cell order of the cell-centric kernels.  On the generated box mesh (cells
// numbered x fastest, then y, then z) workgroup b visits a brick of
// tx x 1 x tz cells (tx tz = 256; m = 128: one whole x row), the bricks after
// the XCD remap in chunks of yc y rows, within a chunk z-slab by z-slab, row
// by row, tile by tile: a cell's neighbours at x, y and z +- 1 lie in bricks
// a few dozen workgroups apart on the same XCD, so a neighbour's values mostly
// come from L2 rather than from HBM a second time (the chunk's two end rows
// excepted).  Other meshes: 256 consecutive cells per workgroup.  Every cell
// is visited exactly once either way.
struct CellWalk {
    int N = 0;
    int nx = 0, ny = 0, nz = 0;   // nx == 0: consecutive cells
    int lgtx = 0, ntx = 0, ntz = 0;
    int yc = 1;                   // y rows per chunk of the block order
};

// Cell of thread t (0..255) of brick b in CellWalk order, or -1.
__device__ __forceinline__ int walk_cell_at(const CellWalk &w, int b, int t)
{
    if (w.nx == 0) {
        const int c = b * 256 + t;
        return c < w.N ? c : -1;
    }
    // blocks in chunks of yc consecutive y rows; within a chunk z-slab by
    // z-slab, each slab row by row, each row tile by tile: a cell's x, y and z
    // neighbours are visited a few blocks apart (L2-resident), except across a
    // chunk's two y ends
    const int xt = b % w.ntx, r = b / w.ntx;
    const int per = w.yc * w.ntz;
    const int ch = r / per, q = r - ch * per;
    const int zt = q / w.yc, y = ch * w.yc + (q - zt * w.yc);
    if (y >= w.ny) return -1;
    const int x = (xt << w.lgtx) + (t & ((1 << w.lgtx) - 1));
    const int z = zt * (256 >> w.lgtx) + (t >> w.lgtx);
    if (x >= w.nx || z >= w.nz) return -1;
    return x + w.nx * (y + w.ny * z);
}

__device__ __forceinline__ int walk_cell(const CellWalk &w)
{
    return walk_cell_at(w, xcd_block(blockIdx.x, gridDim.x), threadIdx.x);
}

// std::min / std::max semantics (FlowSolver.inl uses them on doubles).
__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }

// Store of a kernel output that nothing later in the same launch reads: a
// streaming (non-temporal) store, so the written lines do not displace from
// L2 the neighbour values that other workgroups of the launch still gather.
// Coalesced full-word stores only: scattered or sub-word streaming stores
// (the pressure assembly's FP16 / FP32 copies) measured slower than plain ones.
template <class T>
__device__ __forceinline__ void st_out(T *p, T v)
{
    __builtin_nontemporal_store(v, p);
}

inline int div_up(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

// Workgroups (of 256 threads) that cover every cell in CellWalk order.
inline int walk_blocks(const CellWalk &w)
{
    return w.nx ? div_up(w.ny, w.yc) * w.yc * w.ntx * w.ntz : div_up(w.N, 256);
}

inline CellWalk cell_walk(const Ctx *c)
{
    CellWalk w;
    w.N = c->N;
    if (c->m > 0) {   // generated box: 2m x m x m cells
        w.nx = (int)(2 * c->m); w.ny = (int)c->m; w.nz = (int)c->m;
        // default tile: the widest power of two up to 256 that leaves at most
        // 1/8 of the lanes of a row idle (m = 128: 256, the whole row)
        int tx = c->opt.cell_tile;
        if (tx <= 0) {
            tx = 256;
            while (tx > 1 && (int64_t)8 * (div_up(w.nx, tx) * tx - w.nx) > w.nx) tx >>= 1;
        }
        int lg = 0;
        while ((1 << lg) < tx) lg++;
        w.lgtx = lg; w.ntx = div_up(w.nx, tx); w.ntz = div_up(w.nz, 256 / tx);
        w.yc = c->opt.cell_ychunk > 0 ? c->opt.cell_ychunk : 16;
    }
    return w;
}

// Block-wide sum of NV values over W waves (blockDim.x == 64 W), result valid
// in thread 0; the wave sums are added in wave order.
template <int NV, int W = 4>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *lds /* NV*W */)
{
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double x = v[k];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        v[k] = x;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) lds[k * W + wid] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) {
            double s = lds[k * W];
            for (int w = 1; w < W; w++) s += lds[k * W + w];
            v[k] = s;
        }
    }
}

// Deterministic two-kernel reduction.  Each block of the producing kernel
// stores its NV block sums with plain stores (partials[k * nblocks + block]);
// the end of that kernel makes them visible, and a one-block finishing kernel
// (k_finish) sums them in a fixed order and applies an epilogue.  A "last
// block finishes" scheme inside the producer would need a device-scope
// release fence and a counter atomic in every block, which on a multi-XCD
// part costs far more than the second launch.
template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], double *partials, int bid, int nblocks)
{
    __shared__ double lds[NV * 4];
    block_sum<NV>(v, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) partials[(size_t)k * nblocks + bid] = v[k];
    }
}

// Fixed-order sum of the partials of `nblocks` blocks (one block of
// FINISH_THREADS threads); the totals are valid in thread 0.
constexpr int FINISH_THREADS = 1024;
template <int NV>
__device__ __forceinline__ void sum_partials(const double *partials, int nblocks, double (&v)[NV])
{
    __shared__ double lds[NV * (FINISH_THREADS / 64)];
#pragma unroll
    for (int k = 0; k < NV; k++) {
        const double *pk = partials + (size_t)k * nblocks;
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        const int step = blockDim.x;
        int i = threadIdx.x;
        for (; i + 3 * step < nblocks; i += 4 * step) {
            s0 += pk[i];
            s1 += pk[i + step];
            s2 += pk[i + 2 * step];
            s3 += pk[i + 3 * step];
        }
        for (; i < nblocks; i += step) s0 += pk[i];
        v[k] = (s0 + s1) + (s2 + s3);
    }
    block_sum<NV, FINISH_THREADS / 64>(v, lds);
}

// Finishing kernel: one block of FINISH_THREADS threads; thread 0 runs epi(totals).
template <int NV, class Epi>
__global__ void __launch_bounds__(FINISH_THREADS) k_finish(const double *partials, int nblocks, Epi epi)
{
    double v[NV];
    sum_partials<NV>(partials, nblocks, v);
    if (threadIdx.x == 0) epi(v);
}

template <int NV, class Epi>
inline void launch_finish(hipStream_t s, const double *partials, int nblocks, const Epi &epi)
{
    hipLaunchKernelGGL((k_finish<NV, Epi>), dim3(1), dim3(FINISH_THREADS), 0, s, partials, nblocks, epi);
}

// ------------------------------------------------------------ host utils --
int alloc(Ctx *c, void **p, size_t bytes);
template <typename T>
inline int dalloc(Ctx *c, T **p, size_t count)
{
    return alloc(c, (void **)p, count * sizeof(T) + 16);
}
int dfree_all(Ctx *c);
// Frees every allocation charged to one category (after both streams drain).
int dfree_category(Ctx *c, int cat);
// charge the allocations of a scope to one memory category
struct MemScope {
    Ctx *c;
    int old;
    MemScope(Ctx *c_, int cat) : c(c_), old(c_->mem_cat) { c->mem_cat = cat; }
    ~MemScope() { c->mem_cat = old; }
};

// profiling of kernel classes with HIP events on the context stream: the
// constructor records the start event, the destructor the end event; pairs
// are resolved by profile_fold (which synchronises the stream)
struct KTimer {
    Ctx *c;
    long slot = -1;
    KTimer(Ctx *c_, int kc_, double bytes_);
    ~KTimer() { end(); }
    void end();   // the closing event now (the destructor then records none)
};
static_assert(MC_KC_COUNT + MC_ST_COUNT <= 32, "class and stage bits share one 32-bit mask");
// a stage of the SIMPLE iteration (MC_ST_x), timed when stages are enabled
struct StageTimer : KTimer {
    StageTimer(Ctx *c_, int st) : KTimer(c_, MC_KC_COUNT + st, 0.0) {}
};
int profile_fold(Ctx *c);
int copy_probe(Ctx *c, int64_t bytes, int reps, double *ms_out);   // mc_linalg.hip
void profile_release(Ctx *c);
void graph_release(Ctx *c);   // mc_linalg.hip: the PCG iteration graphs

// resolves Ctx::noalloc_scratch (once per device, with that device current)
int noalloc_scratch_init(Ctx *c);   // mc_couple.hip

// stages (mc_mesh.hip)
int mesh_box_generate(Ctx *c, int64_t m);
int mesh_from_host(Ctx *c, const mc_mesh_desc *d);
int mesh_finish(Ctx *c);
int init_fields(Ctx *c);   // constructor fields and boundary slots   // geometry, init, CSR pattern, face orders

// stages (mc_flow.hip)
int grad_multi(Ctx *c, bool limit);                      // 11 fields, optionally limited
int limit_gradients(Ctx *c);                             // limiter alone, 11 fields
int clip_nonneg(Ctx *c, double *phi, int slot);          // TE / ED repair (slot 0 / 1: pass history)
int grad_single(Ctx *c, int field, bool pressure);
int set_up_field(Ctx *c);
// r0: the assembly also forms b - A x0 for the warm-started solve that
// follows it directly (Ctx::r0buf; the solve must be passed the same flag)
int momentum_assemble(Ctx *c, bool r0 = false);
int momentum_solve(Ctx *c, bool r0 = false);
int calculate_mass_flux_assemble(Ctx *c);               // mass flux + pressure matrix
int calculate_pressure(Ctx *c);
int scalar_assemble(Ctx *c, int type, double *r0 = nullptr);
int scalar_solve(Ctx *c, int type, bool fork_batch = false);
int scalar_batch_solve(Ctx *c);
int flow_workspace(Ctx *c);
int simple_iteration(Ctx *c, int inner);
int pressure_boundary_slots(Ctx *c, int field);
int update_P(Ctx *c);
int precomp_AU(Ctx *c);
int boundary_mass_flux(Ctx *c);

// solvers (mc_linalg.hip)
// diagonal storage holding the matrix assembled as `csr` (A_vals -> A_dia,
// A_vals2 -> A_dia2); nullptr without a diagonal pattern
const double *fine_dia_of(const Ctx *c, const double *csr);
int dia_to_csr(Ctx *c, const double *dia, double *csr);   // array access in CSR order
int csr_to_dia(Ctx *c, const double *csr, double *dia);
// warm: start from the values in x (else x0 = 0)
int bicgstab_solve(Ctx *c, const double *vals, int nr, double *const *b, double *const *x,
                   double rtol, int maxit, mc_solve_status *st, bool warm, double *const *r0 = nullptr);
// the momentum and scalar solves start from the current field
inline bool warm_start(const Ctx *c)
{
    return c->opt.warm_start > 0 || (c->opt.warm_start < 0 && c->opt.mode == MC_MODE_SIMPLE);
}
int pressure_solve(Ctx *c, const double *b, double *x, mc_solve_status *st, double atol = 0.0);
int amg_setup(Ctx *c, const double *vals);
int amg_update_values(Ctx *c, const double *vals);
int spmv(Ctx *c, const double *vals, const double *x, double *y);
int dia_setup_fine(Ctx *c);
// x0 = true: x holds the initial guess (else the solve starts from zero);
// bnorm0 = ||b - mean b|| when starting from a guess
int pcg_amg(Ctx *c, const double *vals, const double *b, double *x, double rtol, int maxit,
            double atol, mc_solve_status *st, bool x0 = false);
// The first refinement residual formed by the deflation start: r = b - mean
// - A x0 in wk[4], mean and ||r|| on the device, ||r|| on the host.
struct IrStart {
    const double *mean, *rn;
    double rnorm;
};
int mixed_ir_solve(Ctx *c, const double *vals, const double *b, double *x, double rtol, int maxit,
                   double atol, mc_solve_status *st, bool x0 = false, double bnorm0 = 0.0,
                   const IrStart *start = nullptr);
int reduce_maxabs_pp(Ctx *c, const double *v, double *out_host);
int mean_project(Ctx *c, const double *b, double *out, int n);

// learned initial guess (mc_mpnn.hip)
int mpnn_guess(Ctx *c, const double *b, double *x0, int *ok);
int mpnn_collect_pair(Ctx *c, const double *b, const double *x);   // data set, when collecting
void mpnn_destroy(Ctx *c);

}  // namespace mc

// the opaque handle of the C-ABI
struct mc_ctx {
    mc::Ctx c;
};

