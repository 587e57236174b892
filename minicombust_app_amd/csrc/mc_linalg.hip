// mc_linalg.hip -- Krylov solvers and the aggregation-AMG hierarchy.
//
// Replaces the AmgX solves of the reference GPU path (gpu/FlowSolver.inl:
// 716-782, 943-969) and PETSc's KSPBCGS / KSPFGMRES+PCGAMG of the CPU path
// (FlowSolver.hpp:513-527, FlowSolver.inl:1341-1375, 2203-2273).
//
// Storage.  Matrices are assembled in the PETSc CSR pattern.  When a pattern
// has exactly 7 distinct diagonals (the box mesh's 7-point coupling) or 27
// (a coarse level of the regular 2x2x2 aggregation of that mesh), the solver
// also keeps a diagonal-indexed copy, v[d*n + i] = A[i, i + off_d] (zero
// where the pattern has no entry): one thread per row, no column indices to
// stream, and all ND loads of a row independent and coalesced.  Any other
// pattern runs the CSR kernels: 8 lanes per row, the row sum reduced with 3
// cross-lane shuffles.  A ND-template drives every kernel through for_rows().
//
// Reductions: fixed-order two-level block sums finished by the last block
// (deterministic, no host round trip).  Solver scalars live in device memory;
// kernels of a converged system return at once, and the host polls a flag
// after a number of iterations predicted from the observed convergence rate.
//
//  * BiCGStab: van der Vorst's unpreconditioned method, x0 = 0 (momentum and
//    scalar equations; the reference's KSPBCGS+PCNONE), up to 4 RHS sharing
//    each matrix read.
//  * Pressure: PCG preconditioned by an unsmoothed-aggregation AMG V-cycle
//    (damped Jacobi nu/nu, scaled coarse correction, dense regularised inverse
//    on the coarsest level) on the singular Neumann system with the RHS
//    projected to mean zero.  Mixed form: FP32 PCG-AMG (SpMV, smoothers,
//    hierarchy) inside FP64 iterative refinement; the refinement residual,
//    the update and the coarsest solve stay FP64.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <set>

#include "mc_internal.hpp"

namespace mc {

// ------------------------------------------------------------ layout --
constexpr int RED_MAX_BLOCKS = 8192;   // partial sums: at most 12 per block (RED_PARTIALS)
static_assert(12 * RED_MAX_BLOCKS <= RED_PARTIALS, "reduction partials buffer");
constexpr int BLOCK = 256;

struct Grid {
    int blocks;
    int chunk;   // rows per block (multiple of BLOCK)
};

static Grid grid_for(int n)
{
    const int64_t g = div_up(n, BLOCK);
    int blocks = (int)std::min<int64_t>(g, RED_MAX_BLOCKS);
    if (blocks < 1) blocks = 1;
    const int64_t per = div_up(g, blocks);
    Grid gr;
    gr.chunk = (int)(per * BLOCK);
    gr.blocks = div_up(n, gr.chunk);
    if (gr.blocks < 1) gr.blocks = 1;
    return gr;
}

// Device scalars of one Krylov solve (NR <= 4 systems).
struct KState {
    double rho_old[4], alpha[4], omega[4], rho[4], beta[4];
    double tol[4], bnorm[4], rnorm[4], sn[4], mean;
    double atol;
    int done[4], half[4], iters[4], conv[4];
    int zero_start[4];   // BiCGStab: the initial guess is dropped for x0 = 0
    int all_done, nr, maxit, pad;
};

// ------------------------------------------------------- matrix view --
template <typename T>
struct Mat {
    int n;
    const int *rp, *col;
    const T *csr;
    const T *dia;          // diagonal storage, nullptr -> CSR kernels
    int nd;
    int off[DIA_MAX];
    int sym;               // symmetric: dia holds diagonals nd/2 .. nd-1 (offsets >= 0) only
};

// Entry (i, i + off[d]) of a diagonal-storage matrix.  Symmetric storage
// (SYM) keeps the main diagonal and the upper ones; a lower entry is the
// stored upper entry of the transposed position (row i + off[d], diagonal
// nd-1-d), zero where that row does not exist.  Reading it there touches the
// words a later row reads as its upper entry, so each value crosses HBM once.
// SYM is a compile-time choice (negative ND in the kernels' dispatch): a
// run-time branch here slows every diagonal-storage kernel.
template <bool SYM, typename T>
__device__ __forceinline__ double dia_at(const Mat<T> &A, int d, int i)
{
    if constexpr (!SYM) {
        return (double)A.dia[(size_t)d * A.n + i];
    } else {
        const int h = A.nd >> 1;
        if (d >= h) return (double)A.dia[(size_t)(d - h) * A.n + i];
        const int j = i + A.off[d];
        return j >= 0 ? (double)A.dia[(size_t)(A.nd - 1 - d - h) * A.n + j] : 0.0;
    }
}

// bytes of one pass over a matrix: the stored diagonals (symmetric storage:
// the main and upper ones), or CSR values + columns + row pointers
template <typename T>
static double mat_bytes(const Mat<T> &A, int64_t nnz, size_t vsize)
{
    if (A.nd) return (double)vsize * (A.sym ? A.nd / 2 + 1 : A.nd) * A.n;
    return (double)(vsize + 4) * nnz + 4.0 * (A.n + 1);
}


__device__ __forceinline__ int clampi(int j, int n) { return j < 0 ? 0 : (j >= n ? n - 1 : j); }

// Visit the rows [r0, r1) handled by this thread and call f(row, A x) once
// per row.  ND > 0: one thread per row over ND diagonals (compile-time, fully
// unrolled); ND == 0: CSR, 8 lanes per row, lane 0 calls f.  X(k, j) yields
// operand k at column j, so fused producers (prolongation, p-update) feed the
// product without a separate pass.
template <int ND, int NR, typename T, typename XF, typename F>
__device__ __forceinline__ void for_rows(const Mat<T> &A, int r0, int r1, XF X, F f)
{
    constexpr bool SYM = ND < 0;
    constexpr int NA = SYM ? -ND : ND;
    if constexpr (NA > 0) {
        for (int i = r0 + (int)threadIdx.x; i < r1; i += BLOCK) {
            double acc[NR];
#pragma unroll
            for (int k = 0; k < NR; k++) acc[k] = 0.0;
            double a[NA];
            int j[NA];
#pragma unroll
            for (int d = 0; d < NA; d++) {
                a[d] = dia_at<SYM>(A, d, i);
                j[d] = clampi(i + A.off[d], A.n);
            }
#pragma unroll
            for (int d = 0; d < NA; d++) {
#pragma unroll
                for (int k = 0; k < NR; k++) acc[k] += a[d] * X(k, j[d]);
            }
            f(i, acc);
        }
    } else {
        const int lane8 = threadIdx.x & 7;
        for (int base = r0; base < r1; base += BLOCK / 8) {
            const int r = base + (threadIdx.x >> 3);
            double acc[NR];
#pragma unroll
            for (int k = 0; k < NR; k++) acc[k] = 0.0;
            if (r < r1) {
                const int beg = A.rp[r], end = A.rp[r + 1];
                for (int e = beg + lane8; e < end; e += 8) {
                    const int jj = A.col[e];
                    const double a = (double)A.csr[e];
#pragma unroll
                    for (int k = 0; k < NR; k++) acc[k] += a * X(k, jj);
                }
            }
#pragma unroll
            for (int k = 0; k < NR; k++) {
                double v = acc[k];
                v += __shfl_xor(v, 1, 64);
                v += __shfl_xor(v, 2, 64);
                v += __shfl_xor(v, 4, 64);
                acc[k] = v;
            }
            if (lane8 == 0 && r < r1) f(r, acc);
        }
    }
}

// Launch a kernel template over the supported diagonal counts; KERNEL names
// the count ND_ (7, 27, or 0 for the CSR path).
// Kernels take the storage as a compile-time ND: the number of diagonals,
// negative for symmetric storage, 0 for CSR.
template <typename T>
inline int nd_code(const Mat<T> &A)
{
    return A.dia ? (A.sym ? -A.nd : A.nd) : 0;
}

#define MC_DISPATCH_ND(code, KERNEL, ...)                                                         \
    do {                                                                                          \
        const int code_ = (code);                                                                 \
        if (code_ == 7) { constexpr int ND_ = 7; hipLaunchKernelGGL(KERNEL, __VA_ARGS__); }        \
        else if (code_ == -7) { constexpr int ND_ = -7; hipLaunchKernelGGL(KERNEL, __VA_ARGS__); } \
        else if (code_ == 27) { constexpr int ND_ = 27; hipLaunchKernelGGL(KERNEL, __VA_ARGS__); } \
        else if (code_ == -27) { constexpr int ND_ = -27; hipLaunchKernelGGL(KERNEL, __VA_ARGS__); } \
        else { constexpr int ND_ = 0; hipLaunchKernelGGL(KERNEL, __VA_ARGS__); }                  \
    } while (0)

// ------------------------------------------ CSR -> diagonal storage --
// Offsets of a pattern; diagonal storage when there are exactly 7 or 27.
static bool dia_analyze(int n, const int *rp, const int *col, std::vector<int> &offs, std::vector<int> &map)
{
    std::set<int> s;
    for (int i = 0; i < n; i++)
        for (int k = rp[i]; k < rp[i + 1]; k++) {
            s.insert(col[k] - i);
            if (s.size() > 27) return false;
        }
    if (s.size() != 7 && s.size() != 27) return false;
    offs.assign(s.begin(), s.end());
    map.resize(rp[n]);
    for (int i = 0; i < n; i++)
        for (int k = rp[i]; k < rp[i + 1]; k++) {
            const int d = (int)(std::lower_bound(offs.begin(), offs.end(), col[k] - i) - offs.begin());
            map[k] = d * n + i;
        }
    return true;
}

__global__ void k_csr_to_dia(int nnz, const int *map, const double *csr, double *dia, float *dia32)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    const double v = csr[k];
    dia[map[k]] = v;
    if (dia32) dia32[map[k]] = (float)v;
}

int dia_setup_fine(Ctx *c)
{
    MemScope ms(c, MC_MEM_MATRIX);
    std::vector<int> rp(c->N + 1), col(c->nnz), offs, map;
    MC_HIP(c, hipMemcpy(rp.data(), c->row_ptr, (size_t)(c->N + 1) * 4, hipMemcpyDeviceToHost));
    MC_HIP(c, hipMemcpy(col.data(), c->col, (size_t)c->nnz * 4, hipMemcpyDeviceToHost));
    c->fine_nd = 0;
    if (!dia_analyze(c->N, rp.data(), col.data(), offs, map)) return MC_OK;
    c->fine_nd = (int)offs.size();
    for (int d = 0; d < c->fine_nd; d++) c->fine_off[d] = offs[d];
    MC_TRY(dalloc(c, &c->fine_map, c->nnz));
    MC_HIP(c, hipMemcpy(c->fine_map, map.data(), (size_t)c->nnz * 4, hipMemcpyHostToDevice));
    MC_TRY(dalloc(c, &c->A_dia, (size_t)c->fine_nd * c->N));
    MC_TRY(dalloc(c, &c->A_dia2, (size_t)c->fine_nd * c->N));
    // entries outside the pattern (missing neighbours) stay zero for good
    MC_HIP(c, hipMemset(c->A_dia, 0, (size_t)c->fine_nd * c->N * 8));
    MC_HIP(c, hipMemset(c->A_dia2, 0, (size_t)c->fine_nd * c->N * 8));
    return MC_OK;
}

const double *fine_dia_of(const Ctx *c, const double *csr)
{
    if (!c->fine_nd) return nullptr;
    if (csr == c->A_vals) return c->A_dia;
    if (csr == c->A_vals2) return c->A_dia2;
    return nullptr;
}

__global__ void k_dia_to_csr(int nnz, const int *map, const double *dia, double *csr)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nnz) csr[k] = dia[map[k]];
}

int dia_to_csr(Ctx *c, const double *dia, double *csr)
{
    hipLaunchKernelGGL(k_dia_to_csr, dim3(div_up(c->nnz, 256)), dim3(256), 0, c->stream, c->nnz, c->fine_map, dia, csr);
    MC_LAUNCH_CHECK(c);
    return MC_OK;
}

static Mat<double> fine_mat(Ctx *c, const double *csr, const double *dia)
{
    Mat<double> A;
    A.n = c->N; A.rp = c->row_ptr; A.col = c->col; A.csr = csr;
    A.dia = (c->fine_nd && dia) ? dia : nullptr;
    A.nd = A.dia ? c->fine_nd : 0;
    A.sym = 0;
    for (int d = 0; d < DIA_MAX; d++) A.off[d] = d < c->fine_nd ? c->fine_off[d] : 0;
    return A;
}

// CSR values (set through the array access) into the diagonal storage.
int csr_to_dia(Ctx *c, const double *csr, double *dia)
{
    if (!c->fine_nd) return MC_OK;
    hipLaunchKernelGGL(k_csr_to_dia, dim3(div_up(c->nnz, 256)), dim3(256), 0, c->stream, c->nnz, c->fine_map, csr, dia,
                       (float *)nullptr);
    MC_LAUNCH_CHECK(c);
    return MC_OK;
}

// ------------------------------------------------------------ SpMV --
template <int ND>
__global__ void __launch_bounds__(BLOCK) k_spmv(Mat<double> A, const double *x, double *y, int chunk)
{
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int r0 = b * chunk, r1 = min(A.n, r0 + chunk);
    for_rows<ND, 1>(A, r0, r1, [&](int, int j) { return x[j]; }, [&](int i, const double (&ax)[1]) { y[i] = ax[0]; });
}

int spmv(Ctx *c, const double *vals, const double *x, double *y)
{
    const Grid g = grid_for(c->N);
    // the momentum / scalar matrices through their diagonal storage, any other
    // (the pressure matrix) through CSR
    const Mat<double> A = fine_mat(c, vals, fine_dia_of(c, vals));
    MC_DISPATCH_ND(nd_code(A), (k_spmv<ND_>), dim3(g.blocks), dim3(BLOCK), 0, c->stream, A, x, y, g.chunk);
    MC_LAUNCH_CHECK(c);
    return MC_OK;
}

// ------------------------------------------------------- BiCGStab --
// One iteration is three streaming kernels, each followed by a one-block
// finishing kernel for its dot products:
//   pv: p = r + beta (p' - omega v'), formed on the fly (also where a row
//       reads it at a neighbour); v = A p; (r^, v)       -> alpha
//   st: s = r - alpha v, formed on the fly; t = A s;
//       (t, s), (t, t), (s, s)                           -> half-step test, omega
//   xr: x += alpha p + omega s; r = s - omega t;
//       (r, r), (r^, r)                                  -> convergence, rho, beta
// p and v alternate between two buffers, since rows read the previous
// iteration's values at their neighbours while others write the new ones.
struct BiVecs {
    double *x[4], *r[4], *s[4], *t[4];
    double *p[4], *v[4];            // this iteration's
    const double *po[4], *vo[4];    // the previous iteration's
    const double *rh[4];            // r^ (= b, never written)
    const double *b[4];
};

template <int NR>
__global__ void __launch_bounds__(BLOCK) k_bi_init(int n, BiVecs w, double *partials, int chunk)
{
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int i0 = b * chunk, i1 = min(n, i0 + chunk);
    double acc[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) acc[k] = 0.0;
    for (int i = i0 + threadIdx.x; i < i1; i += BLOCK) {
#pragma unroll
        for (int k = 0; k < NR; k++) {
            const double bi = w.b[k][i];
            w.x[k][i] = 0.0; w.r[k][i] = bi;
            acc[k] += bi * bi;
        }
    }
    block_partials<NR>(acc, partials, b, gridDim.x);
}

// Warm start: r = b - A x0 with the starting values the caller left in x;
// partials of (b, b), (r, r) and (r^, r) with r^ = b.
template <int ND, int NR>
__global__ void __launch_bounds__(BLOCK) k_bi_init_x0(Mat<double> A, BiVecs w, double *partials, int chunk)
{
    const int bb = xcd_block(blockIdx.x, gridDim.x);
    const int r0 = bb * chunk, r1 = min(A.n, r0 + chunk);
    double acc[3 * NR];
#pragma unroll
    for (int k = 0; k < 3 * NR; k++) acc[k] = 0.0;
    for_rows<ND, NR>(A, r0, r1, [&](int k, int j) { return w.x[k][j]; }, [&](int i, const double (&ax)[NR]) {
#pragma unroll
        for (int k = 0; k < NR; k++) {
            const double bi = w.b[k][i], ri = bi - ax[k];
            w.r[k][i] = ri;
            acc[3 * k] += bi * bi;
            acc[3 * k + 1] += ri * ri;
            acc[3 * k + 2] += bi * ri;
        }
    });
    block_partials<3 * NR>(acc, partials, bb, gridDim.x);
}

// The same partials when the assembly already formed r = b - A x0.
template <int NR>
__global__ void __launch_bounds__(BLOCK) k_bi_init_r0(int n, BiVecs w, double *partials, int chunk)
{
    const int bb = xcd_block(blockIdx.x, gridDim.x);
    const int i0 = bb * chunk, i1 = min(n, i0 + chunk);
    double acc[3 * NR];
#pragma unroll
    for (int k = 0; k < 3 * NR; k++) acc[k] = 0.0;
    for (int i = i0 + threadIdx.x; i < i1; i += BLOCK) {
#pragma unroll
        for (int k = 0; k < NR; k++) {
            const double bi = w.b[k][i], ri = w.r[k][i];
            acc[3 * k] += bi * bi;
            acc[3 * k + 1] += ri * ri;
            acc[3 * k + 2] += bi * ri;
        }
    }
    block_partials<3 * NR>(acc, partials, bb, gridDim.x);
}

// The criterion stays ||r|| <= rtol ||b|| (PETSc's default test, as with a
// zero start).  A guess whose residual is not below ||b|| is dropped: that
// right-hand side restarts from x0 = 0 (k_bi_zero_start), so a warm start
// never converges later than the zero start would.
template <int NR>
struct EpiBiInitX0 {
    KState *ks;
    double rtol;
    int maxit;
    __device__ void operator()(const double (&v)[3 * NR]) const
    {
        int all = 1;
        for (int k = 0; k < NR; k++) {
            const double bn = sqrt(v[3 * k]), rn = sqrt(v[3 * k + 1]);
            const int zs = !(rn < bn);
            ks->zero_start[k] = zs;
            ks->bnorm[k] = bn; ks->tol[k] = rtol * bn; ks->rnorm[k] = zs ? bn : rn;
            ks->rho_old[k] = 1.0; ks->alpha[k] = 1.0; ks->omega[k] = 1.0;
            ks->rho[k] = zs ? v[3 * k] : v[3 * k + 2];   // (r^, r0)
            ks->beta[k] = 0.0;
            ks->iters[k] = 0; ks->half[k] = 0; ks->sn[k] = 0.0;
            const double r0n = zs ? bn : rn;
            ks->done[k] = (bn == 0.0 || r0n <= ks->tol[k]) ? 1 : 0;
            ks->conv[k] = ks->done[k];
            all &= ks->done[k];
        }
        ks->nr = NR; ks->maxit = maxit;
        ks->all_done = all;
    }
};

// x = 0, r = b for the right-hand sides whose guess was dropped (b = 0
// included: its solution is zero)
template <int NR>
__global__ void __launch_bounds__(BLOCK) k_bi_zero_start(int n, BiVecs w, const KState *ks)
{
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int k = 0; k < NR; k++)
        if (ks->zero_start[k] || ks->bnorm[k] == 0.0) { w.x[k][i] = 0.0; w.r[k][i] = w.b[k][i]; }
}

template <int NR>
struct EpiBiInit {
    KState *ks;
    double rtol;
    int maxit;
    __device__ void operator()(const double (&v)[NR]) const
    {
        int all = 1;
        for (int k = 0; k < NR; k++) {
            const double bn = sqrt(v[k]);
            ks->bnorm[k] = bn; ks->tol[k] = rtol * bn; ks->rnorm[k] = bn;
            ks->rho_old[k] = 1.0; ks->alpha[k] = 1.0; ks->omega[k] = 1.0;
            ks->rho[k] = v[k];  // (r^, r) with r^ = r = b
            ks->beta[k] = 0.0;
            ks->iters[k] = 0; ks->half[k] = 0; ks->sn[k] = 0.0;
            ks->done[k] = (bn == 0.0) ? 1 : 0;
            ks->conv[k] = (bn == 0.0) ? 1 : 0;
            all &= ks->done[k];
        }
        ks->nr = NR; ks->maxit = maxit;
        ks->all_done = all;
    }
};

// v = A p with p formed on the fly ; partials of (r^, v)
template <int ND, int NR>
__global__ void __launch_bounds__(BLOCK) k_bi_pv(Mat<double> A, BiVecs w, const KState *ks, double *partials,
                                                 int chunk)
{
    if (ks->all_done) return;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int r0 = b * chunk, r1 = min(A.n, r0 + chunk);
    bool skip[NR], first[NR];
    double beta[NR], om[NR], dsum[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) {
        skip[k] = ks->done[k] != 0;
        first[k] = ks->iters[k] == 0;
        beta[k] = ks->beta[k];
        om[k] = ks->omega[k];
        dsum[k] = 0.0;
    }
    auto P = [&](int k, int j) -> double {
        if (first[k]) return w.r[k][j];
        return fma(beta[k], fma(-om[k], w.vo[k][j], w.po[k][j]), w.r[k][j]);
    };
    for_rows<ND, NR>(A, r0, r1, [&](int k, int j) { return skip[k] ? 0.0 : P(k, j); },
                     [&](int r, const double (&ax)[NR]) {
#pragma unroll
                         for (int k = 0; k < NR; k++) {
                             if (skip[k]) continue;
                             w.p[k][r] = P(k, r);
                             w.v[k][r] = ax[k];
                             dsum[k] += w.rh[k][r] * ax[k];
                         }
                     });
    block_partials<NR>(dsum, partials, b, gridDim.x);
}

template <int NR>
struct EpiBiAlpha {
    KState *ks;
    __device__ void operator()(const double (&v)[NR]) const
    {
        if (ks->all_done) return;
        for (int k = 0; k < NR; k++)
            if (!ks->done[k]) ks->alpha[k] = ks->rho[k] / v[k];
    }
};

// t = A s with s = r - alpha v formed on the fly ; partials of (t,s), (t,t), (s,s)
template <int ND, int NR>
__global__ void __launch_bounds__(BLOCK) k_bi_st(Mat<double> A, BiVecs w, const KState *ks, double *partials,
                                                 int chunk)
{
    if (ks->all_done) return;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int r0 = b * chunk, r1 = min(A.n, r0 + chunk);
    bool skip[NR];
    double al[NR], dsum[3 * NR];
#pragma unroll
    for (int k = 0; k < NR; k++) {
        skip[k] = ks->done[k] != 0;
        al[k] = ks->alpha[k];
        dsum[3 * k] = dsum[3 * k + 1] = dsum[3 * k + 2] = 0.0;
    }
    auto S = [&](int k, int j) -> double { return fma(-al[k], w.v[k][j], w.r[k][j]); };
    for_rows<ND, NR>(A, r0, r1, [&](int k, int j) { return skip[k] ? 0.0 : S(k, j); },
                     [&](int r, const double (&ax)[NR]) {
#pragma unroll
                         for (int k = 0; k < NR; k++) {
                             if (skip[k]) continue;
                             const double sv = S(k, r);
                             w.s[k][r] = sv;
                             w.t[k][r] = ax[k];
                             dsum[3 * k] += ax[k] * sv;
                             dsum[3 * k + 1] += ax[k] * ax[k];
                             dsum[3 * k + 2] += sv * sv;
                         }
                     });
    block_partials<3 * NR>(dsum, partials, b, gridDim.x);
}

template <int NR>
struct EpiBiOmega {
    KState *ks;
    __device__ void operator()(const double (&v)[3 * NR]) const
    {
        if (ks->all_done) return;
        for (int k = 0; k < NR; k++) {
            if (ks->done[k]) continue;
            const double sn = sqrt(v[3 * k + 2]);
            ks->sn[k] = sn;
            ks->half[k] = (sn <= ks->tol[k]) ? 1 : 0;
            if (!ks->half[k]) ks->omega[k] = v[3 * k] / v[3 * k + 1];
        }
    }
};

// x += alpha p + omega s ; r = s - omega t ; partials of (r,r), (r^,r)
template <int NR>
__global__ void __launch_bounds__(BLOCK) k_bi_xr(int n, BiVecs w, const KState *ks, double *partials, int chunk)
{
    if (ks->all_done) return;
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int i0 = b * chunk, i1 = min(n, i0 + chunk);
    bool skip[NR], half[NR];
    double al[NR], om[NR], acc[2 * NR];
#pragma unroll
    for (int k = 0; k < NR; k++) {
        skip[k] = ks->done[k] != 0;
        half[k] = ks->half[k] != 0;
        al[k] = ks->alpha[k];
        om[k] = ks->omega[k];
        acc[2 * k] = acc[2 * k + 1] = 0.0;
    }
    for (int i = i0 + threadIdx.x; i < i1; i += BLOCK) {
#pragma unroll
        for (int k = 0; k < NR; k++) {
            if (skip[k]) continue;
            if (half[k]) {
                w.x[k][i] = fma(al[k], w.p[k][i], w.x[k][i]);
            } else {
                const double sv = w.s[k][i];
                w.x[k][i] = fma(om[k], sv, fma(al[k], w.p[k][i], w.x[k][i]));
                const double r = fma(-om[k], w.t[k][i], sv);
                w.r[k][i] = r;
                acc[2 * k] += r * r;
                acc[2 * k + 1] += w.rh[k][i] * r;
            }
        }
    }
    block_partials<2 * NR>(acc, partials, b, gridDim.x);
}

template <int NR>
struct EpiBiXr {
    KState *ks;
    __device__ void operator()(const double (&v)[2 * NR]) const
    {
        if (ks->all_done) return;
        int all = 1;
        for (int k = 0; k < NR; k++) {
            if (!ks->done[k]) {
                ks->iters[k] += 1;
                if (ks->half[k]) {
                    ks->rnorm[k] = ks->sn[k];
                    ks->done[k] = 1; ks->conv[k] = 1;
                } else {
                    const double rn = sqrt(v[2 * k]);
                    ks->rnorm[k] = rn;
                    ks->rho_old[k] = ks->rho[k];
                    ks->rho[k] = v[2 * k + 1];
                    if (rn <= ks->tol[k]) { ks->done[k] = 1; ks->conv[k] = 1; }
                    else if (ks->rho[k] == 0.0 || ks->iters[k] >= ks->maxit || !(rn == rn)) ks->done[k] = 1;
                    else ks->beta[k] = (ks->rho[k] / ks->rho_old[k]) * (ks->alpha[k] / ks->omega[k]);
                }
            }
            all &= ks->done[k];
        }
        ks->all_done = all;
    }
};

static int poll_done(Ctx *c, KState *ks, KState *host)
{
    MC_HIP(c, hipMemcpyAsync(host, ks, sizeof(KState), hipMemcpyDeviceToHost, c->stream));
    MC_HIP(c, hipStreamSynchronize(c->stream));
    return MC_OK;
}

// Iterations to enqueue before the next host poll: the remaining count
// predicted from the last observed residual reduction rate (at least 1).
static int predict_batch(double rn_prev, double rn, int its, double tol, int cap)
{
    if (its <= 0 || !(rn > 0) || !(rn_prev > 0) || rn >= rn_prev) return 1;
    const double rate = std::pow(rn / rn_prev, 1.0 / its);
    const double need = std::log(tol / rn) / std::log(rate);
    int k = (int)std::floor(need);
    return std::max(1, std::min(k, cap));
}

template <int NR>
static int bicgstab_nr(Ctx *c, const double *vals, double *const *b, double *const *x, double rtol, int maxit,
                       mc_solve_status *st, bool warm, double *const *r0)
{
    static_assert(NR <= 4, "NR");
    const int n = c->N;
    // seven work vectors per RHS (r, s, t and two each of p, v) from the 24
    // FP64 work vectors and the 8 scratch vectors (sized for doubles)
    double *pool[32];
    int np = 0;
    for (int i = 0; i < 24; i++) pool[np++] = c->wk[i];
    for (int i = 0; i < 8; i++) pool[np++] = (double *)c->wk32[i];
    BiVecs w[2] = {};
    int q = 0;
    for (int k = 0; k < NR; k++) {
        double *r = pool[q++], *s = pool[q++], *t = pool[q++];
        double *pa = pool[q++], *pb = pool[q++], *va = pool[q++], *vb = pool[q++];
        for (int e = 0; e < 2; e++) {
            w[e].b[k] = b[k]; w[e].x[k] = x[k]; w[e].rh[k] = b[k];
            w[e].r[k] = r; w[e].s[k] = s; w[e].t[k] = t;
        }
        w[0].p[k] = pa; w[0].v[k] = va; w[0].po[k] = pb; w[0].vo[k] = vb;
        w[1].p[k] = pb; w[1].v[k] = vb; w[1].po[k] = pa; w[1].vo[k] = va;
    }
    const double *dia = fine_dia_of(c, vals);
    if (c->fine_nd && !dia) { c->err = "bicgstab: matrix without its diagonal storage"; return MC_ERR_STATE; }
    const Mat<double> A = fine_mat(c, vals, dia);
    KState *ks = (KState *)c->dscal;
    KState *hs = (KState *)c->hscal;
    const Grid g = grid_for(n);
    if (warm && r0) {
        // the assembly formed r = b - A x0 (same x0, same matrix): it is the r vector
        for (int k = 0; k < NR; k++) w[0].r[k] = w[1].r[k] = r0[k];
        hipLaunchKernelGGL(k_bi_init_r0<NR>, dim3(g.blocks), dim3(BLOCK), 0, c->stream, n, w[0], c->red, g.chunk);
        launch_finish<3 * NR>(c->stream, c->red, g.blocks, EpiBiInitX0<NR>{ks, rtol, maxit});
        hipLaunchKernelGGL(k_bi_zero_start<NR>, dim3(div_up(n, BLOCK)), dim3(BLOCK), 0, c->stream, n, w[0], ks);
    } else if (warm) {
        MC_DISPATCH_ND(nd_code(A), (k_bi_init_x0<ND_, NR>), dim3(g.blocks), dim3(BLOCK), 0, c->stream, A, w[0], c->red,
                       g.chunk);
        launch_finish<3 * NR>(c->stream, c->red, g.blocks, EpiBiInitX0<NR>{ks, rtol, maxit});
        hipLaunchKernelGGL(k_bi_zero_start<NR>, dim3(div_up(n, BLOCK)), dim3(BLOCK), 0, c->stream, n, w[0], ks);
    } else {
        hipLaunchKernelGGL(k_bi_init<NR>, dim3(g.blocks), dim3(BLOCK), 0, c->stream, n, w[0], c->red, g.chunk);
        launch_finish<NR>(c->stream, c->red, g.blocks, EpiBiInit<NR>{ks, rtol, maxit});
    }
    MC_LAUNCH_CHECK(c);
    // A; r, p', v', r^ read and p, v written per RHS
    const double mbytes = mat_bytes(A, c->nnz, 8) + 48.0 * NR * n;
    // first batch: as many iterations as the last solve for these unknowns
    // took (surplus iterations return at once on the device), so a solve that
    // repeats its count needs one host poll
    int hslot = -1;
    for (int k = 0; k < Ctx::BI_HIST && hslot < 0; k++)
        if (c->bi_hist_key[k] == x[0]) hslot = k;
    int it = 0, batch = hslot >= 0 ? std::max(1, std::min(c->bi_hist_its[hslot], 8)) : 1;
    double rn_prev = -1.0;
    for (;;) {
        for (int j = 0; j < batch && it < maxit; j++, it++) {
            const BiVecs &W = w[it & 1];
            {
                KTimer kt(c, MC_KC_BICG, mbytes);
                MC_DISPATCH_ND(nd_code(A), (k_bi_pv<ND_, NR>), dim3(g.blocks), dim3(BLOCK), 0, c->stream, A, W, ks, c->red,
                               g.chunk);
            }
            launch_finish<NR>(c->stream, c->red, g.blocks, EpiBiAlpha<NR>{ks});
            MC_DISPATCH_ND(nd_code(A), (k_bi_st<ND_, NR>), dim3(g.blocks), dim3(BLOCK), 0, c->stream, A, W, ks, c->red,
                           g.chunk);
            launch_finish<3 * NR>(c->stream, c->red, g.blocks, EpiBiOmega<NR>{ks});
            hipLaunchKernelGGL(k_bi_xr<NR>, dim3(g.blocks), dim3(BLOCK), 0, c->stream, n, W, ks, c->red, g.chunk);
            launch_finish<2 * NR>(c->stream, c->red, g.blocks, EpiBiXr<NR>{ks});
        }
        MC_LAUNCH_CHECK(c);
        MC_TRY(poll_done(c, ks, hs));
        c->st.host_polls++;
        if (hs->all_done || it >= maxit) break;
        double rn = 0.0;
        for (int k = 0; k < NR; k++)
            if (!hs->done[k]) rn = std::max(rn, hs->rnorm[k] / std::max(hs->tol[k], 1e-300));
        batch = predict_batch(rn_prev, rn, batch, 1.0, 8);
        rn_prev = rn;
    }
    {
        int its = 0;
        for (int k = 0; k < NR; k++) its = std::max(its, hs->iters[k]);
        if (hslot < 0) {   // a free slot, else the oldest entry's
            hslot = 0;
            while (hslot < Ctx::BI_HIST - 1 && c->bi_hist_key[hslot]) hslot++;
            if (c->bi_hist_key[hslot]) {
                for (int k = 0; k + 1 < Ctx::BI_HIST; k++) {
                    c->bi_hist_key[k] = c->bi_hist_key[k + 1];
                    c->bi_hist_its[k] = c->bi_hist_its[k + 1];
                }
                hslot = Ctx::BI_HIST - 1;
            }
            c->bi_hist_key[hslot] = x[0];
        }
        c->bi_hist_its[hslot] = its;
    }
    if (st) {
        st->iterations = 0; st->converged = 1; st->relres = 0.0; st->outer = 0; st->guess_used = 0; st->bnorm = 0.0;
        for (int k = 0; k < NR; k++) {
            st->iterations += hs->iters[k];
            st->converged &= hs->conv[k];
            const double rr = hs->bnorm[k] > 0 ? hs->rnorm[k] / hs->bnorm[k] : 0.0;
            st->relres = std::max(st->relres, rr);
            st->bnorm = std::max(st->bnorm, hs->bnorm[k]);
        }
    }
    return MC_OK;
}

int bicgstab_solve(Ctx *c, const double *vals, int nr, double *const *b, double *const *x, double rtol, int maxit,
                   mc_solve_status *st, bool warm, double *const *r0)
{
    switch (nr) {
    case 1: return bicgstab_nr<1>(c, vals, b, x, rtol, maxit, st, warm, r0);
    case 2: return bicgstab_nr<2>(c, vals, b, x, rtol, maxit, st, warm, r0);
    case 3: return bicgstab_nr<3>(c, vals, b, x, rtol, maxit, st, warm, r0);
    case 4: return bicgstab_nr<4>(c, vals, b, x, rtol, maxit, st, warm, r0);
    default: c->err = "bicgstab: 1..4 right-hand sides"; return MC_ERR_ARG;
    }
}

// ------------------------------------------------- generic reductions --
__global__ void __launch_bounds__(BLOCK) k_sum(int n, const double *v, double *partials, int chunk)
{
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int i0 = b * chunk, i1 = min(n, i0 + chunk);
    double acc[1] = {0.0};
    for (int i = i0 + threadIdx.x; i < i1; i += BLOCK) acc[0] += v[i];
    block_partials<1>(acc, partials, b, gridDim.x);
}

struct EpiScale {
    double *out;
    double scale;
    __device__ void operator()(const double (&v)[1]) const { *out = v[0] * scale; }
};

// *out = scale * sum(v[0..n)) on the stream
static void sum_scaled(Ctx *c, const double *v, int n, double *out, double scale)
{
    const Grid g = grid_for(n);
    hipLaunchKernelGGL(k_sum, dim3(g.blocks), dim3(BLOCK), 0, c->stream, n, v, c->red, g.chunk);
    launch_finish<1>(c->stream, c->red, g.blocks, EpiScale{out, scale});
}

__global__ void k_sub_scalar(int n, const double *in, const double *s, double *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i] - *s;
}

int mean_project(Ctx *c, const double *b, double *out, int n)
{
    double *mean = c->dscal + 512;
    sum_scaled(c, b, n, mean, 1.0 / (double)n);
    hipLaunchKernelGGL(k_sub_scalar, dim3(div_up(n, BLOCK)), dim3(BLOCK), 0, c->stream, n, b, mean, out);
    MC_LAUNCH_CHECK(c);
    return MC_OK;
}

// max(PP[0], max_{i>=1} |PP[i]|), the reference's Pressure_correction_max
// (FlowSolver.inl:2371-2378): block maxima, then one block combines them
__global__ void __launch_bounds__(BLOCK) k_maxabs(int n, const double *v, double *partials, int chunk)
{
    const int b = xcd_block(blockIdx.x, gridDim.x);
    const int i0 = b * chunk, i1 = min(n, i0 + chunk);
    double m = 0.0;
    for (int i = max(i0, 1) + threadIdx.x; i < i1; i += BLOCK) m = fmax(m, fabs(v[i]));
    __shared__ double lds[BLOCK / 64];
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) partials[b] = fmax(fmax(lds[0], lds[1]), fmax(lds[2], lds[3]));
}

__global__ void __launch_bounds__(BLOCK) k_maxabs_finish(const double *v, const double *partials, int nblocks,
                                                         double *out)
{
    double mm = 0.0;
    for (int i = threadIdx.x; i < nblocks; i += BLOCK) mm = fmax(mm, partials[i]);
    __shared__ double lds[BLOCK / 64];
    for (int off = 32; off > 0; off >>= 1) mm = fmax(mm, __shfl_down(mm, off, 64));
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = mm;
    __syncthreads();
    if (threadIdx.x == 0) {
        mm = fmax(fmax(lds[0], lds[1]), fmax(lds[2], lds[3]));
        const double p0 = v[0];
        *out = (p0 != p0) ? p0 : ((mm > p0) ? mm : p0);
    }
}

int reduce_maxabs_pp(Ctx *c, const double *v, double *out_host)
{
    const Grid g = grid_for(c->N);
    double *d = c->dscal + 520;
    hipLaunchKernelGGL(k_maxabs, dim3(g.blocks), dim3(BLOCK), 0, c->stream, c->N, v, c->red, g.chunk);
    hipLaunchKernelGGL(k_maxabs_finish, dim3(1), dim3(BLOCK), 0, c->stream, v, c->red, g.blocks, d);
    MC_LAUNCH_CHECK(c);
    MC_HIP(c, hipMemcpyAsync(c->hscal + 520, d, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    MC_HIP(c, hipStreamSynchronize(c->stream));
    *out_host = c->hscal[520];
    return MC_OK;
}

// =================================================================== AMG ==
// Host-side aggregation: three passes of greedy pairwise matching (strongest
// unmatched neighbour, ties to the lowest column) per level -> aggregates of
// at most 8 cells.  Done once per context; later matrices on the same pattern
// only recompute the Galerkin values on the GPU.
struct HostCSR {
    int n = 0;
    std::vector<int> rp, col;
    std::vector<double> val;
};

static void pairwise_pass(const HostCSR &A, std::vector<int> &agg, int &nc)
{
    agg.assign(A.n, -1);
    nc = 0;
    for (int i = 0; i < A.n; i++) {
        if (agg[i] >= 0) continue;
        int best = -1;
        double bw = 0.0;
        for (int k = A.rp[i]; k < A.rp[i + 1]; k++) {
            const int j = A.col[k];
            if (j == i || agg[j] >= 0) continue;
            const double w = -A.val[k];
            if (w > bw * (1.0 + 1e-12)) { best = j; bw = w; }
        }
        agg[i] = nc;
        if (best >= 0) agg[best] = nc;
        nc++;
    }
}

// Galerkin product with piecewise-constant prolongation; also returns, for
// every coarse nonzero, the list of fine nonzeros summed into it.
static void galerkin(const HostCSR &A, const std::vector<int> &agg, int nc, HostCSR &Ac, std::vector<int> *gptr,
                     std::vector<int> *gmap)
{
    std::vector<int> mp(nc + 1, 0), mem(A.n);
    for (int i = 0; i < A.n; i++) mp[agg[i] + 1]++;
    for (int I = 0; I < nc; I++) mp[I + 1] += mp[I];
    {
        std::vector<int> pos(mp.begin(), mp.end() - 1);
        for (int i = 0; i < A.n; i++) mem[pos[agg[i]]++] = i;
    }
    Ac.n = nc;
    Ac.rp.assign(nc + 1, 0);
    Ac.col.clear();
    Ac.val.clear();
    std::vector<std::pair<int, int>> ent;  // (coarse col, fine nnz)
    std::vector<int> gp, gm;
    gp.push_back(0);
    for (int I = 0; I < nc; I++) {
        ent.clear();
        for (int t = mp[I]; t < mp[I + 1]; t++) {
            const int i = mem[t];
            for (int k = A.rp[i]; k < A.rp[i + 1]; k++) ent.emplace_back(agg[A.col[k]], k);
        }
        std::sort(ent.begin(), ent.end());
        for (size_t e = 0; e < ent.size();) {
            const int J = ent[e].first;
            double s = 0.0;
            size_t f = e;
            for (; f < ent.size() && ent[f].first == J; f++) {
                s += A.val[ent[f].second];
                gm.push_back(ent[f].second);
            }
            Ac.col.push_back(J);
            Ac.val.push_back(s);
            gp.push_back((int)gm.size());
            e = f;
        }
        Ac.rp[I + 1] = (int)Ac.col.size();
    }
    if (gptr) *gptr = std::move(gp);
    if (gmap) *gmap = std::move(gm);
}

__global__ void k_gather_sum(int nnzc, const int *gptr, const int *gmap, const double *vf, double *vc)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnzc) return;
    double s = 0.0;
    for (int t = gptr[k]; t < gptr[k + 1]; t++) s += vf[gmap[t]];
    vc[k] = s;
}

__global__ void k_diag_inv(int n, const int *rp, const int *col, const double *v, double *dinv, float *dinv32,
                           float *v32, int nnz)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        double d = 0.0;
        for (int k = rp[i]; k < rp[i + 1]; k++) if (col[k] == i) d = v[k];
        const double di = (d != 0.0) ? 1.0 / d : 0.0;
        dinv[i] = di;
        dinv32[i] = (float)di;
    }
    for (int k = i; k < nnz; k += gridDim.x * blockDim.x) v32[k] = (float)v[k];
}

// A level with diagonal storage: one pass over its CSR rows fills the FP64 and
// FP32 diagonal copies and the inverse diagonal.
__global__ void k_level_dia(int n, const int *rp, const int *col, const int *map, const double *v, double *dia,
                            float *dia32, double *dinv, float *dinv32, int sym_skip, _Float16 *hoff, float *diag32)
{
    // sym_skip > 0: the FP64 copy is symmetric storage without the sym_skip
    // lower diagonals; the FP32 copy keeps all of them
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d = 0.0;
    const size_t shift = (size_t)sym_skip * n;
    for (int k = rp[i]; k < rp[i + 1]; k++) {
        const double x = v[k];
        const int m = map[k];
        if (col[k] == i) d = x;
        dia32[m] = (float)x;
        if (!sym_skip || col[k] >= i) dia[(size_t)m - shift] = x;
    }
    const double di = (d != 0.0) ? 1.0 / d : 0.0;
    dinv[i] = di;
    dinv32[i] = (float)di;
    if (hoff) {   // 7 diagonals, the main one in the middle
        diag32[i] = (float)d;
        for (int k = rp[i]; k < rp[i + 1]; k++) {
            if (col[k] == i) continue;
            const int dd = (map[k] - i) / n;
            hoff[(size_t)(dd < 3 ? dd : dd - 1) * n + i] = (_Float16)(v[k] / d);
        }
    }
}

// Dense regularised coarsest operator A + alpha 11^T (alpha = mean diag / n):
// for a mean-zero right-hand side its inverse returns the mean-zero solution
// of the singular Neumann system.  In-place Gauss-Jordan inversion (SPD, no
// pivoting) of n <= 128 rows in one workgroup of 1024 threads, the matrix
// held in registers and padded with zeros to 128 x 128: thread t owns row
// t / 8 and the 16 consecutive columns 16 (t % 8) .. 16 (t % 8) + 15, so no
// element needs a bounds test and the pivot row is read as contiguous words.
// Only the first n pivots are taken; the zero padding then stays zero and the
// leading n x n block holds the inverse.  A step k passes the pivot column and
// the scaled pivot row through LDS (double-buffered by the parity of k, so two
// barriers per step suffice) and updates every other element in place.
__global__ void __launch_bounds__(1024) k_dense_inverse(int n, const int *rp, const int *col, const double *v,
                                                        double *inv)
{
    constexpr int QN = DENSE_MAX / 8;
    __shared__ double colk[2][DENSE_MAX];
    __shared__ __align__(16) double prow[2][DENSE_MAX];
    __shared__ double red[1024 / 64];
    const int tid = threadIdx.x;
    const int i = tid >> 3, j0 = QN * (tid & 7);
    double s = 0.0;
    for (int r = tid; r < n; r += blockDim.x)
        for (int k = rp[r]; k < rp[r + 1]; k++) if (col[k] == r) s += v[k];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if ((tid & 63) == 0) red[tid >> 6] = s;
    __syncthreads();
    double tot = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) tot += red[w];
    const double alpha = (tot / n) / n;
    double m[QN];
#pragma unroll
    for (int q = 0; q < QN; q++) m[q] = (i < n && j0 + q < n) ? alpha : 0.0;
    if (i < n)
        for (int k = rp[i]; k < rp[i + 1]; k++) {
            const int jq = col[k] - j0;
#pragma unroll
            for (int q = 0; q < QN; q++) if (q == jq) m[q] += v[k];
        }
    for (int k = 0; k < n; k++) {
        const int buf = k & 1, kq = k - j0;
        if (kq >= 0 && kq < QN) {
            double mk = 0.0;
#pragma unroll
            for (int q = 0; q < QN; q++) if (q == kq) mk = m[q];
            colk[buf][i] = mk;
        }
        __syncthreads();
        if (i == k) {
            const double rk = 1.0 / colk[buf][k];
#pragma unroll
            for (int q = 0; q < QN; q++) {
                m[q] = ((q == kq) ? 1.0 : m[q]) * rk;
                prow[buf][j0 + q] = m[q];
            }
        }
        __syncthreads();
        if (i != k) {
            const double cik = colk[buf][i];
#pragma unroll
            for (int q = 0; q < QN; q++) {
                const double mij = (q == kq) ? 0.0 : m[q];
                m[q] = mij - cik * prow[buf][j0 + q];
            }
        }
    }
    // stored transposed: k_dense_apply's threads (rows) then read consecutive words
    if (i < n)
#pragma unroll
        for (int q = 0; q < QN; q++) {
            const int j = j0 + q;
            if (j < n) inv[(size_t)j * n + i] = m[q];
        }
}

static int amg_build(Ctx *c, const double *vals);

// Aggregation-AMG hierarchy from the pressure matrix (host aggregation and
// Galerkin pattern, then the first values on the device); the wall time up to
// the first values' completion goes to the statistics
int amg_setup(Ctx *c, const double *vals)
{
    const auto t0 = std::chrono::steady_clock::now();
    MC_TRY(amg_build(c, vals));
    MC_HIP(c, hipStreamSynchronize(c->stream));
    c->st.amg_setups += 1;
    c->st.amg_setup_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return MC_OK;
}

static int amg_build(Ctx *c, const double *vals)
{
    MemScope ms(c, MC_MEM_AMG);
    HostCSR A;
    A.n = c->N;
    A.rp.resize(c->N + 1);
    A.col.resize(c->nnz);
    A.val.resize(c->nnz);
    MC_HIP(c, hipMemcpyAsync(A.rp.data(), c->row_ptr, (size_t)(c->N + 1) * 4, hipMemcpyDeviceToHost, c->stream));
    MC_HIP(c, hipMemcpyAsync(A.col.data(), c->col, (size_t)c->nnz * 4, hipMemcpyDeviceToHost, c->stream));
    MC_HIP(c, hipMemcpyAsync(A.val.data(), vals, (size_t)c->nnz * 8, hipMemcpyDeviceToHost, c->stream));
    MC_HIP(c, hipStreamSynchronize(c->stream));
    // a rebuild (e.g. after amg_coarse_max changed) first releases the old
    // hierarchy: every allocation charged to MC_MEM_AMG is made below
    if (!c->amg.empty()) {
        c->amg.clear();
        c->amg_dense = nullptr;
        c->amg_ready = false;
        MC_TRY(dfree_category(c, MC_MEM_AMG));
    }
    const int coarse_max = std::max(8, std::min(c->opt.amg_coarse_max, DENSE_MAX));
    c->graph_gen++;   // the PCG graphs hold the old levels' buffers
    HostCSR cur = std::move(A);
    bool first = true;
    for (;;) {
        AmgLevel L;
        L.n = cur.n;
        L.nnz = (int)cur.col.size();
        if (first) {
            L.row_ptr = c->row_ptr; L.col = c->col;
        } else {
            MC_TRY(dalloc(c, &L.row_ptr, L.n + 1));
            MC_TRY(dalloc(c, &L.col, L.nnz));
            MC_HIP(c, hipMemcpy(L.row_ptr, cur.rp.data(), (size_t)(L.n + 1) * 4, hipMemcpyHostToDevice));
            MC_HIP(c, hipMemcpy(L.col, cur.col.data(), (size_t)L.nnz * 4, hipMemcpyHostToDevice));
        }
        if (first) L.vals = c->PA_vals;   // the fine level reads the pressure matrix in place
        else MC_TRY(dalloc(c, &L.vals, L.nnz));
        MC_TRY(dalloc(c, &L.dinv, L.n));
        MC_TRY(dalloc(c, &L.dinv32, L.n));
        MC_TRY(dalloc(c, &L.x, L.n)); MC_TRY(dalloc(c, &L.b, L.n)); MC_TRY(dalloc(c, &L.t, L.n));
        MC_TRY(dalloc(c, &L.x32, L.n)); MC_TRY(dalloc(c, &L.b32, L.n)); MC_TRY(dalloc(c, &L.t32, L.n));
        // diagonal storage for this level's pattern
        {
            std::vector<int> offs, map;
            L.nd = 0;
            if (cur.n > coarse_max && dia_analyze(cur.n, cur.rp.data(), cur.col.data(), offs, map)) {
                L.nd = (int)offs.size();
                for (int d = 0; d < L.nd; d++) L.off[d] = offs[d];
                // pressure levels are symmetric: the fine matrix entry for a face is one value
                // written to both rows, and Galerkin levels of it are taken as their upper part
                L.sym = (L.nd % 2 == 1);
                for (int d = 0; d < L.nd && L.sym; d++) L.sym = (L.off[d] == -L.off[L.nd - 1 - d]);
                const int ns = L.sym ? L.nd / 2 + 1 : L.nd;
                MC_TRY(dalloc(c, &L.dia_map, L.nnz));
                MC_HIP(c, hipMemcpy(L.dia_map, map.data(), (size_t)L.nnz * 4, hipMemcpyHostToDevice));
                MC_TRY(dalloc(c, &L.dia, (size_t)ns * L.n));
                MC_TRY(dalloc(c, &L.dia32, (size_t)L.nd * L.n));
                MC_HIP(c, hipMemset(L.dia, 0, (size_t)ns * L.n * 8));
                MC_HIP(c, hipMemset(L.dia32, 0, (size_t)L.nd * L.n * 4));
                if (first && L.nd == 7 && L.sym && c->opt.amg_half) {
                    MC_TRY(dalloc(c, &L.hoff, (size_t)6 * L.n));
                    MC_TRY(dalloc(c, &L.diag32, (size_t)L.n));
                    MC_HIP(c, hipMemset(L.hoff, 0, (size_t)6 * L.n * sizeof(_Float16)));
                }
            }
            if (!L.nd) MC_TRY(dalloc(c, &L.vals32, L.nnz));   // FP32 CSR only where no diagonal copy
        }
        if (cur.n <= coarse_max) {
            c->amg.push_back(L);
            break;
        }
        std::vector<int> agg(cur.n), a1;
        std::iota(agg.begin(), agg.end(), 0);
        HostCSR g = cur;
        int nc = cur.n;
        for (int pass = 0; pass < 3; pass++) {
            int ncn = 0;
            pairwise_pass(g, a1, ncn);
            for (int i = 0; i < cur.n; i++) agg[i] = a1[agg[i]];
            HostCSR gn;
            galerkin(g, a1, ncn, gn, nullptr, nullptr);
            g = std::move(gn);
            nc = ncn;
            if (nc <= coarse_max) break;
        }
        if (nc > (int)(0.75 * cur.n)) { c->err = "aggregation stalled"; return MC_ERR_STATE; }
        HostCSR Ac;
        std::vector<int> gp, gm;
        galerkin(cur, agg, nc, Ac, &gp, &gm);
        L.nc = nc;
        MC_TRY(dalloc(c, &L.agg, L.n));
        MC_HIP(c, hipMemcpy(L.agg, agg.data(), (size_t)L.n * 4, hipMemcpyHostToDevice));
        {
            std::vector<int> mp(nc + 1, 0), mem(L.n);
            for (int i = 0; i < L.n; i++) mp[agg[i] + 1]++;
            for (int I = 0; I < nc; I++) mp[I + 1] += mp[I];
            int maxm = 0;
            for (int I = 0; I < nc; I++) maxm = std::max(maxm, mp[I + 1] - mp[I]);
            if (maxm > 8) { c->err = "aggregate larger than 8"; return MC_ERR_STATE; }
            std::vector<int> pos(mp.begin(), mp.end() - 1);
            for (int i = 0; i < L.n; i++) mem[pos[agg[i]]++] = i;
            MC_TRY(dalloc(c, &L.agg_ptr, nc + 1));
            MC_TRY(dalloc(c, &L.agg_members, L.n));
            MC_HIP(c, hipMemcpy(L.agg_ptr, mp.data(), (size_t)(nc + 1) * 4, hipMemcpyHostToDevice));
            MC_HIP(c, hipMemcpy(L.agg_members, mem.data(), (size_t)L.n * 4, hipMemcpyHostToDevice));
        }
        MC_TRY(dalloc(c, &L.gptr, gp.size()));
        MC_TRY(dalloc(c, &L.gmap, gm.size()));
        MC_HIP(c, hipMemcpy(L.gptr, gp.data(), gp.size() * 4, hipMemcpyHostToDevice));
        MC_HIP(c, hipMemcpy(L.gmap, gm.data(), gm.size() * 4, hipMemcpyHostToDevice));
        c->amg.push_back(L);
        cur = std::move(Ac);
        first = false;
    }
    const int nco = c->amg.back().n;
    c->amg_coarse_n = nco;
    MC_TRY(dalloc(c, &c->amg_dense, (size_t)nco * nco));
    c->amg_ready = true;
    c->amg_updates = 0;
    // level 0's diagonal copies were just allocated and zeroed: whatever an
    // earlier pressure assembly wrote went to the hierarchy released above
    c->pa_dia_fresh = false;
    return amg_update_values(c, vals);
}

// The main stream waits for the coarsest inverse formed on the side stream.
static int dense_join(Ctx *c)
{
    if (!c->dense_pending) return MC_OK;
    MC_HIP(c, hipStreamWaitEvent(c->stream, c->ev_dense, 0));
    c->dense_pending = false;
    return MC_OK;
}

int amg_update_values(Ctx *c, const double *vals)
{
    if (!c->amg_ready) return amg_setup(c, vals);
    MC_TRY(dense_join(c));   // the previous inverse still reads the coarsest values
    const double *vf = vals;
    // the pressure assembly wrote level 0's diagonal copies with the values
    const bool l0_done = c->pa_dia_fresh && vals == c->PA_vals;
    c->pa_dia_fresh = false;
    // between refreshes the coarse levels keep the values of an earlier
    // matrix: the V-cycle stays a fixed symmetric operator for the solve
    const bool coarse =
        c->amg.size() == 1 || c->opt.amg_refresh <= 1 || c->amg_updates % c->opt.amg_refresh == 0;
    c->amg_updates++;
    for (size_t l = 0; l < c->amg.size(); l++) {
        AmgLevel &L = c->amg[l];
        if (L.vals != vf)
            MC_HIP(c, hipMemcpyAsync(L.vals, vf, (size_t)L.nnz * 8, hipMemcpyDeviceToDevice, c->stream));
        if (!(l == 0 && l0_done)) {
            if (L.nd)
                hipLaunchKernelGGL(k_level_dia, dim3(div_up(L.n, 256)), dim3(256), 0, c->stream, L.n, L.row_ptr,
                                   L.col, L.dia_map, L.vals, L.dia, L.dia32, L.dinv, L.dinv32, L.sym ? L.nd / 2 : 0,
                                   L.hoff, L.diag32);
            else
                hipLaunchKernelGGL(k_diag_inv, dim3(div_up(L.n, 256)), dim3(256), 0, c->stream, L.n, L.row_ptr,
                                   L.col, L.vals, L.dinv, L.dinv32, L.vals32, L.nnz);
            MC_LAUNCH_CHECK(c);
        }
        if (l + 1 < c->amg.size()) {
            if (!coarse) return MC_OK;
            AmgLevel &Lc = c->amg[l + 1];
            hipLaunchKernelGGL(k_gather_sum, dim3(div_up(Lc.nnz, 256)), dim3(256), 0, c->stream, Lc.nnz, L.gptr, L.gmap,
                               L.vals, Lc.vals);
            MC_LAUNCH_CHECK(c);
            vf = Lc.vals;
        }
    }
    AmgLevel &Lc = c->amg.back();
    if (Lc.n > DENSE_MAX) { c->err = "coarsest AMG level larger than the dense limit"; return MC_ERR_STATE; }
    // one workgroup for ~0.2 ms: run it beside the work that precedes the
    // first coarsest-level solve (residual, first V-cycle's descent)
    MC_HIP(c, hipEventRecord(c->ev_coarse, c->stream));
    MC_HIP(c, hipStreamWaitEvent(c->side, c->ev_coarse, 0));
    hipLaunchKernelGGL(k_dense_inverse, dim3(1), dim3(1024), 0, c->side, Lc.n, Lc.row_ptr, Lc.col, Lc.vals,
                       c->amg_dense);
    MC_LAUNCH_CHECK(c);
    MC_HIP(c, hipEventRecord(c->ev_dense, c->side));
    c->dense_pending = true;
    return MC_OK;
}

template <typename T>
static Mat<T> level_mat(AmgLevel &L)
{
    Mat<T> A;
    A.n = L.n; A.rp = L.row_ptr; A.col = L.col;
    if constexpr (sizeof(T) == 8) { A.csr = (const T *)L.vals; A.dia = L.nd ? (const T *)L.dia : nullptr; }
    else { A.csr = (const T *)L.vals32; A.dia = L.nd ? (const T *)L.dia32 : nullptr; }
    A.sym = (sizeof(T) == 8) ? L.sym : 0;   // FP32 copies keep every diagonal (faster there, measured)
    A.nd = L.nd;
    for (int d = 0; d < DIA_MAX; d++) A.off[d] = d < L.nd ? L.off[d] : 0;
    return A;
}

// ------------------------------------------------------ V-cycle kernels --
// first two pre-smoothing sweeps from x = 0: x1 = w1 D^-1 b, x2 = x1 + w2 D^-1 (b - A x1)
template <int ND, typename T>
__global__ void __launch_bounds__(BLOCK) k_jac_first2(Mat<T> A, const T *dinv, const T *b, T *xout, T w1, T w2,
                                                      int chunk, int two, const KState *stop)
{
    if (stop && stop->all_done) return;
    const int bb = xcd_block(blockIdx.x, gridDim.x);
    const int r0 = bb * chunk, r1 = min(A.n, r0 + chunk);
    if (!two) {
        for (int i = r0 + threadIdx.x; i < r1; i += BLOCK) xout[i] = w1 * dinv[i] * b[i];
        return;
    }
    for_rows<ND, 1>(A, r0, r1, [&](int, int j) { return (double)(w1 * dinv[j] * b[j]); },
                    [&](int r, const double (&ax)[1]) {
                        const T x1 = w1 * dinv[r] * b[r];
                        xout[r] = (T)(x1 + w2 * dinv[r] * (b[r] - (T)ax[0]));
                    });
}

// Jacobi weights of the nu smoothing sweeps: all amg_omega, or, with
// amg_cheb_ratio > 0, the reciprocals of the Chebyshev nodes of degree nu on
// [2 / ratio, 2] -- a Chebyshev smoother written as nu Jacobi sweeps (the
// spectrum of D^-1 A lies in [0, 2] for the zero-row-sum M-matrices of every
// level).  Pre-smoothing uses them in this order, post-smoothing in reverse,
// so the V-cycle stays symmetric.
static void smoother_weights(const Ctx *c, int nu, double *w)
{
    const double ratio = c->opt.amg_cheb_ratio;
    if (!(ratio > 1.0)) {
        for (int k = 0; k < nu; k++) w[k] = c->opt.amg_omega;
        return;
    }
    const double lo = 2.0 / ratio, hi = 2.0;
    for (int k = 0; k < nu; k++) {
        const double node = 0.5 * (hi + lo) + 0.5 * (hi - lo) * std::cos(M_PI * (2.0 * k + 1.0) / (2.0 * nu));
        w[k] = 1.0 / node;
    }
}

// x_out = x + w D^-1 (b - A x).  With dotv != nullptr the last sweep also
// forms the block partials of (dotv, x_out): the PCG's (r, z), finished into
// rho / beta by EpiRz.
template <int ND, typename T>
__global__ void __launch_bounds__(BLOCK) k_jac(Mat<T> A, const T *dinv, const T *b, const T *x, T *xout, T omega,
                                               int chunk, const T *dotv, double *partials, const KState *stop)
{
    if (stop && stop->all_done) return;
    const int bb = xcd_block(blockIdx.x, gridDim.x);
    const int r0 = bb * chunk, r1 = min(A.n, r0 + chunk);
    double acc[1] = {0.0};
    for_rows<ND, 1>(A, r0, r1, [&](int, int j) { return (double)x[j]; }, [&](int r, const double (&ax)[1]) {
        const T y = x[r];
        const T o = y + omega * dinv[r] * (b[r] - (T)ax[0]);
        xout[r] = o;
        if (dotv) acc[0] += (double)dotv[r] * (double)o;
    });
    if (dotv) block_partials<1>(acc, partials, bb, gridDim.x);
}

// y = x + cs * xc[agg] in place: the prolongation of the coarse correction,
// one pass before the post-smoothing sweeps (folded into the first sweep it
// was 15 us slower per fine-level V-cycle at m = 128: every row gathered
// its seven neighbours' aggregates and corrections)
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_prolong(int n, const int *agg, const T *xc, T cs, T *x, const KState *stop)
{
    if (stop && stop->all_done) return;
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) x[i] = (T)(x[i] + cs * xc[agg[i]]);
}

// rho = (r, z), beta = rho / rho_old of the PCG
struct EpiRz {
    KState *ks;
    __device__ void operator()(const double (&v)[1]) const
    {
        if (ks->all_done) return;
        ks->rho[0] = v[0];
        ks->beta[0] = (ks->iters[0] == 0) ? 0.0 : v[0] / ks->rho_old[0];
        ks->rho_old[0] = v[0];
    }
};

// bc[I] = sum_{i in aggregate I} (b_i - (A x)_i): residual fused with
// restriction.  8 lanes per coarse row; DIA: lane m evaluates member m's row;
// CSR: the 8 lanes stride over each member's entries.
template <int ND, typename T>
__global__ void __launch_bounds__(BLOCK) k_resid_restrict(Mat<T> A, int nc, const int *aptr, const int *amem,
                                                          const T *b, const T *x, T *bc, const KState *stop)
{
    if (stop && stop->all_done) return;
    const int I = blockIdx.x * (BLOCK / 8) + (threadIdx.x >> 3);
    const int lane8 = threadIdx.x & 7;
    double s = 0.0;
    if (I < nc) {
        const int t0 = aptr[I], t1 = aptr[I + 1];
        constexpr bool SYM = ND < 0;
        constexpr int NA = SYM ? -ND : ND;
        if constexpr (NA > 0) {
            if (t0 + lane8 < t1) {
                const int i = amem[t0 + lane8];
                double ax = 0.0;
#pragma unroll
                for (int d = 0; d < NA; d++)
                    ax += dia_at<SYM>(A, d, i) * (double)x[clampi(i + A.off[d], A.n)];
                s = (double)b[i] - ax;
            }
        } else {
            for (int t = t0; t < t1; t++) {
                const int i = amem[t];
                double acc = 0.0;
                for (int e = A.rp[i] + lane8; e < A.rp[i + 1]; e += 8) acc += (double)A.csr[e] * (double)x[A.col[e]];
                if (lane8 == 0) s += (double)b[i];
                s -= acc;
            }
        }
    }
    s += __shfl_xor(s, 1, 64);
    s += __shfl_xor(s, 2, 64);
    s += __shfl_xor(s, 4, 64);
    if (lane8 == 0 && I < nc) bc[I] = (T)s;
}

// r = b - A x in row order (diagonal storage: coalesced), then the
// restriction sums each aggregate's members: two passes that replace the
// member-ordered fused kernel, whose row gathers follow the aggregates.
template <int ND, typename T>
__global__ void __launch_bounds__(BLOCK) k_resid(Mat<T> A, const T *b, const T *x, T *r, int chunk, const KState *stop)
{
    if (stop && stop->all_done) return;
    const int bb = xcd_block(blockIdx.x, gridDim.x);
    const int r0 = bb * chunk, r1 = min(A.n, r0 + chunk);
    for_rows<ND, 1>(A, r0, r1, [&](int, int j) { return (double)x[j]; },
                    [&](int i, const double (&ax)[1]) { r[i] = (T)((double)b[i] - ax[0]); });
}

template <typename T>
__global__ void __launch_bounds__(BLOCK) k_restrict(int nc, const int *aptr, const int *amem, const T *r, T *bc,
                                                    const KState *stop)
{
    if (stop && stop->all_done) return;
    const int I = blockIdx.x * BLOCK + threadIdx.x;
    if (I >= nc) return;
    double s = 0.0;
    for (int t = aptr[I]; t < aptr[I + 1]; t++) s += (double)r[amem[t]];
    bc[I] = (T)s;
}

// k_resid then k_restrict in one launch, for the coarse levels in diagonal
// storage (a launch there costs more than the pass): 8 lanes per coarse row,
// lane m forms member m's residual as k_resid does (rounded to T), and the
// first lane adds the members' residuals in member order as k_restrict does,
// so bc holds the same bits as the two passes give.  Aggregates have at most
// 8 members (the host's aggregation checks it).
template <int ND, typename T>
__global__ void __launch_bounds__(BLOCK) k_resid_restrict_seq(Mat<T> A, int nc, const int *aptr, const int *amem,
                                                              const T *b, const T *x, T *bc, const KState *stop)
{
    if (stop && stop->all_done) return;
    constexpr bool SYM = ND < 0;
    constexpr int NA = SYM ? -ND : ND;
    const int I = blockIdx.x * (BLOCK / 8) + (threadIdx.x >> 3);
    const int lane8 = threadIdx.x & 7;
    int t0 = 0, cnt = 0;
    double r = 0.0;
    if (I < nc) {
        t0 = aptr[I];
        cnt = aptr[I + 1] - t0;
        if (lane8 < cnt) {
            const int i = amem[t0 + lane8];
            double a[NA > 0 ? NA : 1];
            int j[NA > 0 ? NA : 1];
#pragma unroll
            for (int d = 0; d < NA; d++) {
                a[d] = dia_at<SYM>(A, d, i);
                j[d] = clampi(i + A.off[d], A.n);
            }
            double acc = 0.0;
#pragma unroll
            for (int d = 0; d < NA; d++) acc += a[d] * (double)x[j[d]];
            r = (double)(T)((double)b[i] - acc);
        }
    }
    double s = 0.0;
    const int base = threadIdx.x & ~7;
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const double v = __shfl(r, base + m, 64);
        if (m < cnt) s += v;
    }
    if (lane8 == 0 && I < nc) bc[I] = (T)s;
}

// The first post-smoothing sweep with the prolongation folded in: every x
// value it reads is the prolonged x_j + cs xc[agg_j], formed by k_prolong's
// expression, so the sweep's output equals k_prolong then k_jac bit for bit
// and the prolonged vector is never stored.  For the coarse levels, where a
// launch costs more than the gathered aggregate ids (on the finest level the
// separate pass is faster).
template <int ND, typename T>
__global__ void __launch_bounds__(BLOCK) k_jac_prolong(Mat<T> A, const T *dinv, const T *b, const T *x, const int *agg,
                                                       const T *xc, T cs, T *xout, T omega, int chunk,
                                                       const KState *stop)
{
    if (stop && stop->all_done) return;
    const int bb = xcd_block(blockIdx.x, gridDim.x);
    const int r0 = bb * chunk, r1 = min(A.n, r0 + chunk);
    auto xp = [&](int j) -> T { return (T)(x[j] + cs * xc[agg[j]]); };
    for_rows<ND, 1>(A, r0, r1, [&](int, int j) { return (double)xp(j); }, [&](int r, const double (&ax)[1]) {
        const T y = xp(r);
        xout[r] = y + omega * dinv[r] * (b[r] - (T)ax[0]);
    });
}

// coarsest level: x = Ainv b in FP64.  One workgroup of 1024 threads: thread
// t sums row t % 128 over the 16 columns of part t / 128 (consecutive lanes
// read consecutive words of the transposed inverse), then the 8 part sums of
// a row are added in part order.
template <typename T>
__global__ void __launch_bounds__(1024) k_dense_apply(int n, const double *inv, const T *b, T *x, const KState *stop)
{
    if (stop && stop->all_done) return;
    constexpr int PARTS = 1024 / DENSE_MAX, PLEN = DENSE_MAX / PARTS;
    __shared__ double bs[DENSE_MAX];
    __shared__ double part[PARTS][DENSE_MAX];
    const int t = threadIdx.x;
    if (t < n) bs[t] = (double)b[t];
    __syncthreads();
    const int i = t % DENSE_MAX, p = t / DENSE_MAX;
    double s = 0.0;
    if (i < n) {
        const int j0 = p * PLEN;
#pragma unroll
        for (int jj = 0; jj < PLEN; jj++) {
            const int j = j0 + jj;
            if (j < n) s += inv[(size_t)j * n + i] * bs[j];
        }
    }
    part[p][i] = s;
    __syncthreads();
    if (t < n) {
        double y = 0.0;
#pragma unroll
        for (int q = 0; q < PARTS; q++) y += part[q][t];
        x[t] = (T)y;
    }
}

// partials of (r, z) when no smoothing sweep can carry them
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_rz(int n, const T *r, const T *z, const KState *ks, double *partials,
                                              int chunk)
{
    if (ks->all_done) return;
    const int bb = xcd_block(blockIdx.x, gridDim.x);
    const int i0 = bb * chunk, i1 = min(n, i0 + chunk);
    double acc[1] = {0.0};
    for (int i = i0 + threadIdx.x; i < i1; i += BLOCK) acc[0] += (double)r[i] * (double)z[i];
    block_partials<1>(acc, partials, bb, gridDim.x);
}

template <typename T> static T *lv_x(AmgLevel &L);
template <typename T> static T *lv_b(AmgLevel &L);
template <typename T> static T *lv_t(AmgLevel &L);
template <typename T> static T *lv_dinv(AmgLevel &L);
template <> double *lv_x<double>(AmgLevel &L) { return L.x; }
template <> double *lv_b<double>(AmgLevel &L) { return L.b; }
template <> double *lv_t<double>(AmgLevel &L) { return L.t; }
template <> double *lv_dinv<double>(AmgLevel &L) { return L.dinv; }
template <> float *lv_x<float>(AmgLevel &L) { return L.x32; }
template <> float *lv_b<float>(AmgLevel &L) { return L.b32; }
template <> float *lv_t<float>(AmgLevel &L) { return L.t32; }
template <> float *lv_dinv<float>(AmgLevel &L) { return L.dinv32; }

// ----------------------------------------------------- V-cycle tail --
// The levels of at most TAIL_ROWS rows, and the dense coarsest level, in one
// workgroup: each launch of vcycle() on those levels becomes a pass of the
// 1024 threads over the level's rows, the passes separated by barriers, so
// half a dozen launches of a few microseconds each become one.  Every row is
// formed by the expressions of the per-pass kernels (k_jac_first2, k_jac,
// k_resid, k_restrict, k_dense_apply) in the same order, so the result is the
// same.  Measured at m = 128: a tail from the 1,024-row level saves 0.05 ms per
// pressure solve; one from the 8,192-row level costs 0.3 ms more than the
// launches it replaces (one CU's passes over 8K rows are slower).
constexpr int TAIL_LEVELS = 4, TAIL_ROWS = 2048, TAIL_THREADS = 1024;

template <typename T>
struct TailLevel {
    Mat<T> A;
    const T *dinv;
    T *b, *x, *t;
    const int *agg, *aptr, *amem;   // aggregation to the next level
    int nc;
};

template <typename T>
struct TailArgs {
    TailLevel<T> lv[TAIL_LEVELS + 1];   // lv[nl]: the dense level (b, x only)
    int nl;                             // smoothed levels
    const double *inv;                  // dense inverse, transposed
    int nd;                             // dense level size
    int nu;
    T cs;
    double w[8];
};

template <int ND, typename T, typename XF>
__device__ __forceinline__ double tail_ax(const Mat<T> &A, int i, XF X)
{
    constexpr bool SYM = ND < 0;
    constexpr int NA = SYM ? -ND : ND;
    double a[NA > 0 ? NA : 1];
    int j[NA > 0 ? NA : 1];
#pragma unroll
    for (int d = 0; d < NA; d++) {
        a[d] = dia_at<SYM>(A, d, i);
        j[d] = clampi(i + A.off[d], A.n);
    }
    double acc = 0.0;
#pragma unroll
    for (int d = 0; d < NA; d++) acc += a[d] * X(j[d]);
    return acc;
}

template <int ND, typename T>
__global__ void __launch_bounds__(TAIL_THREADS) k_vtail(TailArgs<T> a, const KState *stop)
{
    if constexpr (ND != 7 && ND != -7) {
        return;   // 7-diagonal storage only (the host launches no other)
    } else {
        if (stop && stop->all_done) return;
        __shared__ double bs[DENSE_MAX];
        __shared__ double part[TAIL_THREADS / DENSE_MAX][DENSE_MAX];
        const int tid = threadIdx.x;
        T *curs[TAIL_LEVELS], *oths[TAIL_LEVELS];
        const int nu = a.nu;
#pragma unroll
        for (int l = 0; l < TAIL_LEVELS; l++) {
            if (l >= a.nl) break;
            const TailLevel<T> &L = a.lv[l];
            const int n = L.A.n;
            T *cur = L.x, *oth = L.t;
            const T w1 = (T)a.w[0], w2 = (T)a.w[nu >= 2 ? 1 : 0];
            for (int i = tid; i < n; i += TAIL_THREADS) {
                if (nu >= 2) {
                    const double ax = tail_ax<ND>(L.A, i, [&](int jj) { return (double)(w1 * L.dinv[jj] * L.b[jj]); });
                    const T x1 = w1 * L.dinv[i] * L.b[i];
                    cur[i] = (T)(x1 + w2 * L.dinv[i] * (L.b[i] - (T)ax));
                } else {
                    cur[i] = w1 * L.dinv[i] * L.b[i];
                }
            }
            __syncthreads();
            for (int s = 2; s < nu; s++) {
                const T omega = (T)a.w[s];
                for (int i = tid; i < n; i += TAIL_THREADS) {
                    const double ax = tail_ax<ND>(L.A, i, [&](int jj) { return (double)cur[jj]; });
                    const T y = cur[i];
                    oth[i] = y + omega * L.dinv[i] * (L.b[i] - (T)ax);
                }
                __syncthreads();
                T *sw = cur; cur = oth; oth = sw;
            }
            T *r = (cur == L.x) ? L.t : L.x;
            for (int i = tid; i < n; i += TAIL_THREADS)
                r[i] = (T)((double)L.b[i] - tail_ax<ND>(L.A, i, [&](int jj) { return (double)cur[jj]; }));
            __syncthreads();
            T *bc = a.lv[l + 1].b;
            for (int I = tid; I < L.nc; I += TAIL_THREADS) {
                double sum = 0.0;
                for (int t = L.aptr[I]; t < L.aptr[I + 1]; t++) sum += (double)r[L.amem[t]];
                bc[I] = (T)sum;
            }
            __syncthreads();
            curs[l] = cur;
            oths[l] = oth;
        }
        {   // dense coarsest level
            constexpr int PARTS = TAIL_THREADS / DENSE_MAX, PLEN = DENSE_MAX / PARTS;
            const int nd = a.nd;
            const T *bd = a.lv[a.nl].b;
            T *xd = a.lv[a.nl].x;
            if (tid < nd) bs[tid] = (double)bd[tid];
            __syncthreads();
            const int i = tid % DENSE_MAX, p = tid / DENSE_MAX;
            double sum = 0.0;
            if (i < nd) {
                const int j0 = p * PLEN;
#pragma unroll
                for (int jj = 0; jj < PLEN; jj++) {
                    const int j = j0 + jj;
                    if (j < nd) sum += a.inv[(size_t)j * nd + i] * bs[j];
                }
            }
            part[p][i] = sum;
            __syncthreads();
            if (tid < nd) {
                double y = 0.0;
#pragma unroll
                for (int q = 0; q < PARTS; q++) y += part[q][tid];
                xd[tid] = (T)y;
            }
            __syncthreads();
        }
#pragma unroll
        for (int l = TAIL_LEVELS - 1; l >= 0; l--) {
            if (l >= a.nl) continue;
            const TailLevel<T> &L = a.lv[l];
            const int n = L.A.n;
            const T *xc = a.lv[l + 1].x;
            T *cur = curs[l], *oth = oths[l], *x = L.x, *tmp = L.t;
            for (int s = 0; s < nu; s++) {
                const bool first = s == 0, last = s == nu - 1;
                T *dst = last ? x : oth;
                if (dst == cur) dst = tmp;
                const T omega = (T)a.w[nu - 1 - s];
                auto yv = [&](int jj) -> T { return first ? (T)(cur[jj] + a.cs * xc[L.agg[jj]]) : cur[jj]; };
                for (int i = tid; i < n; i += TAIL_THREADS) {
                    const double ax = tail_ax<ND>(L.A, i, [&](int jj) { return (double)yv(jj); });
                    const T y = yv(i);
                    dst[i] = y + omega * L.dinv[i] * (L.b[i] - (T)ax);
                }
                __syncthreads();
                oth = cur;
                cur = dst;
            }
            if (cur != x) {
                for (int i = tid; i < n; i += TAIL_THREADS) x[i] = cur[i];
                __syncthreads();
            }
        }
    }
}

// ----------------------------------------- FP16 fine-level smoothing --
// The FP32 V-cycle's fine-level passes on an FP16 copy of D^-1 A's six
// off-diagonals (entries in [-1, 0] for the pressure M-matrix) with the FP32
// inverse diagonal and diagonal: x' = x + w (D^-1 b - x - (D^-1 A)_off x),
// r = D (D^-1 b - x - (D^-1 A)_off x).  A row reads 12 bytes of matrix
// instead of 28.  The PCG's own product keeps the FP32 matrix, and the FP64
// refinement residual checks every step, so the rounding of the copy (5e-4
// relative) changes only the preconditioner.
struct HMat {
    int n;
    const _Float16 *h;
    const float *dinv, *diag;
    int off[6];
};

template <typename XF>
__device__ __forceinline__ float hmat_offx(const HMat &H, int i, XF X)
{
    float a[6];
    int j[6];
#pragma unroll
    for (int d = 0; d < 6; d++) {
        a[d] = (float)H.h[(size_t)d * H.n + i];
        j[d] = clampi(i + H.off[d], H.n);
    }
    float acc = 0.0f;
#pragma unroll
    for (int d = 0; d < 6; d++) acc += a[d] * X(j[d]);
    return acc;
}

__global__ void __launch_bounds__(BLOCK) k_jac_first2_h(HMat H, const float *b, float *xout, float w1, float w2,
                                                        int chunk, int two, const KState *stop)
{
    if (stop && stop->all_done) return;
    const int bb = xcd_block(blockIdx.x, gridDim.x);
    const int r0 = bb * chunk, r1 = min(H.n, r0 + chunk);
    for (int i = r0 + threadIdx.x; i < r1; i += BLOCK) {
        const float db = H.dinv[i] * b[i];
        const float x1 = w1 * db;
        if (!two) { xout[i] = x1; continue; }
        const float ax = hmat_offx(H, i, [&](int j) { return w1 * (H.dinv[j] * b[j]); });
        xout[i] = x1 + w2 * (db - x1 - ax);
    }
}

__global__ void __launch_bounds__(BLOCK) k_jac_h(HMat H, const float *b, const float *x, float *xout, float omega,
                                                 int chunk, const float *dotv, double *partials, const KState *stop)
{
    if (stop && stop->all_done) return;
    const int bb = xcd_block(blockIdx.x, gridDim.x);
    const int r0 = bb * chunk, r1 = min(H.n, r0 + chunk);
    double acc[1] = {0.0};
    for (int i = r0 + threadIdx.x; i < r1; i += BLOCK) {
        const float y = x[i];
        const float o = y + omega * (H.dinv[i] * b[i] - y - hmat_offx(H, i, [&](int j) { return x[j]; }));
        xout[i] = o;
        if (dotv) acc[0] += (double)dotv[i] * (double)o;
    }
    if (dotv) block_partials<1>(acc, partials, bb, gridDim.x);
}

__global__ void __launch_bounds__(BLOCK) k_resid_h(HMat H, const float *b, const float *x, float *r, int chunk,
                                                   const KState *stop)
{
    if (stop && stop->all_done) return;
    const int bb = xcd_block(blockIdx.x, gridDim.x);
    const int r0 = bb * chunk, r1 = min(H.n, r0 + chunk);
    for (int i = r0 + threadIdx.x; i < r1; i += BLOCK)
        r[i] = H.diag[i] * (H.dinv[i] * b[i] - x[i] - hmat_offx(H, i, [&](int j) { return x[j]; }));
}

// V-cycle on level l: x_l = M^-1 b_l.  On level 0 with ks != nullptr the
// last sweep also reduces (b, x) into the PCG's rho / beta.
template <typename T>
static int vcycle(Ctx *c, size_t l, const T *b, T *x, KState *ks, const KState *stop)
{
    AmgLevel &L = c->amg[l];
    double wts[8] = {};
    smoother_weights(c, std::max(1, std::min(c->opt.amg_nu, 8)), wts);
    const int nu = std::max(1, c->opt.amg_nu);
    if (l >= 1 && l + 1 < c->amg.size() && c->opt.amg_tail) {
        // the rest of the cycle in one workgroup when every level left is
        // small and in diagonal storage of one kind
        const size_t nl = c->amg.size() - 1 - l;
        const int code = nd_code(level_mat<T>(L));
        bool ok = nl <= (size_t)TAIL_LEVELS && (code == 7 || code == -7) && nu <= 8 && b == lv_b<T>(L) &&
                  x == lv_x<T>(L);
        for (size_t k = l; ok && k + 1 < c->amg.size(); k++)
            ok = c->amg[k].n <= TAIL_ROWS && nd_code(level_mat<T>(c->amg[k])) == code;
        if (ok) {
            TailArgs<T> ta;
            memset(&ta, 0, sizeof(ta));
            ta.nl = (int)nl;
            for (size_t k = 0; k < nl; k++) {
                AmgLevel &Lk = c->amg[l + k];
                TailLevel<T> &t = ta.lv[k];
                t.A = level_mat<T>(Lk);
                t.dinv = lv_dinv<T>(Lk);
                t.b = lv_b<T>(Lk); t.x = lv_x<T>(Lk); t.t = lv_t<T>(Lk);
                t.agg = Lk.agg; t.aptr = Lk.agg_ptr; t.amem = Lk.agg_members; t.nc = Lk.nc;
            }
            AmgLevel &Ld = c->amg.back();
            ta.lv[nl].b = lv_b<T>(Ld); ta.lv[nl].x = lv_x<T>(Ld);
            ta.inv = c->amg_dense; ta.nd = Ld.n;
            ta.nu = nu; ta.cs = (T)c->opt.amg_cscale;
            for (int k = 0; k < 8; k++) ta.w[k] = wts[k];
            MC_TRY(dense_join(c));
            MC_DISPATCH_ND(code, (k_vtail<ND_, T>), dim3(1), dim3(TAIL_THREADS), 0, c->stream, ta, stop);
            MC_LAUNCH_CHECK(c);
            return MC_OK;
        }
    }
    if (l + 1 == c->amg.size()) {
        MC_TRY(dense_join(c));
        hipLaunchKernelGGL(k_dense_apply<T>, dim3(1), dim3(1024), 0, c->stream, L.n, c->amg_dense, b, x, stop);
        if (ks) {  // one-level hierarchy: no smoothing sweep carries the (r, z) reduction
            const Grid g = grid_for(L.n);
            hipLaunchKernelGGL(k_rz<T>, dim3(g.blocks), dim3(BLOCK), 0, c->stream, L.n, b, x, ks, c->red, g.chunk);
            launch_finish<1>(c->stream, c->red, g.blocks, EpiRz{ks});
        }
        return MC_OK;
    }
    const Grid g = grid_for(L.n);
    const Mat<T> A = level_mat<T>(L);
    T *dinv = lv_dinv<T>(L), *tmp = lv_t<T>(L);
    T *cur = x, *oth = tmp;
    // FP32 cycle on a level with the FP16 copy: its passes read that copy
    const bool half = sizeof(T) == 4 && L.hoff != nullptr && c->opt.amg_half;
    HMat H{};
    if (half) {
        H.n = L.n; H.h = L.hoff; H.dinv = L.dinv32; H.diag = L.diag32;
        for (int d = 0, q = 0; d < L.nd; d++) if (L.off[d] != 0) H.off[q++] = L.off[d];
    }
    if constexpr (sizeof(T) == 4) {
        if (half)
            hipLaunchKernelGGL(k_jac_first2_h, dim3(g.blocks), dim3(BLOCK), 0, c->stream, H, b, cur, (float)wts[0],
                               (float)wts[nu >= 2 ? 1 : 0], g.chunk, nu >= 2 ? 1 : 0, stop);
    }
    if (!half)
        MC_DISPATCH_ND(nd_code(A), (k_jac_first2<ND_, T>), dim3(g.blocks), dim3(BLOCK), 0, c->stream, A, dinv, b, cur,
                       (T)wts[0], (T)wts[nu >= 2 ? 1 : 0], g.chunk, nu >= 2 ? 1 : 0, stop);
    for (int s = 2; s < nu; s++) {
        if constexpr (sizeof(T) == 4) {
            if (half)
                hipLaunchKernelGGL(k_jac_h, dim3(g.blocks), dim3(BLOCK), 0, c->stream, H, b, (const float *)cur, oth,
                                   (float)wts[s], g.chunk, (const float *)nullptr, (double *)nullptr, stop);
        }
        if (!half)
            MC_DISPATCH_ND(nd_code(A), (k_jac<ND_, T>), dim3(g.blocks), dim3(BLOCK), 0, c->stream, A, dinv, b, cur,
                           oth, (T)wts[s], g.chunk, (const T *)nullptr, (double *)nullptr, stop);
        std::swap(cur, oth);
    }
    AmgLevel &Lc = c->amg[l + 1];
    T *bc = lv_b<T>(Lc), *xc = lv_x<T>(Lc);
    // below the finest level the residual and the restriction, and the
    // prolongation and the first post-smoothing sweep, share a launch
    const bool fuse = l >= 1 && !half;
    if (fuse && A.nd) {
        MC_DISPATCH_ND(nd_code(A), (k_resid_restrict_seq<ND_, T>), dim3(div_up(L.nc, BLOCK / 8)), dim3(BLOCK), 0,
                       c->stream, A, L.nc, L.agg_ptr, L.agg_members, b, (const T *)cur, bc, stop);
    } else if (A.nd) {
        T *r = (cur == x) ? tmp : x;   // free until post-smoothing
        if constexpr (sizeof(T) == 4) {
            if (half)
                hipLaunchKernelGGL(k_resid_h, dim3(g.blocks), dim3(BLOCK), 0, c->stream, H, b, (const float *)cur, r,
                                   g.chunk, stop);
        }
        if (!half)
            MC_DISPATCH_ND(nd_code(A), (k_resid<ND_, T>), dim3(g.blocks), dim3(BLOCK), 0, c->stream, A, b,
                           (const T *)cur, r, g.chunk, stop);
        hipLaunchKernelGGL(k_restrict<T>, dim3(div_up(L.nc, BLOCK)), dim3(BLOCK), 0, c->stream, L.nc, L.agg_ptr,
                           L.agg_members, (const T *)r, bc, stop);
    } else {
        MC_DISPATCH_ND(nd_code(A), (k_resid_restrict<ND_, T>), dim3(div_up(L.nc, BLOCK / 8)), dim3(BLOCK), 0, c->stream, A,
                       L.nc, L.agg_ptr, L.agg_members, b, cur, bc, stop);
    }
    MC_TRY(vcycle<T>(c, l + 1, bc, xc, nullptr, stop));
    const T cs = (T)c->opt.amg_cscale;
    if (!fuse)
        hipLaunchKernelGGL(k_prolong<T>, dim3(div_up(L.n, BLOCK)), dim3(BLOCK), 0, c->stream, L.n, L.agg,
                           (const T *)xc, cs, cur, stop);
    // post-smoothing: the last sweep writes x (and the fused dot on level 0)
    for (int s = 0; s < nu; s++) {
        const bool last = s == nu - 1;
        T *dst = last ? x : oth;
        if (dst == cur) dst = tmp;
        const T *dotv = (last && ks) ? b : nullptr;
        // A (FP16 copy: its six off-diagonals), dinv, b, x read, x' written
        const double jbytes = (l != 0 || !c->profile) ? 0.0
                              : half ? 28.0 * L.n : mat_bytes(A, L.nnz, sizeof(T)) + 4.0 * sizeof(T) * L.n;
        if (fuse && s == 0) {
            MC_DISPATCH_ND(nd_code(A), (k_jac_prolong<ND_, T>), dim3(g.blocks), dim3(BLOCK), 0, c->stream, A, dinv, b,
                           (const T *)cur, L.agg, (const T *)xc, cs, dst, (T)wts[nu - 1 - s], g.chunk, stop);
        } else {
            KTimer kt(c, (l == 0) ? (sizeof(T) == 8 ? MC_KC_JAC64 : MC_KC_JAC32) : -1, jbytes);
            if constexpr (sizeof(T) == 4) {
                if (half)
                    hipLaunchKernelGGL(k_jac_h, dim3(g.blocks), dim3(BLOCK), 0, c->stream, H, b, (const float *)cur,
                                       dst, (float)wts[nu - 1 - s], g.chunk, dotv, c->red, stop);
            }
            if (!half)
                MC_DISPATCH_ND(nd_code(A), (k_jac<ND_, T>), dim3(g.blocks), dim3(BLOCK), 0, c->stream, A, dinv, b,
                               cur, dst, (T)wts[nu - 1 - s], g.chunk, dotv, c->red, stop);
        }
        if (dotv) launch_finish<1>(c->stream, c->red, g.blocks, EpiRz{ks});
        oth = cur;
        cur = dst;
    }
    if (cur != x) MC_HIP(c, hipMemcpyAsync(x, cur, (size_t)L.n * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
    return MC_OK;
}

// Algorithmic bytes of one V-cycle as vcycle() runs it on the built
// hierarchy, each array counted once per pass.  Per level of n rows (T the
// value size, M one pass over its matrix, D its inverse diagonal):
//   first two sweeps from zero   M + D + 2Tn   (b read, x written)
//   each further sweep           M + D + 3Tn   (b, x read, x' written)
//   residual                     M + 3Tn       (+ the FP32 diagonal and D on the FP16 copy)
//   restriction                  4(n + nc + 1) + T(n + nc)
//   prolongation                 4n + T(nc + 2n)
// Below the finest level the residual and restriction share a pass (r is
// neither written nor read again: M + 2Tn + 4(n + nc + 1) + T nc) and the
// prolongation is folded into the first post-smoothing sweep (that sweep
// reads the aggregate ids and the coarse correction, 4n + T nc, and no
// prolonged vector is written or read).  On the FP16 copy M is its six
// off-diagonals (12n) and D the FP32 inverse diagonal; the dense coarsest
// solve reads its inverse (8 n^2) and b, writes x.
template <typename T>
static double vcycle_bytes(Ctx *c)
{
    const int nu = std::max(1, c->opt.amg_nu);
    const double t = (double)sizeof(T);
    double b = 0.0;
    for (size_t l = 0; l < c->amg.size(); l++) {
        AmgLevel &L = c->amg[l];
        const double n = L.n;
        if (l + 1 == c->amg.size()) {
            b += 8.0 * n * n + 2.0 * t * n;
            break;
        }
        const bool half = sizeof(T) == 4 && L.hoff != nullptr && c->opt.amg_half;
        const double M = half ? 12.0 * n : mat_bytes(level_mat<T>(L), L.nnz, sizeof(T));
        const double D = half ? 4.0 * n : t * n;
        b += (nu >= 2 ? M : 0.0) + D + 2.0 * t * n;
        b += (double)std::max(0, nu - 2) * (M + D + 3.0 * t * n);
        if (l >= 1 && !half) {
            b += M + 2.0 * t * n + 4.0 * (n + L.nc + 1) + t * L.nc;
            b += 4.0 * n + t * L.nc;
        } else {
            b += M + (half ? 8.0 * n : 0.0) + 3.0 * t * n;
            b += 4.0 * (n + L.nc + 1) + t * (n + L.nc);
            b += 4.0 * n + t * (L.nc + 2.0 * n);
        }
        b += (double)nu * (M + D + 3.0 * t * n);
    }
    return b;
}

// ------------------------------------------------------------ PCG --
template <typename T>
struct PcgVecs {
    T *x, *r, *z, *p, *q, *p2;
};

// r = b - mean, x = 0, p = 0; partials of (r, r) and (b - mean, b - mean)
template <typename T>
__global__ void __launch_bounds__(BLOCK) k_pcg_init(int n, const T *b, const double *mean, PcgVecs<T> w,
                                                    double *partials, int chunk)
{
    const int bb = xcd_block(blockIdx.x, gridDim.x);
    const int i0 = bb * chunk, i1 = min(n, i0 + chunk);
    double acc[2] = {0.0, 0.0};
    const double mu = mean ? *mean : 0.0;
    for (int i = i0 + threadIdx.x; i < i1; i += BLOCK) {
        const T r = (T)((double)b[i] - mu);
        w.r[i] = r; w.x[i] = (T)0; w.p[i] = (T)0;
        acc[0] += (double)r * (double)r;
    }
    acc[1] = acc[0];
    block_partials<2>(acc, partials, bb, gridDim.x);
}

// from a given x: r = (b - mean) - A x, p = 0; same partials
template <int ND>
__global__ void __launch_bounds__(BLOCK) k_pcg_init_x0(Mat<double> A, const double *b, const double *mean,
                                                       PcgVecs<double> w, double *partials, int chunk)
{
    const int bb = xcd_block(blockIdx.x, gridDim.x);
    const int r0 = bb * chunk, r1 = min(A.n, r0 + chunk);
    double acc[2] = {0.0, 0.0};
    const double mu = *mean;
    const double *x = w.x;
    for_rows<ND, 1>(A, r0, r1, [&](int, int j) { return x[j]; }, [&](int i, const double (&ax)[1]) {
        const double bt = b[i] - mu, r = bt - ax[0];
        w.r[i] = r; w.p[i] = 0.0;
        acc[0] += r * r;
        acc[1] += bt * bt;
    });
    block_partials<2>(acc, partials, bb, gridDim.x);
}

// tolerance from ||b - mean||, start from ||r||
struct EpiPcgInit {
    KState *ks;
    double rtol, atol;
    int maxit;
    __device__ void operator()(const double (&v)[2]) const
    {
        const double rn = sqrt(v[0]), bn = sqrt(v[1]);
        ks->bnorm[0] = bn; ks->rnorm[0] = rn;
        ks->tol[0] = fmax(rtol * bn, atol);
        ks->rho_old[0] = 0.0; ks->iters[0] = 0; ks->maxit = maxit;
        ks->done[0] = (rn <= ks->tol[0]) ? 1 : 0;
        ks->conv[0] = ks->done[0];
        ks->all_done = ks->done[0];
    }
};

// p_new = z + beta p_old ; q = A p_new ; alpha = rho / (p_new, q).  The
// product reads z + beta p_old at the neighbours, so the p update needs no
// pass of its own (p_old and p_new are separate buffers).
template <int ND, typename T>
__global__ void __launch_bounds__(BLOCK) k_pcg_pq(Mat<T> A, const T *z, const T *pold, T *pnew, T *q,
                                                  const KState *ks, double *partials, int chunk)
{
    if (ks->all_done) return;
    const int bb = xcd_block(blockIdx.x, gridDim.x);
    const int r0 = bb * chunk, r1 = min(A.n, r0 + chunk);
    const T beta = (T)ks->beta[0];
    double dsum[1] = {0.0};
    for_rows<ND, 1>(A, r0, r1, [&](int, int j) { return (double)(z[j] + beta * pold[j]); },
                    [&](int r, const double (&ax)[1]) {
                        const T p = z[r] + beta * pold[r];
                        pnew[r] = p;
                        q[r] = (T)ax[0];
                        dsum[0] += (double)p * ax[0];
                    });
    block_partials<1>(dsum, partials, bb, gridDim.x);
}

struct EpiPcgAlpha {
    KState *ks;
    __device__ void operator()(const double (&v)[1]) const
    {
        if (!ks->all_done) ks->alpha[0] = ks->rho[0] / v[0];
    }
};

template <typename T>
__global__ void __launch_bounds__(BLOCK) k_pcg_xr(int n, T *x, T *r, const T *p, const T *q, const KState *ks,
                                                  double *partials, int chunk)
{
    if (ks->all_done) return;
    const int bb = xcd_block(blockIdx.x, gridDim.x);
    const int i0 = bb * chunk, i1 = min(n, i0 + chunk);
    const T al = (T)ks->alpha[0];
    double acc[1] = {0.0};
    for (int i = i0 + threadIdx.x; i < i1; i += BLOCK) {
        x[i] += al * p[i];
        const T rr = r[i] - al * q[i];
        r[i] = rr;
        acc[0] += (double)rr * (double)rr;
    }
    block_partials<1>(acc, partials, bb, gridDim.x);
}

struct EpiPcgXr {
    KState *ks;
    __device__ void operator()(const double (&v)[1]) const
    {
        if (ks->all_done) return;
        const double rn = sqrt(v[0]);
        ks->rnorm[0] = rn;
        ks->iters[0] += 1;
        if (rn <= ks->tol[0]) { ks->done[0] = 1; ks->conv[0] = 1; }
        else if (ks->iters[0] >= ks->maxit || !(rn == rn)) ks->done[0] = 1;
        ks->all_done = ks->done[0];
    }
};

// One PCG iteration: z = M^-1 r (V-cycle, whose last sweep also forms rho and
// beta), p = z + beta p_old with q = A p (pa: p_old, pb: p), alpha, then x, r
// and ||r||; every step returns at once when the solve is done.
template <typename T>
static int pcg_iteration(Ctx *c, const Mat<T> &A, const Grid &g, PcgVecs<T> w, T *pa, T *pb, KState *ks,
                         double vbytes)
{
    const int n = c->N;
    {
        KTimer kt(c, MC_KC_VCYCLE, c->profile ? vcycle_bytes<T>(c) : 0.0);
        MC_TRY(vcycle<T>(c, 0, w.r, w.z, ks, ks));
    }
    {
        KTimer kt(c, sizeof(T) == 8 ? MC_KC_SPMV64 : MC_KC_SPMV32, vbytes);
        MC_DISPATCH_ND(nd_code(A), (k_pcg_pq<ND_, T>), dim3(g.blocks), dim3(BLOCK), 0, c->stream, A, w.z, pa, pb, w.q,
                       ks, c->red, g.chunk);
    }
    launch_finish<1>(c->stream, c->red, g.blocks, EpiPcgAlpha{ks});
    hipLaunchKernelGGL(k_pcg_xr<T>, dim3(g.blocks), dim3(BLOCK), 0, c->stream, n, w.x, w.r, pb, w.q, ks, c->red,
                       g.chunk);
    launch_finish<1>(c->stream, c->red, g.blocks, EpiPcgXr{ks});
    return MC_OK;
}

void graph_release(Ctx *c)
{
    for (auto &pr : c->pcg_graphs)
        for (auto &G : pr) {
            if (G.exec) (void)hipGraphExecDestroy(G.exec);
            G = Ctx::PcgGraph{};
        }
}

// The iteration above as a HIP graph: captured on first use (and again after
// the hierarchy, the options or the vectors changed), then replayed, so the
// GPU runs the V-cycle's ~30 small launches back to back instead of at the
// host's launch rate.  The capture holds kernel launches only (no memset or
// event nodes, no other stream: the coarsest inverse is joined first).
template <typename T>
static int pcg_graph(Ctx *c, const Mat<T> &A, const Grid &g, PcgVecs<T> w, T *pa, T *pb, KState *ks, double vbytes,
                     hipGraphExec_t *out)
{
    Ctx::PcgGraph &G = c->pcg_graphs[sizeof(T) == 8 ? 1 : 0][pa == w.p ? 0 : 1];
    const void *key[8] = {w.x, w.r, w.z, w.p, w.q, w.p2, ks, pa};
    bool hit = G.exec && G.gen == c->graph_gen;
    for (int k = 0; hit && k < 8; k++) hit = G.key[k] == key[k];
    if (!hit) {
        if (G.exec) (void)hipGraphExecDestroy(G.exec);
        G = Ctx::PcgGraph{};
        MC_TRY(dense_join(c));
        MC_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int rc = pcg_iteration<T>(c, A, g, w, pa, pb, ks, vbytes);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &graph);
        if (rc != MC_OK || e != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            if (rc != MC_OK) return rc;
            c->err = std::string("PCG graph capture: ") + hipGetErrorString(e);
            return MC_ERR_HIP;
        }
        const hipError_t ei = hipGraphInstantiate(&G.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            G.exec = nullptr;
            c->err = std::string("PCG graph instantiation: ") + hipGetErrorString(ei);
            return MC_ERR_HIP;
        }
        G.gen = c->graph_gen;
        for (int k = 0; k < 8; k++) G.key[k] = key[k];
    }
    *out = G.exec;
    return MC_OK;
}

template <typename T>
static int pcg_core(Ctx *c, const T *b, bool project, PcgVecs<T> w, double rtol, double atol, int maxit, KState *ks,
                    KState *hs, int *iters_out, bool x0 = false)
{
    const int n = c->N;
    const Grid g = grid_for(n);
    const Mat<T> A = level_mat<T>(c->amg[0]);
    const double *mean = nullptr;
    if (project) {
        double *dm = c->dscal + 600;
        if constexpr (sizeof(T) == 8) {
            sum_scaled(c, (const double *)b, n, dm, 1.0 / (double)n);
            mean = dm;
        }
    }
    if constexpr (sizeof(T) == 8) {
        if (x0 && mean) {
            MC_DISPATCH_ND(nd_code(A), (k_pcg_init_x0<ND_>), dim3(g.blocks), dim3(BLOCK), 0, c->stream, A,
                           (const double *)b, mean, w, c->red, g.chunk);
        } else {
            hipLaunchKernelGGL(k_pcg_init<T>, dim3(g.blocks), dim3(BLOCK), 0, c->stream, n, b, mean, w, c->red,
                               g.chunk);
        }
    } else {
        hipLaunchKernelGGL(k_pcg_init<T>, dim3(g.blocks), dim3(BLOCK), 0, c->stream, n, b, mean, w, c->red, g.chunk);
    }
    launch_finish<2>(c->stream, c->red, g.blocks, EpiPcgInit{ks, rtol, atol, maxit});
    MC_LAUNCH_CHECK(c);
    // A; z and p_old read, p and q written
    const double vbytes = mat_bytes(A, c->nnz, sizeof(T)) + 4.0 * sizeof(T) * n;
    T *pa = w.p, *pb = w.p2;
    int it = 0, batch = 2;
    double rn_prev = -1.0;
    // the kernel classes an iteration would time: profiling them launches
    // kernel by kernel
    const unsigned timed = (1u << MC_KC_VCYCLE) | (1u << MC_KC_SPMV64) | (1u << MC_KC_SPMV32) |
                           (1u << MC_KC_JAC64) | (1u << MC_KC_JAC32);
    const bool graphs = c->opt.pcg_graph && !(c->profile & timed) && c->stream != nullptr;   // no capture on the null stream
    for (;;) {
        for (int j = 0; j < batch && it < maxit; j++, it++) {
            if (graphs) {
                hipGraphExec_t ge = nullptr;
                MC_TRY(pcg_graph<T>(c, A, g, w, pa, pb, ks, vbytes, &ge));
                MC_TRY(dense_join(c));
                MC_HIP(c, hipGraphLaunch(ge, c->stream));
            } else {
                MC_TRY(pcg_iteration<T>(c, A, g, w, pa, pb, ks, vbytes));
            }
            std::swap(pa, pb);
        }
        MC_LAUNCH_CHECK(c);
        MC_TRY(poll_done(c, ks, hs));
        c->st.host_polls++;
        if (hs->all_done || it >= maxit) break;
        const double rn = hs->rnorm[0] / std::max(hs->tol[0], 1e-300);
        batch = predict_batch(rn_prev, rn, batch, 1.0, 8);
        rn_prev = rn;
    }
    *iters_out = hs->iters[0];
    return MC_OK;
}

__global__ void k_sub_mean(int n, double *x, const double *mean)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] -= *mean;
}

static int project_mean_zero(Ctx *c, double *x)
{
    const int n = c->N;
    double *dm = c->dscal + 610;
    sum_scaled(c, x, n, dm, 1.0 / (double)n);
    hipLaunchKernelGGL(k_sub_mean, dim3(div_up(n, 256)), dim3(256), 0, c->stream, n, x, dm);
    MC_LAUNCH_CHECK(c);
    return MC_OK;
}

int pcg_amg(Ctx *c, const double *vals, const double *b, double *x, double rtol, int maxit, double atol,
            mc_solve_status *st, bool x0)
{
    (void)vals;  // the hierarchy's level 0 holds these values
    KState *ks = (KState *)c->dscal;
    KState *hs = (KState *)c->hscal;
    PcgVecs<double> w{x, c->wk[0], c->wk[1], c->wk[2], c->wk[3], c->wk[5]};
    int iters = 0;
    MC_TRY(pcg_core<double>(c, b, true, w, rtol, atol, maxit, ks, hs, &iters, x0));
    // met at the zero start: x is all +0, already mean-zero
    c->p_zero = !x0 && iters == 0 && hs->conv[0];
    if (!c->p_zero) MC_TRY(project_mean_zero(c, x));
    if (st) {
        st->iterations = iters;
        st->converged = hs->conv[0];
        st->relres = hs->bnorm[0] > 0 ? hs->rnorm[0] / hs->bnorm[0] : 0.0;
        st->outer = 1;
        st->bnorm = hs->bnorm[0];
    }
    return MC_OK;
}

// ---------------------------------------------- mixed: FP32 inside FP64 IR --
template <int ND>
__global__ void __launch_bounds__(BLOCK) k_ir_resid(Mat<double> A, const double *b, const double *mean, const double *x,
                                                    double *r, double *partials, int chunk)
{
    const int bb = xcd_block(blockIdx.x, gridDim.x);
    const int r0 = bb * chunk, r1 = min(A.n, r0 + chunk);
    double dsum[1] = {0.0};
    const double mu = *mean;
    for_rows<ND, 1>(A, r0, r1, [&](int, int j) { return x[j]; }, [&](int row, const double (&ax)[1]) {
        const double rr = (b[row] - mu) - ax[0];
        r[row] = rr;
        dsum[0] += rr * rr;
    });
    block_partials<1>(dsum, partials, bb, gridDim.x);
}

// The first refinement residual of a solve started from zero: r = b - mean
// (the product A 0 adds +0.0, so these are the bits k_ir_resid gives for a
// matrix in diagonal storage, whose rows it visits in the same order) with x
// set to zero in the same pass; A and x are not read.
__global__ void __launch_bounds__(BLOCK) k_ir_resid0(int n, const double *b, const double *mean, double *x,
                                                     double *r, double *partials, int chunk)
{
    const int bb = xcd_block(blockIdx.x, gridDim.x);
    const int r0 = bb * chunk, r1 = min(n, r0 + chunk);
    double dsum[1] = {0.0};
    const double mu = *mean;
    for (int row = r0 + threadIdx.x; row < r1; row += BLOCK) {
        const double rr = b[row] - mu;
        r[row] = rr;
        x[row] = 0.0;
        dsum[0] += rr * rr;
    }
    block_partials<1>(dsum, partials, bb, gridDim.x);
}

struct EpiSqrt {
    double *out;
    __device__ void operator()(const double (&v)[1]) const { *out = sqrt(v[0]); }
};

__global__ void k_to32_scaled(int n, const double *r, const double *nrm, float *r32)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double s = (*nrm > 0.0) ? 1.0 / *nrm : 0.0;
    r32[i] = (float)(r[i] * s);
}

__global__ void k_ir_update(int n, double *x, const float *e32, const double *nrm)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += (double)e32[i] * *nrm;
}

__global__ void k_zero(int n, double *x)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0.0;
}

int mixed_ir_solve(Ctx *c, const double *vals, const double *b, double *x, double rtol, int maxit, double atol,
                   mc_solve_status *st, bool x0, double bnorm0, const IrStart *start)
{
    (void)vals;
    const int n = c->N;
    const Grid g = grid_for(n);
    KState *ks = (KState *)c->dscal;
    KState *hs = (KState *)c->hscal;
    double *rn = c->dscal + 621;
    const double *mean = c->dscal + 620;
    double *r = c->wk[4];
    float *r32 = c->wk32[0], *e32 = c->wk32[1];
    PcgVecs<float> w{e32, c->wk32[2], c->wk32[3], c->wk32[4], c->wk32[5], c->wk32[6]};
    const Mat<double> A = level_mat<double>(c->amg[0]);
    if (start) mean = start->mean;   // the deflation start's mean and first residual
    else sum_scaled(c, b, n, c->dscal + 620, 1.0 / (double)n);
    // diagonal storage: the first residual of a zero start skips A (same bits,
    // the same rows per thread); CSR rows are reduced in another order
    const bool fast0 = !x0 && nd_code(A) != 0;
    if (!x0 && !fast0) hipLaunchKernelGGL(k_zero, dim3(div_up(n, 256)), dim3(256), 0, c->stream, n, x);
    double bnorm = x0 ? bnorm0 : -1.0, rnorm = 0.0;
    int total = 0, outer = 0;
    bool conv = false;
    const double inner = c->opt.mixed_inner_rtol > 0 ? c->opt.mixed_inner_rtol : 1e-5;
    const double rbytes = mat_bytes(A, c->nnz, 8) + 24.0 * n;   // A; b, x read, r written
    for (outer = 0; outer < 50; outer++) {
        const double *rnd = rn;   // ||r|| on the device, the scale of this step's FP32 system
        if (outer == 0 && start) {
            rnd = start->rn;
            rnorm = start->rnorm;
        } else {
            if (outer == 0 && fast0) {
                hipLaunchKernelGGL(k_ir_resid0, dim3(g.blocks), dim3(BLOCK), 0, c->stream, n, b, mean, x, r, c->red,
                                   g.chunk);
            } else {
                KTimer kt(c, MC_KC_IR_RESID, rbytes);
                MC_DISPATCH_ND(nd_code(A), (k_ir_resid<ND_>), dim3(g.blocks), dim3(BLOCK), 0, c->stream, A, b, mean, x,
                               r, c->red, g.chunk);
            }
            launch_finish<1>(c->stream, c->red, g.blocks, EpiSqrt{rn});
            MC_LAUNCH_CHECK(c);
            MC_HIP(c, hipMemcpyAsync(c->hscal + 621, rn, sizeof(double), hipMemcpyDeviceToHost, c->stream));
            MC_HIP(c, hipStreamSynchronize(c->stream));
            c->st.host_polls++;
            rnorm = c->hscal[621];
        }
        if (bnorm < 0) bnorm = rnorm;
        if (rnorm <= std::max(rtol * bnorm, atol) || !(rnorm == rnorm)) { conv = rnorm == rnorm; break; }
        if (total >= maxit) break;
        hipLaunchKernelGGL(k_to32_scaled, dim3(div_up(n, 256)), dim3(256), 0, c->stream, n, r, rnd, r32);
        // the last refinement step needs only the remaining reduction
        const double need = std::max(rtol * bnorm, atol) / rnorm;
        int it = 0;
        MC_TRY(pcg_core<float>(c, r32, false, w, std::max(inner, 0.5 * need), 0.0, maxit - total, ks, hs, &it));
        total += it;
        hipLaunchKernelGGL(k_ir_update, dim3(div_up(n, 256)), dim3(256), 0, c->stream, n, x, e32, rnd);
        MC_LAUNCH_CHECK(c);
    }
    // met at the zero start: x is all +0, already mean-zero
    c->p_zero = !x0 && conv && total == 0;
    if (!c->p_zero) MC_TRY(project_mean_zero(c, x));
    if (st) {
        st->iterations = total;
        st->converged = conv ? 1 : 0;
        st->relres = bnorm > 0 ? rnorm / bnorm : 0.0;
        st->outer = outer;
        st->bnorm = bnorm;
    }
    return MC_OK;
}

// ------------------------------------------------ deflation-space start --
// A loop-1 pressure solve starts from the Galerkin projection of its solution
// onto the deflation space W = [g, x_1, ..., x_s]: g the learned net's guess
// (when the learned form uses it for this solve), x_1 .. x_s the last
// s = p_recycle loop-1 solutions (newest first), every vector mean-zero.
// x0 = W y with (W^T A W) y = W^T b is the vector of span W closest to the
// solution in the A-norm, so the error left to the PCG is A-orthogonal to W.
// The K x K system (K <= 5) is solved by Cholesky on one thread; a vector
// whose pivot falls to 1e-10 of its diagonal entry (numerically in the span
// of the earlier ones: successive solutions are nearly parallel), or is not a
// number (a net output with NaN / Inf fails every comparison), is dropped.
// The start is kept only if ||b - A x0|| < ||b - mean b||, else x0 = 0, so the
// solve never starts worse than the reference's zero start.  The PCG-AMG and
// its residual test are unchanged: W can change the iteration count, never
// the tolerance met.
constexpr int DEFL_W = DEFL_MAX + 1;   // the learned guess and the recycled solutions

struct DeflVecs {
    const double *w[DEFL_W];
    double *aw[DEFL_W];
};

template <int K>
constexpr int defl_nv() { return K * (K + 3) / 2 + 1; }   // E upper triangle, W^T b, sum b

// aw_k = A w_k for the K vectors in one pass over the matrix, with the block
// partials of (w_a, aw_c) for a <= c, (w_a, b) and sum b formed from the
// products' rows as they are made (each row read once, not again by a
// separate reduction pass)
template <int ND, int K>
__global__ void __launch_bounds__(BLOCK) k_defl_aw_dots(Mat<double> A, DeflVecs v, const double *b, double *partials,
                                                        int chunk)
{
    constexpr int NV = defl_nv<K>();
    const int bb = xcd_block(blockIdx.x, gridDim.x);
    const int r0 = bb * chunk, r1 = min(A.n, r0 + chunk);
    double acc[NV];
#pragma unroll
    for (int q = 0; q < NV; q++) acc[q] = 0.0;
    for_rows<ND, K>(A, r0, r1, [&](int k, int j) { return v.w[k][j]; }, [&](int row, const double (&ax)[K]) {
        double w[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            v.aw[k][row] = ax[k];
            w[k] = v.w[k][row];
        }
        const double bi = b[row];
        int q = 0;
#pragma unroll
        for (int a = 0; a < K; a++)
#pragma unroll
            for (int c = a; c < K; c++) acc[q++] += w[a] * ax[c];
#pragma unroll
        for (int a = 0; a < K; a++) acc[q++] += w[a] * bi;
        acc[q] += bi;
    });
    block_partials<NV>(acc, partials, bb, gridDim.x);
}

// sc[0] = mean b, sc[1 .. K] = y, sc[K + 1] = vectors kept
template <int K>
struct EpiDefl {
    double *sc;
    double inv_n;
    __device__ void operator()(const double (&v)[defl_nv<K>()]) const
    {
        double E[K][K], f[K], L[K][K], z[K], y[K];
        bool keep[K];
        int q = 0;
        for (int a = 0; a < K; a++)
            for (int c = a; c < K; c++) { E[a][c] = v[q]; E[c][a] = v[q]; q++; }
        for (int a = 0; a < K; a++) f[a] = v[q++];
        sc[0] = v[q] * inv_n;
        int kept = 0;
        for (int j = 0; j < K; j++) {
            for (int i = 0; i < K; i++) L[i][j] = 0.0;
            double d = E[j][j];
            for (int k = 0; k < j; k++)
                if (keep[k]) d -= L[j][k] * L[j][k];
            keep[j] = E[j][j] > 0.0 && d > 1e-10 * E[j][j];
            if (!keep[j]) continue;
            kept++;
            L[j][j] = sqrt(d);
            for (int i = j + 1; i < K; i++) {
                double t = E[i][j];
                for (int k = 0; k < j; k++)
                    if (keep[k]) t -= L[i][k] * L[j][k];
                L[i][j] = t / L[j][j];
            }
        }
        for (int i = 0; i < K; i++) {
            z[i] = 0.0;
            if (!keep[i]) continue;
            double t = f[i];
            for (int k = 0; k < i; k++)
                if (keep[k]) t -= L[i][k] * z[k];
            z[i] = t / L[i][i];
        }
        for (int i = K - 1; i >= 0; i--) {
            y[i] = 0.0;
            if (!keep[i]) continue;
            double t = z[i];
            for (int k = i + 1; k < K; k++)
                if (keep[k]) t -= L[k][i] * y[k];
            y[i] = t / L[i][i];
        }
        for (int i = 0; i < K; i++) sc[1 + i] = y[i];
        sc[K + 1] = kept;
    }
};

// x = W y and r = (b - mean) - A W y (from the stored products), with the
// partials of ||b - mean||^2 and ||r||^2: the start, its residual (the first
// residual of the refinement that follows) and the two norms of its
// acceptance test in one pass.  A dropped vector has y_k = 0 and is skipped
// (it may hold NaN).  x and r are written whatever the test decides: a
// rejected start leaves the solve to set x = 0 itself.
template <int K>
__global__ void __launch_bounds__(BLOCK) k_defl_start(int n, DeflVecs v, const double *b, const double *sc, double *x,
                                                      double *r, double *partials, int chunk)
{
    const int bb = xcd_block(blockIdx.x, gridDim.x);
    const int i0 = bb * chunk, i1 = min(n, i0 + chunk);
    double y[K];
#pragma unroll
    for (int k = 0; k < K; k++) y[k] = sc[1 + k];
    const double mu = sc[0];
    double acc[2] = {0.0, 0.0};
    for (int i = i0 + threadIdx.x; i < i1; i += BLOCK) {
        const double bt = b[i] - mu;
        double rr = bt, s = 0.0;
#pragma unroll
        for (int k = 0; k < K; k++)
            if (y[k] != 0.0) {
                rr -= y[k] * v.aw[k][i];
                s += y[k] * v.w[k][i];
            }
        x[i] = s;
        r[i] = rr;
        acc[0] += bt * bt;
        acc[1] += rr * rr;
    }
    block_partials<2>(acc, partials, bb, gridDim.x);
}

// sc[K + 2] = ||b - mean||, sc[K + 3] = ||r0||, sc[K + 4] = 1 if the start is kept
struct EpiDeflAccept {
    double *sc;
    int K;
    __device__ void operator()(const double (&v)[2]) const
    {
        const double bn = sqrt(v[0]), rn = sqrt(v[1]);
        sc[K + 2] = bn;
        sc[K + 3] = rn;
        sc[K + 4] = (sc[K + 1] > 0.0 && rn < bn) ? 1.0 : 0.0;
    }
};

// A grid for the deflation reductions: at most 21 partials per block.
static Grid defl_grid(int n)
{
    Grid g = grid_for(n);
    while (g.blocks > RED_PARTIALS / defl_nv<DEFL_W>() && g.chunk < (1 << 28)) {
        g.chunk *= 2;
        g.blocks = div_up(n, g.chunk);
    }
    return g;
}

static int defl_buffers(Ctx *c, int nrec)
{
    MemScope ms(c, MC_MEM_WORK);
    for (int k = 0; k < nrec; k++)
        if (!c->defl_x[k]) MC_TRY(dalloc(c, &c->defl_x[k], (size_t)c->N));
    for (int k = 0; k <= nrec; k++)
        if (!c->defl_aw[k]) MC_TRY(dalloc(c, &c->defl_aw[k], (size_t)c->N));
    return MC_OK;
}

template <int K>
static void defl_launch(Ctx *c, const DeflVecs &v, const double *b, double *x, double *r, double *sc)
{
    const int n = c->N;
    const Grid g = defl_grid(n);
    const Mat<double> A = level_mat<double>(c->amg[0]);
    MC_DISPATCH_ND(nd_code(A), (k_defl_aw_dots<ND_, K>), dim3(g.blocks), dim3(BLOCK), 0, c->stream, A, v, b, c->red,
                   g.chunk);
    launch_finish<defl_nv<K>()>(c->stream, c->red, g.blocks, EpiDefl<K>{sc, 1.0 / n});
    hipLaunchKernelGGL(k_defl_start<K>, dim3(g.blocks), dim3(BLOCK), 0, c->stream, n, v, b, (const double *)sc, x, r,
                       c->red, g.chunk);
    launch_finish<2>(c->stream, c->red, g.blocks, EpiDeflAccept{sc, K});
}

// The start of a loop-1 solve: *used = the start is kept (x = W y, and *ir
// the refinement's first residual: r = b - mean b - A x in wk[4], its norm
// and the mean), *bnorm = ||b - mean b||; *guess = the learned guess is in W.
// A start not kept leaves x to the solve, which starts from zero.
static int deflation_start(Ctx *c, const double *b, double *x, bool want_guess, bool *used, double *bnorm,
                           bool *guess, IrStart *ir)
{
    *used = false;
    *guess = false;
    *bnorm = 0.0;
    const int n = c->N;
    const int nrec = (c->defl_cap == c->opt.p_recycle) ? std::min(c->opt.p_recycle, c->defl_count) : 0;
    if (c->opt.p_recycle > 0 || want_guess) MC_TRY(defl_buffers(c, std::max(c->opt.p_recycle, 0)));
    DeflVecs v{};
    int K = 0;
    if (want_guess && c->mpnn) {
        int ok = 0;
        double *gv = c->wk[6];
        MC_TRY(mpnn_guess(c, b, gv, &ok));
        if (ok) {
            MC_TRY(project_mean_zero(c, gv));
            v.w[K] = gv;
            v.aw[K] = c->defl_aw[K];
            K++;
            *guess = true;
        }
    }
    for (int k = 0; k < nrec; k++, K++) {
        v.w[K] = c->defl_x[(c->defl_head - k + c->defl_cap) % c->defl_cap];
        v.aw[K] = c->defl_aw[K];
    }
    if (K == 0) return MC_OK;
    double *sc = c->dscal + 640, *r = c->wk[4];
    switch (K) {
    case 1: defl_launch<1>(c, v, b, x, r, sc); break;
    case 2: defl_launch<2>(c, v, b, x, r, sc); break;
    case 3: defl_launch<3>(c, v, b, x, r, sc); break;
    case 4: defl_launch<4>(c, v, b, x, r, sc); break;
    default: defl_launch<5>(c, v, b, x, r, sc); break;
    }
    MC_LAUNCH_CHECK(c);
    MC_HIP(c, hipMemcpyAsync(c->hscal + 640, sc, (K + 5) * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    MC_HIP(c, hipStreamSynchronize(c->stream));
    const double *h = c->hscal + 640;
    *used = h[K + 4] > 0.0;
    *bnorm = h[K + 2];
    // the guess is w[0]; a dropped vector (NaN / Inf output, or in the span
    // of the others) has coefficient exactly 0 and contributes nothing to x0
    if (*guess && h[1] == 0.0) *guess = false;
    if (*used) {
        ir->mean = sc;
        ir->rn = sc + K + 3;
        ir->rnorm = h[K + 3];
        const double ratio = h[K + 2] > 0.0 ? h[K + 3] / h[K + 2] : 0.0;
        c->st.p_defl_starts += 1;
        c->st.p_defl_ratio_sum += ratio;
        c->st.p_defl_vectors += (int64_t)h[K + 1];
        if (*guess) c->st.p_guess_ratio_sum += ratio;
    } else {
        *guess = false;
    }
    return MC_OK;
}

// A converged loop-1 solution (mean-zero) joins the recycled solutions: a
// ring of p_recycle slots (restarted when p_recycle changes).
static int defl_push(Ctx *c, const double *x)
{
    const int cap = c->opt.p_recycle;
    if (cap <= 0) return MC_OK;
    MC_TRY(defl_buffers(c, cap));
    if (c->defl_cap != cap) {
        c->defl_cap = cap;
        c->defl_count = 0;
        c->defl_head = cap - 1;
    }
    const int slot = (c->defl_head + 1) % cap;
    MC_HIP(c, hipMemcpyAsync(c->defl_x[slot], x, (size_t)c->N * 8, hipMemcpyDeviceToDevice, c->stream));
    c->defl_head = slot;
    c->defl_count = std::min(c->defl_count + 1, cap);
    return MC_OK;
}

static double ema(double old, double v) { return old < 0.0 ? v : 0.7 * old + 0.3 * v; }

int pressure_solve(Ctx *c, const double *b, double *x, mc_solve_status *st, double atol)
{
    const auto t0 = std::chrono::steady_clock::now();
    mc_solve_status s{};
    int rc;
    bool used = false, tried = false;
    double bn0 = 0.0;
    // a state set through the C-ABI can reach a solve before any assembly
    // built the hierarchy: build it from the stored pressure matrix
    if (!c->amg_ready) MC_TRY(amg_setup(c, c->PA_vals));
    // learned = 2: every loop-1 solve starts from the guess; learned = 1: only
    // while its measured saving (iterations x time per iteration) exceeds its
    // cost, re-checking the other choice after 16 solves, then after twice as
    // many each time a re-check confirms the choice (at most 1024)
    bool recheck = false, guess = false;
    IrStart ir{};
    if (c->opt.learned && c->mpnn && atol == 0.0) {
        tried = true;
        if (c->opt.learned == 1) {
            if (c->lg_it_on < 0.0) tried = true;
            else if (c->lg_it_off < 0.0) tried = false;
            else {
                recheck = c->lg_count - c->lg_last >= c->lg_period;
                tried = recheck ? !c->lg_on : c->lg_on;
            }
        }
    }
    // loop-1 solves (atol == 0) start from the deflation space when it holds
    // a vector; later loops solve the small non-orthogonal remainder from zero
    if (atol == 0.0 && (tried || c->opt.p_recycle > 0))
        MC_TRY(deflation_start(c, b, x, tried, &used, &bn0, &guess, &ir));
    const auto tg = std::chrono::steady_clock::now();
    if (c->opt.psolve == MC_PSOLVE_MIXED)
        rc = mixed_ir_solve(c, c->PA_vals, b, x, c->opt.p_rtol, c->opt.p_maxit, atol, &s, used, bn0,
                            used ? &ir : nullptr);
    else
        rc = pcg_amg(c, c->PA_vals, b, x, c->opt.p_rtol, c->opt.p_maxit, atol, &s, used);
    s.guess_used = guess ? 1 : 0;
    const auto t1 = std::chrono::steady_clock::now();
    c->st.p_solve_ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
    if (rc) return rc;
    if (c->opt.learned && c->mpnn && atol == 0.0) {
        // the guess's cost: the start's time with it over the start's time
        // without it (the recycled vectors cost the same in both)
        const double solve_ms = std::chrono::duration<double, std::milli>(t1 - tg).count();
        const double start_ms = std::chrono::duration<double, std::milli>(tg - t0).count();
        if (s.iterations > 0) c->lg_iter_ms = ema(c->lg_iter_ms, solve_ms / s.iterations);
        if (tried) {
            c->lg_cost_ms = ema(c->lg_cost_ms, start_ms);
            c->lg_it_on = ema(c->lg_it_on, guess ? s.iterations : (c->lg_it_off < 0 ? s.iterations : c->lg_it_off));
        } else {
            c->lg_base_ms = ema(c->lg_base_ms, start_ms);
            c->lg_it_off = ema(c->lg_it_off, s.iterations);
        }
        const bool was_on = c->lg_on;
        if (c->lg_it_on >= 0.0 && c->lg_it_off >= 0.0)
            c->lg_on = (c->lg_it_off - c->lg_it_on) * c->lg_iter_ms > c->lg_cost_ms - std::max(c->lg_base_ms, 0.0);
        if (recheck) {
            c->lg_last = c->lg_count;
            c->lg_period = (c->lg_on == was_on) ? std::min(2 * c->lg_period, 1024) : 16;
        }
        c->lg_count++;
    }
    c->st.p_guesses += guess ? 1 : 0;
    if (c->mpnn) MC_TRY(mpnn_collect_pair(c, b, x));
    if (atol == 0.0 && s.converged) MC_TRY(defl_push(c, x));
    c->st.p_iters += s.iterations;
    c->st.p_solves += 1;
    c->st.p_outer += s.outer;
    c->st.last_p_relres = s.relres;
    {
        const double ref = std::max(s.bnorm, atol / c->opt.p_rtol);
        const double eff = ref > 0 ? s.relres * s.bnorm / ref : 0.0;
        c->st.p_relres_max = std::max(c->st.p_relres_max, eff);
    }
    if (st) *st = s;
    if (!s.converged) {
        c->err = "pressure solve did not converge";
        return MC_ERR_NOCONV;
    }
    return MC_OK;
}

// ------------------------------------------------------ copy probe --
// Streaming copy, the known-size kernel for HBM bandwidth and counter
// calibration: one 16-byte element per lane, one lane per element (a grid of
// n / 256 workgroups), streaming (non-temporal) loads and stores; every
// element is read once and written once.  Of the forms measured on the
// MI355X (several elements per lane at a 256-lane stride, grid-stride loops,
// plain or streaming accesses) this one moves the most bytes per second.
typedef double dvec2 __attribute__((ext_vector_type(2)));
__global__ void __launch_bounds__(256) k_copy_probe(int64_t n16, const dvec2 *__restrict__ src,
                                                   dvec2 *__restrict__ dst)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n16) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

int copy_probe(Ctx *c, int64_t bytes, int reps, double *ms_out)
{
    void *src = nullptr, *dst = nullptr;
    MC_HIP(c, hipMalloc(&src, (size_t)bytes));
    if (hipMalloc(&dst, (size_t)bytes) != hipSuccess) {
        (void)hipFree(src);
        c->err = "copy probe: hipMalloc failed";
        return MC_ERR_NOMEM;
    }
    int rc = MC_OK;
    hipEvent_t a = nullptr, b = nullptr;
    const int64_t n16 = bytes / 16;
    if (div_up(n16, 256) > INT32_MAX) {
        c->err = "copy probe: size beyond one launch";
        return MC_ERR_ARG;
    }
    const int blocks = div_up(n16, 256);
    if (hipMemsetAsync(src, 0, (size_t)bytes, c->stream) != hipSuccess || hipEventCreate(&a) != hipSuccess ||
        hipEventCreate(&b) != hipSuccess) {
        rc = MC_ERR_HIP;
    } else {
        hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, c->stream, n16, (const dvec2 *)src,
                           (dvec2 *)dst);   // untimed first touch
        (void)hipEventRecord(a, c->stream);
        for (int r = 0; r < reps; r++)
            hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, c->stream, n16, (const dvec2 *)src,
                               (dvec2 *)dst);
        (void)hipEventRecord(b, c->stream);
        float ms = 0.f;
        if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess) rc = MC_ERR_HIP;
        *ms_out = (double)ms / reps;
    }
    if (rc) c->err = "copy probe: HIP error";
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    (void)hipFree(src);
    (void)hipFree(dst);
    return rc;
}

}  // namespace mc
