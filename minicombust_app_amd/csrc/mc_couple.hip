// mc_couple.hip -- the flow side of the particle/flow coupling, without its
// MPI transport: particle source terms gathered into their cells, and flow
// values interpolated to the nodes of the cells the particle side works in
// (update_flow_field, FlowSolver.inl:496-700; GPU reference
// gpu/FlowSolver.inl:361-601 with kernel_process_particle_fields and
// kernel_interpolate_phi_to_nodes, gpu_kernels.cu:131-148, 2132-2191).
//
// Both are deterministic: the reference GPU kernels add with atomics (terms)
// or unguarded += (nodes); here the terms of a cell are added in the order
// they were received (a stable sort by cell, then one thread per run), and a
// node sums its cells in ascending cell order, which is the CPU reference's
// result for a cell set iterated in that order.  Compiled without FP
// contraction, like the reference's CPU arithmetic.
#include <hipcub/hipcub.hpp>

#include <mutex>
#include <vector>

#include "mc_internal.hpp"

#pragma clang fp contract(off)

namespace mc {

__global__ void k_terms_add(int n, const unsigned *cell_sorted, const int *idx_sorted, const double *terms, double *pt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned c = cell_sorted[i];
    if (i > 0 && cell_sorted[i - 1] == c) return;   // not the first entry of its cell's run
    double *d = pt + 5 * (size_t)c;
    double a0 = d[0], a1 = d[1], a2 = d[2], a3 = d[3], a4 = d[4];
    for (int k = i; k < n && cell_sorted[k] == c; k++) {
        const double *t = terms + 5 * (size_t)idx_sorted[k];
        a0 = a0 + t[0]; a1 = a1 + t[1]; a2 = a2 + t[2];
        a3 += t[3];
        a4 += t[4];
    }
    d[0] = a0; d[1] = a1; d[2] = a2; d[3] = a3; d[4] = a4;
}

__global__ void k_flag_cells(int n, const unsigned *cells, uint8_t *flag)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[cells[i]] = 1;
}

struct InterpArgs {
    const int *pptr, *pcell;           // point -> cells, ascending cell ids
    const uint8_t *flag, *cpp;         // interpolation cells; cells per point
    const double *px, *py, *pz, *cx, *cy, *cz;
    const double *phi[5];              // U V W P TEM
    const double *gx[5], *gy[5], *gz[5];
};

// FlowSolver.inl:326-333 (dummy gas state of the missing cells, Mesh.hpp:114-116)
// then :478-491 for each interpolation cell of the node.
__global__ void k_interp_nodes(int nn, const unsigned *nodes, InterpArgs a, double *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const unsigned nd = nodes[i];
    const double node_neighbours = 8;
    const double b = node_neighbours - a.cpp[nd];
    double v[5];
    v[0] = 50. * (b / node_neighbours);
    v[1] = 0.00 * (b / node_neighbours);
    v[2] = 0.00 * (b / node_neighbours);
    v[3] = 4000.0 * (b / node_neighbours);
    v[4] = 2000.0 * (b / node_neighbours);
    const double xn = a.px[nd], yn = a.py[nd], zn = a.pz[nd];
    for (int k = a.pptr[nd]; k < a.pptr[nd + 1]; k++) {
        const int c = a.pcell[k];
        if (!a.flag[c]) continue;
        const double dx = xn - a.cx[c], dy = yn - a.cy[c], dz = zn - a.cz[c];
#pragma unroll
        for (int q = 0; q < 5; q++)
            v[q] += (a.phi[q][c] + (a.gx[q][c] * dx + a.gy[q][c] * dy + a.gz[q][c] * dz)) / node_neighbours;
    }
    double *o = out + 5 * (size_t)i;
#pragma unroll
    for (int q = 0; q < 5; q++) o[q] = v[q];
}

// grow-only device scratch of the coupling calls
static int couple_scratch(Ctx *c, size_t bytes, void **p)
{
    if (bytes > c->cpl_bytes) {
        MemScope ms(c, MC_MEM_WORK);
        MC_TRY(alloc(c, &c->cpl_buf, bytes));
        c->cpl_bytes = bytes;
    }
    *p = c->cpl_buf;
    return MC_OK;
}

static int ensure_particle_terms(Ctx *c)
{
    if (c->pterms) return MC_OK;
    MemScope ms(c, MC_MEM_FIELDS);
    MC_TRY(dalloc(c, &c->pterms, 5 * (size_t)c->N));
    MC_HIP(c, hipMemsetAsync(c->pterms, 0, 5 * (size_t)c->N * 8, c->stream));
    return MC_OK;
}

int particle_terms_reset(Ctx *c)
{
    MC_TRY(ensure_particle_terms(c));
    MC_HIP(c, hipMemsetAsync(c->pterms, 0, 5 * (size_t)c->N * 8, c->stream));
    return MC_OK;
}

static size_t up16(size_t b) { return (b + 255) & ~(size_t)255; }

int particle_terms_add(Ctx *c, const uint64_t *cells, const double *terms, int64_t n)
{
    if (n < 0 || n > (int64_t)INT32_MAX / 8 || (n > 0 && (!cells || !terms))) {
        c->err = "particle terms: bad arguments";
        return MC_ERR_ARG;
    }
    MC_TRY(ensure_particle_terms(c));
    if (n == 0) return MC_OK;
    std::vector<unsigned> key(n);
    std::vector<int> idx(n);
    for (int64_t i = 0; i < n; i++) {
        if (cells[i] >= (uint64_t)c->N) { c->err = "particle terms: cell index out of range"; return MC_ERR_ARG; }
        key[i] = (unsigned)cells[i];
        idx[i] = (int)i;
    }
    int bits = 1;
    while (bits < 32 && ((uint64_t)1 << bits) < (uint64_t)c->N) bits++;
    size_t tb = 0;
    MC_HIP(c, hipcub::DeviceRadixSort::SortPairs(nullptr, tb, (const unsigned *)nullptr, (unsigned *)nullptr,
                                                 (const int *)nullptr, (int *)nullptr, (int)n, 0, bits, c->stream));
    const size_t b4 = up16((size_t)n * 4), b8 = up16((size_t)n * 40);
    void *base = nullptr;
    MC_TRY(couple_scratch(c, 4 * b4 + b8 + up16(tb), &base));
    char *p = (char *)base;
    unsigned *k_in = (unsigned *)p, *k_out = (unsigned *)(p + b4);
    int *i_in = (int *)(p + 2 * b4), *i_out = (int *)(p + 3 * b4);
    double *t_dev = (double *)(p + 4 * b4);
    void *tmp = p + 4 * b4 + b8;
    MC_HIP(c, hipMemcpyAsync(k_in, key.data(), (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
    MC_HIP(c, hipMemcpyAsync(i_in, idx.data(), (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
    MC_HIP(c, hipMemcpyAsync(t_dev, terms, (size_t)n * 40, hipMemcpyHostToDevice, c->stream));
    MC_HIP(c, hipcub::DeviceRadixSort::SortPairs(tmp, tb, k_in, k_out, i_in, i_out, (int)n, 0, bits, c->stream));
    hipLaunchKernelGGL(k_terms_add, dim3(div_up(n, 256)), dim3(256), 0, c->stream, (int)n, k_out, i_out, t_dev,
                       c->pterms);
    MC_LAUNCH_CHECK(c);
    // the host arrays above must outlive the copies
    MC_HIP(c, hipStreamSynchronize(c->stream));
    return MC_OK;
}

// ---- allocation-free variant of particle_terms_add -------------------------
// Fixed device scratch of mc_particle_terms_add_noalloc: the staged cell ids
// and term rows, the sort's keys and indices (in and out) and its temporary
// storage.  Each GPU the library runs on holds its own copy until the
// process ends.  Host code never passes `scratch` itself to a kernel or a
// copy (in host code the name is only a stand-in for the device symbol): it
// uses the address hipGetSymbolAddress returned for the context's device.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
constexpr int SCRATCH_DEVICES = 64;
static void *g_scratch_dev[SCRATCH_DEVICES];   // per device, resolved once

int noalloc_scratch_init(Ctx *c)
{
    if (c->device < 0 || c->device >= SCRATCH_DEVICES) {
        c->err = "device id out of range for the fixed scratch table";
        return MC_ERR_STATE;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[c->device]) {
        void *p = nullptr;
        // the context's device is current here (context creation)
        MC_HIP(c, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_dev[c->device] = p;
    }
    c->noalloc_scratch = g_scratch_dev[c->device];
    return MC_OK;
}

__global__ void k_keys_of_cells(int n, const uint64_t *cells, unsigned *key, int *idx)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    key[i] = (unsigned)cells[i];
    idx[i] = i;
}

// particle_terms_add with every device buffer in the fixed scratch: no
// device allocation, no allocating library call.  Calls that overlap on one
// GPU (two contexts on two streams) share the array and race.
int particle_terms_add_noalloc(Ctx *c, const uint64_t *cells, const double *terms, int64_t n)
{
    if (n < 0 || n > (int64_t)INT32_MAX / 8 || (n > 0 && (!cells || !terms))) {
        c->err = "particle terms: bad arguments";
        return MC_ERR_ARG;
    }
    if (!c->pterms) {
        c->err = "particle terms not allocated: call mc_particle_terms_reset once first (this entry point "
                 "allocates nothing)";
        return MC_ERR_STATE;
    }
    if (!c->noalloc_scratch) { c->err = "fixed scratch not resolved"; return MC_ERR_STATE; }
    if (n == 0) return MC_OK;
    for (int64_t i = 0; i < n; i++)
        if (cells[i] >= (uint64_t)c->N) { c->err = "particle terms: cell index out of range"; return MC_ERR_ARG; }
    int bits = 1;
    while (bits < 32 && ((uint64_t)1 << bits) < (uint64_t)c->N) bits++;
    size_t tb = 0;   // the size query allocates nothing
    MC_HIP(c, hipcub::DeviceRadixSort::SortPairs(nullptr, tb, (const unsigned *)nullptr, (unsigned *)nullptr,
                                                 (const int *)nullptr, (int *)nullptr, (int)n, 0, bits, c->stream));
    const size_t b8 = up16((size_t)n * 8), b4 = up16((size_t)n * 4), b40 = up16((size_t)n * 40);
    const size_t need = b8 + 4 * b4 + b40 + up16(tb);
    if (need > NOALLOC_SCRATCH_BYTES) {
        char msg[200];
        snprintf(msg, sizeof msg, "particle terms (no allocation): %lld rows need %zu bytes of scratch, above the "
                 "fixed %zu", (long long)n, need, NOALLOC_SCRATCH_BYTES);
        c->err = msg;
        return MC_ERR_ARG;
    }
    char *p = (char *)c->noalloc_scratch;
    uint64_t *c_dev = (uint64_t *)p;
    unsigned *k_in = (unsigned *)(p + b8), *k_out = (unsigned *)(p + b8 + b4);
    int *i_in = (int *)(p + b8 + 2 * b4), *i_out = (int *)(p + b8 + 3 * b4);
    double *t_dev = (double *)(p + b8 + 4 * b4);
    void *tmp = p + b8 + 4 * b4 + b40;
    MC_HIP(c, hipMemcpyAsync(c_dev, cells, (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
    MC_HIP(c, hipMemcpyAsync(t_dev, terms, (size_t)n * 40, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_keys_of_cells, dim3(div_up(n, 256)), dim3(256), 0, c->stream, (int)n, c_dev, k_in, i_in);
    MC_HIP(c, hipcub::DeviceRadixSort::SortPairs(tmp, tb, k_in, k_out, i_in, i_out, (int)n, 0, bits, c->stream));
    hipLaunchKernelGGL(k_terms_add, dim3(div_up(n, 256)), dim3(256), 0, c->stream, (int)n, k_out, i_out, t_dev,
                       c->pterms);
    MC_LAUNCH_CHECK(c);
    // the caller's host arrays must outlive the copies
    MC_HIP(c, hipStreamSynchronize(c->stream));
    return MC_OK;
}

// point -> cells in ascending cell order, built once from the cells' vertex lists
static int ensure_point_cells(Ctx *c)
{
    if (c->p2c_ptr) return MC_OK;
    const int64_t N = c->N, NP = c->NP;
    std::vector<uint64_t> cv((size_t)8 * N);
    MC_HIP(c, hipStreamSynchronize(c->stream));
    MC_HIP(c, hipMemcpy(cv.data(), c->cells, (size_t)8 * N * 8, hipMemcpyDeviceToHost));
    std::vector<int> ptr(NP + 1, 0), lst((size_t)8 * N);
    for (int64_t k = 0; k < 8 * N; k++) {
        if (cv[k] >= (uint64_t)NP) { c->err = "cell vertex out of range"; return MC_ERR_STATE; }
        ptr[cv[k] + 1]++;
    }
    for (int64_t p = 0; p < NP; p++) ptr[p + 1] += ptr[p];
    std::vector<int> fill(ptr.begin(), ptr.end() - 1);
    for (int64_t cc = 0; cc < N; cc++)
        for (int n = 0; n < 8; n++) lst[fill[cv[8 * cc + n]]++] = (int)cc;
    MemScope ms(c, MC_MEM_MESH);
    MC_TRY(dalloc(c, &c->p2c_ptr, NP + 1));
    MC_TRY(dalloc(c, &c->p2c_cell, 8 * N));
    MC_TRY(dalloc(c, &c->cell_flag, N));
    MC_HIP(c, hipMemcpy(c->p2c_ptr, ptr.data(), (size_t)(NP + 1) * 4, hipMemcpyHostToDevice));
    MC_HIP(c, hipMemcpy(c->p2c_cell, lst.data(), (size_t)8 * N * 4, hipMemcpyHostToDevice));
    return MC_OK;
}

int interpolate_to_nodes(Ctx *c, const uint64_t *cells, int64_t ncells, const uint64_t *nodes, int64_t nnodes,
                         double *out)
{
    if (ncells < 0 || nnodes < 0 || ncells > (int64_t)INT32_MAX / 8 || nnodes > (int64_t)INT32_MAX / 40 ||
        (ncells > 0 && !cells) || (nnodes > 0 && (!nodes || !out))) {
        c->err = "interpolation: bad arguments";
        return MC_ERR_ARG;
    }
    if (nnodes == 0) return MC_OK;
    std::vector<unsigned> cl(ncells), nl(nnodes);
    for (int64_t i = 0; i < ncells; i++) {
        if (cells[i] >= (uint64_t)c->N) { c->err = "interpolation: cell index out of range"; return MC_ERR_ARG; }
        cl[i] = (unsigned)cells[i];
    }
    for (int64_t i = 0; i < nnodes; i++) {
        if (nodes[i] >= (uint64_t)c->NP) { c->err = "interpolation: node index out of range"; return MC_ERR_ARG; }
        nl[i] = (unsigned)nodes[i];
    }
    MC_TRY(ensure_point_cells(c));
    const size_t bc = up16((size_t)ncells * 4), bn = up16((size_t)nnodes * 4), bo = up16((size_t)nnodes * 40);
    void *base = nullptr;
    MC_TRY(couple_scratch(c, bc + bn + bo, &base));
    char *p = (char *)base;
    unsigned *c_dev = (unsigned *)p, *n_dev = (unsigned *)(p + bc);
    double *o_dev = (double *)(p + bc + bn);
    MC_HIP(c, hipMemsetAsync(c->cell_flag, 0, (size_t)c->N, c->stream));
    if (ncells > 0) {
        MC_HIP(c, hipMemcpyAsync(c_dev, cl.data(), (size_t)ncells * 4, hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_flag_cells, dim3(div_up(ncells, 256)), dim3(256), 0, c->stream, (int)ncells, c_dev,
                           c->cell_flag);
    }
    MC_HIP(c, hipMemcpyAsync(n_dev, nl.data(), (size_t)nnodes * 4, hipMemcpyHostToDevice, c->stream));
    InterpArgs a;
    a.pptr = c->p2c_ptr; a.pcell = c->p2c_cell; a.flag = c->cell_flag; a.cpp = c->cells_per_point;
    a.px = c->px; a.py = c->py; a.pz = c->pz; a.cx = c->cx; a.cy = c->cy; a.cz = c->cz;
    static const int fl[5] = {U, V, W, P, TEM};
    for (int q = 0; q < 5; q++) {
        a.phi[q] = c->phi[fl[q]]; a.gx[q] = c->gx[fl[q]]; a.gy[q] = c->gy[fl[q]]; a.gz[q] = c->gz[fl[q]];
    }
    hipLaunchKernelGGL(k_interp_nodes, dim3(div_up(nnodes, 256)), dim3(256), 0, c->stream, (int)nnodes, n_dev, a,
                       o_dev);
    MC_LAUNCH_CHECK(c);
    MC_HIP(c, hipMemcpyAsync(out, o_dev, (size_t)nnodes * 40, hipMemcpyDeviceToHost, c->stream));
    MC_HIP(c, hipStreamSynchronize(c->stream));
    return MC_OK;
}

}  // namespace mc

extern "C" {

int mc_particle_terms_reset(mc_ctx *x)
{
    if (!x) return MC_ERR_ARG;
    return mc::particle_terms_reset(&x->c);
}

int mc_particle_terms_add(mc_ctx *x, const uint64_t *cells, const double *terms, int64_t n)
{
    if (!x) return MC_ERR_ARG;
    return mc::particle_terms_add(&x->c, cells, terms, n);
}

int mc_particle_terms_add_noalloc(mc_ctx *x, const uint64_t *cells, const double *terms, int64_t n)
{
    if (!x) return MC_ERR_ARG;
    return mc::particle_terms_add_noalloc(&x->c, cells, terms, n);
}

int mc_interpolate_to_nodes(mc_ctx *x, const uint64_t *cells, int64_t ncells, const uint64_t *nodes, int64_t nnodes,
                            double *node_fields)
{
    if (!x) return MC_ERR_ARG;
    return mc::interpolate_to_nodes(&x->c, cells, ncells, nodes, nnodes, node_fields);
}

}  // extern "C"
