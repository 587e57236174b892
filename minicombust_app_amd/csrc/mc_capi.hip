// mc_capi.hip -- the extern "C" entry points declared in include/mcflow.h.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

#include "mc_internal.hpp"


namespace mc {

static std::vector<void *> &allocs_of(Ctx *c) { return c->allocs; }

int alloc(Ctx *c, void **p, size_t bytes)
{
    void *q = nullptr;
    const hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) {
        c->err = std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e);
        *p = nullptr;
        return MC_ERR_NOMEM;
    }
    allocs_of(c).push_back(q);
    c->alloc_cat.push_back(c->mem_cat);
    c->alloc_bytes.push_back((int64_t)bytes);
    c->mem_bytes[c->mem_cat] += (int64_t)bytes;
    *p = q;
    return MC_OK;
}

int dfree_all(Ctx *c)
{
    for (void *p : allocs_of(c)) (void)hipFree(p);
    allocs_of(c).clear();
    c->alloc_cat.clear();
    c->alloc_bytes.clear();
    return MC_OK;
}

int dfree_category(Ctx *c, int cat)
{
    MC_HIP(c, hipStreamSynchronize(c->stream));
    if (c->side) MC_HIP(c, hipStreamSynchronize(c->side));
    std::vector<void *> &a = allocs_of(c);
    // a failing hipFree is recorded and the compaction finishes, so that no
    // entry is left twice in the list (mc_ctx_destroy frees what remains)
    hipError_t first_err = hipSuccess;
    size_t k = 0;
    for (size_t i = 0; i < a.size(); i++) {
        if (c->alloc_cat[i] == cat) {
            const hipError_t e = hipFree(a[i]);
            if (e != hipSuccess && first_err == hipSuccess) first_err = e;
            c->mem_bytes[cat] -= c->alloc_bytes[i];
            continue;
        }
        a[k] = a[i];
        c->alloc_cat[k] = c->alloc_cat[i];
        c->alloc_bytes[k] = c->alloc_bytes[i];
        k++;
    }
    a.resize(k);
    c->alloc_cat.resize(k);
    c->alloc_bytes.resize(k);
    if (first_err != hipSuccess) {
        c->err = std::string("hipFree: ") + hipGetErrorString(first_err);
        return MC_ERR_HIP;
    }
    return MC_OK;
}

static constexpr size_t EV_PAIRS_MAX = 4096;   // pending pairs before a fold

KTimer::KTimer(Ctx *c_, int kc_, double bytes_) : c(c_)
{
    if (kc_ < 0 || !(c->profile & (1u << kc_))) return;   // kc_ < 0: not timed
    if (c->ev_used == EV_PAIRS_MAX && profile_fold(c) != MC_OK) return;
    const size_t i = c->ev_used;
    if (2 * i + 2 > c->ev_pool.size()) {
        hipEvent_t a = nullptr, b = nullptr;
        if (hipEventCreate(&a) != hipSuccess) return;
        if (hipEventCreate(&b) != hipSuccess) { (void)hipEventDestroy(a); return; }
        c->ev_pool.push_back(a);
        c->ev_pool.push_back(b);
        c->ev_kc.push_back(0);
        c->ev_bytes.push_back(0.0);
    }
    if (hipEventRecord(c->ev_pool[2 * i], c->stream) != hipSuccess) return;
    c->ev_kc[i] = kc_;
    c->ev_bytes[i] = bytes_;
    c->ev_used++;
    slot = (long)i;
}

void KTimer::end()
{
    if (slot >= 0) (void)hipEventRecord(c->ev_pool[2 * slot + 1], c->stream);
    slot = -1;
}

int profile_fold(Ctx *c)
{
    if (!c->ev_used) return MC_OK;
    MC_HIP(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < c->ev_used; i++) {
        float ms = 0.f;
        MC_HIP(c, hipEventElapsedTime(&ms, c->ev_pool[2 * i], c->ev_pool[2 * i + 1]));
        c->kt_ms[c->ev_kc[i]] += ms;
        c->kt_calls[c->ev_kc[i]] += 1;
        c->kt_bytes[c->ev_kc[i]] += c->ev_bytes[i];
    }
    c->ev_used = 0;
    return MC_OK;
}

void profile_release(Ctx *c)
{
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    c->ev_pool.clear();
    c->ev_kc.clear();
    c->ev_bytes.clear();
    c->ev_used = 0;
}

static void set_constants(Ctx *c)
{
    Consts &k = c->k;
    // Mesh.hpp:114-120
    k.gas_vel_x = 50. * c->opt.inlet_scale; k.gas_vel_y = 0.00; k.gas_vel_z = 0.00;
    k.gas_pre = 4000.0; k.gas_tem = 2000.0; k.gas_te = 0.0001; k.gas_ed = 0.0; k.gas_ful = 0.1; k.gas_pro = 0.001;
    // FlowSolver.hpp:296-297, 421-432 (host libm, the reference's expressions)
    k.ev = 0.000014;
    const double velmag2 = pow(k.gas_vel_x, 2) + pow(k.gas_vel_y, 2) + pow(k.gas_vel_z, 2);
    k.inlet_ev = k.ev + 1.2 * 0.09 * (3.0 / 2.0 * ((0.1 * 0.1) * velmag2)) /
                            ((pow(0.09, 0.75) * pow((3.0 / 2.0 * ((0.1 * 0.1) * velmag2)), 1.5)) + 0.00000000000000000001);
    k.inlet_te = 3.0 / 2.0 * ((0.1 * 0.1) * velmag2);
    k.inlet_ed = pow(0.09, 0.75) * pow(k.inlet_te, 1.5);
    k.cmu75 = pow(0.09, 0.75);
    k.rdelta = 1.0 / c->opt.delta;
}

static int alloc_workspace(Ctx *c)
{
    MemScope ms(c, MC_MEM_WORK);
    const int N = c->N;
    for (int i = 0; i < 24; i++) MC_TRY(dalloc(c, &c->wk[i], N));
    for (int i = 0; i < 8; i++) {
        double *p = nullptr;
        MC_TRY(dalloc(c, &p, N));  // sized for doubles, used as float or double scratch
        c->wk32[i] = (float *)p;
    }
    MC_TRY(dalloc(c, &c->red, RED_PARTIALS));
    MC_TRY(dalloc(c, &c->dscal, 1024));
    MC_HIP(c, hipMemset(c->dscal, 0, 1024 * sizeof(double)));
    MC_HIP(c, hipHostMalloc((void **)&c->hscal, 1024 * sizeof(double), hipHostMallocDefault));
    memset(c->hscal, 0, 1024 * sizeof(double));
    MC_TRY(flow_workspace(c));
    return MC_OK;
}

// The constructor's first assemblies (FlowSolver.hpp:538-539) run on the
// zero face_fields and zero mass fluxes: they leave A_phi.U = 0, S_phi.U = 0,
// an all-zero momentum matrix and an all-zero Pressure_A, which is exactly
// the state mesh_finish leaves.
static int finish_ctor(Ctx *c)
{
    MC_TRY(mesh_finish(c));
    MC_TRY(alloc_workspace(c));
    MC_TRY(dia_setup_fine(c));
    c->first_pressure = true;
    c->timestep_count = 0;
    MC_HIP(c, hipStreamSynchronize(c->stream));
    return MC_OK;
}

}  // namespace mc

using namespace mc;

extern "C" {

int mc_options_default(mc_options *o)
{
    if (!o) return MC_ERR_ARG;
    memset(o, 0, sizeof(*o));
    o->mode = MC_MODE_COMPAT;
    o->psolve = MC_PSOLVE_FP64;
    o->learned = 0;
    o->solve_scalars = 1;
    o->delta = 1.0e-5;
    o->mom_rtol = 1.0e-5;
    o->mom_maxit = 10000;
    o->p_rtol = 1.0e-8;
    o->p_maxit = 2000;
    o->amg_nu = 3;             // m = 128 SIMPLE iter/s, 2 -> 3 sweeps: mixed 105.5 -> 107.0, recycled_mixed 115.3 -> 117.9
    o->amg_omega = 0.8;
    o->amg_cscale = 2.0;       // PCG iterations at 1e-8, 1.8 -> 2.0: m = 64 9.0 -> 8.3, m = 128 9.4 -> 8.1, m = 200 15.2 -> 13.0
    o->amg_coarse_max = 128;
    o->mixed_inner_rtol = 1.0e-5;  // m = 128 SIMPLE iter/s, 1e-3 -> 1e-5: mixed 103.7 -> 105.5, recycled 114.7 -> 115.4
    o->check_every = 2;
    o->inlet_scale = 1.0;
    o->amg_cheb_ratio = 4.0;   // Chebyshev weights on [0.5, 2]: 12.8 -> 9.8 PCG iterations at m = 128
    o->cell_tile = 0;
    o->limit_grad = 0;
    o->amg_refresh = 4;   // m = 128: same PCG iteration counts as 1, pressure solve 5% faster (measured)
    o->cell_ychunk = 0;
    o->warm_start = -1;
    o->amg_half = 1;
    o->amg_tail = 1;
    o->p_recycle = 0;
    o->pcg_graph = 0;
    return MC_OK;
}

static int validate(const mc_options *o, std::string &err)
{
    if (o->mode != MC_MODE_COMPAT && o->mode != MC_MODE_SIMPLE) { err = "mode"; return MC_ERR_ARG; }
    if (o->psolve != MC_PSOLVE_FP64 && o->psolve != MC_PSOLVE_MIXED) { err = "psolve"; return MC_ERR_ARG; }
    if (!(o->delta > 0) || !(o->mom_rtol > 0) || !(o->p_rtol > 0) || o->mom_maxit < 1 || o->p_maxit < 1) {
        err = "tolerances / iteration limits";
        return MC_ERR_ARG;
    }
    if (o->amg_coarse_max < 8 || o->amg_coarse_max > 128) { err = "amg_coarse_max must be in [8, 128]"; return MC_ERR_ARG; }
    if (o->amg_nu < 1 || o->amg_nu > 8) { err = "amg_nu must be in [1, 8]"; return MC_ERR_ARG; }
    if (!(o->inlet_scale > 0.0)) { err = "inlet_scale must be positive"; return MC_ERR_ARG; }
    if (o->cell_tile < 0 || o->cell_tile > 256 || (o->cell_tile & (o->cell_tile - 1))) {
        err = "cell_tile must be 0 or a power of two in [1, 256]";
        return MC_ERR_ARG;
    }
    if (o->limit_grad != 0 && o->limit_grad != 1) { err = "limit_grad must be 0 or 1"; return MC_ERR_ARG; }
    if (o->amg_refresh < 1 || o->amg_refresh > 1000) { err = "amg_refresh must be in [1, 1000]"; return MC_ERR_ARG; }
    if (o->amg_tail != 0 && o->amg_tail != 1) { err = "amg_tail must be 0 or 1"; return MC_ERR_ARG; }
    if (o->cell_ychunk < 0 || o->cell_ychunk > 4096) { err = "cell_ychunk must be in [0, 4096]"; return MC_ERR_ARG; }
    if (o->warm_start < -1 || o->warm_start > 1) { err = "warm_start must be -1, 0 or 1"; return MC_ERR_ARG; }
    if (o->amg_half != 0 && o->amg_half != 1) { err = "amg_half must be 0 or 1"; return MC_ERR_ARG; }
    if (o->p_recycle < 0 || o->p_recycle > DEFL_MAX) { err = "p_recycle must be in [0, 4]"; return MC_ERR_ARG; }
    if (o->pcg_graph != 0 && o->pcg_graph != 1) { err = "pcg_graph must be 0 or 1"; return MC_ERR_ARG; }
    return MC_OK;
}

static int create_common(const mc_options *opt, mc_ctx **out, mc_ctx **made)
{
    if (!out) return MC_ERR_ARG;
    *out = nullptr;
    // options first: a bad option is an argument error with or without a device
    std::string why;
    if (opt && validate(opt, why)) return MC_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MC_ERR_NODEV;
    mc_ctx *x = new (std::nothrow) mc_ctx();
    if (!x) return MC_ERR_NOMEM;
    Ctx *c = &x->c;
    if (opt) c->opt = *opt; else mc_options_default(&c->opt);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete x; return MC_ERR_HIP; }
    c->own_stream = true;
    if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_coarse, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_dense, hipEventDisableTiming) != hipSuccess) {
        (void)mc_ctx_destroy(x);
        return MC_ERR_HIP;
    }
    set_constants(c);
    *made = x;
    if (hipGetDevice(&c->device) != hipSuccess) {
        c->err = "hipGetDevice failed";
        *out = x;
        return MC_ERR_HIP;
    }
    if (int rc = noalloc_scratch_init(c)) {
        *out = x;
        return rc;
    }
    return MC_OK;
}

int mc_ctx_create_box(int64_t modifier, const mc_options *opt, mc_ctx **out)
{
    mc_ctx *x = nullptr;
    MC_TRY(create_common(opt, out, &x));
    Ctx *c = &x->c;
    int rc = mesh_box_generate(c, modifier);
    if (!rc) rc = finish_ctor(c);
    *out = x;  // returned even on failure so mc_last_error can report
    return rc;
}

int mc_ctx_create(const mc_mesh_desc *mesh, const mc_options *opt, mc_ctx **out)
{
    mc_ctx *x = nullptr;
    MC_TRY(create_common(opt, out, &x));
    Ctx *c = &x->c;
    int rc = mesh_from_host(c, mesh);
    if (!rc) rc = finish_ctor(c);
    *out = x;
    return rc;
}

int mc_ctx_destroy(mc_ctx *x)
{
    if (!x) return MC_OK;
    Ctx *c = &x->c;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->side) (void)hipStreamSynchronize(c->side);
    mpnn_destroy(c);
    profile_release(c);
    graph_release(c);
    dfree_all(c);
    if (c->hscal) (void)hipHostFree(c->hscal);
    if (c->stream && c->own_stream) (void)hipStreamDestroy(c->stream);
    if (c->ev_coarse) (void)hipEventDestroy(c->ev_coarse);
    if (c->ev_dense) (void)hipEventDestroy(c->ev_dense);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->side) (void)hipStreamDestroy(c->side);
    delete x;
    return MC_OK;
}

const char *mc_last_error(mc_ctx *x) { return x ? x->c.err.c_str() : "null context"; }

int mc_ctx_sizes(mc_ctx *x, int64_t *n_cells, int64_t *n_faces, int64_t *nnz, int64_t *n_points)
{
    if (!x) return MC_ERR_ARG;
    if (n_cells) *n_cells = x->c.N;
    if (n_faces) *n_faces = x->c.F;
    if (nnz) *nnz = x->c.nnz;
    if (n_points) *n_points = x->c.NP;
    return MC_OK;
}

int mc_set_options(mc_ctx *x, const mc_options *opt)
{
    if (!x || !opt) return MC_ERR_ARG;
    Ctx *c = &x->c;
    MC_TRY(validate(opt, c->err));
    // the hierarchy is rebuilt for a new coarse bound, and when the FP16
    // smoothing copy is switched on (it is allocated with the hierarchy)
    const bool amg_change = opt->amg_coarse_max != c->opt.amg_coarse_max || (opt->amg_half && !c->opt.amg_half);
    c->opt = *opt;
    set_constants(c);
    if (amg_change) c->amg_ready = false;
    c->graph_gen++;   // smoother weights and level forms are baked into the PCG graphs
    return MC_OK;
}

int mc_reset_fields(mc_ctx *x)
{
    if (!x) return MC_ERR_ARG;
    Ctx *c = &x->c;
    MC_TRY(init_fields(c));
    MC_HIP(c, hipMemsetAsync(c->mflux, 0, (size_t)c->F * 8, c->stream));
    MC_HIP(c, hipMemsetAsync(c->ff0, 0, (size_t)c->F * 8, c->stream));
    MC_HIP(c, hipMemsetAsync(c->ff1, 0, (size_t)c->F * 8, c->stream));
    MC_HIP(c, hipMemsetAsync(c->fcorr, 0, (size_t)c->F * 8, c->stream));
    MC_HIP(c, hipMemsetAsync(c->A_vals, 0, (size_t)c->nnz * 8, c->stream));
    if (c->fine_nd) {
        MC_HIP(c, hipMemsetAsync(c->A_dia, 0, (size_t)c->fine_nd * c->N * 8, c->stream));
        MC_HIP(c, hipMemsetAsync(c->A_dia2, 0, (size_t)c->fine_nd * c->N * 8, c->stream));
    }
    MC_HIP(c, hipMemsetAsync(c->PA_vals, 0, (size_t)c->nnz * 8, c->stream));
    MC_HIP(c, hipMemsetAsync(c->Pb, 0, (size_t)c->N * 8, c->stream));
    MC_HIP(c, hipStreamSynchronize(c->stream));
    c->pa_dia_fresh = false;
    c->first_pressure = true;
    c->field_set_up = false;
    c->timestep_count = 0;
    c->defl_count = 0;   // the recycled solutions belong to the previous run
    c->defl_head = 0;
    return MC_OK;
}

#define CTX_OR_FAIL \
    if (!x) return MC_ERR_ARG; \
    Ctx *c = &x->c;

int mc_timestep(mc_ctx *x)
{
    CTX_OR_FAIL
    for (int i = 0; i < 3; i++) MC_TRY(simple_iteration(c, i));
    c->timestep_count++;
    MC_HIP(c, hipStreamSynchronize(c->stream));
    return MC_OK;
}

int mc_simple_iteration(mc_ctx *x, int inner)
{
    CTX_OR_FAIL
    MC_TRY(simple_iteration(c, inner));
    return MC_OK;
}

int mc_get_phi_gradients(mc_ctx *x) { CTX_OR_FAIL return grad_multi(c, false); }
int mc_limit_phi_gradients(mc_ctx *x) { CTX_OR_FAIL return limit_gradients(c); }

int mc_repair_nonneg(mc_ctx *x, int field)
{
    CTX_OR_FAIL
    if (field < 0 || field >= MC_NFIELDS) return MC_ERR_ARG;
    return clip_nonneg(c, c->phi[field], field == MC_ED ? 1 : 0);
}

int mc_get_phi_gradient(mc_ctx *x, int field, int pressure)
{
    CTX_OR_FAIL
    if (field < 0 || field >= MC_NFIELDS) return MC_ERR_ARG;
    return grad_single(c, field, pressure != 0);
}

int mc_set_up_field(mc_ctx *x) { CTX_OR_FAIL return set_up_field(c); }
int mc_momentum_assemble(mc_ctx *x) { CTX_OR_FAIL return momentum_assemble(c); }

int mc_momentum_solve(mc_ctx *x, mc_solve_status *st)
{
    CTX_OR_FAIL
    double *b[3] = {c->SU, c->SV, c->SW};
    double *xx[3] = {c->phi[MC_U], c->phi[MC_V], c->phi[MC_W]};
    mc_solve_status s{};
    MC_TRY(bicgstab_solve(c, c->A_vals, 3, b, xx, c->opt.mom_rtol, c->opt.mom_maxit, &s, warm_start(c)));
    c->st.mom_iters += s.iterations;
    if (st) *st = s;
    if (!s.converged) { c->err = "momentum BiCGStab did not converge"; return MC_ERR_NOCONV; }
    return MC_OK;
}

int mc_pressure_assemble(mc_ctx *x) { CTX_OR_FAIL return calculate_mass_flux_assemble(c); }

int mc_pressure_solve(mc_ctx *x, mc_solve_status *st)
{
    CTX_OR_FAIL
    if (!c->amg_ready) MC_TRY(amg_setup(c, c->PA_vals));
    else MC_TRY(amg_update_values(c, c->PA_vals));
    return pressure_solve(c, c->Pb, c->phi[MC_PP], st);
}

int mc_calculate_pressure(mc_ctx *x) { CTX_OR_FAIL return calculate_pressure(c); }

int mc_scalar_assemble(mc_ctx *x, int type)
{
    CTX_OR_FAIL
    if (type < 0 || type > 6) return MC_ERR_ARG;
    return scalar_assemble(c, type);
}

int mc_scalar_solve(mc_ctx *x, int type)
{
    CTX_OR_FAIL
    if (type < 0 || type > 6) return MC_ERR_ARG;
    return scalar_solve(c, type);
}

int mc_scalar_batch_solve(mc_ctx *x) { CTX_OR_FAIL return scalar_batch_solve(c); }

int mc_update_boundary_slots(mc_ctx *x, int field)
{
    CTX_OR_FAIL
    if (field < 0 || field >= MC_NFIELDS) return MC_ERR_ARG;
    return pressure_boundary_slots(c, field);
}

int mc_update_P(mc_ctx *x) { CTX_OR_FAIL return update_P(c); }

int mc_spmv(mc_ctx *x, int which, const double *xh, double *yh)
{
    CTX_OR_FAIL
    if (!xh || !yh) return MC_ERR_ARG;
    const size_t nb = (size_t)c->N * 8;
    MC_HIP(c, hipMemcpyAsync(c->wk[8], xh, nb, hipMemcpyHostToDevice, c->stream));
    MC_TRY(spmv(c, which ? c->PA_vals : c->A_vals, c->wk[8], c->wk[9]));
    MC_HIP(c, hipMemcpyAsync(yh, c->wk[9], nb, hipMemcpyDeviceToHost, c->stream));
    MC_HIP(c, hipStreamSynchronize(c->stream));
    return MC_OK;
}

int mc_solve_pressure_host(mc_ctx *x, const double *bh, double *xh, mc_solve_status *st)
{
    CTX_OR_FAIL
    if (!bh || !xh) return MC_ERR_ARG;
    const size_t nb = (size_t)c->N * 8;
    if (!c->amg_ready) MC_TRY(amg_setup(c, c->PA_vals));
    else MC_TRY(amg_update_values(c, c->PA_vals));
    MC_HIP(c, hipMemcpyAsync(c->wk[10], bh, nb, hipMemcpyHostToDevice, c->stream));
    const int rc = pressure_solve(c, c->wk[10], c->wk[11], st);
    MC_HIP(c, hipMemcpyAsync(xh, c->wk[11], nb, hipMemcpyDeviceToHost, c->stream));
    MC_HIP(c, hipStreamSynchronize(c->stream));
    return rc;
}

// ------------------------------------------------------------ arrays --
struct ArrDesc {
    void *ptr;
    int64_t count;   // elements
    int kind;        // 0 f64 plain, 1 f64 x3 from SoA, 2 i32, 3 u64 from i32 (faces pairs), 4 u64 from i32 (-1 -> max),
                     // 5 u64 plain, 6 u8, 7 points x3 SoA, 8 u64 cell_faces from i32
    const double *sx, *sy, *sz;
    const int *a, *b;
    bool face;       // one entry per face, stored in the context's face order
};

// reference face id <-> storage position, host copies made on first use
static int face_maps(Ctx *c)
{
    if ((int64_t)c->h_face_pos.size() == c->F) return MC_OK;
    c->h_face_pos.resize(c->F);
    c->h_face_ref.resize(c->F);
    MC_HIP(c, hipStreamSynchronize(c->stream));
    MC_HIP(c, hipMemcpy(c->h_face_pos.data(), c->face_pos, (size_t)c->F * 4, hipMemcpyDeviceToHost));
    MC_HIP(c, hipMemcpy(c->h_face_ref.data(), c->face_ref, (size_t)c->F * 4, hipMemcpyDeviceToHost));
    return MC_OK;
}

static int describe(Ctx *c, int id, ArrDesc &d)
{
    memset(&d, 0, sizeof(d));
    const int64_t NS = c->N + 6;
    if (id >= MC_ARR_PHI && id < MC_ARR_PHI + MC_NFIELDS) { d.ptr = c->phi[id]; d.count = NS; return MC_OK; }
    if (id >= MC_ARR_GRAD && id < MC_ARR_GRAD + MC_NFIELDS) {
        const int f = id - MC_ARR_GRAD;
        d.kind = 1; d.count = NS * 3; d.sx = c->gx[f]; d.sy = c->gy[f]; d.sz = c->gz[f];
        return MC_OK;
    }
    if (id >= MC_ARR_PDEFL && id < MC_ARR_PDEFL + DEFL_MAX) {
        const int k = id - MC_ARR_PDEFL;
        if (k >= c->defl_count || c->defl_cap != c->opt.p_recycle) {
            c->err = "no recycled pressure solution in that slot";
            return MC_ERR_STATE;
        }
        d.kind = 5; d.ptr = c->defl_x[(c->defl_head - k + c->defl_cap) % c->defl_cap]; d.count = c->N;
        return MC_OK;
    }
    switch (id) {
    case MC_ARR_AU: d.ptr = c->AU; d.count = NS; return MC_OK;
    case MC_ARR_AV: d.ptr = c->AV; d.count = NS; return MC_OK;
    case MC_ARR_AW: d.ptr = c->AW; d.count = NS; return MC_OK;
    case MC_ARR_SU: d.ptr = c->SU; d.count = NS; return MC_OK;
    case MC_ARR_SV: d.ptr = c->SV; d.count = NS; return MC_OK;
    case MC_ARR_SW: d.ptr = c->SW; d.count = NS; return MC_OK;
    case MC_ARR_FF0: d.ptr = c->ff0; d.count = c->F; d.face = true; return MC_OK;
    case MC_ARR_FF1: d.ptr = c->ff1; d.count = c->F; d.face = true; return MC_OK;
    case MC_ARR_MFLUX: d.ptr = c->mflux; d.count = c->F; d.face = true; return MC_OK;
    case MC_ARR_FACE_AREA: d.ptr = c->farea; d.count = c->F; d.face = true; return MC_OK;
    case MC_ARR_FACE_LAMBDA: d.ptr = c->flam; d.count = c->F; d.face = true; return MC_OK;
    case MC_ARR_FACE_RLENCOS: d.ptr = c->frlc; d.count = c->F; d.face = true; return MC_OK;
    case MC_ARR_FACE_NORMAL: d.kind = 1; d.count = (int64_t)c->F * 3; d.sx = c->fnx; d.sy = c->fny; d.sz = c->fnz; d.face = true; return MC_OK;
    case MC_ARR_FACE_CENTER: d.kind = 1; d.count = (int64_t)c->F * 3; d.sx = c->fcx; d.sy = c->fcy; d.sz = c->fcz; d.face = true; return MC_OK;
    case MC_ARR_CELL_CENTER: d.kind = 1; d.count = (int64_t)c->N * 3; d.sx = c->cx; d.sy = c->cy; d.sz = c->cz; return MC_OK;
    case MC_ARR_CELL_VOLUME: d.ptr = c->vol; d.count = c->N; return MC_OK;
    case MC_ARR_CELL_DENSITY: d.ptr = c->rho; d.count = c->N; return MC_OK;
    case MC_ARR_ROW_PTR: d.kind = 2; d.ptr = c->row_ptr; d.count = c->N + 1; return MC_OK;
    case MC_ARR_COL: d.kind = 2; d.ptr = c->col; d.count = c->nnz; return MC_OK;
    case MC_ARR_A_VALS: d.ptr = c->A_vals; d.count = c->nnz; return MC_OK;
    case MC_ARR_PA_VALS: d.ptr = c->PA_vals; d.count = c->nnz; return MC_OK;
    case MC_ARR_PB: d.ptr = c->Pb; d.count = c->N; return MC_OK;
    case MC_ARR_FACES: d.kind = 3; d.count = (int64_t)c->F * 2; d.a = c->fc0; d.b = c->fc1; d.face = true; return MC_OK;
    case MC_ARR_CELL_FACES: d.kind = 8; d.count = (int64_t)c->N * 6; d.a = c->cell_faces; return MC_OK;
    case MC_ARR_CELL_NEIGHBOURS: d.kind = 4; d.count = (int64_t)c->N * 6; d.a = c->nbr; return MC_OK;
    case MC_ARR_CELLS: d.kind = 5; d.ptr = c->cells; d.count = (int64_t)c->N * 8; return MC_OK;
    case MC_ARR_POINTS: d.kind = 7; d.count = (int64_t)c->NP * 3; d.sx = c->px; d.sy = c->py; d.sz = c->pz; return MC_OK;
    case MC_ARR_PARTICLE_TERMS:
        if (!c->pterms) { c->err = "no particle terms yet (mc_particle_terms_reset / _add)"; return MC_ERR_STATE; }
        d.ptr = c->pterms; d.count = (int64_t)c->N * 5; return MC_OK;
    case MC_ARR_CELLS_PER_POINT: d.kind = 6; d.ptr = c->cells_per_point; d.count = c->NP; return MC_OK;
    default: c->err = "unknown array id"; return MC_ERR_ARG;
    }
}

int mc_array_get(mc_ctx *x, int id, void *host, int64_t count)
{
    CTX_OR_FAIL
    ArrDesc d;
    MC_TRY(describe(c, id, d));
    if (!host || count != d.count) { c->err = "array size mismatch (expected " + std::to_string(d.count) + ")"; return MC_ERR_ARG; }
    // a matrix held in diagonal storage is handed out in CSR order
    if (id == MC_ARR_A_VALS && fine_dia_of(c, c->A_vals)) MC_TRY(dia_to_csr(c, fine_dia_of(c, c->A_vals), c->A_vals));
    MC_HIP(c, hipStreamSynchronize(c->stream));
    if (d.face || d.kind == 8) MC_TRY(face_maps(c));
    const int *pos = d.face ? c->h_face_pos.data() : nullptr;   // reference id -> stored entry
    if (d.kind == 0 && !d.face) {
        MC_HIP(c, hipMemcpy(host, d.ptr, (size_t)count * 8, hipMemcpyDeviceToHost));
    } else if (d.kind == 0) {
        std::vector<double> t(count);
        MC_HIP(c, hipMemcpy(t.data(), d.ptr, (size_t)count * 8, hipMemcpyDeviceToHost));
        double *o = (double *)host;
        for (int64_t r = 0; r < count; r++) o[r] = t[pos[r]];
    } else if (d.kind == 1 || d.kind == 7) {
        const int64_t n = count / 3;
        std::vector<double> t(n);
        double *o = (double *)host;
        const double *src[3] = {d.sx, d.sy, d.sz};
        for (int k = 0; k < 3; k++) {
            MC_HIP(c, hipMemcpy(t.data(), src[k], (size_t)n * 8, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < n; i++) o[3 * i + k] = t[pos ? pos[i] : i];
        }
    } else if (d.kind == 2) {
        MC_HIP(c, hipMemcpy(host, d.ptr, (size_t)count * 4, hipMemcpyDeviceToHost));
    } else if (d.kind == 3) {
        const int64_t n = count / 2;
        std::vector<int> a(n), b(n);
        MC_HIP(c, hipMemcpy(a.data(), d.a, (size_t)n * 4, hipMemcpyDeviceToHost));
        MC_HIP(c, hipMemcpy(b.data(), d.b, (size_t)n * 4, hipMemcpyDeviceToHost));
        uint64_t *o = (uint64_t *)host;
        for (int64_t i = 0; i < n; i++) {
            const int64_t j = pos ? pos[i] : i;
            o[2 * i] = (uint64_t)a[j];
            o[2 * i + 1] = (uint64_t)b[j];
        }
    } else if (d.kind == 4 || d.kind == 8) {
        std::vector<int> a(count);
        MC_HIP(c, hipMemcpy(a.data(), d.a, (size_t)count * 4, hipMemcpyDeviceToHost));
        uint64_t *o = (uint64_t *)host;
        if (d.kind == 8)   // storage positions -> reference face ids
            for (int64_t i = 0; i < count; i++) o[i] = (uint64_t)c->h_face_ref[a[i]];
        else
            for (int64_t i = 0; i < count; i++) o[i] = a[i] < 0 ? UINT64_MAX : (uint64_t)a[i];
    } else if (d.kind == 5) {
        MC_HIP(c, hipMemcpy(host, d.ptr, (size_t)count * 8, hipMemcpyDeviceToHost));
    } else if (d.kind == 6) {
        MC_HIP(c, hipMemcpy(host, d.ptr, (size_t)count, hipMemcpyDeviceToHost));
    }
    return MC_OK;
}

int mc_array_set(mc_ctx *x, int id, const void *host, int64_t count)
{
    CTX_OR_FAIL
    c->pa_dia_fresh = false;   // the AMG fine level is refreshed from whatever the arrays now hold
    ArrDesc d;
    MC_TRY(describe(c, id, d));
    if (!host || count != d.count) { c->err = "array size mismatch (expected " + std::to_string(d.count) + ")"; return MC_ERR_ARG; }
    MC_HIP(c, hipStreamSynchronize(c->stream));
    if (d.face) MC_TRY(face_maps(c));
    const int *pos = d.face ? c->h_face_pos.data() : nullptr;
    if (d.kind == 0 && !d.face) {
        MC_HIP(c, hipMemcpy(d.ptr, host, (size_t)count * 8, hipMemcpyHostToDevice));
    } else if (d.kind == 0) {
        std::vector<double> t(count);
        const double *in = (const double *)host;
        for (int64_t r = 0; r < count; r++) t[pos[r]] = in[r];
        MC_HIP(c, hipMemcpy(d.ptr, t.data(), (size_t)count * 8, hipMemcpyHostToDevice));
    } else if (d.kind == 1) {
        const int64_t n = count / 3;
        std::vector<double> t(n);
        const double *in = (const double *)host;
        const double *dst[3] = {d.sx, d.sy, d.sz};
        for (int k = 0; k < 3; k++) {
            for (int64_t i = 0; i < n; i++) t[pos ? pos[i] : i] = in[3 * i + k];
            MC_HIP(c, hipMemcpy((void *)dst[k], t.data(), (size_t)n * 8, hipMemcpyHostToDevice));
        }
    } else {
        c->err = "array is read-only";
        return MC_ERR_ARG;
    }
    if (id == MC_ARR_A_VALS && fine_dia_of(c, c->A_vals))
        MC_TRY(csr_to_dia(c, c->A_vals, (double *)fine_dia_of(c, c->A_vals)));
    return MC_OK;
}

int mc_get_stats(mc_ctx *x, mc_stats *out)
{
    CTX_OR_FAIL
    if (!out) return MC_ERR_ARG;
    *out = c->st;
    return MC_OK;
}

int mc_state_flags(mc_ctx *x, int *first_pressure, int64_t *timestep_count)
{
    CTX_OR_FAIL
    if (first_pressure) *first_pressure = c->first_pressure ? 1 : 0;
    if (timestep_count) *timestep_count = c->timestep_count;
    return MC_OK;
}

int mc_set_state_flags(mc_ctx *x, int first_pressure, int64_t timestep_count)
{
    CTX_OR_FAIL
    c->first_pressure = first_pressure != 0;
    c->timestep_count = timestep_count;
    return MC_OK;
}

int mc_sync(mc_ctx *x)
{
    CTX_OR_FAIL
    MC_HIP(c, hipStreamSynchronize(c->stream));
    return MC_OK;
}

int mc_profile_enable(mc_ctx *x, int mask)
{
    CTX_OR_FAIL
    MC_TRY(profile_fold(c));
    const unsigned kc = (1u << MC_KC_COUNT) - 1u;
    c->profile = (c->profile & ~kc) | ((unsigned)mask & kc);
    return MC_OK;
}

int mc_stage_enable(mc_ctx *x, int on)
{
    CTX_OR_FAIL
    MC_TRY(profile_fold(c));
    const unsigned st = ((1u << MC_ST_COUNT) - 1u) << MC_KC_COUNT;
    c->profile = on ? (c->profile | st) : (c->profile & ~st);
    return MC_OK;
}

int mc_stage_read(mc_ctx *x, double *ms, int64_t *calls)
{
    CTX_OR_FAIL
    MC_TRY(profile_fold(c));
    for (int k = 0; k < MC_ST_COUNT; k++) {
        if (ms) ms[k] = c->kt_ms[MC_KC_COUNT + k];
        if (calls) calls[k] = c->kt_calls[MC_KC_COUNT + k];
    }
    return MC_OK;
}

int mc_stage_reset(mc_ctx *x)
{
    CTX_OR_FAIL
    MC_TRY(profile_fold(c));
    for (int k = MC_KC_COUNT; k < MC_KC_COUNT + MC_ST_COUNT; k++) { c->kt_ms[k] = 0; c->kt_calls[k] = 0; }
    return MC_OK;
}

int mc_profile_read(mc_ctx *x, double *ms, int64_t *calls, double *bytes)
{
    CTX_OR_FAIL
    MC_TRY(profile_fold(c));
    for (int k = 0; k < MC_KC_COUNT; k++) {
        if (ms) ms[k] = c->kt_ms[k];
        if (calls) calls[k] = c->kt_calls[k];
        if (bytes) bytes[k] = c->kt_bytes[k];
    }
    return MC_OK;
}

int mc_profile_reset(mc_ctx *x)
{
    CTX_OR_FAIL
    MC_TRY(profile_fold(c));
    for (int k = 0; k < MC_KC_COUNT; k++) { c->kt_ms[k] = 0; c->kt_calls[k] = 0; c->kt_bytes[k] = 0; }
    memset(&c->st, 0, sizeof(c->st));
    return MC_OK;
}

int mc_ctx_memory(mc_ctx *x, int64_t *bytes)
{
    CTX_OR_FAIL
    if (!bytes) return MC_ERR_ARG;
    for (int k = 0; k < MC_MEM_COUNT; k++) bytes[k] = c->mem_bytes[k];
    return MC_OK;
}

int mc_copy_probe(mc_ctx *x, int64_t bytes, int reps, double *ms_out)
{
    CTX_OR_FAIL
    if (bytes < 16 || bytes % 16 || reps < 1 || !ms_out) {
        c->err = "copy probe: bytes a positive multiple of 16, reps >= 1, ms_out set";
        return MC_ERR_ARG;
    }
    MC_TRY(copy_probe(c, bytes, reps, ms_out));
    return MC_OK;
}

}  // extern "C"
