// mc_flow.hip -- the SIMPLE-step kernels and their host drivers.
//
// Every reference face loop that scatters into two cells (S_phi[cell0] -=,
// S_phi[cell1] += ...) is recast as a cell-centric gather: one thread per
// cell walks its six faces in ascending face id -- the order in which the
// sequential face loop reaches them -- and performs the same additions in the
// same order, so sums round exactly as in the reference, with no atomics.  The
// owner (cell0) of an interior face writes the face's own results
// (face_fields, mass flux).  Boundary-side slots (phi[N + side]) are written
// by the owner of the highest-numbered face of that side: the reference's
// last writer.  Compiled without FP contraction so each expression rounds as
// written in FlowSolver.inl.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

#include "mc_internal.hpp"

#pragma clang fp contract(off)

namespace mc {

struct Geo {
    int N, F;
    const int *fc0, *fc1, *forder, *fother, *cell_faces, *nbr, *row_ptr, *diag_idx;
    const double *cx, *cy, *cz, *fcx, *fcy, *fcz, *fnx, *fny, *fnz, *farea, *flam, *frlc, *vol, *rho;
    const uint8_t *is_wall;
    int btype[6];
    int last_face[6];
    CellWalk w;
    int dnd;                 // diagonals of the momentum / scalar matrix storage (0: CSR)
    int doff[DIA_MAX];       // their column offsets, ascending
};

static Geo geo(Ctx *c)
{
    Geo g;
    g.N = c->N; g.F = c->F;
    g.fc0 = c->fc0; g.fc1 = c->fc1; g.forder = c->forder; g.fother = c->fother; g.cell_faces = c->cell_faces;
    g.nbr = c->nbr;
    g.row_ptr = c->row_ptr; g.diag_idx = c->diag_idx;
    g.cx = c->cx; g.cy = c->cy; g.cz = c->cz; g.fcx = c->fcx; g.fcy = c->fcy; g.fcz = c->fcz;
    g.fnx = c->fnx; g.fny = c->fny; g.fnz = c->fnz; g.farea = c->farea; g.flam = c->flam; g.frlc = c->frlc;
    g.vol = c->vol; g.rho = c->rho; g.is_wall = c->is_wall_cell;
    for (int s = 0; s < 6; s++) { g.btype[s] = c->btype[s]; g.last_face[s] = c->last_face[s]; }
    g.w = cell_walk(c);
    g.dnd = c->fine_nd;
    for (int d = 0; d < DIA_MAX; d++) g.doff[d] = d < c->fine_nd ? c->fine_off[d] : 0;
    return g;
}

struct dv3 { double x, y, z; };
__device__ __forceinline__ dv3 v3(double x, double y, double z) { dv3 r; r.x = x; r.y = y; r.z = z; return r; }
__device__ __forceinline__ dv3 operator+(dv3 a, dv3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ dv3 operator-(dv3 a, dv3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ dv3 operator*(dv3 a, double b) { return v3(a.x * b, a.y * b, a.z * b); }
__device__ __forceinline__ dv3 operator/(dv3 a, double b) { return v3(a.x / b, a.y / b, a.z / b); }
__device__ __forceinline__ double dot(dv3 a, dv3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double mag(dv3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ dv3 ld3(const double *x, const double *y, const double *z, int i) { return v3(x[i], y[i], z[i]); }

// Position of each entry of row c in the sorted CSR row {c} u {nb_k}.
__device__ __forceinline__ void row_positions(int c, int n_int, const int *nb, int *pos, int &dpos)
{
    dpos = 0;
    for (int k = 0; k < n_int; k++) dpos += nb[k] < c;
    for (int k = 0; k < n_int; k++) {
        int p = (c < nb[k]);
        for (int t = 0; t < n_int; t++) p += nb[t] < nb[k];
        pos[k] = p;
    }
}

// Diagonal of column offset `off` (present in every row's pattern).
__device__ __forceinline__ int dslot(const Geo &g, int off)
{
    int d = 0;
    while (d < g.dnd - 1 && g.doff[d] != off) d++;
    return d;
}

// Row c of a momentum / scalar matrix: with a diagonal pattern each entry
// goes to the diagonal of its column offset (the solvers read that storage);
// otherwise to the sorted-column CSR positions.
__device__ __forceinline__ void store_row(const Geo &g, int c, int nint, const int *nb, const double *v, double diag,
                                          double *vals, double *dia)
{
    if (dia) {
        const size_t N = (size_t)g.N;
        st_out(&dia[(size_t)dslot(g, 0) * N + c], diag);
        for (int k = 0; k < nint; k++) st_out(&dia[(size_t)dslot(g, nb[k] - c) * N + c], v[k]);
    } else {
        int pos[6], dpos;
        row_positions(c, nint, nb, pos, dpos);
        const int rp = g.row_ptr[c];
        vals[rp + dpos] = diag;
        for (int k = 0; k < nint; k++) vals[rp + pos[k]] = v[k];
    }
}

// A cell's six face ids (reference face order) and their fother codes,
// loaded ahead of the face loop so that each face's gathers wait on one
// memory round trip instead of two (7-9% faster assembly at m = 128).
struct CellFaces {
    int f[6], o[6];
};
__device__ __forceinline__ CellFaces cell_faces_of(const Geo &g, int c)
{
    CellFaces r;
#pragma unroll
    for (int k = 0; k < 6; k++) { r.f[k] = g.forder[6 * c + k]; r.o[k] = g.fother[6 * c + k]; }
    return r;
}

// ------------------------------------------------- algorithmic bytes --
// Bytes of one launch with every array the kernel reads or writes counted
// once (a neighbour's value read again is free): FP64 values, int32 indices,
// per-cell arrays x N, arrays of all faces x F, of interior faces only x F_i,
// of boundary faces only x F_b, matrix entries x nnz (the entries a row
// stores: the PETSc pattern).  Each kernel's table lists its arrays.
struct Sizes {
    double N, F, Fi, Fb, nnz;
};
static Sizes sizes_of(const Ctx *c)
{
    return Sizes{(double)c->N, (double)c->F, (double)(c->F - c->Fb), (double)c->Fb, (double)c->nnz};
}

// ========================================================== A4 / A5 ==
struct GradArgs {
    const double *phi[11];
    double *gx[11], *gy[11], *gz[11];
    int zero_boundary[11];
};

// limit_phi_gradient (FlowSolver.inl:804-873), Venkatakrishnan's limiter
// for one cell: pmax / pmin are phi's extremes over the cell and its six face
// neighbours (boundary slots at the boundary), ds[s] = x_face - x_c for the
// faces in side order; the gradient is scaled by the least
// (r^2 + 2r) / (r^2 + r + 2) over the faces.  pow(r, 2) is evaluated as r * r.
__device__ __forceinline__ void venkat(double phic, double pmax, double pmin, const dv3 (&ds)[6], double &gx,
                                       double &gy, double &gz)
{
    const double deltamax = pmax - phic, deltamin = pmin - phic;
    double alpha = 1.0;
    for (int s = 0; s < 6; s++) {
        const double delta_face = gx * ds[s].x + gy * ds[s].y + gz * ds[s].z;
        double r;
        if (fabs(delta_face) < 0.000006) r = 1000.0;
        else if (delta_face > 0.0) r = deltamax / delta_face;
        else r = deltamin / delta_face;
        alpha = dmin(alpha, (r * r + 2.0 * r) / (r * r + r + 2.0));
    }
    gx = gx * alpha; gy = gy * alpha; gz = gz * alpha;
}

// get_phi_gradients / get_phi_gradient (FlowSolver.inl:738-803, 917-1063):
// A = sum dX dX^T, b_k = sum dX dphi_k over the six faces in side order,
// then the 3x3 Cramer solve of FlowSolver.inl:895-915.  For an interior face
// mask*(cc1-cc0) and mask*(phi1-phi0) equal cc[nb]-cc[c] and phi[nb]-phi[c]
// exactly (IEEE subtraction is antisymmetric).  LIMIT: limit_phi_gradients
// (FlowSolver.inl:875-893) follows in the same pass, from the gradient just
// formed and the neighbour values this cell has already read.
template <int NF, bool LIMIT = false>
__global__ void __launch_bounds__(256) k_grad(Geo g, GradArgs a)
{
    const int c = walk_cell(g.w);
    if (c < 0) return;
    const dv3 cc = ld3(g.cx, g.cy, g.cz, c);
    double ph[NF];
#pragma unroll
    for (int k = 0; k < NF; k++) ph[k] = a.phi[k][c];
    double a00 = 0.0, a10 = 0.0, a20 = 0.0, a11 = 0.0, a21 = 0.0, a22 = 0.0;
    double bx[NF], by[NF], bz[NF], pmax[LIMIT ? NF : 1], pmin[LIMIT ? NF : 1];
#pragma unroll
    for (int k = 0; k < NF; k++) { bx[k] = 0.0; by[k] = 0.0; bz[k] = 0.0; }
    if constexpr (LIMIT) {
#pragma unroll
        for (int k = 0; k < NF; k++) { pmax[k] = ph[k]; pmin[k] = ph[k]; }
    }
    int nbs[6];   // loaded ahead: each side's gathers then wait on one round trip
#pragma unroll
    for (int s = 0; s < 6; s++) nbs[s] = g.nbr[6 * c + s];
    // one side per trip: unrolling by 2 or 6 (more loads in flight, two waves
    // per SIMD instead of three) measured 4% / 9% slower
    for (int s = 0; s < 6; s++) {
        const int nb = nbs[s];
        dv3 dX;
        double d[NF];
        if (nb >= 0) {
            dX = ld3(g.cx, g.cy, g.cz, nb) - cc;
#pragma unroll
            for (int k = 0; k < NF; k++) {
                const double v = a.phi[k][nb];
                d[k] = v - ph[k];
                if constexpr (LIMIT) { pmax[k] = dmax(pmax[k], v); pmin[k] = dmin(pmin[k], v); }
            }
        } else {
            const int f = g.cell_faces[6 * c + s];
            const int slot = g.fc1[f];
            dX = ld3(g.fcx, g.fcy, g.fcz, f) - cc;
#pragma unroll
            for (int k = 0; k < NF; k++) {
                if constexpr (LIMIT) {
                    const double v = a.phi[k][slot];
                    pmax[k] = dmax(pmax[k], v); pmin[k] = dmin(pmin[k], v);
                }
                d[k] = a.zero_boundary[k] ? 0.0 : a.phi[k][slot] - ph[k];
            }
        }
        a00 += (dX.x * dX.x); a10 += (dX.x * dX.y); a20 += (dX.x * dX.z);
        a11 += (dX.y * dX.y); a21 += (dX.y * dX.z); a22 += (dX.z * dX.z);
#pragma unroll
        for (int k = 0; k < NF; k++) {
            bx[k] += (dX.x * d[k]); by[k] += (dX.y * d[k]); bz[k] += (dX.z * d[k]);
        }
    }
    // A[i][j] == A[j][i] bit for bit (commutative products, same order)
    const double A00 = a00, A01 = a10, A02 = a20, A10 = a10, A11 = a11, A12 = a21, A20 = a20, A21 = a21, A22 = a22;
    const double det = (A00 * (A11 * A22 - A21 * A12)) - (A01 * (A10 * A22 - A20 * A12)) +
                       (A02 * (A10 * A21 - A20 * A11));
    double invdet = 0.0;
    if (det != 0.0) invdet = 1 / det;
    const double c00 = (A11 * A22 - A21 * A12), c01 = (A21 * A02 - A01 * A22), c02 = (A01 * A12 - A11 * A02);
    const double c10 = (A12 * A20 - A10 * A22), c11 = (A00 * A22 - A20 * A02), c12 = (A10 * A02 - A00 * A12);
    const double c20 = (A10 * A21 - A20 * A11), c21 = (A20 * A01 - A00 * A21), c22 = (A00 * A11 - A10 * A01);
    dv3 ds[6];
    if constexpr (LIMIT) {
        for (int s = 0; s < 6; s++) {
            const int f = g.cell_faces[6 * c + s];
            ds[s] = ld3(g.fcx, g.fcy, g.fcz, f) - cc;
        }
    }
#pragma unroll
    for (int k = 0; k < NF; k++) {
        double gx = (bx[k] * invdet * c00) + (by[k] * invdet * c01) + (bz[k] * invdet * c02);
        double gy = (bx[k] * invdet * c10) + (by[k] * invdet * c11) + (bz[k] * invdet * c12);
        double gz = (bx[k] * invdet * c20) + (by[k] * invdet * c21) + (bz[k] * invdet * c22);
        if constexpr (LIMIT) venkat(ph[k], pmax[k], pmin[k], ds, gx, gy, gz);
        st_out(&a.gx[k][c], gx); st_out(&a.gy[k][c], gy); st_out(&a.gz[k][c], gz);
    }
}

// limit_phi_gradients on its own (FlowSolver.inl:875-893): the stored
// gradients of the NF fields are limited in place.
template <int NF>
__global__ void __launch_bounds__(256) k_limit(Geo g, GradArgs a)
{
    const int c = walk_cell(g.w);
    if (c < 0) return;
    const dv3 cc = ld3(g.cx, g.cy, g.cz, c);
    dv3 ds[6];
    for (int s = 0; s < 6; s++) {
        const int f = g.cell_faces[6 * c + s];
        ds[s] = ld3(g.fcx, g.fcy, g.fcz, f) - cc;
    }
    for (int k = 0; k < NF; k++) {
        const double phic = a.phi[k][c];
        double pmax = phic, pmin = phic;
        for (int s = 0; s < 6; s++) {
            const int nb = g.nbr[6 * c + s];
            const double v = (nb >= 0) ? a.phi[k][nb] : a.phi[k][g.fc1[g.cell_faces[6 * c + s]]];
            pmax = dmax(pmax, v);
            pmin = dmin(pmin, v);
        }
        double gx = a.gx[k][c], gy = a.gy[k][c], gz = a.gz[k][c];
        venkat(phic, pmax, pmin, ds, gx, gy, gz);
        a.gx[k][c] = gx; a.gy[k][c] = gy; a.gz[k][c] = gz;
    }
}

// k_grad<NF>: per cell phi (8 NF) read, gradients (24 NF) written, centre
// (24), neighbour list (24); per boundary face cell_faces and fc1 entries (8)
// and the face centre (24).
static double grad_bytes(Ctx *c, int nf)
{
    const Sizes z = sizes_of(c);
    return (32.0 * nf + 48.0) * z.N + 32.0 * z.Fb;
}

static GradArgs grad11_args(Ctx *c)
{
    static const int fields[11] = {U, V, W, P, TE, ED, TEM, FUL, PRO, VARF, VARP};
    GradArgs a;
    for (int k = 0; k < 11; k++) {
        a.phi[k] = c->phi[fields[k]];
        a.gx[k] = c->gx[fields[k]]; a.gy[k] = c->gy[fields[k]]; a.gz[k] = c->gz[fields[k]];
        a.zero_boundary[k] = fields[k] == P;
    }
    return a;
}

int grad_multi(Ctx *c, bool limit)
{
    const GradArgs a = grad11_args(c);
    KTimer kt(c, MC_KC_GRAD, grad_bytes(c, 11));
    if (limit)
        hipLaunchKernelGGL((k_grad<11, true>), dim3(walk_blocks(cell_walk(c))), dim3(256), 0, c->stream, geo(c), a);
    else
        hipLaunchKernelGGL((k_grad<11, false>), dim3(walk_blocks(cell_walk(c))), dim3(256), 0, c->stream, geo(c), a);
    MC_LAUNCH_CHECK(c);
    return MC_OK;
}

int limit_gradients(Ctx *c)
{
    hipLaunchKernelGGL(k_limit<11>, dim3(walk_blocks(cell_walk(c))), dim3(256), 0, c->stream, geo(c), grad11_args(c));
    MC_LAUNCH_CHECK(c);
    return MC_OK;
}

int grad_single(Ctx *c, int field, bool pressure)
{
    GradArgs a;
    memset(&a, 0, sizeof(a));
    a.phi[0] = c->phi[field];
    a.gx[0] = c->gx[field]; a.gy[0] = c->gy[field]; a.gz[0] = c->gz[field];
    a.zero_boundary[0] = pressure ? 1 : 0;
    KTimer kt(c, MC_KC_GRAD1, grad_bytes(c, 1));
    hipLaunchKernelGGL(k_grad<1>, dim3(walk_blocks(cell_walk(c))), dim3(256), 0, c->stream, geo(c), a);
    MC_LAUNCH_CHECK(c);
    return MC_OK;
}

// ============================================================== A6 ==
// precomp_AU (FlowSolver.inl:1065-1104)
__global__ void __launch_bounds__(256) k_precomp_AU(Geo g, const double *mflux, double *AU, double ev, double rdelta)
{
    const int c = walk_cell(g.w);
    if (c < 0) return;
    double A = 0.0;
    for (int k = 0; k < 6; k++) {
        const int f = g.forder[6 * c + k];
        const int c1 = g.fc1[f];
        if (c1 < g.N) continue;
        const int bt = g.btype[c1 - g.N];
        if (bt == INLET || bt == OUTLET) {
            const double VisFace = ev * g.frlc[f];
            const double fb = -VisFace + dmin(mflux[f], 0.0);
            A = A - fb;
        }
    }
    const double f = g.rho[c] * g.vol[c] * rdelta;
    A += f;
    AU[c] = A;
}

int precomp_AU(Ctx *c)
{
    hipLaunchKernelGGL(k_precomp_AU, dim3(walk_blocks(cell_walk(c))), dim3(256), 0, c->stream, geo(c), c->mflux, c->AU, c->k.ev,
                       c->k.rdelta);
    MC_LAUNCH_CHECK(c);
    return MC_OK;
}

// ====================================================== A7 + A8 + A9 ==
struct MomArgs {
    double *U, *V, *W;
    const double *gUx, *gUy, *gUz, *gVx, *gVy, *gVz, *gWx, *gWy, *gWz, *gPx, *gPy, *gPz, *TEM;
    double *mflux, *ff0, *ff1, *AU, *SU, *SV, *SW, *vals, *dia;
    double *r0[3];   // non-null: also b - A x0 of the three systems, x0 = the current velocity
    double ev, inlet_ev, rdelta, urf, gas_tem, vel_x, vel_y, vel_z;
};

// calculate_flux_UVW (FlowSolver.inl:1377-1653), the cell sources of
// calculate_UVW (:1689-1721) and setup/update_sparse_matrix (:1186-1339).
// A_phi.U, .V and .W receive identical updates, so one diagonal serves all
// three momentum systems.
__global__ void __launch_bounds__(256) k_momentum(Geo g, MomArgs a)
{
    const int c = walk_cell(g.w);
    if (c < 0) return;
    const int N = g.N;
    double SU = 0.0, SV = 0.0, SW = 0.0, A = 0.0;
    double xU = 0.0, xV = 0.0, xW = 0.0;   // off-diagonal part of A x0 (warm-started solve)
    double own_ff[6];
    int own_nb[6];
    int nint = 0;
    // the cell's own values, read once (each face reads the other cell's)
    const dv3 gUc = ld3(a.gUx, a.gUy, a.gUz, c), gVc = ld3(a.gVx, a.gVy, a.gVz, c), gWc = ld3(a.gWx, a.gWy, a.gWz, c);
    const dv3 cc = ld3(g.cx, g.cy, g.cz, c);
    const double Uc = a.U[c], Vc = a.V[c], Wc = a.W[c];
    const CellFaces cf = cell_faces_of(g, c);
    // two faces per trip: their gathers overlap (2% faster at m = 128 than
    // one face per trip; unrolling all six is slower)
#pragma unroll 2
    for (int k = 0; k < 6; k++) {
        const int f = cf.f[k];
        const int o = cf.o[k];
        const bool own0 = o >= 0;          // c is the face's cell0
        const int nb = o & 0x7fffffff;     // cell across the face, or N + side
        const double mf = a.mflux[f];
        if (nb < N) {
            const double lambda0 = g.flam[f], lambda1 = 1.0 - lambda0;
            const dv3 gUn = ld3(a.gUx, a.gUy, a.gUz, nb), gVn = ld3(a.gVx, a.gVy, a.gVz, nb),
                      gWn = ld3(a.gWx, a.gWy, a.gWz, nb);
            const dv3 dUdXac = (own0 ? gUc : gUn) * lambda0 + (own0 ? gUn : gUc) * lambda1;
            const dv3 dVdXac = (own0 ? gVc : gVn) * lambda0 + (own0 ? gVn : gVc) * lambda1;
            const dv3 dWdXac = (own0 ? gWc : gWn) * lambda0 + (own0 ? gWn : gWc) * lambda1;
            const double Visac = a.ev * lambda0 + a.ev * lambda1;
            const double VisFace = Visac * g.frlc[f];
            const dv3 cn = ld3(g.cx, g.cy, g.cz, nb);
            const dv3 Xpn = own0 ? (cn - cc) : (cc - cn);
            const double Un = a.U[nb], Vn = a.V[nb], Wn = a.W[nb];
            const double U0 = own0 ? Uc : Un, U1 = own0 ? Un : Uc;
            const double V0 = own0 ? Vc : Vn, V1 = own0 ? Vn : Vc;
            const double W0 = own0 ? Wc : Wn, W1 = own0 ? Wn : Wc;
            const double sx = g.fnx[f], sy = g.fny[f], sz = g.fnz[f];
            const double fude = Visac * ((dUdXac.x + dUdXac.x) * sx + (dUdXac.y + dVdXac.x) * sy + (dUdXac.z + dWdXac.x) * sz);
            const double fvde = Visac * ((dUdXac.y + dVdXac.x) * sx + (dVdXac.y + dVdXac.y) * sy + (dVdXac.z + dWdXac.y) * sz);
            const double fwde = Visac * ((dUdXac.z + dWdXac.x) * sx + (dWdXac.y + dVdXac.z) * sy + (dWdXac.z + dWdXac.z) * sz);
            const double fudi = VisFace * dot(dUdXac, Xpn);
            const double fvdi = VisFace * dot(dVdXac, Xpn);
            const double fwdi = VisFace * dot(dWdXac, Xpn);
            const double ff0 = -VisFace - dmax(mf, 0.0);
            const double ff1 = -VisFace + dmin(mf, 0.0);
            // The convective deferred correction GammaBlend * (fce - fci)
            // (FlowSolver.inl:1471-1473) is a signed zero for finite operands
            // (GammaBlend = 0, :1387); adding or subtracting a signed zero
            // leaves a sum that starts at +0 unchanged (it never becomes -0),
            // so the term is left out.
            if (own0) {
                SU = SU + fude - fudi;
                SV = SV + fvde - fvdi;
                SW = SW + fwde - fwdi;
                own_ff[nint] = ff1;
                st_out(&a.ff0[f], ff0); st_out(&a.ff1[f], ff1);
            } else {
                SU = SU - fude + fudi;
                SV = SV - fvde + fvdi;
                SW = SW - fwde + fwdi;
                own_ff[nint] = ff0;
            }
            xU += own_ff[nint] * Un; xV += own_ff[nint] * Vn; xW += own_ff[nint] * Wn;
            own_nb[nint] = nb;
            nint++;
        } else {
            const int side = nb - N;
            const int bt = g.btype[side];
            const dv3 Xpn = ld3(g.fcx, g.fcy, g.fcz, f) - cc;
            const bool last = (f == g.last_face[side]);
            if (bt == INLET || bt == OUTLET) {
                const dv3 dUdXac = gUc, dVdXac = gVc, dWdXac = gWc;
                double UFace, VFace, WFace, Visac;
                if (bt == INLET) {
                    UFace = a.vel_x; VFace = a.vel_y; WFace = a.vel_z; Visac = a.inlet_ev;
                } else {
                    Visac = a.ev; UFace = Uc; VFace = Vc; WFace = Wc;
                }
                const double VisFace = Visac * g.frlc[f];
                const double sx = g.fnx[f], sy = g.fny[f], sz = g.fnz[f];
                const double fude = Visac * ((dUdXac.x + dUdXac.x) * sx + (dUdXac.y + dVdXac.x) * sy + (dUdXac.z + dWdXac.x) * sz);
                const double fvde = Visac * ((dUdXac.y + dVdXac.x) * sx + (dVdXac.y + dVdXac.y) * sy + (dVdXac.z + dWdXac.y) * sz);
                const double fwde = Visac * ((dUdXac.z + dWdXac.x) * sx + (dWdXac.y + dVdXac.z) * sy + (dWdXac.z + dWdXac.z) * sz);
                const double fudi = VisFace * dot(dUdXac, Xpn);
                const double fvdi = VisFace * dot(dVdXac, Xpn);
                const double fwdi = VisFace * dot(dWdXac, Xpn);
                double m = mf;
                if (bt == OUTLET && m < 0.0) { m = 1e-15; a.mflux[f] = m; }   // FlowSolver.inl:1576-1580
                const double fb = -VisFace + dmin(m, 0.0);
                A = A - fb;
                SU = SU - fb * UFace + fude - fudi;
                SV = SV - fb * VFace + fvde - fvdi;
                SW = SW - fb * WFace + fwde - fwdi;
                if (last) { a.U[N + side] = UFace; a.V[N + side] = VFace; a.W[N + side] = WFace; }
            } else if (bt == WALL) {
                const double Visac = a.ev;
                const double coef = Visac * g.frlc[f];
                const double U0 = Uc, V0 = Vc, W0 = Wc;
                const dv3 Up = v3(U0 - 0., V0 - 0., W0 - 0.);
                const dv3 n = ld3(g.fnx, g.fny, g.fnz, f);
                const dv3 nh = n / mag(n);
                const double dp = dot(Up, nh);
                const dv3 Ut = Up - nh * dp;
                const double Uvel = fabs(Ut.x) + fabs(Ut.y) + fabs(Ut.z);
                dv3 force = v3(0.0, 0.0, 0.0);
                if (Uvel > 0.0) {
                    const double distance_to_face = mag(Xpn);
                    force = (Ut * (g.farea[f] * Visac)) / distance_to_face;
                }
                A = A + coef;
                SU = SU + coef * U0 - force.x;
                SV = SV + coef * V0 - force.y;
                SW = SW + coef * W0 - force.z;
                if (last) { a.U[N + side] = 0.; a.V[N + side] = 0.; a.W[N + side] = 0.; }
            }
        }
    }
    // buoyancy, pressure gradient, transient (FlowSolver.inl:1689-1721)
    const double vol = g.vol[c], rho = g.rho[c];
    const double BodyForce = -0.001 * rho * vol * (a.TEM[c] - a.gas_tem);
    SU += 0.0 * BodyForce;
    SV += -9.81 * BodyForce;
    SW += 0.0 * BodyForce;
    SU -= a.gPx[c] * vol;
    SV -= a.gPy[c] * vol;
    SW -= a.gPz[c] * vol;
    const double fv = rho * vol * a.rdelta;
    SU += fv * Uc; SV += fv * Vc; SW += fv * Wc;
    A += fv;
    // setup_sparse_matrix / update_sparse_matrix
    for (int k = 0; k < nint; k++) A -= own_ff[k];
    A *= (1. / a.urf);
    SU = SU + (1.0 - a.urf) * A * Uc;
    SV = SV + (1.0 - a.urf) * A * Vc;
    SW = SW + (1.0 - a.urf) * A * Wc;
    st_out(&a.AU[c], A);
    st_out(&a.SU[c], SU); st_out(&a.SV[c], SV); st_out(&a.SW[c], SW);
    if (a.r0[0]) {
        st_out(&a.r0[0][c], SU - (A * Uc + xU));
        st_out(&a.r0[1][c], SV - (A * Vc + xV));
        st_out(&a.r0[2][c], SW - (A * Wc + xW));
    }
    store_row(g, c, nint, own_nb, own_ff, A, a.vals, a.dia);
}

static MomArgs mom_args(Ctx *c)
{
    MomArgs a;
    a.U = c->phi[U]; a.V = c->phi[V]; a.W = c->phi[W];
    a.gUx = c->gx[U]; a.gUy = c->gy[U]; a.gUz = c->gz[U];
    a.gVx = c->gx[V]; a.gVy = c->gy[V]; a.gVz = c->gz[V];
    a.gWx = c->gx[W]; a.gWy = c->gy[W]; a.gWz = c->gz[W];
    a.gPx = c->gx[P]; a.gPy = c->gy[P]; a.gPz = c->gz[P];
    a.TEM = c->phi[TEM];
    a.mflux = c->mflux; a.ff0 = c->ff0; a.ff1 = c->ff1;
    a.AU = c->AU; a.SU = c->SU; a.SV = c->SV; a.SW = c->SW; a.vals = c->A_vals;
    a.dia = (double *)fine_dia_of(c, c->A_vals);
    a.ev = c->k.ev; a.inlet_ev = c->k.inlet_ev; a.rdelta = c->k.rdelta; a.urf = 0.5; a.gas_tem = c->k.gas_tem;
    a.vel_x = c->k.gas_vel_x; a.vel_y = c->k.gas_vel_y; a.vel_z = c->k.gas_vel_z;
    a.r0[0] = a.r0[1] = a.r0[2] = nullptr;
    return a;
}

// Buffers for the starting residuals b - A x0 that the assembly kernels form
// for warm-started solves (allocated on first use).
static int r0_buffers(Ctx *c)
{
    if (c->r0buf[0]) return MC_OK;
    MemScope ms(c, MC_MEM_WORK);
    for (int k = 0; k < 3; k++) MC_TRY(dalloc(c, &c->r0buf[k], (size_t)c->N));
    return MC_OK;
}

// k_momentum: per cell U, V, W (24), their gradients (72), grad P (24), TEM
// (8), centre (24), volume and density (16), forder and fother (48) read;
// A_U, S_U, S_V, S_W (32) and, warm-started, b - A x0 of the three systems
// (24) written; per face mass flux, lambda, rlencos, normal (48) read; per
// interior face ff0, ff1 (16) written; per boundary face centre and area
// (32) read; the matrix entries (8 nnz) written.
static double momentum_bytes(Ctx *c, bool r0)
{
    const Sizes z = sizes_of(c);
    return (248.0 + (r0 ? 24.0 : 0.0)) * z.N + 48.0 * z.F + 16.0 * z.Fi + 32.0 * z.Fb + 8.0 * z.nnz;
}

int momentum_assemble(Ctx *c, bool r0)
{
    MomArgs a = mom_args(c);
    if (r0 && warm_start(c)) {
        MC_TRY(r0_buffers(c));
        for (int k = 0; k < 3; k++) a.r0[k] = c->r0buf[k];
    }
    KTimer kt(c, MC_KC_MOM_ASM, momentum_bytes(c, a.r0[0] != nullptr));
    hipLaunchKernelGGL(k_momentum, dim3(walk_blocks(cell_walk(c))), dim3(256), 0, c->stream, geo(c), a);
    MC_LAUNCH_CHECK(c);
    return MC_OK;
}

int momentum_solve(Ctx *c, bool r0)
{
    double *b[3] = {c->SU, c->SV, c->SW};
    double *x[3] = {c->phi[U], c->phi[V], c->phi[W]};
    mc_solve_status st{};
    const bool warm = warm_start(c);
    MC_TRY(bicgstab_solve(c, c->A_vals, 3, b, x, c->opt.mom_rtol, c->opt.mom_maxit, &st, warm,
                          (r0 && warm) ? c->r0buf : nullptr));
    c->st.mom_iters += st.iterations;
    if (!st.converged) { c->err = "momentum BiCGStab did not converge"; return MC_ERR_NOCONV; }
    return MC_OK;
}

// ============================================================= A11 ==
// Boundary mass fluxes (FlowSolver.inl:1957-2006): fixed inlet velocity at
// inlet and outlet, zero at walls.
__global__ void k_boundary_mflux(Geo g, const int *bfaces, int Fb, double *mflux, double *TE, double *ED, double *TEM,
                                 double vx, double vy, double vz)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Fb) return;
    const int f = bfaces[i];
    const int side = g.fc1[f] - g.N, c0 = g.fc0[f];
    const int bt = g.btype[side];
    const double Din = 1.2;
    if (bt == INLET) {
        mflux[f] = Din * (vx * g.fnx[f] + vy * g.fny[f] + vz * g.fnz[f]);
    } else if (bt == OUTLET) {
        double m = Din * (vx * g.fnx[f] + vy * g.fny[f] + vz * g.fnz[f]);
        if (m < 0.0) {
            m = 1e-15;
            if (f == g.last_face[side]) { TE[g.N + side] = TE[c0]; ED[g.N + side] = ED[c0]; TEM[g.N + side] = TEM[c0]; }
        }
        mflux[f] = m;
    } else if (bt == WALL) {
        mflux[f] = 0.0;
    }
}

// FlowIn / FlowOut / areaout and the FlowOut < 1e-10 rescale
// (FlowSolver.inl:2010-2079), one workgroup.
// The outlet balance (FlowSolver.inl:1957-2006) as a deterministic
// two-kernel reduction over the boundary faces (inflow, outflow, outlet
// area), then the rescaling pass when the outflow vanishes.
__global__ void __launch_bounds__(256) k_outlet_sums(Geo g, const int *bfaces, int Fb, const double *mflux,
                                                     double *partials, int chunk)
{
    const int b = blockIdx.x;
    const int i0 = b * chunk, i1 = min(Fb, i0 + chunk);
    double v[3] = {0.0, 0.0, 0.0};
    for (int i = i0 + (int)threadIdx.x; i < i1; i += 256) {
        const int f = bfaces[i];
        const int bt = g.btype[g.fc1[f] - g.N];
        if (bt == INLET) v[0] += mflux[f];
        else if (bt == OUTLET) { v[1] += mflux[f]; v[2] += g.farea[f]; }
    }
    block_partials<3>(v, partials, b, gridDim.x);
}

struct EpiOutlet {
    double *out;   // FlowIn, FlowOut, areaout
    __device__ void operator()(const double (&v)[3]) const { out[0] = v[0]; out[1] = v[1]; out[2] = v[2]; }
};

__global__ void k_outlet_apply(Geo g, const int *bfaces, int Fb, const double *sums, double *mflux, double *U,
                               double *V, double *W, double vx, double vy, double vz)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Fb) return;
    const double FlowIn = sums[0], FlowOut = sums[1], areaout = sums[2];
    if (!(FlowOut < 0.0000000001)) return;
    const double ratearea = -FlowIn / areaout;
    const int f = bfaces[i];
    const int side = g.fc1[f] - g.N;
    if (g.btype[side] == OUTLET) {
        mflux[f] = ratearea * g.farea[f];
        if (f == g.last_face[side]) { U[g.N + side] = vx; V[g.N + side] = vy; W[g.N + side] = vz; }
    }
}

int boundary_mass_flux(Ctx *c)
{
    if (c->Fb == 0) return MC_OK;
    hipLaunchKernelGGL(k_boundary_mflux, dim3(div_up(c->Fb, 256)), dim3(256), 0, c->stream, geo(c), c->bfaces, c->Fb,
                       c->mflux, c->phi[TE], c->phi[ED], c->phi[TEM], c->k.gas_vel_x, c->k.gas_vel_y, c->k.gas_vel_z);
    if (c->Fb > 0) {
        const int Fb = c->Fb;
        const int blocks = std::max(1, std::min(256, div_up(Fb, 2048)));
        const int chunk = div_up(Fb, blocks);
        double *sums = c->dscal + 760;
        hipLaunchKernelGGL(k_outlet_sums, dim3(blocks), dim3(256), 0, c->stream, geo(c), c->bfaces, Fb, c->mflux, c->red,
                           chunk);
        launch_finish<3>(c->stream, c->red, blocks, EpiOutlet{sums});
        hipLaunchKernelGGL(k_outlet_apply, dim3(div_up(Fb, 256)), dim3(256), 0, c->stream, geo(c), c->bfaces, Fb, sums,
                           c->mflux, c->phi[U], c->phi[V], c->phi[W], c->k.gas_vel_x, c->k.gas_vel_y, c->k.gas_vel_z);
    }
    MC_LAUNCH_CHECK(c);
    return MC_OK;
}

struct MassArgs {
    const double *U, *V, *W, *P, *AU;
    const double *gUx, *gUy, *gUz, *gVx, *gVy, *gVz, *gWx, *gWy, *gWz, *gPx, *gPy, *gPz;
    double *mflux, *ff0, *ff1, *SU, *AV, *Pb, *PAvals;
    int assemble;      // 1: setup_pressure_matrix too
    int into_pressure; // write the pressure matrix and RHS
    // with into_pressure: the AMG fine level's diagonal copies of the matrix
    // (FP64 without the pskip lower diagonals, FP32 in full) and inverse
    // diagonal, written here instead of by a pass over the CSR values
    double *pdia, *pdinv;
    float *pdia32, *pdinv32;
    int pskip;
    _Float16 *phoff;   // with pdia: the FP16 copy of D^-1 A's off-diagonals
    float *pdiag32;
};

// The per-cell values the Rhie-Chow flux reads on either side of a face.
struct RCCell {
    dv3 gU, gV, gW, gP, cc;
    double U, V, W, P, rho, AU, vol;
};

__device__ __forceinline__ RCCell rc_cell(const Geo &g, const MassArgs &a, int i)
{
    RCCell q;
    q.gU = ld3(a.gUx, a.gUy, a.gUz, i); q.gV = ld3(a.gVx, a.gVy, a.gVz, i); q.gW = ld3(a.gWx, a.gWy, a.gWz, i);
    q.gP = ld3(a.gPx, a.gPy, a.gPz, i); q.cc = ld3(g.cx, g.cy, g.cz, i);
    q.U = a.U[i]; q.V = a.V[i]; q.W = a.W[i]; q.P = a.P[i];
    q.rho = g.rho[i]; q.AU = a.AU[i]; q.vol = g.vol[i];
    return q;
}

// Rhie-Chow interior mass flux (FlowSolver.inl:1893-1950); q0, q1: the
// face's cell0 and cell1.
__device__ __forceinline__ double rhie_chow(const Geo &g, int f, const RCCell &q0, const RCCell &q1, double &ApV_out)
{
    const double lambda0 = g.flam[f], lambda1 = 1.0 - lambda0;
    const dv3 dUdXac = q0.gU * lambda0 + q1.gU * lambda1;
    const dv3 dVdXac = q0.gV * lambda0 + q1.gV * lambda1;
    const dv3 dWdXac = q0.gW * lambda0 + q1.gW * lambda1;
    const dv3 cc0 = q0.cc, cc1 = q1.cc;
    const dv3 Xac = cc1 * lambda1 + cc0 * lambda0;
    const dv3 fc = ld3(g.fcx, g.fcy, g.fcz, f);
    const dv3 delta = fc - Xac;
    const double UFace = q1.U * lambda1 + q0.U * lambda0 + dot(dUdXac, delta);
    const double VFace = q1.V * lambda1 + q0.V * lambda0 + dot(dVdXac, delta);
    const double WFace = q1.W * lambda1 + q0.W * lambda0 + dot(dWdXac, delta);
    const double rho0 = q0.rho, rho1 = q1.rho;
    const double densityf = rho0 * lambda0 + rho1 * lambda1;
    const dv3 n = ld3(g.fnx, g.fny, g.fnz, f);
    double m = densityf * (UFace * n.x + VFace * n.y + WFace * n.z);
    const dv3 nh = n / mag(n);
    const dv3 Xpac = fc - nh * dot(fc - cc0, nh);
    const dv3 Xnac = fc - nh * dot(fc - cc1, nh);
    const dv3 delp = Xpac - cc0, deln = Xnac - cc1;
    const dv3 gP0 = q0.gP, gP1 = q1.gP;
    const double cell0_P = q0.P + dot(gP0, delp);
    const double cell1_P = q1.P + dot(gP1, deln);
    const dv3 Xpn = Xnac - Xpac;
    const dv3 Xpn2 = cc1 - cc0;
    const double A0 = q0.AU, A1 = q1.AU;
    const double ApV0 = (A0 != 0.0) ? 1.0 / A0 : 0.0;
    const double ApV1 = (A1 != 0.0) ? 1.0 / A1 : 0.0;
    double ApV = rho0 * ApV0 * lambda0 + rho1 * ApV1 * lambda1;
    const double volume_avg = q0.vol * lambda0 + q1.vol * lambda1;
    ApV = ApV * g.farea[f] * volume_avg / dot(Xpn2, nh);
    const double dpx = (gP1.x * lambda1 + gP0.x * lambda0) * Xpn.x;
    const double dpy = (gP1.y * lambda1 + gP0.y * lambda0) * Xpn.y;
    const double dpz = (gP1.z * lambda1 + gP0.z * lambda0) * Xpn.z;
    m -= ApV * ((cell1_P - cell0_P) - dpx - dpy - dpz);
    ApV_out = ApV;
    return m;
}

// calculate_mass_flux (FlowSolver.inl:1880-2104), interior faces: one
// thread per cell over the faces it owns (it is their cell0), so each face's
// Rhie-Chow flux is evaluated once; the cell's own values are read once.
__global__ void __launch_bounds__(256) k_mass_faces(Geo g, MassArgs a)
{
    const int c = walk_cell(g.w);
    if (c < 0) return;
    const CellFaces cf = cell_faces_of(g, c);
    const RCCell own = rc_cell(g, a, c);
    for (int k = 0; k < 6; k++) {
        const int c1 = cf.o[k];
        if (c1 < 0 || c1 >= g.N) continue;   // not this cell's face, or a boundary face
        const int f = cf.f[k];
        const RCCell other = rc_cell(g, a, c1);
        double ApV;
        const double m = rhie_chow(g, f, own, other, ApV);
        st_out(&a.mflux[f], m);
        st_out(&a.ff0[f], -ApV); st_out(&a.ff1[f], -ApV);
    }
}

// The cell rows of calculate_mass_flux + setup_pressure_matrix
// (:2106-2201), from the face fluxes and coefficients k_mass_faces stored.
// S_phi.U receives, in the reference's order, the inlet fluxes of the main
// face loop, the outlet fluxes of the balance loop, then the interior
// divergence of setup_pressure_matrix.
__global__ void __launch_bounds__(256) k_mass_rows(Geo g, MassArgs a)
{
    const int c = walk_cell(g.w);
    if (c < 0) return;
    const int N = g.N;
    double S = 0.0, AV = 0.0;
    const CellFaces cf = cell_faces_of(g, c);
    for (int k = 0; k < 6; k++) {
        const int nb = cf.o[k] & 0x7fffffff;
        if (nb >= N && g.btype[nb - N] == INLET) S = S - a.mflux[cf.f[k]];
    }
    for (int k = 0; k < 6; k++) {
        const int nb = cf.o[k] & 0x7fffffff;
        if (nb >= N && g.btype[nb - N] == OUTLET) S -= a.mflux[cf.f[k]];
    }
    if (!a.assemble) { a.SU[c] = S; return; }
    double own_v[6];
    int own_nb[6];
    int nint = 0;
    for (int k = 0; k < 6; k++) {
        const int f = cf.f[k];
        const bool own0 = cf.o[k] >= 0;
        const int nb = cf.o[k] & 0x7fffffff;
        if (nb >= N) continue;
        const double m = a.mflux[f], v = a.ff0[f];   // v = -ApV
        if (own0) S -= m;
        else S += m;
        AV -= v;
        own_nb[nint] = nb;
        own_v[nint] = v;
        nint++;
    }
    a.SU[c] = S;
    a.AV[c] = AV;
    if (a.into_pressure) {
        a.Pb[c] = S;
        int pos[6], dpos;
        row_positions(c, nint, own_nb, pos, dpos);
        const int rp = g.row_ptr[c];
        a.PAvals[rp + dpos] = AV;
        for (int k = 0; k < nint; k++) a.PAvals[rp + pos[k]] = own_v[k];
        if (a.pdia) {   // the values and order of k_level_dia's pass
            const size_t n = (size_t)N;
            const int d0 = dslot(g, 0);
            a.pdia[(size_t)(d0 - a.pskip) * n + c] = AV;
            a.pdia32[(size_t)d0 * n + c] = (float)AV;
            for (int k = 0; k < nint; k++) {
                const int d = dslot(g, own_nb[k] - c);
                a.pdia32[(size_t)d * n + c] = (float)own_v[k];
                if (d >= a.pskip) a.pdia[(size_t)(d - a.pskip) * n + c] = own_v[k];
            }
            const double di = (AV != 0.0) ? 1.0 / AV : 0.0;
            a.pdinv[c] = di;
            a.pdinv32[c] = (float)di;
            if (a.phoff) {   // as k_level_dia writes them
                a.pdiag32[c] = (float)AV;
                for (int k = 0; k < nint; k++) {
                    const int d = dslot(g, own_nb[k] - c);
                    a.phoff[(size_t)(d < d0 ? d : d - 1) * n + c] = (_Float16)(own_v[k] / AV);
                }
            }
        }
    }
}

static void launch_mass(Ctx *c, const MassArgs &a)
{
    const dim3 grid(walk_blocks(cell_walk(c)));
    hipLaunchKernelGGL(k_mass_faces, grid, dim3(256), 0, c->stream, geo(c), a);
    hipLaunchKernelGGL(k_mass_rows, grid, dim3(256), 0, c->stream, geo(c), a);
}

static MassArgs mass_args(Ctx *c, int assemble, int into_pressure)
{
    MassArgs a;
    a.U = c->phi[U]; a.V = c->phi[V]; a.W = c->phi[W]; a.P = c->phi[P]; a.AU = c->AU;
    a.gUx = c->gx[U]; a.gUy = c->gy[U]; a.gUz = c->gz[U];
    a.gVx = c->gx[V]; a.gVy = c->gy[V]; a.gVz = c->gz[V];
    a.gWx = c->gx[W]; a.gWy = c->gy[W]; a.gWz = c->gz[W];
    a.gPx = c->gx[P]; a.gPy = c->gy[P]; a.gPz = c->gz[P];
    a.mflux = c->mflux; a.ff0 = c->ff0; a.ff1 = c->ff1; a.SU = c->SU; a.AV = c->AV; a.Pb = c->Pb;
    a.PAvals = c->PA_vals;
    a.assemble = assemble; a.into_pressure = into_pressure;
    a.pdia = nullptr; a.pdinv = nullptr; a.pdia32 = nullptr; a.pdinv32 = nullptr; a.pskip = 0;
    a.phoff = nullptr; a.pdiag32 = nullptr;
    return a;
}

int set_up_field(Ctx *c)
{
    // precomp_AU reads the mass fluxes left by the constructor (all zero)
    MC_TRY(precomp_AU(c));
    MC_TRY(boundary_mass_flux(c));
    launch_mass(c, mass_args(c, 0, 0));
    MC_LAUNCH_CHECK(c);
    return MC_OK;
}

// The two launches of launch_mass, each array counted once per launch.
// k_mass_faces: per cell gradients of U, V, W, P (96), centre (24), U, V, W,
// P (32), density, A_U, volume (24), forder and fother (48) read; per
// interior face lambda, centre, normal, area (64) read and mass flux, ff0,
// ff1 (24) written.  k_mass_rows: per cell forder and fother (48) read, S_U,
// A_V, Pressure_b (24) written; per face mass flux (8) and per interior face
// ff0 (8) read; the CSR values (8 nnz) written; with the AMG fine level's
// copies also the symmetric FP64 diagonals (8 (N + F_i)), the FP32 ones
// (4 nnz), both inverse diagonals (12 N), and with the FP16 smoothing copy
// the FP32 diagonal (4 N) and the FP16 off-diagonals (2 * 2 F_i).
static double mass_bytes(Ctx *c, const MassArgs &a)
{
    const Sizes z = sizes_of(c);
    double b = 224.0 * z.N + 88.0 * z.Fi;                  // k_mass_faces
    b += 72.0 * z.N + 8.0 * z.F + 8.0 * z.Fi + 8.0 * z.nnz;  // k_mass_rows
    if (a.pdia) b += 8.0 * (z.N + z.Fi) + 4.0 * z.nnz + 12.0 * z.N;
    if (a.phoff) b += 4.0 * z.N + 4.0 * z.Fi;
    return b;
}

int calculate_mass_flux_assemble(Ctx *c)
{
    const bool into = c->first_pressure || c->opt.mode == MC_MODE_SIMPLE;
    MC_TRY(boundary_mass_flux(c));
    MassArgs a = mass_args(c, 1, into ? 1 : 0);
    // an AMG hierarchy built on this pattern takes its fine level's
    // diagonal copies from the assembly
    bool l0 = into && c->amg_ready && !c->amg.empty();
    if (l0) {
        const AmgLevel &L = c->amg[0];
        l0 = L.vals == c->PA_vals && L.nd > 0 && L.nd == c->fine_nd && L.dia && L.dia32;
        for (int d = 0; l0 && d < L.nd; d++) l0 = L.off[d] == c->fine_off[d];
        if (l0) {
            a.pdia = L.dia; a.pdia32 = L.dia32; a.pdinv = L.dinv; a.pdinv32 = L.dinv32;
            a.pskip = L.sym ? L.nd / 2 : 0;
            a.phoff = L.hoff; a.pdiag32 = L.diag32;
        }
    }
    {
        KTimer kt(c, MC_KC_P_ASM, mass_bytes(c, a));
        launch_mass(c, a);
    }
    MC_LAUNCH_CHECK(c);
    c->pa_dia_fresh = l0;
    return MC_OK;
}

// ============================================================= A14 ==
// mass-flux correction (FlowSolver.inl:2402-2412) fused with the
// non-orthogonal update_mass_flux (:1812-1878): both add to the flux in that
// order; the latter's face correction is kept for the RHS gather.
__global__ void __launch_bounds__(256) k_flux_correct(Geo g, const double *PP, const double *gPPx, const double *gPPy,
                                                      const double *gPPz, const double *ff0, double *mflux, double *fcorr)
{
    // one thread per cell, over the interior faces it owns (it is their
    // cell0): the cell's values are read once for its faces
    const int c = walk_cell(g.w);
    if (c < 0) return;
    const dv3 cc0 = ld3(g.cx, g.cy, g.cz, c), gP0 = ld3(gPPx, gPPy, gPPz, c);
    const double P0 = PP[c];
    const CellFaces cf = cell_faces_of(g, c);
    for (int k = 0; k < 6; k++) {
        const int c1 = cf.o[k];
        if (c1 < 0 || c1 >= g.N) continue;   // not this cell's face, or a boundary face
        const int f = cf.f[k];
        const double fact = ff0[f];
        double m = mflux[f];
        m += (fact * (PP[c1] - P0));
        const dv3 fc = ld3(g.fcx, g.fcy, g.fcz, f), cc1 = ld3(g.cx, g.cy, g.cz, c1);
        const dv3 n = ld3(g.fnx, g.fny, g.fnz, f);
        const dv3 nh = n / mag(n);
        const dv3 Xpac = fc - nh * dot(fc - cc0, nh);
        const dv3 Xnac = fc - nh * dot(fc - cc1, nh);
        dv3 Xn = Xnac - cc1, Xp = Xpac - cc0;
        if (fabs(Xn.x) < 0.0000000000000003) Xn.x = 0.0;
        if (fabs(Xn.y) < 0.0000000000000003) Xn.y = 0.0;
        if (fabs(Xn.z) < 0.0000000000000003) Xn.z = 0.0;
        if (fabs(Xp.x) < 0.0000000000000003) Xp.x = 0.0;
        if (fabs(Xp.y) < 0.0000000000000003) Xp.y = 0.0;
        if (fabs(Xp.z) < 0.0000000000000003) Xp.z = 0.0;
        const dv3 gP1 = ld3(gPPx, gPPy, gPPz, c1);
        const double dpx = gP1.x * Xn.x - gP0.x * Xp.x;
        const double dpy = gP1.y * Xn.y - gP0.y * Xp.y;
        const double dpz = gP1.z * Xn.z - gP0.z * Xp.z;
        const double fcv = fact * (dpx + dpy + dpz) * 0.8;
        m += fcv;
        st_out(&mflux[f], m);
        st_out(&fcorr[f], fcv);
    }
}

// cell correction (FlowSolver.inl:2416-2433), S_phi.U reset + gather of the
// non-orthogonal correction (:2437-2444, 1875-1876), PP reset (:2447-2450).
// which: 0 none, 1 cell 0 only (compat first loop), 2 every cell.
__global__ void __launch_bounds__(256) k_cell_correct(Geo g, const double *fcorr, double *SU, double *P, double *PP,
                                                      double *Uf, double *Vf, double *Wf, const double *gPPx,
                                                      const double *gPPy, const double *gPPz, const double *AU,
                                                      int which)
{
    const int c = walk_cell(g.w);
    if (c < 0) return;
    if (c < 6) PP[g.N + c] = 0.0;
    double S = 0.0;
    const CellFaces cf = cell_faces_of(g, c);
    for (int k = 0; k < 6; k++) {
        const int f = cf.f[k];
        const int o = cf.o[k];
        if ((o & 0x7fffffff) >= g.N) continue;
        if (o >= 0) S -= fcorr[f];   // c is the face's cell0
        else S += fcorr[f];
    }
    SU[c] = S;
    if (which == 2 || (which == 1 && c == 0)) {
        const double Ar = (AU[c] != 0.0) ? 1.0 / AU[c] : 0.0;
        const double fact = g.vol[c] * Ar;
        P[c] += 0.2 * PP[c];
        Uf[c] -= gPPx[c] * fact;
        Vf[c] -= gPPy[c] * fact;
        Wf[c] -= gPPz[c] * fact;
    }
    PP[c] = 0.0;
}

// Update_P_at_boundaries (FlowSolver.inl:2275-2290): slot <- last writer.
__global__ void k_slots_copy(int N, int c0, int c1, int c2, int c3, int c4, int c5, double *phi)
{
    const int s = threadIdx.x;
    if (s >= 6) return;
    const int cells[6] = {c0, c1, c2, c3, c4, c5};
    if (cells[s] >= 0) phi[N + s] = phi[cells[s]];
}

int pressure_boundary_slots(Ctx *c, int field)
{
    const int *lc = c->last_cell;
    hipLaunchKernelGGL(k_slots_copy, dim3(1), dim3(64), 0, c->stream, c->N, lc[0], lc[1], lc[2], lc[3], lc[4], lc[5],
                       c->phi[field]);
    MC_LAUNCH_CHECK(c);
    return MC_OK;
}

// update_P (FlowSolver.inl:2292-2315): outlet copy, else extrapolate.
__global__ void k_update_P(Geo g, double *P, const double *gx, const double *gy, const double *gz)
{
    const int s = threadIdx.x;
    if (s >= 6) return;
    const int f = g.last_face[s];
    if (f < 0) return;
    const int c0 = g.fc0[f];
    if (g.btype[s] == OUTLET) {
        P[g.N + s] = P[c0];
    } else {
        const dv3 ds = ld3(g.fcx, g.fcy, g.fcz, f) - ld3(g.cx, g.cy, g.cz, c0);
        P[g.N + s] = P[c0] + dot(ld3(gx, gy, gz, c0), ds);
    }
}

int update_P(Ctx *c)
{
    hipLaunchKernelGGL(k_update_P, dim3(1), dim3(64), 0, c->stream, geo(c), c->phi[P], c->gx[P], c->gy[P], c->gz[P]);
    MC_LAUNCH_CHECK(c);
    return MC_OK;
}

__global__ void k_copy(int n, const double *a, double *b)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// calculate_pressure (FlowSolver.inl:2317-2463)
__global__ void k_zero4(int n, double *a, double *b, double *c, double *d)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) { a[i] = 0.0; b[i] = 0.0; c[i] = 0.0; d[i] = 0.0; }
}

int calculate_pressure(Ctx *c)
{
    const int N = c->N;
    {
        StageTimer t(c, MC_ST_PRES_FLUX);
        MC_TRY(calculate_mass_flux_assemble(c));
    }
    const bool simple = c->opt.mode == MC_MODE_SIMPLE;
    // AMG values follow the matrix that will be solved
    if (c->first_pressure || simple) {
        StageTimer t(c, MC_ST_PRES_SOLVE);
        if (!c->amg_ready) MC_TRY(amg_setup(c, c->PA_vals));
        else MC_TRY(amg_update_values(c, c->PA_vals));
    }
    int loop = 0;
    double pmax = 0.0, pref = 0.0;
    double first_bnorm = 0.0;
    for (;;) {
        loop++;
        // solve (compat: Pressure_b / Pressure_A are frozen after the first
        // assembly, so every later solve reproduces the stored solution)
        const bool fresh = simple || c->first_pressure || c->u == nullptr;
        StageTimer ts(c, MC_ST_PRES_SOLVE);
        c->p_zero = false;
        if (simple) {
            MC_HIP(c, hipMemcpyAsync(c->Pb, c->SU, (size_t)N * 8, hipMemcpyDeviceToDevice, c->stream));
        }
        if (fresh || !c->st.p_solves) {
            mc_solve_status st{};
            // loops after the first solve only the O(1e-16) non-orthogonal
            // remainder of update_mass_flux: their tolerance is an absolute
            // floor relative to the first loop's RHS
            const double atol = (loop == 1) ? 0.0 : c->opt.p_rtol * first_bnorm;
            MC_TRY(pressure_solve(c, c->Pb, c->phi[PP], &st, atol));
            if (loop == 1) first_bnorm = st.bnorm;
            if (!simple) MC_HIP(c, hipMemcpyAsync(c->u, c->phi[PP], (size_t)N * 8, hipMemcpyDeviceToDevice, c->stream));
        } else {
            MC_HIP(c, hipMemcpyAsync(c->phi[PP], c->u, (size_t)N * 8, hipMemcpyDeviceToDevice, c->stream));
        }
        ts.end();
        StageTimer tc(c, MC_ST_PRES_CORRECT);
        if (c->p_zero) pmax = 0.0;   // the solve left PP = +0 everywhere
        else MC_TRY(reduce_maxabs_pp(c, c->phi[PP], &pmax));
        if (loop == 1) pref = pmax;
        if (loop > 1 && pmax == 0.0) {
            // PP is zero everywhere: this loop's right-hand side (the previous
            // loop's non-orthogonal remainder, zero where the face normals
            // pass through both cell centres, as on the box) was zero and its
            // solve returned x = 0.  The passes below would then add only
            // signed zeros to the fluxes, P and the velocities and leave PP
            // zero; their other outputs are written directly: grad PP and the
            // gathered correction S_U are zero
            hipLaunchKernelGGL(k_zero4, dim3(div_up(N, 256)), dim3(256), 0, c->stream, N, c->gx[PP], c->gy[PP],
                               c->gz[PP], c->SU);
            MC_LAUNCH_CHECK(c);
            c->st.p_loops_skipped += 1;
            c->first_pressure = false;
            break;
        }
        MC_TRY(pressure_boundary_slots(c, PP));
        MC_TRY(grad_single(c, PP, true));
        {
            {
                // per cell centre, grad PP (48), PP (8), forder and fother (48)
                // read; per interior face ff0, mass flux, centre, normal (64)
                // read and mass flux, correction (16) written
                KTimer kt(c, MC_KC_CORRECT, 104.0 * c->N + 80.0 * (double)(c->F - c->Fb));
                hipLaunchKernelGGL(k_flux_correct, dim3(walk_blocks(cell_walk(c))), dim3(256), 0, c->stream, geo(c),
                                   c->phi[PP], c->gx[PP], c->gy[PP], c->gz[PP], c->ff0, c->mflux, c->fcorr);
            }
            const int which = simple ? 2 : (c->first_pressure ? 1 : 0);
            hipLaunchKernelGGL(k_cell_correct, dim3(walk_blocks(cell_walk(c))), dim3(256), 0, c->stream, geo(c), c->fcorr, c->SU,
                               c->phi[P], c->phi[PP], c->phi[U], c->phi[V], c->phi[W], c->gx[PP], c->gy[PP],
                               c->gz[PP], c->AU, which);
            MC_LAUNCH_CHECK(c);
        }
        c->first_pressure = false;
        if (loop >= 4 || pmax <= 0.25 * pref) break;
    }
    c->st.p_loops += loop;
    c->st.last_loops = loop;
    c->st.last_p_max = pref;
    StageTimer tc(c, MC_ST_PRES_CORRECT);
    MC_TRY(pressure_boundary_slots(c, P));
    MC_TRY(grad_single(c, P, true));
    MC_TRY(update_P(c));
    return MC_OK;
}

// ============================================================= A16 ==
struct ScalArgs {
    double *phi;
    const double *gx, *gy, *gz;
    const double *U, *V, *W, *gUx, *gUy, *gUz, *gVx, *gVy, *gVz, *gWx, *gWy, *gWz;
    double *TE, *ED, *TP;
    const double *mflux;
    double *ff0, *ff1, *AV, *SU, *vals, *dia;
    double *r0;      // non-null: also b - A x0, x0 = the current field
    int type;
    double ev, inlet_ev, rdelta, urf, inlet_phi, cmu75;
};

__device__ __forceinline__ double scalar_visac(int type, double ev, double Visac)
{
    if (type == TEMP) return (ev + Visac / 0.9) / 0.6905;
    if (type == TERBTE) return ev + Visac;
    if (type == TERBED) return ev + Visac / 1.219;
    return (ev + Visac / 0.9) / 0.9;
}

// FluxScalar (FlowSolver.inl:2516-2744), the unsteady term (:2908-2916),
// solveTurbulenceModels (:2746-2883) and setup_sparse_matrix (:1186-1297).
__global__ void __launch_bounds__(256) k_scalar(Geo g, ScalArgs a)
{
    const int c = walk_cell(g.w);
    if (c < 0) return;
    const int N = g.N;
    const double ev = a.ev;
    const bool wallcell = g.is_wall[c] != 0;
    const bool ed = a.type == TERBED;
    double S = 0.0, A = 0.0;
    double own_ff[6];
    int own_nb[6];
    int nint = 0;
    // face ids and fother codes loaded ahead of the face loop (a plain loop
    // here: the unrolled load of cell_faces_of makes this kernel 2-9% slower)
    int fo[6], ot[6];
    for (int k = 0; k < 6; k++) { fo[k] = g.forder[6 * c + k]; ot[k] = g.fother[6 * c + k]; }
    // the cell's own values, read once: each face takes the other cell's
    // from memory and these for its own side
    const dv3 gc = ld3(a.gx, a.gy, a.gz, c), cc = ld3(g.cx, g.cy, g.cz, c);
    const double pc = a.phi[c];
    double xo = 0.0;   // off-diagonal part of A x0 (warm-started solve)
    for (int k = 0; k < 6; k++) {
        const int f = fo[k];
        const bool own0 = ot[k] >= 0;            // c is the face's cell0
        const int nb = ot[k] & 0x7fffffff;       // cell across the face, or N + side
        const double mf = a.mflux[f];
        if (nb < N) {
            const double lambda0 = g.flam[f], lambda1 = 1.0 - lambda0;
            double Visac = ev * lambda0 + ev * lambda1;
            Visac -= ev;
            Visac = scalar_visac(a.type, ev, Visac);
            const dv3 gn = ld3(a.gx, a.gy, a.gz, nb), cn = ld3(g.cx, g.cy, g.cz, nb);
            const dv3 dPhiXac = (own0 ? gc : gn) * lambda0 + (own0 ? gn : gc) * lambda1;
            const dv3 Xpn = own0 ? (cn - cc) : (cc - cn);
            const double VisFace = Visac * g.frlc[f];
            const double pn = a.phi[nb];
            const double fde1 = Visac * dot(dPhiXac, ld3(g.fnx, g.fny, g.fnz, f));
            const double fdi = VisFace * dot(dPhiXac, Xpn);
            double ff0 = -VisFace - dmax(mf, 0.0);
            double ff1 = -VisFace + dmin(mf, 0.0);
            // the convective deferred correction GammaBlend * (fce - fci)
            // (FlowSolver.inl:2600-2604) is a signed zero for finite operands
            // (GammaBlend = 0, :2523) and leaves S, which starts at +0 and so
            // never becomes -0, unchanged: it is left out
            if (ed) {  // FlowSolver.inl:2855-2875: rows of wall cells lose their off-diagonals
                const bool wn = g.is_wall[nb] != 0;
                if (own0 ? wallcell : wn) ff1 = 0.0;
                if (own0 ? wn : wallcell) ff0 = 0.0;
            }
            if (own0) {
                S = S + fde1 - fdi;
                own_ff[nint] = ff1;
                st_out(&a.ff0[f], ff0); st_out(&a.ff1[f], ff1);
            } else {
                S = S - fde1 + fdi;
                own_ff[nint] = ff0;
            }
            xo += own_ff[nint] * pn;
            own_nb[nint] = nb;
            nint++;
        } else {
            const int side = nb - N;
            const int bt = g.btype[side];
            const bool last = (f == g.last_face[side]);
            const dv3 Xpn = ld3(g.fcx, g.fcy, g.fcz, f) - cc;
            const dv3 gP = gc;
            if (bt == INLET) {
                const double PhiFace = a.inlet_phi;
                double Visac = a.inlet_ev;
                Visac -= ev;
                Visac = scalar_visac(a.type, ev, Visac);
                const double VisFace = Visac * g.frlc[f];
                const double fde = Visac * dot(gP, ld3(g.fnx, g.fny, g.fnz, f));
                const double fdi = VisFace * dot(gP, Xpn);
                const double fb = -VisFace + dmin(mf, 0.0);
                A = A - fb;
                S = S - fb * PhiFace + fde - fdi;
                if (last) a.phi[N + side] = PhiFace;
            } else if (bt == OUTLET) {
                double Visac = ev;
                Visac -= ev;
                Visac = scalar_visac(a.type, ev, Visac);
                const double PhiFace = pc + dot(gP, Xpn);
                const double VisFace = Visac * g.frlc[f];
                const double fde = Visac * dot(gP, ld3(g.fnx, g.fny, g.fnz, f));
                const double fdi = VisFace * dot(gP, Xpn);
                S = S + fde - fdi;
                if (last) a.phi[N + side] = PhiFace;
            } else if (bt == WALL) {
                if (last && !ed) a.phi[N + side] = pc;
            }
        }
    }
    const double vol = g.vol[c], rho = g.rho[c];
    double phic = pc;
    const double fv = rho * vol * a.rdelta;
    S += fv * phic;
    A += fv;
    if (a.type == TERBTE) {
        const dv3 dUdXp = ld3(a.gUx, a.gUy, a.gUz, c), dVdXp = ld3(a.gVx, a.gVy, a.gVz, c), dWdXp = ld3(a.gWx, a.gWy, a.gWz, c);
        const double s1 = (dUdXp.x + dUdXp.x) * dUdXp.x + (dUdXp.y + dVdXp.x) * dUdXp.y + (dUdXp.z + dWdXp.x) * dUdXp.z;
        const double s2 = (dVdXp.x + dUdXp.y) * dVdXp.x + (dVdXp.y + dVdXp.y) * dVdXp.y + (dVdXp.z + dWdXp.y) * dVdXp.z;
        const double s3 = (dWdXp.x + dUdXp.z) * dWdXp.x + (dWdXp.y + dVdXp.z) * dWdXp.y + (dWdXp.z + dWdXp.z) * dWdXp.z;
        const double VisT = ev - ev;
        double TP = VisT * (s1 + s2 + s3);
        const double Dis = rho * a.ED[c];
        const double TEc = a.TE[c];
        S = S + TP * vol;
        A = A + Dis / (TEc + 0.000000000000000001) * vol;
        if (wallcell) {
            for (int k = 0; k < 6; k++) {
                const int f = fo[k];
                const int c1 = ot[k] & 0x7fffffff;
                if (c1 < N || g.btype[c1 - N] != WALL) continue;
                S = S - TP * vol;
                const dv3 Xpn = ld3(g.fcx, g.fcy, g.fcz, f) - cc;
                const dv3 Up = v3(a.U[c] - 0., a.V[c] - 0., a.W[c] - 0.);
                const dv3 n = ld3(g.fnx, g.fny, g.fnz, f);
                const dv3 nh = n / mag(n);
                const double dp = dot(Up, nh);
                const dv3 Ut = Up - nh * dp;
                const double Uvel = sqrt(dot(Ut, Ut));
                const double distance_to_face = mag(Xpn);
                const double rkapdn = 1.0 / (0.419 * distance_to_face);
                const double Tau_w = ev * Uvel / distance_to_face;
                const double Utau = sqrt(Tau_w / rho);
                TP = Tau_w * Utau * rkapdn;
                if (f == g.last_face[c1 - N]) a.TE[N + c1 - N] = TEc;
                S = S + TP * vol;
                const double DisP = a.cmu75 * sqrt(TEc) * rkapdn;
                A = A + rho * DisP * vol;
            }
        }
        st_out(&a.TP[c], TP);
    } else if (ed) {
        const double TEc = a.TE[c];
        const double fact = a.ED[c] / (TEc + 0.000000000000000001) * vol;
        S = S + 1.44 * fact * a.TP[c];
        A = A + 1.92 * fact * rho;
        if (wallcell) {
            for (int k = 0; k < 6; k++) {
                const int f = fo[k];
                const int c1 = ot[k] & 0x7fffffff;
                if (c1 < N || g.btype[c1 - N] != WALL) continue;
                const double distance = mag(ld3(g.fcx, g.fcy, g.fcz, f) - cc);
                const double Dis = a.cmu75 * pow(TEc, 1.5) / (distance * 0.419);
                phic = Dis;
                S = Dis;
                A = 1;
                if (f == g.last_face[c1 - N]) a.ED[N + c1 - N] = Dis;
            }
            // The reference stores Dis into phi.ED[c] here (FlowSolver.inl:2876),
            // after every face flux has read the old values; the solve that
            // follows (x0 = 0) overwrites it, so the value only enters through
            // phic below, and neighbours keep reading the old ED[c].
        }
    }
    for (int k = 0; k < nint; k++) A -= own_ff[k];
    A *= (1. / a.urf);
    S = S + (1.0 - a.urf) * A * phic;
    st_out(&a.AV[c], A);
    st_out(&a.SU[c], S);
    if (a.r0) st_out(&a.r0[c], S - (A * pc + xo));   // x0 = phi[c], also where phic is ED's wall value
    store_row(g, c, nint, own_nb, own_ff, A, a.vals, a.dia);
}

// The passive scalars FUL, PRO, VARF and VARP share one Visac rule
// (FlowSolver.inl:2557-2560), the same under-relaxation (0.95) and no
// turbulence terms, so their Scalar_solve systems have bit-identical matrices
// and differ only in the right-hand side: one kernel assembles the shared
// matrix once and the four right-hand sides, each with exactly the additions
// of the single-field kernel above.
constexpr int SB = 4;
struct ScalBatchArgs {
    double *phi[SB];
    const double *gx[SB], *gy[SB], *gz[SB];
    double *S[SB];
    double inlet_phi[SB];
    const double *mflux;
    double *ff0, *ff1, *AV, *vals, *dia;
    int type;
    double ev, inlet_ev, rdelta, urf;
    // the temperature equation in the same pass (its own Visac and matrix)
    double *tphi;
    const double *tgx, *tgy, *tgz;
    double *tS, *tvals, *tdia;
    double tinlet;
};

// TEM: with its own diffusion coefficient, every expression as in k_scalar
// for TEMP, so its matrix and right-hand side are the ones k_scalar forms.
// Three waves per SIMD (168 VGPRs, a 44-byte spill) run it 10% faster than
// the two its natural 180 VGPRs allow (measured at m = 128).
template <bool TEM>
__global__ void __launch_bounds__(256, 3) k_scalar_batch(Geo g, ScalBatchArgs a)
{
    const int c = walk_cell(g.w);
    if (c < 0) return;
    const int N = g.N;
    const double ev = a.ev;
    double S[SB], A = 0.0;
#pragma unroll
    for (int s = 0; s < SB; s++) S[s] = 0.0;
    double own_ff[6];
    int own_nb[6];
    int nint = 0;
    double St = 0.0, At = 0.0, own_fft[6];
    // the cell's own values, read once (each face reads the other cell's)
    const dv3 cc = ld3(g.cx, g.cy, g.cz, c);
    dv3 gs[SB];
    double ps[SB];
#pragma unroll
    for (int s = 0; s < SB; s++) { gs[s] = ld3(a.gx[s], a.gy[s], a.gz[s], c); ps[s] = a.phi[s][c]; }
    dv3 gt = v3(0.0, 0.0, 0.0);
    double pt = 0.0;
    if constexpr (TEM) { gt = ld3(a.tgx, a.tgy, a.tgz, c); pt = a.tphi[c]; }
    const CellFaces cf = cell_faces_of(g, c);
    for (int k = 0; k < 6; k++) {
        const int f = cf.f[k];
        const int o = cf.o[k];
        const bool own0 = o >= 0;          // c is the face's cell0
        const int nb = o & 0x7fffffff;     // cell across the face, or N + side
        const double mf = a.mflux[f];
        if (nb < N) {
            const double lambda0 = g.flam[f], lambda1 = 1.0 - lambda0;
            double Visac = ev * lambda0 + ev * lambda1;
            Visac -= ev;
            Visac = scalar_visac(a.type, ev, Visac);
            const dv3 cn = ld3(g.cx, g.cy, g.cz, nb);
            const dv3 Xpn = own0 ? (cn - cc) : (cc - cn);
            const dv3 nrm = ld3(g.fnx, g.fny, g.fnz, f);
            const double VisFace = Visac * g.frlc[f];
            const double ff0 = -VisFace - dmax(mf, 0.0);
            const double ff1 = -VisFace + dmin(mf, 0.0);
            // The convective deferred correction GammaBlend * (fce - fci) is a
            // signed zero for finite operands (GammaBlend = 0, FlowSolver.inl:
            // 2523), and adding or subtracting a signed zero leaves a source
            // sum that is never -0 unchanged (it starts at +0 and x - x = +0):
            // the term, and the neighbour values that only it reads, are left
            // out here (k_scalar keeps them; the batch equals it bit for bit).
            if constexpr (TEM) {
                double Vt = ev * lambda0 + ev * lambda1;
                Vt -= ev;
                Vt = scalar_visac(TEMP, ev, Vt);
                const dv3 gn = ld3(a.tgx, a.tgy, a.tgz, nb);
                const dv3 dPhiXac = (own0 ? gt : gn) * lambda0 + (own0 ? gn : gt) * lambda1;
                const double VisFaceT = Vt * g.frlc[f];
                const double fde1 = Vt * dot(dPhiXac, nrm);
                const double fdi = VisFaceT * dot(dPhiXac, Xpn);
                const double t0 = -VisFaceT - dmax(mf, 0.0);
                const double t1 = -VisFaceT + dmin(mf, 0.0);
                if (own0) { St = St + fde1 - fdi; own_fft[nint] = t1; }
                else { St = St - fde1 + fdi; own_fft[nint] = t0; }
            }
#pragma unroll
            for (int s = 0; s < SB; s++) {
                const dv3 gn = ld3(a.gx[s], a.gy[s], a.gz[s], nb);
                const dv3 dPhiXac = (own0 ? gs[s] : gn) * lambda0 + (own0 ? gn : gs[s]) * lambda1;
                const double fde1 = Visac * dot(dPhiXac, nrm);
                const double fdi = VisFace * dot(dPhiXac, Xpn);
                if (own0) S[s] = S[s] + fde1 - fdi;
                else S[s] = S[s] - fde1 + fdi;
            }
            if (own0) {
                own_ff[nint] = ff1;
                st_out(&a.ff0[f], ff0); st_out(&a.ff1[f], ff1);
            } else {
                own_ff[nint] = ff0;
            }
            own_nb[nint] = nb;
            nint++;
        } else {
            const int side = nb - N;
            const int bt = g.btype[side];
            const bool last = (f == g.last_face[side]);
            const dv3 Xpn = ld3(g.fcx, g.fcy, g.fcz, f) - cc;
            const dv3 nrm = ld3(g.fnx, g.fny, g.fnz, f);
            if (bt == INLET) {
                if constexpr (TEM) {
                    double Vt = a.inlet_ev;
                    Vt -= ev;
                    Vt = scalar_visac(TEMP, ev, Vt);
                    const double VisFaceT = Vt * g.frlc[f];
                    const dv3 gP = gt;
                    const double fde = Vt * dot(gP, nrm);
                    const double fdi = VisFaceT * dot(gP, Xpn);
                    const double fb = -VisFaceT + dmin(mf, 0.0);
                    At = At - fb;
                    St = St - fb * a.tinlet + fde - fdi;
                    if (last) a.tphi[N + side] = a.tinlet;
                }
                double Visac = a.inlet_ev;
                Visac -= ev;
                Visac = scalar_visac(a.type, ev, Visac);
                const double VisFace = Visac * g.frlc[f];
                const double fb = -VisFace + dmin(mf, 0.0);
                A = A - fb;
#pragma unroll
                for (int s = 0; s < SB; s++) {
                    const dv3 gP = gs[s];
                    const double fde = Visac * dot(gP, nrm);
                    const double fdi = VisFace * dot(gP, Xpn);
                    S[s] = S[s] - fb * a.inlet_phi[s] + fde - fdi;
                    if (last) a.phi[s][N + side] = a.inlet_phi[s];
                }
            } else if (bt == OUTLET) {
                if constexpr (TEM) {
                    double Vt = ev;
                    Vt -= ev;
                    Vt = scalar_visac(TEMP, ev, Vt);
                    const dv3 gP = gt;
                    const double PhiFace = pt + dot(gP, Xpn);
                    const double VisFaceT = Vt * g.frlc[f];
                    const double fde = Vt * dot(gP, nrm);
                    const double fdi = VisFaceT * dot(gP, Xpn);
                    St = St + fde - fdi;
                    if (last) a.tphi[N + side] = PhiFace;
                }
                double Visac = ev;
                Visac -= ev;
                Visac = scalar_visac(a.type, ev, Visac);
                const double VisFace = Visac * g.frlc[f];
#pragma unroll
                for (int s = 0; s < SB; s++) {
                    const dv3 gP = gs[s];
                    const double PhiFace = ps[s] + dot(gP, Xpn);
                    const double fde = Visac * dot(gP, nrm);
                    const double fdi = VisFace * dot(gP, Xpn);
                    S[s] = S[s] + fde - fdi;
                    if (last) a.phi[s][N + side] = PhiFace;
                }
            } else if (bt == WALL) {
                if (last) {
                    if constexpr (TEM) a.tphi[N + side] = pt;
#pragma unroll
                    for (int s = 0; s < SB; s++) a.phi[s][N + side] = ps[s];
                }
            }
        }
    }
    const double fv = g.rho[c] * g.vol[c] * a.rdelta;
    if constexpr (TEM) {
        St += fv * pt;
        At += fv;
        for (int k = 0; k < nint; k++) At -= own_fft[k];
        At *= (1. / a.urf);
        St = St + (1.0 - a.urf) * At * pt;
        st_out(&a.tS[c], St);
    }
    double phic[SB];
#pragma unroll
    for (int s = 0; s < SB; s++) {
        phic[s] = ps[s];
        S[s] += fv * phic[s];
    }
    A += fv;
    for (int k = 0; k < nint; k++) A -= own_ff[k];
    A *= (1. / a.urf);
#pragma unroll
    for (int s = 0; s < SB; s++) {
        S[s] = S[s] + (1.0 - a.urf) * A * phic[s];
        st_out(&a.S[s][c], S[s]);
    }
    st_out(&a.AV[c], A);
    store_row(g, c, nint, own_nb, own_ff, A, a.vals, a.dia);
    if constexpr (TEM) store_row(g, c, nint, own_nb, own_fft, At, a.tvals, a.tdia);
}

// FUL, PRO, VARF, VARP (FlowSolver.inl:3243-3260) as one 4-RHS solve; with
// tem, the temperature equation (:3237-3242) is assembled in the same pass
// into its own matrix and solved first, as the reference orders them (none of
// the four reads the temperature, and TEM's assembly reads nothing the TE /
// ED solves change).
// The batched pass's arguments.  side = true: it writes the alternate matrix,
// diagonal and face-coefficient arrays and keeps VARP's right-hand side in
// srhs[3] (S_U is then the ED system's), so it can run beside the ED solve.
static ScalBatchArgs batch_args(Ctx *c, bool side)
{
    static const int types[SB] = {FUEL, PROG, VARFU, VARPR};
    static const int fields[SB] = {FUL, PRO, VARF, VARP};
    ScalBatchArgs a;
    for (int s = 0; s < SB; s++) {
        a.phi[s] = c->phi[fields[s]];
        a.gx[s] = c->gx[fields[s]]; a.gy[s] = c->gy[fields[s]]; a.gz[s] = c->gz[fields[s]];
        a.S[s] = (s == SB - 1 && !side) ? c->SU : c->srhs[s];   // S_phi.U ends holding VARP's, as in the reference
        a.inlet_phi[s] = (types[s] == FUEL || types[s] == VARFU) ? c->k.gas_ful : 0.0;
    }
    a.mflux = c->mflux;
    a.ff0 = side ? c->ff0b : c->ff0; a.ff1 = side ? c->ff1b : c->ff1; a.AV = side ? c->AV2 : c->AV;
    a.vals = c->A_vals;
    a.dia = side ? c->A_dia3 : (double *)fine_dia_of(c, c->A_vals);
    a.type = FUEL;
    a.ev = c->k.ev; a.inlet_ev = c->k.inlet_ev; a.rdelta = c->k.rdelta; a.urf = 0.95;
    a.tphi = c->phi[TEM]; a.tgx = c->gx[TEM]; a.tgy = c->gy[TEM]; a.tgz = c->gz[TEM];
    a.tS = c->srhs[SB]; a.tvals = c->A_vals2; a.tdia = (double *)fine_dia_of(c, c->A_vals2); a.tinlet = c->k.gas_tem;
    return a;
}

static void batch_launch(Ctx *c, const ScalBatchArgs &a, bool tem, hipStream_t s)
{
    // per cell and field phi and gradient (32) read and S (8) written;
    // per cell centre (24), volume and density (16), forder and fother
    // (48) read, A (8) written; per face mass flux, lambda, rlencos,
    // normal (48) read; per interior face ff0, ff1 (16) written; per
    // boundary face centre (24) read; one matrix per equation system
    // (8 nnz each) written
    const int nsc = SB + (tem ? 1 : 0);
    const Sizes z = sizes_of(c);
    const double bytes = (40.0 * nsc + 96.0) * z.N + 48.0 * z.F + 16.0 * z.Fi + 24.0 * z.Fb +
                         8.0 * z.nnz * (tem ? 2 : 1);
    KTimer kt(c, s == c->stream ? MC_KC_SCALAR_BATCH : -1, bytes);
    if (tem)
        hipLaunchKernelGGL(k_scalar_batch<true>, dim3(walk_blocks(cell_walk(c))), dim3(256), 0, s, geo(c), a);
    else
        hipLaunchKernelGGL(k_scalar_batch<false>, dim3(walk_blocks(cell_walk(c))), dim3(256), 0, s, geo(c), a);
}

// The solves of the batched pass's systems: TEM (its own matrix), then FUL,
// PRO, VARF, VARP as one 4-RHS solve.  The warm-started solves form b - A x0
// in their own pass here: in the assembly kernel the extra accumulators cost
// more (168 VGPRs, +0.2 ms) than the pass.
static int batch_solves(Ctx *c, const ScalBatchArgs &a, bool tem)
{
    static const int fields[SB] = {FUL, PRO, VARF, VARP};
    const bool warm = warm_start(c);
    StageTimer ts(c, MC_ST_SCA_SOLVE);
    if (tem) {
        double *bt[1] = {c->srhs[SB]}, *xt[1] = {c->phi[TEM]};
        mc_solve_status stt{};
        MC_TRY(bicgstab_solve(c, c->A_vals2, 1, bt, xt, c->opt.mom_rtol, c->opt.mom_maxit, &stt, warm));
        c->st.scalar_iters += stt.iterations;
        if (!stt.converged) { c->err = "scalar BiCGStab did not converge"; return MC_ERR_NOCONV; }
    }
    double *b[SB], *x[SB];
    for (int s = 0; s < SB; s++) { b[s] = a.S[s]; x[s] = c->phi[fields[s]]; }
    mc_solve_status st{};
    MC_TRY(bicgstab_solve(c, c->A_vals, SB, b, x, c->opt.mom_rtol, c->opt.mom_maxit, &st, warm));
    c->st.scalar_iters += st.iterations;
    if (!st.converged) { c->err = "scalar BiCGStab did not converge"; return MC_ERR_NOCONV; }
    return MC_OK;
}

static int batch_solve(Ctx *c, bool tem)
{
    const ScalBatchArgs a = batch_args(c, false);
    {
        StageTimer t(c, MC_ST_SCA_FLUX);
        batch_launch(c, a, tem, c->stream);
        MC_LAUNCH_CHECK(c);
    }
    return batch_solves(c, a, tem);
}

// The batched pass beside the ED solve: it reads nothing the TE / ED solves
// write (scalar fields and gradients, mass fluxes, geometry) and writes its
// own arrays, so it can run on the second stream from the end of ED's
// assembly to the end of ED's solve and repair (not beside an assembly
// kernel: TE / ED's `k_scalar` keeps the device to itself).  Unused while a
// stage, the batch kernel or the BiCGStab kernel is timed: their events are
// on the main stream and would take in the other stream's work.
static bool batch_beside(const Ctx *c)
{
    const unsigned stages = ((1u << MC_ST_COUNT) - 1u) << MC_KC_COUNT;
    return c->side && c->fine_nd && !(c->profile & (stages | (1u << MC_KC_SCALAR_BATCH) | (1u << MC_KC_BICG)));
}

// allocated at the first fork; cleared on the main stream, so the fork's
// event orders the clearing before the batch kernel on the other stream
static int batch_buffers(Ctx *c)
{
    if (c->A_dia3) return MC_OK;
    {
        MemScope ms(c, MC_MEM_MATRIX);
        MC_TRY(dalloc(c, &c->A_dia3, (size_t)c->fine_nd * c->N));   // entries outside the pattern stay zero
    }
    {
        MemScope ms(c, MC_MEM_FIELDS);
        MC_TRY(dalloc(c, &c->AV2, c->N + 6));   // boundary slots zero, as AV's
        MC_TRY(dalloc(c, &c->ff0b, c->F));      // boundary faces zero, as ff0's and ff1's
        MC_TRY(dalloc(c, &c->ff1b, c->F));
    }
    MC_HIP(c, hipMemsetAsync(c->A_dia3, 0, (size_t)c->fine_nd * c->N * 8, c->stream));
    MC_HIP(c, hipMemsetAsync(c->AV2, 0, (size_t)(c->N + 6) * 8, c->stream));
    MC_HIP(c, hipMemsetAsync(c->ff0b, 0, (size_t)c->F * 8, c->stream));
    MC_HIP(c, hipMemsetAsync(c->ff1b, 0, (size_t)c->F * 8, c->stream));
    if (!c->ev_fork) MC_HIP(c, hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    if (!c->ev_join) MC_HIP(c, hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    return MC_OK;
}

static int batch_fork(Ctx *c)
{
    MC_TRY(batch_buffers(c));
    MC_HIP(c, hipEventRecord(c->ev_fork, c->stream));
    MC_HIP(c, hipStreamWaitEvent(c->side, c->ev_fork, 0));
    batch_launch(c, batch_args(c, true), true, c->side);
    MC_LAUNCH_CHECK(c);
    MC_HIP(c, hipEventRecord(c->ev_join, c->side));
    return MC_OK;
}

// Join: the alternate arrays hold the batched systems (the last assembled, as
// in the reference's order) and become the current ones; VARP's right-hand
// side goes to S_U
static int batch_join(Ctx *c)
{
    MC_HIP(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
    std::swap(c->A_dia, c->A_dia3);
    std::swap(c->AV, c->AV2);
    std::swap(c->ff0, c->ff0b);
    std::swap(c->ff1, c->ff1b);
    const ScalBatchArgs a = batch_args(c, true);
    MC_TRY(batch_solves(c, a, true));
    MC_HIP(c, hipMemcpyAsync(c->SU, c->srhs[SB - 1], (size_t)c->N * 8, hipMemcpyDeviceToDevice, c->stream));
    return MC_OK;
}

int scalar_batch_solve(Ctx *c) { return batch_solve(c, false); }

static int scalar_field(int type)
{
    static const int f[7] = {TE, ED, TEM, FUL, PRO, VARF, VARP};
    return f[type];
}

int scalar_assemble(Ctx *c, int type, double *r0)
{
    const int field = scalar_field(type);
    ScalArgs a;
    a.r0 = r0;
    a.phi = c->phi[field];
    a.gx = c->gx[field]; a.gy = c->gy[field]; a.gz = c->gz[field];
    a.U = c->phi[U]; a.V = c->phi[V]; a.W = c->phi[W];
    a.gUx = c->gx[U]; a.gUy = c->gy[U]; a.gUz = c->gz[U];
    a.gVx = c->gx[V]; a.gVy = c->gy[V]; a.gVz = c->gz[V];
    a.gWx = c->gx[W]; a.gWy = c->gy[W]; a.gWz = c->gz[W];
    a.TE = c->phi[TE]; a.ED = c->phi[ED]; a.TP = c->phi[TP];
    a.mflux = c->mflux; a.ff0 = c->ff0; a.ff1 = c->ff1; a.AV = c->AV; a.SU = c->SU; a.vals = c->A_vals;
    a.dia = (double *)fine_dia_of(c, c->A_vals);
    a.type = type;
    a.ev = c->k.ev; a.inlet_ev = c->k.inlet_ev; a.rdelta = c->k.rdelta;
    a.urf = (type == TERBTE || type == TERBED) ? 0.5 : 0.95;
    a.cmu75 = c->k.cmu75;
    switch (type) {
    case TEMP: a.inlet_phi = c->k.gas_tem; break;
    case TERBTE: a.inlet_phi = c->k.inlet_te; break;
    case TERBED: a.inlet_phi = c->k.inlet_ed; break;
    case FUEL: case VARFU: a.inlet_phi = c->k.gas_ful; break;
    default: a.inlet_phi = 0.0; break;
    }
    // per cell phi and gradient (32), centre (24), volume and density (16),
    // forder and fother (48) read, A and S (16) written (r0: 8 more); TE also
    // reads ED (8) and the velocity gradients (72) and writes TP (8); ED also
    // reads TE, TP (16) and the wall flags (1); per face mass flux, lambda,
    // rlencos, normal (48) read; per interior face ff0, ff1 (16) written; per
    // boundary face centre (24) read; the matrix entries (8 nnz) written
    const Sizes z = sizes_of(c);
    double cell = 136.0 + (r0 ? 8.0 : 0.0);
    if (type == TERBTE) cell += 88.0;
    if (type == TERBED) cell += 17.0;
    KTimer kt(c, MC_KC_SCALAR_ASM, cell * z.N + 48.0 * z.F + 16.0 * z.Fi + 24.0 * z.Fb + 8.0 * z.nnz);
    hipLaunchKernelGGL(k_scalar, dim3(walk_blocks(cell_walk(c))), dim3(256), 0, c->stream, geo(c), a);
    MC_LAUNCH_CHECK(c);
    return MC_OK;
}

// Non-negativity repair of TE/ED (FlowSolver.inl:2956-2985).  The reference
// sweeps cells in ascending order and reads already-repaired lower
// neighbours.  A negative cell's new value depends only on lower-numbered
// negative cells (a DAG), so the sequential result is the unique fixed point
//   v_c = max(mean_j (j < c ? v_j : orig_j), 1e-12)   over the negative cells.
// The negative cells are compacted once (ascending); passes relax them in
// place until a pass changes nothing -- that state is the fixed point, hence
// the reference's values exactly, whatever order the threads ran in.
//
// Compaction of the negative cells in ascending order: block b owns
// [b * span, (b + 1) * span), span = 256 * rounds; k_neg_count counts its
// negatives, k_neg_scan turns the counts into exclusive offsets (one
// workgroup, fixed order) and the total, k_neg_write writes each negative
// cell at its block offset plus its rank in index order (wave ballots).
__device__ __forceinline__ bool is_neg(const double *v, int i, int n) { return i < n && v[i] < 0.0; }

__global__ void __launch_bounds__(256) k_neg_count(int n, const double *v, int rounds, int *bcount)
{
    __shared__ int ws[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int base = blockIdx.x * 256 * rounds;
    int cnt = 0;
    for (int r = 0; r < rounds; r++) {
        const unsigned long long b = __ballot(is_neg(v, base + r * 256 + (int)threadIdx.x, n));
        cnt += __popcll(b);
    }
    if (lane == 0) ws[wave] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) bcount[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

// in place: counts -> exclusive offsets; *total = the sum
__global__ void __launch_bounds__(1024) k_neg_scan(int nb, int *bcount, int *total)
{
    __shared__ int part[1024];
    const int t = threadIdx.x, per = (nb + 1023) / 1024;
    const int i0 = min(nb, t * per), i1 = min(nb, i0 + per);
    int s = 0;
    for (int i = i0; i < i1; i++) s += bcount[i];
    part[t] = s;
    __syncthreads();
    if (t == 0) {
        int run = 0;
        for (int k = 0; k < 1024; k++) { const int x = part[k]; part[k] = run; run += x; }
        *total = run;
    }
    __syncthreads();
    int run = part[t];
    for (int i = i0; i < i1; i++) { const int x = bcount[i]; bcount[i] = run; run += x; }
}

__global__ void __launch_bounds__(256) k_neg_write(int n, const double *v, int rounds, const int *boff, int *list)
{
    __shared__ int ws[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int base = blockIdx.x * 256 * rounds;
    int off = boff[blockIdx.x];
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int r = 0; r < rounds; r++) {
        const int i = base + r * 256 + (int)threadIdx.x;
        const bool neg = is_neg(v, i, n);
        const unsigned long long b = __ballot(neg);
        if (lane == 0) ws[wave] = __popcll(b);
        __syncthreads();
        int pre = off;
        for (int w = 0; w < wave; w++) pre += ws[w];
        if (neg) list[pre + __popcll(b & below)] = i;
        off += ws[0] + ws[1] + ws[2] + ws[3];
        __syncthreads();
    }
}

// One relaxation pass.  After the first (all != 0), a cell is evaluated only
// when a lower neighbour changed in the previous pass (dirty_in), the only
// way its inputs can change; a cell that changes marks its higher neighbours
// in dirty_out for the next pass.
__global__ void __launch_bounds__(256) k_clip_list(Geo g, const int *list, int count, const double *orig, double *phi,
                                                   const int *prev, int *changed, uint8_t *dirty_in,
                                                   uint8_t *dirty_out, int all)
{
    if (prev && *prev == 0) return;   // the previous pass changed nothing: fixed point reached
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    const int c = list[i];
    if (!all) {
        if (!dirty_in[c]) return;
        dirty_in[c] = 0;
    }
    int cnt = 0;
    double phisum = 0.0;
    for (int j = 0; j < 6; j++) {
        const int nb = g.nbr[6 * c + j];
        if (nb < 0) continue;
        cnt++;
        phisum += (nb < c) ? phi[nb] : orig[nb];
    }
    phisum = (phisum / cnt);
    const double out = dmax(phisum, 0.000000000001);
    const double cur = phi[c];
    if (out != cur && !(out != out && cur != cur)) {
        phi[c] = out;
        atomicAdd(changed, 1);
        for (int j = 0; j < 6; j++) {
            const int nb = g.nbr[6 * c + j];
            if (nb > c) dirty_out[nb] = 1;
        }
    }
}

int flow_workspace(Ctx *c)
{
    {
        MemScope ms(c, MC_MEM_MATRIX);
        MC_TRY(dalloc(c, &c->A_vals2, c->nnz));
    }
    for (int s = 0; s < 5; s++) MC_TRY(dalloc(c, &c->srhs[s], c->N));
    MC_TRY(dalloc(c, &c->neg_list, c->N));
    return MC_OK;
}

// Passes are launched in batches with one host check per batch; a pass
// whose predecessor changed nothing returns at once, so a batch longer than
// needed costs only launches.  The batch length follows the passes the same
// field needed last time.
constexpr int CLIP_BATCH_MAX = 64;

int clip_nonneg(Ctx *c, double *phi, int slot)
{
    const int N = c->N;
    int *dcount = (int *)(c->dscal + 700);
    int *hcount = (int *)(c->hscal + 700);
    int *changed = (int *)(c->dscal + 704);   // CLIP_BATCH_MAX pass counters
    int *hchanged = (int *)(c->hscal + 704);
    uint8_t *flag = (uint8_t *)c->wk[13];
    {
        // block counts in the reduction partials buffer (as ints)
        const int cap = 2 * RED_PARTIALS;
        int rounds = 4;
        while (div_up(N, 256 * rounds) > cap) rounds *= 2;
        const int nb = div_up(N, 256 * rounds);
        int *bc = (int *)c->red;
        hipLaunchKernelGGL(k_neg_count, dim3(nb), dim3(256), 0, c->stream, N, (const double *)phi, rounds, bc);
        hipLaunchKernelGGL(k_neg_scan, dim3(1), dim3(1024), 0, c->stream, nb, bc, dcount);
        hipLaunchKernelGGL(k_neg_write, dim3(nb), dim3(256), 0, c->stream, N, (const double *)phi, rounds,
                           (const int *)bc, c->neg_list);
        MC_LAUNCH_CHECK(c);
    }
    MC_HIP(c, hipMemcpyAsync(hcount, dcount, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    MC_HIP(c, hipStreamSynchronize(c->stream));
    const int K = *hcount;
    c->st.clip_cells += K;
    if (K == 0) return MC_OK;
    double *orig = c->wk[14];
    MC_HIP(c, hipMemcpyAsync(orig, phi, (size_t)N * 8, hipMemcpyDeviceToDevice, c->stream));
    uint8_t *dirty[2] = {flag + N, flag + 2 * (size_t)N};   // inside the same 8N-byte work vector
    MC_HIP(c, hipMemsetAsync(dirty[0], 0, 2 * (size_t)N, c->stream));
    int B = std::min(std::max(c->clip_batch[slot], 4), CLIP_BATCH_MAX);
    int used = 0, pass = 0;
    for (int batch = 0; batch < 100000; batch++) {
        MC_HIP(c, hipMemsetAsync(changed, 0, B * sizeof(int), c->stream));
        for (int p = 0; p < B; p++, pass++)
            hipLaunchKernelGGL(k_clip_list, dim3(div_up(K, 256)), dim3(256), 0, c->stream, geo(c), c->neg_list, K,
                               orig, phi, p ? changed + p - 1 : (const int *)nullptr, changed + p, dirty[pass & 1],
                               dirty[(pass + 1) & 1], pass == 0 ? 1 : 0);
        MC_LAUNCH_CHECK(c);
        MC_HIP(c, hipMemcpyAsync(hchanged, changed, B * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        MC_HIP(c, hipStreamSynchronize(c->stream));
        int ran = B;   // passes that did work: up to and including the first that changed nothing
        for (int p = 0; p < B; p++)
            if (hchanged[p] == 0) { ran = p + 1; break; }
        used += ran;
        if (hchanged[B - 1] == 0) break;
        B = CLIP_BATCH_MAX;
    }
    c->st.clip_passes += used;
    c->clip_batch[slot] = used + 1;
    return MC_OK;
}

int scalar_solve(Ctx *c, int type, bool fork_batch)
{
    const int field = scalar_field(type);
    const bool warm = warm_start(c);
    // TE and ED: the assembly forms b - A x0 for the warm start; the scalars
    // the batched kernel assembles take the solver's own pass, as there, so a
    // single solve and the batch give the same bits
    const bool turb = type == TERBTE || type == TERBED;
    const bool fused = warm && turb;
    if (fused) MC_TRY(r0_buffers(c));
    {
        StageTimer t(c, turb ? MC_ST_TURB_FLUX : MC_ST_SCA_FLUX);
        MC_TRY(scalar_assemble(c, type, fused ? c->r0buf[0] : nullptr));
    }
    if (fork_batch) MC_TRY(batch_fork(c));
    StageTimer ts(c, turb ? MC_ST_TURB_SOLVE : MC_ST_SCA_SOLVE);
    double *b[1] = {c->SU};
    double *x[1] = {c->phi[field]};
    mc_solve_status st{};
    MC_TRY(bicgstab_solve(c, c->A_vals, 1, b, x, c->opt.mom_rtol, c->opt.mom_maxit, &st, warm,
                          fused ? c->r0buf : nullptr));
    c->st.scalar_iters += st.iterations;
    if (!st.converged) { c->err = "scalar BiCGStab did not converge"; return MC_ERR_NOCONV; }
    if (type == TERBTE || type == TERBED) MC_TRY(clip_nonneg(c, c->phi[field], type == TERBTE ? 0 : 1));
    return MC_OK;
}

// ============================================================= A17 ==
int simple_iteration(Ctx *c, int inner)
{
    {
        StageTimer t(c, MC_ST_GRAD);
        MC_TRY(grad_multi(c, c->opt.limit_grad != 0));   // + limit_phi_gradients, FlowSolver.inl:3197-3198
    }
    // FlowSolver.inl:3204-3208: the first pass of the first timestep
    if (c->timestep_count == 0 && inner == 0 && !c->field_set_up) {
        StageTimer t(c, MC_ST_PRES_FLUX);
        MC_TRY(set_up_field(c));
        c->field_set_up = true;
    }
    {
        StageTimer t(c, MC_ST_VEL_FLUX);
        MC_TRY(momentum_assemble(c, true));
    }
    {
        StageTimer t(c, MC_ST_VEL_SOLVE);
        MC_TRY(momentum_solve(c, true));
    }
    MC_TRY(calculate_pressure(c));
    if (c->opt.solve_scalars) {
        MC_TRY(scalar_solve(c, TERBTE));
        // TEM, then FUL, PRO, VARF, VARP: assembled beside the ED solve when
        // nothing is being timed
        const bool beside = batch_beside(c);
        const int rc = scalar_solve(c, TERBED, beside);
        if (rc != MC_OK) {
            // a failed ED solve returns with the batch kernel perhaps still
            // running: later work on the main stream is ordered after it
            if (beside && c->ev_join) (void)hipStreamWaitEvent(c->stream, c->ev_join, 0);
            return rc;
        }
        MC_TRY(beside ? batch_join(c) : batch_solve(c, true));
    }
    c->st.simple_iterations++;
    c->global_iter++;
    return MC_OK;
}

}  // namespace mc
