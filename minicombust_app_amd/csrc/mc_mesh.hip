// mc_mesh.hip -- mesh generation, geometry, initial state and CSR pattern.
//
// A1: the box mesh of src/examples/mesh_examples.cpp:120-571 at one flow rank,
//     generated in parallel.  The reference numbers faces in the order the
//     sequential cell/side loop first meets them (:510-561).  For the box that
//     order has a closed form: cell c = (x,y,z) creates its BACK, RIGHT and UP
//     faces, plus FRONT/LEFT/DOWN when they lie on the boundary, in side
//     order, and the faces created before c number
//         3c + min(c, nx*ny) + ceil(c/nx) + (z*nx + min(c mod nx*ny, nx)).
//     Each thread therefore writes its own cell_faces row and the faces it
//     creates without any serial pass; tests compare every array bit for bit
//     with the reference generator.
// A2/A3: face/cell geometry and initial fields (FlowSolver.hpp:252-485), the
//     same operations in the same order, compiled without FP contraction.
#include <hipcub/hipcub.hpp>

#include "mc_internal.hpp"

#pragma clang fp contract(off)

namespace mc {

static const int CFVM[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 2, 4, 6}, {1, 3, 5, 7}, {0, 1, 4, 5}, {2, 3, 6, 7}};
__constant__ int d_cfvm[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 2, 4, 6}, {1, 3, 5, 7}, {0, 1, 4, 5}, {2, 3, 6, 7}};

// ------------------------------------------------------------------ A1 --
__global__ void k_box_points(int64_t NP, int64_t px_, int64_t py_, int64_t pz_, double edx, double edy, double edz,
                             double *px, double *py, double *pz, uint8_t *cpp)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NP) return;
    const int64_t x = i % px_, y = (i / px_) % py_, z = i / (px_ * py_);
    px[i] = (double)x * edx;
    py[i] = (double)y * edy;
    pz[i] = (double)z * edz;
    const int first = (x == 0) + (y == 0) + (z == 0);
    const int last = (x == px_ - 1) + (y == py_ - 1) + (z == pz_ - 1);
    cpp[i] = (uint8_t)(1u << (3 - first - last));
}

__device__ __forceinline__ int64_t faces_before(int64_t c, int64_t nx, int64_t nxny)
{
    const int64_t z = c / nxny, r = c - z * nxny;
    const int64_t z0 = c < nxny ? c : nxny;
    const int64_t x0 = (c + nx - 1) / nx;
    const int64_t y0 = z * nx + (r < nx ? r : nx);
    return 3 * c + z0 + x0 + y0;
}

// Storage position of a face in the side-major order: the RIGHT, UP and BACK
// faces of cell c at c, N + c and 2N + c, then the LEFT (x = 0), DOWN (y = 0)
// and FRONT (z = 0) boundary faces by their cell's position on that side.
__device__ __forceinline__ int64_t box_face_pos(int side, int64_t c, int64_t x, int64_t y, int64_t z, int64_t nx,
                                                int64_t ny, int64_t nz, int64_t N)
{
    const int64_t nxny = nx * ny;
    switch (side) {
    case 3: return c;                                       // RIGHT
    case 5: return N + c;                                   // UP
    case 1: return 2 * N + c;                               // BACK
    case 2: return x > 0 ? c - 1 : 3 * N + y + ny * z;      // LEFT: RIGHT face of x - 1
    case 4: return y > 0 ? N + c - nx : 3 * N + ny * nz + x + nx * z;                   // DOWN
    default: return z > 0 ? 2 * N + c - nxny : 3 * N + ny * nz + nx * nz + x + nx * y;  // FRONT
    }
}

__global__ void k_box_cells(int64_t N, int64_t nx, int64_t ny, int64_t nz, uint64_t *cells, int *nbr,
                            int *cell_faces, int *fc0, int *fc1, int *face_ref, int *face_pos)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    const int64_t nxny = nx * ny;
    const int64_t z = c / nxny, y = (c / nx) % ny, x = c % nx;
    const int64_t px_ = nx + 1, pxy = (nx + 1) * (ny + 1);
    const uint64_t gp = x + y * px_ + z * pxy;
    uint64_t *cv = cells + 8 * c;
    cv[0] = gp; cv[7] = gp + pxy + px_ + 1; cv[1] = gp + 1; cv[2] = gp + px_;
    cv[3] = gp + px_ + 1; cv[4] = gp + pxy; cv[5] = gp + pxy + 1; cv[6] = gp + pxy + px_;
    int64_t nb[6];
    nb[0] = (z == 0) ? -1 : c - nxny;
    nb[1] = (z == nz - 1) ? -1 : c + nxny;
    nb[2] = (x == 0) ? -1 : c - 1;
    nb[3] = (x == nx - 1) ? -1 : c + 1;
    nb[4] = (y == 0) ? -1 : c - nx;
    nb[5] = (y == ny - 1) ? -1 : c + nx;
    bool created[6] = {z == 0, true, x == 0, true, y == 0, true};
    const int64_t base = faces_before(c, nx, nxny);
    int64_t id[6];
    int k = 0;
    for (int f = 0; f < 6; f++) if (created[f]) id[f] = base + k++;
    if (z > 0) { const int64_t cp = c - nxny; id[0] = faces_before(cp, nx, nxny) + ((z - 1) == 0); }
    if (x > 0) { const int64_t cp = c - 1; id[2] = faces_before(cp, nx, nxny) + (z == 0) + 1 + ((x - 1) == 0); }
    if (y > 0) {
        const int64_t cp = c - nx;
        id[4] = faces_before(cp, nx, nxny) + (z == 0) + 1 + (x == 0) + 1 + ((y - 1) == 0);
    }
    for (int f = 0; f < 6; f++) {
        const int64_t pos = box_face_pos(f, c, x, y, z, nx, ny, nz, N);
        nbr[6 * c + f] = (int)nb[f];
        cell_faces[6 * c + f] = (int)pos;
        if (created[f]) {
            const int64_t other = nb[f] < 0 ? N + f : nb[f];
            fc0[pos] = (int)(c < other ? c : other);
            fc1[pos] = (int)(c < other ? other : c);
            face_ref[pos] = (int)id[f];
            face_pos[id[f]] = (int)pos;
        }
    }
}

// --------------------------------------------------- host mesh upload --
__global__ void k_from_u64(int64_t n, const uint64_t *in, int *out, int64_t boundary_sub)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t v = in[i];
    out[i] = (v == UINT64_MAX) ? -1 : (int)v;
    (void)boundary_sub;
}

__global__ void k_split_faces(int64_t F, const uint64_t *faces, int *fc0, int *fc1)
{
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    fc0[f] = (int)faces[2 * f];
    fc1[f] = (int)faces[2 * f + 1];
}

__global__ void k_split_points(int64_t NP, const double *p, double *px, double *py, double *pz)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NP) return;
    px[i] = p[3 * i]; py[i] = p[3 * i + 1]; pz[i] = p[3 * i + 2];
}

// ------------------------------------------------------------ A2 / A3 --
struct dv3 { double x, y, z; };
__device__ __forceinline__ dv3 v3(double x, double y, double z) { dv3 r; r.x = x; r.y = y; r.z = z; return r; }
__device__ __forceinline__ dv3 operator+(dv3 a, dv3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ dv3 operator-(dv3 a, dv3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ dv3 operator*(dv3 a, double b) { return v3(a.x * b, a.y * b, a.z * b); }
__device__ __forceinline__ dv3 operator/(dv3 a, double b) { return v3(a.x / b, a.y / b, a.z / b); }
__device__ __forceinline__ double dot(dv3 a, dv3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ double mag(dv3 a) { return sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
__device__ __forceinline__ dv3 cross(dv3 a, dv3 b)
{
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double cosangle(dv3 a, dv3 b) { return dot(a, b) / (mag(a) * mag(b)); }

// Mesh.hpp:54-67 and the cell loop of FlowSolver.hpp:301-347.
__global__ void k_cell_geom(int N, const uint64_t *cells, const double *px, const double *py, const double *pz,
                            double *cx, double *cy, double *cz, double *vol, double *rho)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    const uint64_t *cv = cells + 8 * (int64_t)c;
    dv3 s = v3(0.0, 0.0, 0.0);
    for (int i = 0; i < 8; i++) s = s + v3(px[cv[i]], py[cv[i]], pz[cv[i]]);
    s = s / 8.0;
    cx[c] = s.x; cy[c] = s.y; cz[c] = s.z;
    const dv3 a = v3(px[cv[0]], py[cv[0]], pz[cv[0]]);
    const double width = mag(v3(px[cv[1]], py[cv[1]], pz[cv[1]]) - a);
    const double height = mag(v3(px[cv[2]], py[cv[2]], pz[cv[2]]) - a);
    const double length = mag(v3(px[cv[4]], py[cv[4]], pz[cv[4]]) - a);
    vol[c] = width * height * length;
    rho[c] = 1.2;
}

// Interior faces: FlowSolver.hpp:252-294.  Boundary faces: :349-397.
__global__ void k_face_geom(int F, int N, const int *fc0, const int *fc1, const uint64_t *cells, const double *px,
                            const double *py, const double *pz, const double *cx, const double *cy, const double *cz,
                            const int *btype, double *farea, double *flam, double *frlc, double *fnx, double *fny,
                            double *fnz, double *fcx, double *fcy, double *fcz, double *mflux)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const int c0 = fc0[f], c1 = fc1[f];
    const dv3 cc0 = v3(cx[c0], cy[c0], cz[c0]);
    if (c1 < N) {
        const uint64_t *a = cells + 8 * (int64_t)c0, *b = cells + 8 * (int64_t)c1;
        uint64_t ids[4] = {0, 0, 0, 0};
        int n = 0;
        for (int n0 = 0; n0 < 8; n0++)
            for (int n1 = 0; n1 < 8; n1++)
                if (a[n0] == b[n1] && n < 4) ids[n++] = a[n0];
        const dv3 fn0 = v3(px[ids[0]], py[ids[0]], pz[ids[0]]), fn1 = v3(px[ids[1]], py[ids[1]], pz[ids[1]]);
        const dv3 fn2 = v3(px[ids[2]], py[ids[2]], pz[ids[2]]), fn3 = v3(px[ids[3]], py[ids[3]], pz[ids[3]]);
        const dv3 cc1 = v3(cx[c1], cy[c1], cz[c1]);
        const dv3 c01 = cc1 - cc0;
        const double area = mag(fn2 - fn0) * mag(fn1 - fn0);
        dv3 fc = ((fn0 + fn1) + fn2) + fn3;
        fc = fc / 4.0;
        const double lam = mag(fc - cc0) / mag(c01);
        dv3 nrm = cross(fn2 - fn0, fn1 - fn0);
        if (dot(nrm, cc1 - cc0) < 0) nrm = nrm * -1.;
        farea[f] = area; flam[f] = lam;
        frlc[f] = area / mag(c01) / cosangle(nrm, c01);
        fnx[f] = nrm.x; fny[f] = nrm.y; fnz[f] = nrm.z;
        fcx[f] = fc.x; fcy[f] = fc.y; fcz[f] = fc.z;
        mflux[f] = 0.0;
    } else {
        const int side = c1 - N;
        const uint64_t *cv = cells + 8 * (int64_t)c0;
        dv3 fn[4];
        for (int k = 0; k < 4; k++) {
            const uint64_t p = cv[d_cfvm[side][k]];
            fn[k] = v3(px[p], py[p], pz[p]);
        }
        const double area = mag(fn[2] - fn[0]) * mag(fn[1] - fn[0]);
        const dv3 fc = (((fn[0] + fn[1]) + fn[2]) + fn[3]) / 4.0;
        dv3 nrm = cross(fn[2] - fn[0], fn[1] - fn[0]);
        const int bt = btype[side];
        if (bt == INLET || bt == OUTLET || bt == WALL) {
            mflux[f] = 0.0;
            if (dot(nrm, cc0 - fc) > 0) nrm = nrm * -1.;
        }
        const dv3 cf = fc - cc0;
        farea[f] = area; flam[f] = 1.0;
        frlc[f] = area / mag(cf) / cosangle(nrm, cf);
        fnx[f] = nrm.x; fny[f] = nrm.y; fnz[f] = nrm.z;
        fcx[f] = fc.x; fcy[f] = fc.y; fcz[f] = fc.z;
    }
}

struct InitVals { double v[NFIELDS]; };

__global__ void k_init_fields(int N, InitVals iv, double *const *phi, double *const *gx, double *const *gy,
                              double *const *gz, double *AU, double *AV, double *AW, double *SU, double *SV, double *SW)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    for (int f = 0; f < NFIELDS; f++) {
        phi[f][c] = iv.v[f];
        gx[f][c] = 0.0; gy[f][c] = 0.0; gz[f][c] = 0.0;
    }
    AU[c] = 0.0; AV[c] = 0.0; AW[c] = 0.0; SU[c] = 0.0; SV[c] = 0.0; SW[c] = 0.0;
}

// --------------------------------------------------------------- CSR --
__global__ void k_row_count(int N, const int *nbr, int *cnt)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > N) return;
    if (c == N) { cnt[c] = 0; return; }
    int n = 1;
    for (int s = 0; s < 6; s++) n += nbr[6 * c + s] >= 0;
    cnt[c] = n;
}

// Row = {c} u {interior neighbours}, columns ascending (MatSeqAIJ after
// MatSetValue of FlowSolver.inl:1221-1266).  forder = the cell's faces in
// ascending face id: the order in which the reference's face loops visit them.
__global__ void k_row_fill(int N, const int *nbr, const int *cell_faces, const int *face_ref, const int *row_ptr,
                           int *col, int *diag_idx, int *forder)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N) return;
    int cols[7];
    int n = 0;
    cols[n++] = c;
    for (int s = 0; s < 6; s++) if (nbr[6 * c + s] >= 0) cols[n++] = nbr[6 * c + s];
    for (int a = 1; a < n; a++) {
        const int v = cols[a];
        int b = a - 1;
        while (b >= 0 && cols[b] > v) { cols[b + 1] = cols[b]; b--; }
        cols[b + 1] = v;
    }
    const int rp = row_ptr[c];
    for (int k = 0; k < n; k++) {
        col[rp + k] = cols[k];
        if (cols[k] == c) diag_idx[c] = rp + k;
    }
    int fo[6], id[6];
    for (int s = 0; s < 6; s++) { fo[s] = cell_faces[6 * c + s]; id[s] = face_ref[fo[s]]; }
    for (int a = 1; a < 6; a++) {   // by reference face id
        const int v = fo[a], iv = id[a];
        int b = a - 1;
        while (b >= 0 && id[b] > iv) { fo[b + 1] = fo[b]; id[b + 1] = id[b]; b--; }
        fo[b + 1] = v; id[b + 1] = iv;
    }
    for (int s = 0; s < 6; s++) forder[6 * c + s] = fo[s];
}

// fother[6c + k]: the cell across face forder[6c + k] (boundary faces:
// N + side), with bit 31 set when c is that face's cell1.
__global__ void k_face_other(int N, const int *forder, const int *fc0, const int *fc1, int *fother)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 6 * N) return;
    const int c = i / 6, f = forder[i];
    const int c0 = fc0[f], c1 = fc1[f];
    fother[i] = (c0 == c) ? c1 : (int)((unsigned)c0 | 0x80000000u);
}

// flags and last faces by reference face id r (the selection below then lists
// boundary faces in reference order, as storage positions)
__global__ void k_boundary_flags(int F, int N, const int *fc1, const int *face_pos, uint8_t *flag, int *last_ref)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= F) return;
    const int f = face_pos[r];
    const int b = fc1[f] >= N;
    flag[r] = (uint8_t)b;
    if (b) atomicMax(&last_ref[fc1[f] - N], r);
}

__global__ void k_wall_cells(int Fb, int N, const int *bfaces, const int *fc0, const int *fc1, const int *btype,
                             uint8_t *is_wall)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Fb) return;
    const int f = bfaces[i];
    if (btype[fc1[f] - N] == WALL) is_wall[fc0[f]] = 1;
}

__global__ void k_iota(int n, int *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = i;
}

// ------------------------------------------------------------------ host --
int mesh_box_generate(Ctx *c, int64_t m)
{
    if (m < 1 || m > 400) { c->err = "modifier out of range"; return MC_ERR_ARG; }
    c->m = m;
    // src/minicombust.cpp:123-124 and mesh_examples.cpp:125-126, host doubles
    const double mdx = 0.05 * (double)m * 2, mdy = 0.05 * (double)m, mdz = 0.05 * (double)m;
    const int64_t ex = 2 * m, ey = m, ez = m;
    const double edx = mdx / (double)ex, edy = mdy / (double)ey, edz = mdz / (double)ez;
    const int64_t N = ex * ey * ez, NP = (ex + 1) * (ey + 1) * (ez + 1);
    const int64_t F = (ex + 1) * ey * ez + (ey + 1) * ex * ez + (ez + 1) * ex * ey;
    if (F >= (int64_t)INT32_MAX - 8) { c->err = "mesh too large for int32 indices"; return MC_ERR_ARG; }
    c->N = (int)N; c->F = (int)F; c->NP = (int)NP;
    c->btype[0] = WALL; c->btype[1] = WALL; c->btype[2] = INLET;
    c->btype[3] = OUTLET; c->btype[4] = WALL; c->btype[5] = WALL;
    MC_TRY(dalloc(c, &c->px, NP)); MC_TRY(dalloc(c, &c->py, NP)); MC_TRY(dalloc(c, &c->pz, NP));
    MC_TRY(dalloc(c, &c->cells_per_point, NP));
    MC_TRY(dalloc(c, &c->cells, 8 * N));
    MC_TRY(dalloc(c, &c->nbr, 6 * N));
    MC_TRY(dalloc(c, &c->cell_faces, 6 * N));
    MC_TRY(dalloc(c, &c->fc0, F)); MC_TRY(dalloc(c, &c->fc1, F));
    MC_TRY(dalloc(c, &c->face_ref, F)); MC_TRY(dalloc(c, &c->face_pos, F));
    hipLaunchKernelGGL(k_box_points, dim3(div_up(NP, 256)), dim3(256), 0, c->stream, NP, ex + 1, ey + 1, ez + 1, edx, edy,
                       edz, c->px, c->py, c->pz, c->cells_per_point);
    MC_LAUNCH_CHECK(c);
    hipLaunchKernelGGL(k_box_cells, dim3(div_up(N, 256)), dim3(256), 0, c->stream, N, ex, ey, ez, c->cells, c->nbr,
                       c->cell_faces, c->fc0, c->fc1, c->face_ref, c->face_pos);
    MC_LAUNCH_CHECK(c);
    return MC_OK;
}

int mesh_from_host(Ctx *c, const mc_mesh_desc *d)
{
    if (!d || d->n_cells <= 0 || d->n_faces <= 0 || d->n_points <= 0 || !d->points || !d->cells || !d->faces ||
        !d->cell_faces || !d->cell_neighbours) {
        c->err = "incomplete mesh descriptor";
        return MC_ERR_ARG;
    }
    if (d->n_faces >= (int64_t)INT32_MAX - 8) { c->err = "mesh too large for int32 indices"; return MC_ERR_ARG; }
    const int64_t N = d->n_cells, F = d->n_faces, NP = d->n_points;
    c->N = (int)N; c->F = (int)F; c->NP = (int)NP; c->m = 0;
    for (int s = 0; s < 6; s++) c->btype[s] = (int)d->boundary_types[s];
    MC_TRY(dalloc(c, &c->px, NP)); MC_TRY(dalloc(c, &c->py, NP)); MC_TRY(dalloc(c, &c->pz, NP));
    MC_TRY(dalloc(c, &c->cells_per_point, NP));
    MC_TRY(dalloc(c, &c->cells, 8 * N));
    MC_TRY(dalloc(c, &c->nbr, 6 * N));
    MC_TRY(dalloc(c, &c->cell_faces, 6 * N));
    MC_TRY(dalloc(c, &c->fc0, F)); MC_TRY(dalloc(c, &c->fc1, F));
    MC_TRY(dalloc(c, &c->face_ref, F)); MC_TRY(dalloc(c, &c->face_pos, F));
    hipLaunchKernelGGL(k_iota, dim3(div_up(F, 256)), dim3(256), 0, c->stream, (int)F, c->face_ref);
    hipLaunchKernelGGL(k_iota, dim3(div_up(F, 256)), dim3(256), 0, c->stream, (int)F, c->face_pos);
    void *tmp = nullptr;
    const size_t tb = std::max<size_t>({(size_t)NP * 24, (size_t)F * 16, (size_t)N * 48});
    MC_HIP(c, hipMalloc(&tmp, tb));
    auto up = [&](const void *src, size_t bytes) { return hipMemcpyAsync(tmp, src, bytes, hipMemcpyHostToDevice, c->stream); };
    int rc = MC_OK;
    do {
        if (up(d->points, NP * 24) != hipSuccess) { rc = MC_ERR_HIP; break; }
        hipLaunchKernelGGL(k_split_points, dim3(div_up(NP, 256)), dim3(256), 0, c->stream, NP, (const double *)tmp, c->px,
                           c->py, c->pz);
        if (up(d->faces, F * 16) != hipSuccess) { rc = MC_ERR_HIP; break; }
        hipLaunchKernelGGL(k_split_faces, dim3(div_up(F, 256)), dim3(256), 0, c->stream, F, (const uint64_t *)tmp, c->fc0,
                           c->fc1);
        if (up(d->cell_faces, N * 48) != hipSuccess) { rc = MC_ERR_HIP; break; }
        hipLaunchKernelGGL(k_from_u64, dim3(div_up(6 * N, 256)), dim3(256), 0, c->stream, 6 * N, (const uint64_t *)tmp,
                           c->cell_faces, (int64_t)0);
        if (up(d->cell_neighbours, N * 48) != hipSuccess) { rc = MC_ERR_HIP; break; }
        hipLaunchKernelGGL(k_from_u64, dim3(div_up(6 * N, 256)), dim3(256), 0, c->stream, 6 * N, (const uint64_t *)tmp,
                           c->nbr, (int64_t)0);
        if (hipMemcpyAsync(c->cells, d->cells, N * 64, hipMemcpyHostToDevice, c->stream) != hipSuccess) { rc = MC_ERR_HIP; break; }
        if (hipStreamSynchronize(c->stream) != hipSuccess) { rc = MC_ERR_HIP; break; }
    } while (0);
    (void)hipFree(tmp);
    if (rc) { c->err = "mesh upload failed"; return rc; }
    MC_HIP(c, hipMemsetAsync(c->cells_per_point, 0, NP, c->stream));
    return MC_OK;
}

// Initial fields and boundary slots of the FlowSolver constructor
// (FlowSolver.hpp:301-485, values Mesh.hpp:114-120); used at construction and
// by mc_reset_fields.
int init_fields(Ctx *c)
{
    const int N = c->N;
    double *const *dptr = c->field_ptrs;
    MC_HIP(c, hipStreamSynchronize(c->stream));   // nothing in flight touches the fields
    const Consts &k = c->k;
    InitVals iv;
    iv.v[U] = k.gas_vel_x; iv.v[V] = k.gas_vel_y; iv.v[W] = k.gas_vel_z; iv.v[P] = k.gas_pre; iv.v[PP] = 0.0;
    iv.v[TE] = k.gas_te; iv.v[ED] = k.gas_ed; iv.v[TP] = 0.0; iv.v[TEM] = k.gas_tem; iv.v[FUL] = k.gas_ful;
    iv.v[PRO] = k.gas_pro; iv.v[VARF] = k.gas_ful; iv.v[VARP] = k.gas_pro;
    hipLaunchKernelGGL(k_init_fields, dim3(div_up(N, 256)), dim3(256), 0, c->stream, N, iv, dptr, dptr + NFIELDS,
                       dptr + 2 * NFIELDS, dptr + 3 * NFIELDS, c->AU, c->AV, c->AW, c->SU, c->SV, c->SW);
    MC_LAUNCH_CHECK(c);
    // boundary slots, FlowSolver.hpp:402-485 (host values, 6 per field)
    {
        double slot[NFIELDS][6];
        for (int s = 0; s < 6; s++) {
            const int bt = c->btype[s];
            double v[NFIELDS] = {0.0, 0.0, 0.0, k.gas_pre, 0.0, k.gas_te, k.gas_ed, 0.0, k.gas_tem, k.gas_ful, k.gas_pro,
                                 k.gas_ful, k.gas_pro};
            if (bt == INLET) {
                v[U] = k.gas_vel_x; v[V] = 0.0; v[W] = 0.0; v[P] = k.gas_pre; v[PP] = 0.0; v[TE] = k.inlet_te;
                v[ED] = k.inlet_ed; v[TP] = 0.0; v[TEM] = k.gas_tem; v[FUL] = k.gas_ful; v[PRO] = 0.0;
                v[VARF] = k.gas_ful; v[VARP] = 0.0;
            } else if (bt == OUTLET) {
                v[U] = k.gas_vel_x; v[V] = 0.0; v[W] = 0.0;
            } else if (bt == WALL) {
                v[U] = 0.0; v[V] = 0.0; v[W] = 0.0; v[FUL] = 0.0; v[PRO] = 0.0; v[VARF] = 0.0; v[VARP] = 0.0;
            }
            for (int f = 0; f < NFIELDS; f++) slot[f][s] = v[f];
        }
        static double hslot[NFIELDS][6];
        memcpy(hslot, slot, sizeof(slot));
        double zeros[18] = {0};
        for (int f = 0; f < NFIELDS; f++) {
            MC_HIP(c, hipMemcpy(c->phi[f] + N, hslot[f], 6 * sizeof(double), hipMemcpyHostToDevice));
            MC_HIP(c, hipMemcpy(c->gx[f] + N, zeros, 6 * sizeof(double), hipMemcpyHostToDevice));
            MC_HIP(c, hipMemcpy(c->gy[f] + N, zeros, 6 * sizeof(double), hipMemcpyHostToDevice));
            MC_HIP(c, hipMemcpy(c->gz[f] + N, zeros, 6 * sizeof(double), hipMemcpyHostToDevice));
        }
        for (double *a : {c->AU, c->AV, c->AW, c->SU, c->SV, c->SW})
            MC_HIP(c, hipMemcpy(a + N, zeros, 6 * sizeof(double), hipMemcpyHostToDevice));
    }

    return MC_OK;
}

int mesh_finish(Ctx *c)
{
    const int N = c->N, F = c->F;
    const int NS = N + 6;
    int *d_btype = nullptr;
    MC_TRY(dalloc(c, &d_btype, 6));
    MC_HIP(c, hipMemcpyAsync(d_btype, c->btype, sizeof(c->btype), hipMemcpyHostToDevice, c->stream));
    // geometry
    MC_TRY(dalloc(c, &c->cx, N)); MC_TRY(dalloc(c, &c->cy, N)); MC_TRY(dalloc(c, &c->cz, N));
    MC_TRY(dalloc(c, &c->vol, N)); MC_TRY(dalloc(c, &c->rho, N));
    MC_TRY(dalloc(c, &c->fcx, F)); MC_TRY(dalloc(c, &c->fcy, F)); MC_TRY(dalloc(c, &c->fcz, F));
    MC_TRY(dalloc(c, &c->fnx, F)); MC_TRY(dalloc(c, &c->fny, F)); MC_TRY(dalloc(c, &c->fnz, F));
    MC_TRY(dalloc(c, &c->farea, F)); MC_TRY(dalloc(c, &c->flam, F)); MC_TRY(dalloc(c, &c->frlc, F));
    MC_TRY(dalloc(c, &c->ff0, F)); MC_TRY(dalloc(c, &c->ff1, F)); MC_TRY(dalloc(c, &c->mflux, F));
    MC_TRY(dalloc(c, &c->fcorr, F));
    hipLaunchKernelGGL(k_cell_geom, dim3(div_up(N, 256)), dim3(256), 0, c->stream, N, c->cells, c->px, c->py, c->pz,
                       c->cx, c->cy, c->cz, c->vol, c->rho);
    MC_LAUNCH_CHECK(c);
    hipLaunchKernelGGL(k_face_geom, dim3(div_up(F, 256)), dim3(256), 0, c->stream, F, N, c->fc0, c->fc1, c->cells,
                       c->px, c->py, c->pz, c->cx, c->cy, c->cz, d_btype, c->farea, c->flam, c->frlc, c->fnx, c->fny,
                       c->fnz, c->fcx, c->fcy, c->fcz, c->mflux);
    MC_LAUNCH_CHECK(c);
    MC_HIP(c, hipMemsetAsync(c->ff0, 0, (size_t)F * 8, c->stream));
    MC_HIP(c, hipMemsetAsync(c->ff1, 0, (size_t)F * 8, c->stream));

    // fields
    MemScope ms_fields(c, MC_MEM_FIELDS);
    for (int f = 0; f < NFIELDS; f++) {
        MC_TRY(dalloc(c, &c->phi[f], NS));
        MC_TRY(dalloc(c, &c->gx[f], NS)); MC_TRY(dalloc(c, &c->gy[f], NS)); MC_TRY(dalloc(c, &c->gz[f], NS));
    }
    MC_TRY(dalloc(c, &c->AU, NS)); MC_TRY(dalloc(c, &c->AV, NS)); MC_TRY(dalloc(c, &c->AW, NS));
    MC_TRY(dalloc(c, &c->SU, NS)); MC_TRY(dalloc(c, &c->SV, NS)); MC_TRY(dalloc(c, &c->SW, NS));
    MC_TRY(dalloc(c, &c->field_ptrs, 4 * NFIELDS));
    {
        double *hp[4 * NFIELDS];
        for (int f = 0; f < NFIELDS; f++) {
            hp[f] = c->phi[f]; hp[NFIELDS + f] = c->gx[f]; hp[2 * NFIELDS + f] = c->gy[f]; hp[3 * NFIELDS + f] = c->gz[f];
        }
        MC_HIP(c, hipMemcpyAsync(c->field_ptrs, hp, sizeof(hp), hipMemcpyHostToDevice, c->stream));
        MC_HIP(c, hipStreamSynchronize(c->stream));
    }
    MC_TRY(init_fields(c));

    // CSR pattern, face orders
    MemScope ms_csr(c, MC_MEM_MATRIX);
    int *cnt = nullptr;
    MC_TRY(dalloc(c, &c->row_ptr, N + 1));
    MC_TRY(dalloc(c, &c->diag_idx, N));
    MC_TRY(dalloc(c, &c->forder, 6 * (int64_t)N));
    MC_HIP(c, hipMalloc(&cnt, (size_t)(N + 1) * sizeof(int)));
    hipLaunchKernelGGL(k_row_count, dim3(div_up(N + 1, 256)), dim3(256), 0, c->stream, N, c->nbr, cnt);
    MC_LAUNCH_CHECK(c);
    size_t tmpb = 0;
    MC_HIP(c, hipcub::DeviceScan::ExclusiveSum(nullptr, tmpb, cnt, c->row_ptr, N + 1, c->stream));
    void *tmps = nullptr;
    MC_HIP(c, hipMalloc(&tmps, tmpb + 16));
    MC_HIP(c, hipcub::DeviceScan::ExclusiveSum(tmps, tmpb, cnt, c->row_ptr, N + 1, c->stream));
    int nnz = 0;
    MC_HIP(c, hipMemcpyAsync(&nnz, c->row_ptr + N, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    MC_HIP(c, hipStreamSynchronize(c->stream));
    MC_HIP(c, hipFree(tmps));
    MC_HIP(c, hipFree(cnt));
    c->nnz = nnz;
    MC_TRY(dalloc(c, &c->col, nnz));
    MC_TRY(dalloc(c, &c->A_vals, nnz));
    MC_TRY(dalloc(c, &c->PA_vals, nnz));
    MC_TRY(dalloc(c, &c->Pb, N));
    MC_TRY(dalloc(c, &c->u, N));
    MC_HIP(c, hipMemsetAsync(c->A_vals, 0, (size_t)nnz * 8, c->stream));
    MC_HIP(c, hipMemsetAsync(c->PA_vals, 0, (size_t)nnz * 8, c->stream));
    hipLaunchKernelGGL(k_row_fill, dim3(div_up(N, 256)), dim3(256), 0, c->stream, N, c->nbr, c->cell_faces, c->face_ref,
                       c->row_ptr, c->col, c->diag_idx, c->forder);
    MC_TRY(dalloc(c, &c->fother, 6 * (int64_t)N));
    hipLaunchKernelGGL(k_face_other, dim3(div_up(6 * (int64_t)N, 256)), dim3(256), 0, c->stream, N, c->forder, c->fc0,
                       c->fc1, c->fother);
    MC_LAUNCH_CHECK(c);

    // boundary faces, last face per side, wall cells
    uint8_t *flag = nullptr;
    int *d_last = nullptr, *d_cnt = nullptr;
    MC_HIP(c, hipMalloc(&flag, F));
    MC_HIP(c, hipMalloc(&d_last, 6 * sizeof(int)));
    MC_HIP(c, hipMalloc(&d_cnt, sizeof(int)));
    int neg[6] = {-1, -1, -1, -1, -1, -1};
    MC_HIP(c, hipMemcpyAsync(d_last, neg, sizeof(neg), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_boundary_flags, dim3(div_up(F, 256)), dim3(256), 0, c->stream, F, N, c->fc1, c->face_pos, flag,
                       d_last);
    MC_LAUNCH_CHECK(c);
    MC_TRY(dalloc(c, &c->bfaces, F));
    tmpb = 0;
    MC_HIP(c, hipcub::DeviceSelect::Flagged(nullptr, tmpb, c->face_pos, flag, c->bfaces, d_cnt, F, c->stream));
    MC_HIP(c, hipMalloc(&tmps, tmpb + 16));
    MC_HIP(c, hipcub::DeviceSelect::Flagged(tmps, tmpb, c->face_pos, flag, c->bfaces, d_cnt, F, c->stream));
    int last_ref[6];
    MC_HIP(c, hipMemcpyAsync(&c->Fb, d_cnt, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    MC_HIP(c, hipMemcpyAsync(last_ref, d_last, sizeof(neg), hipMemcpyDeviceToHost, c->stream));
    MC_HIP(c, hipStreamSynchronize(c->stream));
    for (int s = 0; s < 6; s++) {
        c->last_face[s] = -1;
        c->last_cell[s] = -1;
        if (last_ref[s] >= 0) {
            MC_HIP(c, hipMemcpy(&c->last_face[s], c->face_pos + last_ref[s], sizeof(int), hipMemcpyDeviceToHost));
            MC_HIP(c, hipMemcpy(&c->last_cell[s], c->fc0 + c->last_face[s], sizeof(int), hipMemcpyDeviceToHost));
        }
    }
    MC_TRY(dalloc(c, &c->is_wall_cell, N));
    MC_HIP(c, hipMemsetAsync(c->is_wall_cell, 0, N, c->stream));
    if (c->Fb > 0) {
        hipLaunchKernelGGL(k_wall_cells, dim3(div_up(c->Fb, 256)), dim3(256), 0, c->stream, c->Fb, N, c->bfaces, c->fc0,
                           c->fc1, d_btype, c->is_wall_cell);
        MC_LAUNCH_CHECK(c);
    }
    MC_HIP(c, hipStreamSynchronize(c->stream));
    MC_HIP(c, hipFree(tmps));
    MC_HIP(c, hipFree(flag));
    MC_HIP(c, hipFree(d_last));
    MC_HIP(c, hipFree(d_cnt));
    (void)CFVM;
    return MC_OK;
}

}  // namespace mc
