// mc_mpnn.hip -- learned initial guess for the pressure-correction solve: a
// message-passing net over the cell graph, trained and run on the GPU.
//
// No counterpart in the reference (its pressure solves start from zero,
// FlowSolver.inl:2203-2273).  Definition (H = 64, 3 layers):
//
//   features f_i (4): (b_i - mean b) / s_b, A_ii / s_d, V_i / mean V, q_i / s_q
//     b: pressure-correction RHS, A: its matrix, q: net mass outflow of the
//     cell, s_b = rms(b - mean b), s_q = rms(q), s_d = mean(A_ii); edge feature e_is =
//     A_{i, nbr(i,s)} / s_d for the six sides s (0 where there is no neighbour)
//   h = relu(f Wenc + benc)
//   per layer:  [P | Q] = h [Wa | Wb]
//               s_i = sum_j relu(P_i + Q_j + e_ij we + b1)
//               a_i = s_i W2 + deg_i b2
//               u_i = relu([h_i, a_i] Wn1 + bn1)
//               h_i = h_i + u_i Wn2 + bn2
//   y = h wdec + bdec = the guess g, up to scale (the solve applies the
//   Galerkin factor); training target: mean-zero solution / its rms
//
// The edge MLP's second linear layer is applied after the neighbour sum
// (sum_j W2 r_ij = W2 sum_j r_ij), so every matrix product is node-level: an
// N x K by K x M product on MFMA (v_mfma_f32_16x16x16_bf16: BF16 operands,
// FP32 accumulation).  Edge work is a 64-channel gather; the backward pass
// turns the scatters into gathers through the reverse-slot map (the graph is
// symmetric), so nothing uses atomics.  Weight gradients are block partial
// products reduced in a fixed order.
//
// Parameters: one FP32 vector in the order Wenc(4x64), benc, then per layer
// Wa, Wb (64x64), we, b1 (64), W2 (64x64), b2, Wn1 (128x64), bn1, Wn2 (64x64),
// bn2, then wdec (64), bdec (1); matrices row-major (in x out).  Training is
// Adam on FP32 master weights; BF16 copies feed the MFMA products.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "mc_internal.hpp"

namespace mc {

constexpr int MH = 64;          // hidden width
constexpr int ML = 3;           // message-passing layers
constexpr int MFIN = 4;         // node features
constexpr int LSZ = 6 * MH * MH + 5 * MH;   // parameters per layer
constexpr int OFF_WENC = 0, OFF_BENC = MFIN * MH, OFF_L0 = OFF_BENC + MH;
constexpr int OFF_WDEC = OFF_L0 + ML * LSZ, OFF_BDEC = OFF_WDEC + MH;
constexpr int NPARAMS = OFF_BDEC + 1;
// within a layer
constexpr int L_WA = 0, L_WB = MH * MH, L_WE = 2 * MH * MH, L_B1 = L_WE + MH, L_W2 = L_B1 + MH,
              L_B2 = L_W2 + MH * MH, L_WN1 = L_B2 + MH, L_BN1 = L_WN1 + 2 * MH * MH, L_WN2 = L_BN1 + MH,
              L_BN2 = L_WN2 + MH * MH;
static_assert(L_BN2 + MH == LSZ, "layer layout");
// BF16 copies per layer: products need B^T (KP-contiguous rows)
constexpr int P_ABT = 0, P_W2T = P_ABT + 2 * MH * MH, P_N1T = P_W2T + MH * MH, P_N2T = P_N1T + 2 * MH * MH,
              P_ABR = P_N2T + MH * MH, P_W2R = P_ABR + 2 * MH * MH, P_N1R = P_W2R + MH * MH,
              P_N2R = P_N1R + 2 * MH * MH, PSZ = P_N2R + MH * MH;
constexpr int P_ENCT = ML * PSZ, PTOTAL = P_ENCT + MH * 16;

constexpr int TN_ROWS = 1024;   // rows per block of a weight-gradient product

typedef unsigned short bf16_t;
typedef short short4v __attribute__((ext_vector_type(4)));
typedef float float4v __attribute__((ext_vector_type(4)));

struct Mpnn {
    int n = 0;
    bool ready = false;           // weights loaded
    bool train_bufs = false;
    float *w = nullptr, *g = nullptr, *am = nullptr, *av = nullptr;   // master, grad, Adam moments
    int64_t adam_t = 0;
    bf16_t *wb = nullptr;         // BF16 copies
    bf16_t *wf = nullptr;         // BF16 A fragments of the fused inference layers
    bf16_t *hb[2] = {};           // inference activations (BF16, fragment order)
    int cus = 0;                  // compute units of the device
    int *nbr = nullptr, *rev = nullptr, *pmap = nullptr;
    float *deg = nullptr;
    float *f = nullptr, *e = nullptr, *y = nullptr, *dy = nullptr;
    float *H[ML + 1] = {}, *PQ[ML] = {}, *S[ML] = {}, *A[ML] = {}, *U[ML] = {};
    float *dH = nullptr, *dH2 = nullptr, *dU = nullptr, *dHA = nullptr, *dS = nullptr, *dPQ = nullptr;
    float *partial = nullptr;     // block partial products
    int64_t partial_cap = 0;
    double *sc = nullptr;         // device scalars: mean b, s_b, s_d, mean V, loss, ...
    // data set
    int64_t cap = 0, count = 0;
    float *ds_f = nullptr, *ds_e = nullptr, *ds_t = nullptr;
};

static Mpnn *M(Ctx *c) { return (Mpnn *)c->mpnn; }

// ------------------------------------------------------------- helpers --
__device__ __forceinline__ bf16_t to_bf16(float x)
{
    unsigned u = __float_as_uint(x);
    u = (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
    return (bf16_t)u;
}
__device__ __forceinline__ float from_bf16(bf16_t b) { return __uint_as_float(((unsigned)b) << 16); }

// ------------------------------------------------------ weight packing --
// dst[m * kp + k] = bf16(src[k * cols + m]) for k < rows, else 0 (B^T of a
// row-major rows x cols matrix, rows padded to kp)
__global__ void k_pack_t(bf16_t *dst, int kp, const float *src, int rows, int cols)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cols * kp) return;
    const int m = i / kp, k = i % kp;
    dst[i] = (k < rows) ? to_bf16(src[k * cols + m]) : (bf16_t)0;
}
// dst[r * ld + koff + k] = bf16(src[r * cols + k])  (row-major copy into a column block)
__global__ void k_pack_rows(bf16_t *dst, int ld, int koff, const float *src, int rows, int cols)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows * cols) return;
    const int r = i / cols, k = i % cols;
    dst[r * ld + koff + k] = to_bf16(src[i]);
}

// ----------------------------------------------------- MFMA products --
enum { EPI_BIAS = 1, EPI_RELU = 2, EPI_RESID = 4, EPI_MASK = 8, EPI_ROWSCALE = 16 };

// Y[n x M] = epi([X1 | X2] B): X1 n x K1 (leading dim ld1), X2 n x K2 (ld2),
// B given as Bt[m * KP + k] (BF16).  One wave computes 16 rows x M columns
// with M/16 accumulators of v_mfma_f32_16x16x16_bf16:
//   A fragment: lane l holds A[l % 16][4 (l / 16) + t], t = 0..3
//   B fragment: lane l holds B[4 (l / 16) + t][l % 16]
//   result:     lane l holds C[4 (l / 16) + t][l % 16]
// Epilogue order: v = acc (+ R) (+ bias * rowscale) (relu) (zero where mask <= 0).
template <int K1, int K2, int KP, int MM, int EPI>
__global__ void __launch_bounds__(256) k_mm(int n, const float *X1, int ld1, const float *X2, int ld2,
                                           const bf16_t *Bt, const float *bias, const float *rowscale,
                                           const float *R, int ldr, const float *mask, int ldm, float *Y, int ldy)
{
    static_assert(KP % 16 == 0 && MM % 16 == 0 && K1 % 4 == 0 && K2 % 4 == 0 && K1 + K2 <= KP, "shape");
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row0 = (blockIdx.x * 4 + wave) * 16;
    if (row0 >= n) return;
    const int ar = row0 + (lane & 15);
    const int kq = 4 * (lane >> 4);
    float4v acc[MM / 16];
#pragma unroll
    for (int j = 0; j < MM / 16; j++) acc[j] = float4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < KP / 16; kk++) {
        const int k = kk * 16 + kq;
        float4 xv = make_float4(0.f, 0.f, 0.f, 0.f);
        if (ar < n) {
            if (k < K1) xv = *(const float4 *)(X1 + (size_t)ar * ld1 + k);
            else if (k < K1 + K2) xv = *(const float4 *)(X2 + (size_t)ar * ld2 + (k - K1));
        }
        short4v a;
        a[0] = (short)to_bf16(xv.x); a[1] = (short)to_bf16(xv.y);
        a[2] = (short)to_bf16(xv.z); a[3] = (short)to_bf16(xv.w);
#pragma unroll
        for (int j = 0; j < MM / 16; j++) {
            const short4v b = *(const short4v *)(Bt + (size_t)(j * 16 + (lane & 15)) * KP + k);
            acc[j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[j], 0, 0, 0);
        }
    }
#pragma unroll
    for (int j = 0; j < MM / 16; j++) {
        const int col = j * 16 + (lane & 15);
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int row = row0 + kq + t;
            if (row >= n) continue;
            float v = acc[j][t];
            if (EPI & EPI_RESID) v = R[(size_t)row * ldr + col] + v;
            if (EPI & EPI_BIAS) v += (EPI & EPI_ROWSCALE) ? bias[col] * rowscale[row] : bias[col];
            if (EPI & EPI_RELU) v = fmaxf(v, 0.f);
            if (EPI & EPI_MASK) v = (mask[(size_t)row * ldm + col] > 0.f) ? v : 0.f;
            Y[(size_t)row * ldy + col] = v;
        }
    }
}

// Block partial of G = [X1 | X2]^T D (KP x MM) over rows [b*TN_ROWS, ...),
// plus row KP = column sums of D (weighted by rowscale when given):
// partial[b][(KP + 1) * MM].  MFMA with the row index as the reduction
// dimension: A[k][r] = X[r][k], B[r][m] = D[r][m].
template <int K1, int K2, int KP, int MM>
__global__ void __launch_bounds__(256) k_mm_tn(int n, const float *X1, int ld1, const float *X2, int ld2,
                                              const float *D, int ldd, const float *rowscale, float *partial)
{
    constexpr int TI = KP / 16, TJ = MM / 16, T = TI * TJ, TPW = (T + 3) / 4;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r0 = blockIdx.x * TN_ROWS, r1 = min(n, r0 + TN_ROWS);
    float *out = partial + (size_t)blockIdx.x * (KP + 1) * MM;
    float4v acc[TPW];
#pragma unroll
    for (int u = 0; u < TPW; u++) acc[u] = float4v{0.f, 0.f, 0.f, 0.f};
    const int kq = 4 * (lane >> 4);
    for (int rr = r0; rr < r1; rr += 16) {
#pragma unroll
        for (int u = 0; u < TPW; u++) {
            const int tile = wave + 4 * u;
            if (tile >= T) break;
            const int ti = tile / TJ, tj = tile % TJ;
            const int kcol = ti * 16 + (lane & 15), mcol = tj * 16 + (lane & 15);
            short4v a, b;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int r = rr + kq + t;
                float xv = 0.f, dv = 0.f;
                if (r < r1) {
                    if (kcol < K1) xv = X1[(size_t)r * ld1 + kcol];
                    else if (kcol < K1 + K2) xv = X2[(size_t)r * ld2 + (kcol - K1)];
                    dv = D[(size_t)r * ldd + mcol];
                }
                a[t] = (short)to_bf16(xv);
                b[t] = (short)to_bf16(dv);
            }
            acc[u] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[u], 0, 0, 0);
        }
    }
#pragma unroll
    for (int u = 0; u < TPW; u++) {
        const int tile = wave + 4 * u;
        if (tile >= T) break;
        const int ti = tile / TJ, tj = tile % TJ;
#pragma unroll
        for (int t = 0; t < 4; t++) out[(size_t)(ti * 16 + kq + t) * MM + tj * 16 + (lane & 15)] = acc[u][t];
    }
    // column sums in FP32
    __shared__ float cs[256];
    constexpr int PER = 256 / MM;
    const int col = threadIdx.x % MM, part = threadIdx.x / MM;
    float s = 0.f;
    for (int r = r0 + part; r < r1; r += PER) s += D[(size_t)r * ldd + col] * (rowscale ? rowscale[r] : 1.f);
    cs[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < MM) {
        float t = 0.f;
        for (int p = 0; p < PER; p++) t += cs[p * MM + threadIdx.x];
        out[(size_t)KP * MM + threadIdx.x] = t;
    }
}

// Sum the block partials (fixed order) into the gradient vector: element
// (k, m) of the KP x MM product goes to wsplit destinations (m < msplit:
// g[off1 + k * msplit + m], else g[off2 + k * msplit + m - msplit]) for k < K;
// the column-sum row goes to g[offb + m] when offb >= 0.
__global__ void k_tn_reduce(const float *partial, int nblocks, int KP, int MM, int K, int msplit, int off1, int off2,
                            int offb, float *g)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (KP + 1) * MM) return;
    const int k = i / MM, m = i % MM;
    if (k < KP && k >= K) return;
    if (k == KP && offb < 0) return;
    float s = 0.f;
    for (int b = 0; b < nblocks; b++) s += partial[(size_t)b * (KP + 1) * MM + i];
    if (k == KP) g[offb + m] = s;
    else if (m < msplit) g[off1 + k * msplit + m] = s;
    else g[off2 + k * msplit + (m - msplit)] = s;
}

// ------------------------------------------------------------ edges --
// s_i = sum_j relu(P_i + Q_j + e_ij we + b1); 16 threads per node, 4 channels each
__global__ void __launch_bounds__(256) k_edge_fwd(int n, const float *PQ, const int *nbr, const float *e,
                                                 const float *we, const float *b1, float *S)
{
    const int idx = blockIdx.x * 256 + threadIdx.x;
    const int i = idx >> 4, c4 = (idx & 15) * 4;
    if (i >= n) return;
    const float4 P = *(const float4 *)(PQ + (size_t)i * 2 * MH + c4);
    const float4 w = *(const float4 *)(we + c4), bb = *(const float4 *)(b1 + c4);
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int s = 0; s < 6; s++) {
        const int j = nbr[6 * i + s];
        if (j < 0) continue;
        const float4 Q = *(const float4 *)(PQ + (size_t)j * 2 * MH + MH + c4);
        const float es = e[6 * i + s];
        acc.x += fmaxf(((P.x + Q.x) + es * w.x) + bb.x, 0.f);
        acc.y += fmaxf(((P.y + Q.y) + es * w.y) + bb.y, 0.f);
        acc.z += fmaxf(((P.z + Q.z) + es * w.z) + bb.z, 0.f);
        acc.w += fmaxf(((P.w + Q.w) + es * w.w) + bb.w, 0.f);
    }
    *(float4 *)(S + (size_t)i * MH + c4) = acc;
}

// dP_i = sum_{j in nb(i)} ds_i [z_ij > 0],  dQ_i = sum_{k in nb(i)} ds_k [z_ki > 0]
// (z_ki uses k's slot rev pointing back to i); block partials of
// dwe = sum ds_i [z_ij > 0] e_ij and db1 = sum ds_i [z_ij > 0] (partial[b][128]).
__global__ void __launch_bounds__(256) k_edge_bwd(int n, const float *PQ, const int *nbr, const int *rev,
                                                 const float *e, const float *we, const float *b1, const float *dS,
                                                 float *dPQ, float *partial)
{
    const int idx = blockIdx.x * 256 + threadIdx.x;
    const int i = idx >> 4, c4 = (idx & 15) * 4;
    float4 gwe = make_float4(0.f, 0.f, 0.f, 0.f), gb1 = gwe;
    if (i < n) {
        const float4 P = *(const float4 *)(PQ + (size_t)i * 2 * MH + c4);
        const float4 Qi = *(const float4 *)(PQ + (size_t)i * 2 * MH + MH + c4);
        const float4 w = *(const float4 *)(we + c4), bb = *(const float4 *)(b1 + c4);
        const float4 di = *(const float4 *)(dS + (size_t)i * MH + c4);
        float4 dP = make_float4(0.f, 0.f, 0.f, 0.f), dQ = dP;
        for (int s = 0; s < 6; s++) {
            const int j = nbr[6 * i + s];
            if (j < 0) continue;
            const float es = e[6 * i + s];
            const float4 Qj = *(const float4 *)(PQ + (size_t)j * 2 * MH + MH + c4);
            float4 gz;
            gz.x = (((P.x + Qj.x) + es * w.x) + bb.x > 0.f) ? di.x : 0.f;
            gz.y = (((P.y + Qj.y) + es * w.y) + bb.y > 0.f) ? di.y : 0.f;
            gz.z = (((P.z + Qj.z) + es * w.z) + bb.z > 0.f) ? di.z : 0.f;
            gz.w = (((P.w + Qj.w) + es * w.w) + bb.w > 0.f) ? di.w : 0.f;
            dP.x += gz.x; dP.y += gz.y; dP.z += gz.z; dP.w += gz.w;
            gwe.x += gz.x * es; gwe.y += gz.y * es; gwe.z += gz.z * es; gwe.w += gz.w * es;
            gb1.x += gz.x; gb1.y += gz.y; gb1.z += gz.z; gb1.w += gz.w;
            // the message i -> j seen from j: z_ji = P_j + Q_i + e_{j, rev} we + b1
            const int rs = rev[6 * i + s];
            if (rs < 0) continue;   // j does not list i (not a symmetric graph)
            const float ej = e[6 * j + rs];
            const float4 Pj = *(const float4 *)(PQ + (size_t)j * 2 * MH + c4);
            const float4 dj = *(const float4 *)(dS + (size_t)j * MH + c4);
            dQ.x += (((Pj.x + Qi.x) + ej * w.x) + bb.x > 0.f) ? dj.x : 0.f;
            dQ.y += (((Pj.y + Qi.y) + ej * w.y) + bb.y > 0.f) ? dj.y : 0.f;
            dQ.z += (((Pj.z + Qi.z) + ej * w.z) + bb.z > 0.f) ? dj.z : 0.f;
            dQ.w += (((Pj.w + Qi.w) + ej * w.w) + bb.w > 0.f) ? dj.w : 0.f;
        }
        *(float4 *)(dPQ + (size_t)i * 2 * MH + c4) = dP;
        *(float4 *)(dPQ + (size_t)i * 2 * MH + MH + c4) = dQ;
    }
    // block partials over the block's 16 nodes (fixed order)
    __shared__ float4 sw[256], sb[256];
    sw[threadIdx.x] = gwe;
    sb[threadIdx.x] = gb1;
    __syncthreads();
    if (threadIdx.x < 16) {
        float4 a = make_float4(0.f, 0.f, 0.f, 0.f), b = a;
        for (int q = 0; q < 16; q++) {
            const float4 x = sw[q * 16 + threadIdx.x], y = sb[q * 16 + threadIdx.x];
            a.x += x.x; a.y += x.y; a.z += x.z; a.w += x.w;
            b.x += y.x; b.y += y.y; b.z += y.z; b.w += y.w;
        }
        float *out = partial + (size_t)blockIdx.x * 2 * MH;
        *(float4 *)(out + threadIdx.x * 4) = a;
        *(float4 *)(out + MH + threadIdx.x * 4) = b;
    }
}

// sum over blocks of the (we, b1) partials into the gradient
__global__ void k_edge_reduce(const float *partial, int nblocks, int off_we, int off_b1, float *g)
{
    const int i = threadIdx.x;   // 128 threads
    if (i >= 2 * MH) return;
    float s = 0.f;
    for (int b = 0; b < nblocks; b++) s += partial[(size_t)b * 2 * MH + i];
    g[(i < MH) ? off_we + i : off_b1 + i - MH] = s;
}

// ------------------------------------------------- decoder and loss --
// y = h wdec + bdec for 64 rows per workgroup: the rows are staged through LDS
// with coalesced loads, then each row is summed in column order
__global__ void __launch_bounds__(256) k_decode(int n, const float *Hf, const float *w, float *y)
{
    __shared__ float tile[64][MH + 1];
    const int r0 = blockIdx.x * 64;
    for (int q = threadIdx.x; q < 64 * (MH / 4); q += 256) {
        const int r = q / (MH / 4), c4 = 4 * (q % (MH / 4));
        const float4 v = (r0 + r < n) ? *(const float4 *)(Hf + (size_t)(r0 + r) * MH + c4)
                                      : make_float4(0.f, 0.f, 0.f, 0.f);
        tile[r][c4] = v.x; tile[r][c4 + 1] = v.y; tile[r][c4 + 2] = v.z; tile[r][c4 + 3] = v.w;
    }
    __syncthreads();
    const int i = r0 + (int)threadIdx.x;
    if (threadIdx.x >= 64 || i >= n) return;
    float s = 0.f;
    for (int c = 0; c < MH; c++)
        s += from_bf16(to_bf16(tile[threadIdx.x][c])) * from_bf16(to_bf16(w[OFF_WDEC + c]));
    y[i] = s + w[OFF_BDEC];
}

// dH = dy wdec^T ; block partials of (H^T dy, sum dy): partial[b][MH + 1]
__global__ void __launch_bounds__(256) k_decode_bwd(int n, const float *Hf, const float *w, const float *dy, float *dH,
                                                   float *partial)
{
    const int idx = blockIdx.x * 256 + threadIdx.x;
    const int i = idx >> 4, c4 = (idx & 15) * 4;
    float4 gw = make_float4(0.f, 0.f, 0.f, 0.f);
    float gb = 0.f;   // dy of the node, counted once (by its channel-0 thread)
    if (i < n) {
        const float d = dy[i];
        const float4 wd = *(const float4 *)(w + OFF_WDEC + c4);
        *(float4 *)(dH + (size_t)i * MH + c4) = make_float4(d * wd.x, d * wd.y, d * wd.z, d * wd.w);
        const float4 h = *(const float4 *)(Hf + (size_t)i * MH + c4);
        gw = make_float4(h.x * d, h.y * d, h.z * d, h.w * d);
        gb = (c4 == 0) ? d : 0.f;
    }
    __shared__ float4 s4[256];
    __shared__ float s1[256];
    s4[threadIdx.x] = gw;
    s1[threadIdx.x] = gb;
    __syncthreads();
    float *out = partial + (size_t)blockIdx.x * (MH + 1);
    if (threadIdx.x < 16) {
        float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
        for (int q = 0; q < 16; q++) {
            const float4 x = s4[q * 16 + threadIdx.x];
            a.x += x.x; a.y += x.y; a.z += x.z; a.w += x.w;
        }
        out[threadIdx.x * 4 + 0] = a.x; out[threadIdx.x * 4 + 1] = a.y;
        out[threadIdx.x * 4 + 2] = a.z; out[threadIdx.x * 4 + 3] = a.w;
    } else if (threadIdx.x == 16) {
        float t = 0.f;
        for (int q = 0; q < 16; q++) t += s1[q * 16];
        out[MH] = t;
    }
}

__global__ void k_decode_reduce(const float *partial, int nblocks, float *g)
{
    const int i = threadIdx.x;   // MH + 1 threads
    if (i > MH) return;
    float s = 0.f;
    for (int b = 0; b < nblocks; b++) s += partial[(size_t)b * (MH + 1) + i];
    g[(i < MH) ? OFF_WDEC + i : OFF_BDEC] = s;
}

// block partials of sum(y)
__global__ void __launch_bounds__(256) k_fsum(int n, const float *y, double *partials)
{
    double acc[1] = {0.0};
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) acc[0] += (double)y[i];
    block_partials<1>(acc, partials, blockIdx.x, gridDim.x);
}
struct EpiMeanTo {
    double *out;
    double inv_n;
    __device__ void operator()(const double (&v)[1]) const { *out = v[0] * inv_n; }
};
// d = y - mean(y) - t ; dy = 2 d / n ; partials of d^2
__global__ void __launch_bounds__(256) k_loss(int n, const float *y, const float *t, const double *ymean, float *dy,
                                             double *partials)
{
    double acc[1] = {0.0};
    const float ym = (float)*ymean;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const float d = (y[i] - ym) - t[i];
        dy[i] = (2.0f / (float)n) * d;
        acc[0] += (double)d * (double)d;
    }
    block_partials<1>(acc, partials, blockIdx.x, gridDim.x);
}

// residual add of a column block: dst[i][c] += src[i * lds + c]
__global__ void k_add_block(int n, float *dst, const float *src, int lds)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * MH) return;
    const int i = idx / MH, c = idx % MH;
    dst[idx] += src[(size_t)i * lds + c];
}

// Adam (Kingma & Ba), FP32 master weights
__global__ void k_adam(int np, float *w, const float *g, float *m, float *v, float lr, float b1, float b2, float eps,
                       float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ------------------------------------------------- fused inference --
// Inference runs one kernel per message-passing layer on BF16 activations
// (128 bytes a cell), with v_mfma_f32_16x16x32_bf16 (gfx950, K = 32) in the
// transposed form C = W^T X: weights are the A operand (rows = output
// channels), a tile of 16 cells the B operand (columns).  Lane l then holds,
// for output row tile j, channels 16 j + 4 g + t (g = l / 16, t < 4) of cell
// l % 16, and those 16 channels are exactly the ones it needs as its B
// fragment for the next product when the K dimension is taken in the order
// frag_ch below: every product chains into the next without moving data
// between lanes.  Activations are stored per cell in that fragment order, so
// a lane's B fragment (8 channels) is one 16-byte load.
//
// The edge term Q_j = h_j Wb is recomputed for each of a cell's six
// neighbours from the neighbour's stored activations (six 16 x 64 x 64
// products per tile) instead of being written and gathered: a layer reads
// each cell's 128 bytes, the neighbours' (from L2), its graph row, and
// writes 128 bytes.  Cells go in the assembly kernels' brick order, each
// XCD's workgroups over one contiguous range of bricks, so a neighbour's
// activations were read shortly before by the same L2.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// channel at fragment position (ks, g, t) of a 64-channel vector
__host__ __device__ constexpr int frag_ch(int ks, int g, int t) { return 16 * (2 * ks + (t >> 2)) + 4 * g + (t & 3); }

// packed A fragments of one layer (BF16 elements): W^T of Wa, Wb, W2 (K = 64),
// Wn1 (K = 128: h part, then a part), Wn2; fragment (j, ks) of lane l at
// ((j * KS + ks) * 64 + l) * 8
constexpr int FR_WA = 0, FR_WB = 4096, FR_W2 = 8192, FR_N1 = 12288, FR_N2 = 20480, FR_SZ = 24576;

// dst[((j * KS + ks) * 64 + l) * 8 + t] = bf16(W[64 (ks / 2) + frag_ch(ks % 2, l / 16, t)][16 j + l % 16])
__global__ void k_pack_frag(bf16_t *dst, const float *W, int KS)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 4 * KS * 64 * 8) return;
    const int t = i & 7, l = (i >> 3) & 63, q = i >> 9, ks = q % KS, j = q / KS;
    const int kin = 64 * (ks >> 1) + frag_ch(ks & 1, l >> 4, t);
    dst[i] = to_bf16(W[kin * MH + 16 * j + (l & 15)]);
}

__device__ __forceinline__ bf16x8 ld_frag(const bf16_t *h, int cell, int ks, int g)
{
    return ((const bf16x8 *)(h + (size_t)cell * MH))[ks * 4 + g];
}

__device__ __forceinline__ bf16x8 zero_frag()
{
    bf16x8 z;
#pragma unroll
    for (int t = 0; t < 8; t++) z[t] = (__bf16)0.0f;
    return z;
}

// gfx950 converts in hardware (v_cvt_pk_bf16_f32, round to nearest even, two
// values per instruction): the same bits as to_bf16 for every finite value
__device__ __forceinline__ __bf16 bf(float x) { return (__bf16)x; }
__device__ __forceinline__ float fl(__bf16 b) { return (float)b; }

// C layout (4 row tiles) -> the two B fragments of the next product
__device__ __forceinline__ void to_frags(const float4v (&c)[4], bf16x8 (&b)[2])
{
#pragma unroll
    for (int ks = 0; ks < 2; ks++)
#pragma unroll
        for (int t = 0; t < 8; t++) b[ks][t] = bf(c[2 * ks + (t >> 2)][t & 3]);
}

// out[j] += sum_ks A(j, ks) B(ks), A fragments from LDS (out holds the
// starting values: zero or a bias, so the epilogue add rides in the MFMA)
template <int KS>
__device__ __forceinline__ void mm_lds(const bf16x8 *A, const bf16x8 (&B)[KS], float4v (&out)[4], int lane)
{
#pragma unroll
    for (int j = 0; j < 4; j++) {
#pragma unroll
        for (int ks = 0; ks < KS; ks++)
            out[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[(j * KS + ks) * 64 + lane], B[ks], out[j], 0, 0, 0);
    }
}
// the 16 values of a 64-vector v that lane group g holds in C layout
__device__ __forceinline__ void c_layout(const float *v, int g, float4v (&out)[4])
{
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
        for (int t = 0; t < 4; t++) out[j][t] = v[16 * j + 4 * g + t];
}

// Wenc^T as the A fragments of one 16x16x32 product per output row tile j:
// dst[(j * 64 + l) * 8 + t] = bf16(Wenc[t][16 j + l % 16]) for l < 16, t < 4
// (the features are the first 4 of the 32 K rows), else 0
__global__ void k_pack_enc(bf16_t *dst, const float *W)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 4 * 64 * 8) return;
    const int t = i & 7, l = (i >> 3) & 63, j = i >> 9;
    dst[i] = (l < 16 && t < MFIN) ? to_bf16(W[OFF_WENC + t * MH + 16 * j + (l & 15)]) : (bf16_t)0;
}

// h0 = relu(f Wenc + benc) in fragment order: one wave per 16 cells, one
// product per output row tile (the 4 features as the K rows 0..3 of lane
// group 0, the other 28 zero), Wenc^T from k_pack_enc, benc as the start.
__global__ void __launch_bounds__(256) k_mp_encode(int n, const float *f, const bf16_t *wenc, const float *W,
                                                  bf16_t *h)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    const int cell = (blockIdx.x * 4 + wave) * 16 + r;
    const bool valid = cell < n;
    bf16x8 fb = zero_frag();
    if (valid && g == 0) {
        const float4 v = *(const float4 *)(f + (size_t)4 * cell);
        fb[0] = bf(v.x); fb[1] = bf(v.y); fb[2] = bf(v.z); fb[3] = bf(v.w);
    }
    float4v H[4];
    c_layout(W + OFF_BENC, g, H);
#pragma unroll
    for (int j = 0; j < 4; j++) {
        H[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(((const bf16x8 *)wenc)[j * 64 + lane], fb, H[j], 0, 0, 0);
#pragma unroll
        for (int t = 0; t < 4; t++) H[j][t] = fmaxf(H[j][t], 0.f);
    }
    bf16x8 o[2];
    to_frags(H, o);
    if (valid) {
#pragma unroll
        for (int ks = 0; ks < 2; ks++)
            __builtin_nontemporal_store(o[ks], (bf16x8 *)(h + (size_t)cell * MH) + ks * 4 + g);
    }
}

// One message-passing layer for every cell (LAST: the decoder too, y only).
// Persistent: workgroup b serves the bricks of XCD b % 8's contiguous range
// (the grid is a multiple of 8); its eight waves take the brick's 16 tiles of
// 16 cells in turn.  Two resident workgroups of 512 threads per compute unit
// (four waves per SIMD: at most 128 registers; 50 KB of LDS each); three of
// 256 threads (three waves per SIMD) measured 2.5% slower, and holding Wb's
// fragments in registers (two waves per SIMD) 25% slower.
constexpr int MP_THREADS = 512, MP_LB = 2;
template <bool LAST>
__global__ void __launch_bounds__(MP_THREADS, MP_LB) k_mp_layer(CellWalk w, int nbricks, const bf16_t *hin, bf16_t *hout,
                                                    const int *nbr, const float *e, const float *deg,
                                                    const bf16_t *wf, const float *Wl, const float *Wall, float *y)
{
    __shared__ bf16x8 sW[FR_SZ / 8];
    __shared__ float sB[6][MH];   // we, b1, b2, bn1, bn2, bf16(wdec)
    for (int i = threadIdx.x; i < FR_SZ / 8; i += MP_THREADS) sW[i] = ((const bf16x8 *)wf)[i];
    for (int i = threadIdx.x; i < MH; i += MP_THREADS) {
        sB[0][i] = Wl[L_WE + i]; sB[1][i] = Wl[L_B1 + i]; sB[2][i] = Wl[L_B2 + i];
        sB[3][i] = Wl[L_BN1 + i]; sB[4][i] = Wl[L_BN2 + i];
        sB[5][i] = from_bf16(to_bf16(Wall[OFF_WDEC + i]));
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    __syncthreads();
    const int nxcd = 8, xcd = blockIdx.x % nxcd, local = blockIdx.x / nxcd, nloc = gridDim.x / nxcd;
    const int per = (nbricks + nxcd - 1) / nxcd;
    const int bend = min(nbricks, (xcd + 1) * per);
    for (int bi = xcd * per + local; bi < bend; bi += nloc) {
        for (int tile = wave; tile < 16; tile += MP_THREADS / 64) {
            const int cell = walk_cell_at(w, bi, tile * 16 + r);
            const bool valid = cell >= 0;
            // an offset the compiler cannot see through: the weight fragments
            // are loop-invariant LDS loads, and hoisted out of the tile loop
            // they would hold 352 registers; read at each use instead
            int z0 = 0;
            asm volatile("" : "+s"(z0));
            const bf16x8 *fW = sW + z0;
            const float(*vB)[MH] = (const float(*)[MH])(&sB[0][0] + z0);
            // a cell's stored activations as B fragments (zeros where cc < 0)
            auto input = [&](int cc, bf16x8 (&o)[2]) {
#pragma unroll
                for (int ks = 0; ks < 2; ks++) o[ks] = cc >= 0 ? ld_frag(hin, cc, ks, g) : zero_frag();
            };
            bf16x8 hb[2];
            input(cell, hb);
            // P = h Wa + b1 (b1 as the accumulator's start)
            float4v P[4];
            c_layout(vB[1], g, P);
            mm_lds<2>(fW + FR_WA / 8, hb, P, lane);
            float4v we[4];
            c_layout(vB[0], g, we);
            // s = sum over neighbours of relu(P + Q_j + e we); the six
            // neighbours' gathers are all in flight before the first product
            float4v S[4];
#pragma unroll
            for (int j = 0; j < 4; j++) S[j] = float4v{0.f, 0.f, 0.f, 0.f};
            int nb[6];
            float es[6];
            bf16x8 hn[6][2];
#pragma unroll
            for (int q = 0; q < 6; q++) {
                nb[q] = valid ? nbr[6 * cell + q] : -1;
                es[q] = valid ? e[6 * cell + q] : 0.f;
            }
#pragma unroll
            for (int q = 0; q < 6; q++) input(nb[q], hn[q]);
            // output row tile by row tile: the tile's two Wb fragments are
            // read from LDS once for all six neighbours (not once per
            // neighbour); each Q_j starts from P and the MFMAs add h_j Wb in
            // the same order, and every element of S adds its six terms in
            // neighbour order, so the bits are those of neighbour-major order
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const bf16x8 wb0 = fW[FR_WB / 8 + (2 * j) * 64 + lane];
                const bf16x8 wb1 = fW[FR_WB / 8 + (2 * j + 1) * 64 + lane];
#pragma unroll
                for (int q = 0; q < 6; q++) {
                    float4v Q = P[j];
                    Q = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wb0, hn[q][0], Q, 0, 0, 0);
                    Q = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wb1, hn[q][1], Q, 0, 0, 0);
                    if (nb[q] >= 0) {
#pragma unroll
                        for (int t = 0; t < 4; t++) S[j][t] += fmaxf(Q[t] + es[q] * we[j][t], 0.f);
                    }
                }
            }
            // a = s W2 + deg b2
            bf16x8 sb[2];
            to_frags(S, sb);
            float4v A[4];
            c_layout(vB[2], g, A);
            const float dg = valid ? deg[cell] : 0.f;
#pragma unroll
            for (int j = 0; j < 4; j++) A[j] *= dg;   // deg b2 as the start
            mm_lds<2>(fW + FR_W2 / 8, sb, A, lane);
            // u = relu([h, a] Wn1 + bn1)
            bf16x8 ha[4];
            ha[0] = hb[0]; ha[1] = hb[1];
            {
                bf16x8 ab[2];
                to_frags(A, ab);
                ha[2] = ab[0]; ha[3] = ab[1];
            }
            float4v U[4];
            c_layout(vB[3], g, U);   // bn1 as the start
            mm_lds<4>(fW + FR_N1 / 8, ha, U, lane);
#pragma unroll
            for (int j = 0; j < 4; j++)
#pragma unroll
                for (int t = 0; t < 4; t++) U[j][t] = fmaxf(U[j][t], 0.f);
            // h' = h + u Wn2 + bn2
            bf16x8 ub[2];
            to_frags(U, ub);
            // h' = h + u Wn2 + bn2: the accumulator starts from h + bn2
            float4v O[4];
            c_layout(vB[4], g, O);
#pragma unroll
            for (int j = 0; j < 4; j++)
#pragma unroll
                for (int t = 0; t < 4; t++) O[j][t] += fl(hb[j >> 1][4 * (j & 1) + t]);
            mm_lds<2>(fW + FR_N2 / 8, ub, O, lane);
            if constexpr (LAST) {
                // y = bf16(h') . bf16(wdec) + bdec, the four lane groups' partial sums added
                float part = 0.f;
#pragma unroll
                for (int j = 0; j < 4; j++)
#pragma unroll
                    for (int t = 0; t < 4; t++) part += fl(bf(O[j][t])) * vB[5][16 * j + 4 * g + t];
                part += __shfl_xor(part, 16, 64);
                part += __shfl_xor(part, 32, 64);
                if (valid && g == 0) y[cell] = part + Wall[OFF_BDEC];
            } else {
                bf16x8 ho[2];
                to_frags(O, ho);
                if (valid) {
#pragma unroll
                    for (int ks = 0; ks < 2; ks++)   // streaming: read by the next launch only
                        __builtin_nontemporal_store(ho[ks], (bf16x8 *)(hout + (size_t)cell * MH) + ks * 4 + g);
                }
            }
        }
    }
}

// ------------------------------------------------------- graph set-up --
__global__ void k_graph(int N, const int *nbr6, int *nbr, int *rev, float *deg)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int d = 0;
    for (int s = 0; s < 6; s++) {
        const int j = nbr6[6 * i + s];
        nbr[6 * i + s] = j;
        int r = -1;
        if (j >= 0) {
            d++;
            for (int t = 0; t < 6; t++)
                if (nbr6[6 * j + t] == i) { r = t; break; }
        }
        rev[6 * i + s] = r;
    }
    deg[i] = (float)d;
}
// CSR index of (i, nbr(i, s)) in the pressure pattern
__global__ void k_pmap(int N, const int *nbr, const int *row_ptr, const int *col, int *pmap)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    for (int s = 0; s < 6; s++) {
        const int j = nbr[6 * i + s];
        int p = -1;
        if (j >= 0)
            for (int q = row_ptr[i]; q < row_ptr[i + 1]; q++)
                if (col[q] == j) { p = q; break; }
        pmap[6 * i + s] = p;
    }
}

// --------------------------------------------------------- features --
// partials of (b, diag, vol) then of (b - mean)^2
__global__ void __launch_bounds__(256) k_feat_sums(int N, const double *b, const double *vals, const int *diag_idx,
                                                  const double *vol, double *partials)
{
    double acc[3] = {0.0, 0.0, 0.0};
    for (int i = blockIdx.x * 256 + threadIdx.x; i < N; i += gridDim.x * 256) {
        acc[0] += b[i];
        acc[1] += vals[diag_idx[i]];
        acc[2] += vol[i];
    }
    block_partials<3>(acc, partials, blockIdx.x, gridDim.x);
}
struct EpiFeat1 {
    double *sc;
    double inv_n;
    __device__ void operator()(const double (&v)[3]) const
    {
        sc[0] = v[0] * inv_n;   // mean b
        sc[2] = v[1] * inv_n;   // s_d
        sc[3] = v[2] * inv_n;   // mean V
    }
};
// net mass outflow of cell i
__device__ __forceinline__ double cell_outflow(int i, const double *mflux, const int *forder, const int *fother)
{
    double q = 0.0;
    for (int k = 0; k < 6; k++) {
        const double m = mflux[forder[6 * i + k]];
        q += (fother[6 * i + k] >= 0) ? m : -m;
    }
    return q;
}
// partials of (b - mean)^2 and q^2
__global__ void __launch_bounds__(256) k_feat_var(int N, const double *b, const double *mflux, const int *forder,
                                                 const int *fother, const double *sc, double *partials)
{
    double acc[2] = {0.0, 0.0};
    const double mu = sc[0];
    for (int i = blockIdx.x * 256 + threadIdx.x; i < N; i += gridDim.x * 256) {
        const double d = b[i] - mu, q = cell_outflow(i, mflux, forder, fother);
        acc[0] += d * d;
        acc[1] += q * q;
    }
    block_partials<2>(acc, partials, blockIdx.x, gridDim.x);
}
struct EpiFeat2 {
    double *sc;
    double inv_n;
    __device__ void operator()(const double (&v)[2]) const
    {
        const double rb = sqrt(v[0] * inv_n), rq = sqrt(v[1] * inv_n);
        sc[1] = (rb > 0.0) ? rb : 1.0;   // s_b
        sc[7] = (rq > 0.0) ? rq : 1.0;   // s_q
    }
};
__global__ void k_features(int N, const double *b, const double *vals, const int *diag_idx, const int *pmap,
                           const double *vol, const double *mflux, const int *forder, const int *fother,
                           const double *sc, float *f, float *e)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double mu = sc[0], sb = sc[1], sd = sc[2], mv = sc[3], sq = sc[7];
    const double q = cell_outflow(i, mflux, forder, fother);
    *(float4 *)(f + 4 * (size_t)i) = make_float4((float)((b[i] - mu) / sb), (float)(vals[diag_idx[i]] / sd),
                                                 (float)(vol[i] / mv), (float)(q / sq));
    float ev[6];
#pragma unroll
    for (int s = 0; s < 6; s++) {
        const int p = pmap[6 * i + s];
        ev[s] = (p >= 0) ? (float)(vals[p] / sd) : 0.f;
    }
    float2 *e2 = (float2 *)(e + 6 * (size_t)i);
    e2[0] = make_float2(ev[0], ev[1]); e2[1] = make_float2(ev[2], ev[3]); e2[2] = make_float2(ev[4], ev[5]);
}
// The solve rescales the guess by its Galerkin factor, so the net learns the
// solution's shape: the target is the mean-zero solution over its rms.
__global__ void __launch_bounds__(256) k_sumsq(int n, const double *x, double *partials)
{
    double acc[1] = {0.0};
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) acc[0] += x[i] * x[i];
    block_partials<1>(acc, partials, blockIdx.x, gridDim.x);
}
struct EpiRms {
    double *out;
    double inv_n;
    __device__ void operator()(const double (&v)[1]) const
    {
        const double r = sqrt(v[0] * inv_n);
        *out = (r > 0.0) ? r : 1.0;
    }
};
__global__ void k_target(int N, const double *x, const double *sc, float *t)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) t[i] = (float)(x[i] / sc[6]);
}
__global__ void k_guess_out(int N, const float *y, double *g)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) g[i] = (double)y[i];
}

static void pack_weights(Ctx *c)
{
    Mpnn *m = M(c);
    auto T = [&](bf16_t *dst, int kp, const float *src, int rows, int cols) {
        hipLaunchKernelGGL(k_pack_t, dim3(div_up(cols * kp, 256)), dim3(256), 0, c->stream, dst, kp, src, rows, cols);
    };
    auto R = [&](bf16_t *dst, int ld, int koff, const float *src, int rows, int cols) {
        hipLaunchKernelGGL(k_pack_rows, dim3(div_up(rows * cols, 256)), dim3(256), 0, c->stream, dst, ld, koff, src,
                           rows, cols);
    };
    for (int l = 0; l < ML; l++) {
        const float *Wl = m->w + OFF_L0 + l * LSZ;
        bf16_t *P = m->wb + l * PSZ;
        T(P + P_ABT, MH, Wl + L_WA, MH, MH);               // rows m < 64: Wa^T
        T(P + P_ABT + MH * MH, MH, Wl + L_WB, MH, MH);     // rows m >= 64: Wb^T
        T(P + P_W2T, MH, Wl + L_W2, MH, MH);
        T(P + P_N1T, 2 * MH, Wl + L_WN1, 2 * MH, MH);
        T(P + P_N2T, MH, Wl + L_WN2, MH, MH);
        R(P + P_ABR, 2 * MH, 0, Wl + L_WA, MH, MH);        // [Wa | Wb] row-major (64 x 128)
        R(P + P_ABR, 2 * MH, MH, Wl + L_WB, MH, MH);
        R(P + P_W2R, MH, 0, Wl + L_W2, MH, MH);
        R(P + P_N1R, MH, 0, Wl + L_WN1, 2 * MH, MH);
        R(P + P_N2R, MH, 0, Wl + L_WN2, MH, MH);
    }
    T(m->wb + P_ENCT, 16, m->w + OFF_WENC, MFIN, MH);
    auto Fr = [&](bf16_t *dst, const float *src, int KS) {
        hipLaunchKernelGGL(k_pack_frag, dim3(div_up(4 * KS * 64 * 8, 256)), dim3(256), 0, c->stream, dst, src, KS);
    };
    for (int l = 0; l < ML; l++) {
        const float *Wl = m->w + OFF_L0 + l * LSZ;
        bf16_t *P = m->wf + l * FR_SZ;
        Fr(P + FR_WA, Wl + L_WA, 2);
        Fr(P + FR_WB, Wl + L_WB, 2);
        Fr(P + FR_W2, Wl + L_W2, 2);
        Fr(P + FR_N1, Wl + L_WN1, 4);
        Fr(P + FR_N2, Wl + L_WN2, 2);
    }
    hipLaunchKernelGGL(k_pack_enc, dim3(div_up(4 * 64 * 8, 256)), dim3(256), 0, c->stream, m->wf + ML * FR_SZ, m->w);
}

// --------------------------------------------------------------- host --
static int ensure(Ctx *c)
{
    if (c->mpnn) return MC_OK;
    MemScope ms(c, MC_MEM_MPNN);
    Mpnn *m = new (std::nothrow) Mpnn();
    if (!m) return MC_ERR_NOMEM;
    c->mpnn = m;
    const int N = c->N;
    m->n = N;
    MC_TRY(dalloc(c, &m->w, NPARAMS)); MC_TRY(dalloc(c, &m->g, NPARAMS));
    MC_TRY(dalloc(c, &m->am, NPARAMS)); MC_TRY(dalloc(c, &m->av, NPARAMS));
    MC_TRY(dalloc(c, &m->wb, PTOTAL));
    MC_TRY(dalloc(c, &m->nbr, 6 * (int64_t)N)); MC_TRY(dalloc(c, &m->rev, 6 * (int64_t)N));
    MC_TRY(dalloc(c, &m->pmap, 6 * (int64_t)N)); MC_TRY(dalloc(c, &m->deg, N));
    MC_TRY(dalloc(c, &m->f, 4 * (int64_t)N)); MC_TRY(dalloc(c, &m->e, 6 * (int64_t)N));
    MC_TRY(dalloc(c, &m->y, N)); MC_TRY(dalloc(c, &m->dy, N));
    MC_TRY(dalloc(c, &m->wf, ML * FR_SZ + 4 * 64 * 8));   // the layers, then Wenc^T
    MC_TRY(dalloc(c, &m->hb[0], (int64_t)N * MH)); MC_TRY(dalloc(c, &m->hb[1], (int64_t)N * MH));
    MC_TRY(dalloc(c, &m->sc, 16));
    {
        int dev = 0;
        MC_HIP(c, hipGetDevice(&dev));
        MC_HIP(c, hipDeviceGetAttribute(&m->cus, hipDeviceAttributeMultiprocessorCount, dev));
    }
    MC_HIP(c, hipMemsetAsync(m->am, 0, NPARAMS * sizeof(float), c->stream));
    MC_HIP(c, hipMemsetAsync(m->av, 0, NPARAMS * sizeof(float), c->stream));
    hipLaunchKernelGGL(k_graph, dim3(div_up(N, 256)), dim3(256), 0, c->stream, N, c->nbr, m->nbr, m->rev, m->deg);
    hipLaunchKernelGGL(k_pmap, dim3(div_up(N, 256)), dim3(256), 0, c->stream, N, m->nbr, c->row_ptr, c->col, m->pmap);
    MC_LAUNCH_CHECK(c);
    return MC_OK;
}

static int ensure_train(Ctx *c)
{
    Mpnn *m = M(c);
    if (m->train_bufs) return MC_OK;
    MemScope ms(c, MC_MEM_MPNN);
    const int64_t N = m->n;
    for (int l = 0; l <= ML; l++) MC_TRY(dalloc(c, &m->H[l], N * MH));
    for (int l = 0; l < ML; l++) {
        MC_TRY(dalloc(c, &m->PQ[l], N * 2 * MH)); MC_TRY(dalloc(c, &m->S[l], N * MH));
        MC_TRY(dalloc(c, &m->A[l], N * MH)); MC_TRY(dalloc(c, &m->U[l], N * MH));
    }
    MC_TRY(dalloc(c, &m->dH, N * MH)); MC_TRY(dalloc(c, &m->dH2, N * MH)); MC_TRY(dalloc(c, &m->dU, N * MH));
    MC_TRY(dalloc(c, &m->dHA, N * 2 * MH)); MC_TRY(dalloc(c, &m->dS, N * MH)); MC_TRY(dalloc(c, &m->dPQ, N * 2 * MH));
    const int64_t nb = div_up(N, TN_ROWS);
    m->partial_cap = std::max<int64_t>(nb * (2 * MH + 1) * 2 * MH, div_up(N * 16, 256) * (2 * MH + 1));
    MC_TRY(dalloc(c, &m->partial, m->partial_cap));
    m->train_bufs = true;
    return MC_OK;
}

// Training forward pass on features f / edge features e: FP32 activations,
// every layer's intermediates kept for the backward pass.
static void forward_train(Ctx *c, const float *f, const float *e)
{
    Mpnn *m = M(c);
    const int n = m->n;
    const dim3 gmm(div_up(n, 64)), blk(256);
    const float *W = m->w;
    hipLaunchKernelGGL((k_mm<MFIN, 0, 16, MH, EPI_BIAS | EPI_RELU>), gmm, blk, 0, c->stream, n, f, MFIN,
                       (const float *)nullptr, 0, m->wb + P_ENCT, W + OFF_BENC, (const float *)nullptr,
                       (const float *)nullptr, 0, (const float *)nullptr, 0, m->H[0], MH);
    for (int l = 0; l < ML; l++) {
        const float *Wl = W + OFF_L0 + l * LSZ;
        const bf16_t *Pk = m->wb + l * PSZ;
        float *hin = m->H[l], *hout = m->H[l + 1];
        hipLaunchKernelGGL((k_mm<MH, 0, MH, 2 * MH, 0>), gmm, blk, 0, c->stream, n, hin, MH, (const float *)nullptr, 0,
                           Pk + P_ABT, (const float *)nullptr, (const float *)nullptr, (const float *)nullptr, 0,
                           (const float *)nullptr, 0, m->PQ[l], 2 * MH);
        hipLaunchKernelGGL(k_edge_fwd, dim3(div_up((int64_t)n * 16, 256)), blk, 0, c->stream, n, m->PQ[l], m->nbr, e,
                           Wl + L_WE, Wl + L_B1, m->S[l]);
        hipLaunchKernelGGL((k_mm<MH, 0, MH, MH, EPI_BIAS | EPI_ROWSCALE>), gmm, blk, 0, c->stream, n, m->S[l], MH,
                           (const float *)nullptr, 0, Pk + P_W2T, Wl + L_B2, m->deg, (const float *)nullptr, 0,
                           (const float *)nullptr, 0, m->A[l], MH);
        hipLaunchKernelGGL((k_mm<MH, MH, 2 * MH, MH, EPI_BIAS | EPI_RELU>), gmm, blk, 0, c->stream, n, hin, MH, m->A[l],
                           MH, Pk + P_N1T, Wl + L_BN1, (const float *)nullptr, (const float *)nullptr, 0,
                           (const float *)nullptr, 0, m->U[l], MH);
        hipLaunchKernelGGL((k_mm<MH, 0, MH, MH, EPI_RESID | EPI_BIAS>), gmm, blk, 0, c->stream, n, m->U[l], MH,
                           (const float *)nullptr, 0, Pk + P_N2T, Wl + L_BN2, (const float *)nullptr, hin, MH,
                           (const float *)nullptr, 0, hout, MH);
    }
    hipLaunchKernelGGL(k_decode, dim3(div_up(n, 64)), blk, 0, c->stream, n, m->H[ML], W, m->y);
}

// Inference forward pass into m->y: the encoder, then one fused kernel per
// layer on BF16 activations (the last with the decoder).  (The encoder folded
// into the first layer, recomputed for each cell's six neighbours, measured
// 0.13 ms slower than storing h0 once.)
static void forward_infer(Ctx *c, const float *f, const float *e)
{
    Mpnn *m = M(c);
    const int n = m->n;
    const CellWalk w = cell_walk(c);
    const int nbricks = walk_blocks(w);
    // MP_LB resident workgroups per compute unit, a multiple of the 8 XCDs
    const int per_xcd = std::max(1, std::min(div_up(nbricks, 8), MP_LB * std::max(m->cus, 8) / 8));
    const dim3 grid(8 * per_xcd), blk(256), lblk(MP_THREADS);
    const float *W = m->w;
    hipLaunchKernelGGL(k_mp_encode, dim3(div_up(n, 64)), blk, 0, c->stream, n, f, (const bf16_t *)(m->wf + ML * FR_SZ),
                       W, m->hb[0]);
    for (int l = 0; l < ML; l++) {
        const bf16_t *hin = m->hb[l & 1];
        const bf16_t *wf = m->wf + l * FR_SZ;
        const float *Wl = W + OFF_L0 + l * LSZ;
        if (l < ML - 1)
            hipLaunchKernelGGL(k_mp_layer<false>, grid, lblk, 0, c->stream, w, nbricks, hin, m->hb[(l + 1) & 1], m->nbr,
                               e, m->deg, wf, Wl, W, (float *)nullptr);
        else
            hipLaunchKernelGGL(k_mp_layer<true>, grid, lblk, 0, c->stream, w, nbricks, hin, (bf16_t *)nullptr, m->nbr,
                               e, m->deg, wf, Wl, W, m->y);
    }
}

// MFMA work of one training forward pass (flops)
static double forward_flops(int n)
{
    return (double)n * (2.0 * 16 * MH + ML * 2.0 * (MH * 2 * MH + MH * MH + 2 * MH * MH + MH * MH));
}

// MFMA work of one fused inference pass (flops, as executed): per cell and
// layer the products with Wa, W2, Wn2 (64 x 64), Wn1 (128 x 64) and Wb once
// for each of the six neighbour slots (recomputed, not stored), and the
// encoder (K padded to 32)
static double infer_flops(int n)
{
    return (double)n * (ML * 2.0 * MH * MH * (1 + 6 + 1 + 2 + 1) + 2.0 * MH * 32);
}

template <int K1, int K2, int KP, int MM>
static void tn(Ctx *c, const float *X1, int ld1, const float *X2, int ld2, const float *D, int ldd,
               const float *rowscale, int K, int msplit, int off1, int off2, int offb)
{
    Mpnn *m = M(c);
    const int n = m->n;
    const int nb = div_up(n, TN_ROWS);
    hipLaunchKernelGGL((k_mm_tn<K1, K2, KP, MM>), dim3(nb), dim3(256), 0, c->stream, n, X1, ld1, X2, ld2, D, ldd,
                       rowscale, m->partial);
    hipLaunchKernelGGL(k_tn_reduce, dim3(div_up((KP + 1) * MM, 256)), dim3(256), 0, c->stream, m->partial, nb, KP, MM,
                       K, msplit, off1, off2, offb, m->g);
}

// loss (into sc[4]) and dy for target t
static void loss(Ctx *c, const float *t)
{
    Mpnn *m = M(c);
    const int n = m->n;
    const int nb = std::min(div_up(n, 256), 1024);
    hipLaunchKernelGGL(k_fsum, dim3(nb), dim3(256), 0, c->stream, n, m->y, c->red);
    launch_finish<1>(c->stream, c->red, nb, EpiMeanTo{m->sc + 5, 1.0 / n});
    hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, c->stream, n, m->y, t, m->sc + 5, m->dy, c->red);
    launch_finish<1>(c->stream, c->red, nb, EpiMeanTo{m->sc + 4, 1.0 / n});
}

static int backward(Ctx *c, const float *f, const float *e)
{
    Mpnn *m = M(c);
    const int n = m->n;
    const dim3 gmm(div_up(n, 64)), blk(256);
    const float *W = m->w;
    MC_HIP(c, hipMemsetAsync(m->g, 0, NPARAMS * sizeof(float), c->stream));
    const int nb16 = div_up((int64_t)n * 16, 256);
    hipLaunchKernelGGL(k_decode_bwd, dim3(nb16), blk, 0, c->stream, n, m->H[ML], W, m->dy, m->dH, m->partial);
    hipLaunchKernelGGL(k_decode_reduce, dim3(1), dim3(128), 0, c->stream, m->partial, nb16, m->g);
    for (int l = ML - 1; l >= 0; l--) {
        const int base = OFF_L0 + l * LSZ;
        const float *Wl = W + base;
        const bf16_t *Pk = m->wb + l * PSZ;
        // h' = h + u Wn2 + bn2
        tn<MH, 0, MH, MH>(c, m->U[l], MH, nullptr, 0, m->dH, MH, nullptr, MH, MH, base + L_WN2, 0, base + L_BN2);
        hipLaunchKernelGGL((k_mm<MH, 0, MH, MH, EPI_MASK>), gmm, blk, 0, c->stream, n, m->dH, MH,
                           (const float *)nullptr, 0, Pk + P_N2R, (const float *)nullptr, (const float *)nullptr,
                           (const float *)nullptr, 0, m->U[l], MH, m->dU, MH);
        // u = relu([h, a] Wn1 + bn1)
        tn<MH, MH, 2 * MH, MH>(c, m->H[l], MH, m->A[l], MH, m->dU, MH, nullptr, 2 * MH, MH, base + L_WN1, 0,
                               base + L_BN1);
        hipLaunchKernelGGL((k_mm<MH, 0, MH, 2 * MH, 0>), gmm, blk, 0, c->stream, n, m->dU, MH, (const float *)nullptr,
                           0, Pk + P_N1R, (const float *)nullptr, (const float *)nullptr, (const float *)nullptr, 0,
                           (const float *)nullptr, 0, m->dHA, 2 * MH);
        // a = s W2 + deg b2
        tn<MH, 0, MH, MH>(c, m->S[l], MH, nullptr, 0, m->dHA + MH, 2 * MH, m->deg, MH, MH, base + L_W2, 0,
                          base + L_B2);
        hipLaunchKernelGGL((k_mm<MH, 0, MH, MH, 0>), gmm, blk, 0, c->stream, n, m->dHA + MH, 2 * MH,
                           (const float *)nullptr, 0, Pk + P_W2R, (const float *)nullptr, (const float *)nullptr,
                           (const float *)nullptr, 0, (const float *)nullptr, 0, m->dS, MH);
        // edges
        hipLaunchKernelGGL(k_edge_bwd, dim3(nb16), blk, 0, c->stream, n, m->PQ[l], m->nbr, m->rev, e, Wl + L_WE,
                           Wl + L_B1, m->dS, m->dPQ, m->partial);
        hipLaunchKernelGGL(k_edge_reduce, dim3(1), dim3(128), 0, c->stream, m->partial, nb16, base + L_WE,
                           base + L_B1, m->g);
        // [P | Q] = h [Wa | Wb]
        tn<MH, 0, MH, 2 * MH>(c, m->H[l], MH, nullptr, 0, m->dPQ, 2 * MH, nullptr, MH, MH, base + L_WA, base + L_WB,
                              -1);
        hipLaunchKernelGGL(k_add_block, dim3(div_up((int64_t)n * MH, 256)), blk, 0, c->stream, n, m->dH, m->dHA,
                           2 * MH);
        if (l > 0) {
            hipLaunchKernelGGL((k_mm<2 * MH, 0, 2 * MH, MH, EPI_RESID>), gmm, blk, 0, c->stream, n, m->dPQ, 2 * MH,
                               (const float *)nullptr, 0, Pk + P_ABR, (const float *)nullptr, (const float *)nullptr,
                               m->dH, MH, (const float *)nullptr, 0, m->dH2, MH);
        } else {   // into the encoder: mask by its ReLU
            hipLaunchKernelGGL((k_mm<2 * MH, 0, 2 * MH, MH, EPI_RESID | EPI_MASK>), gmm, blk, 0, c->stream, n, m->dPQ,
                               2 * MH, (const float *)nullptr, 0, Pk + P_ABR, (const float *)nullptr,
                               (const float *)nullptr, m->dH, MH, m->H[0], MH, m->dH2, MH);
        }
        std::swap(m->dH, m->dH2);
    }
    tn<MFIN, 0, 16, MH>(c, f, MFIN, nullptr, 0, m->dH, MH, nullptr, MFIN, MH, OFF_WENC, 0, OFF_BENC);
    (void)e;
    return MC_OK;
}

// Features of the pressure system (b, current PA values) into f / e, with
// the scale factors in sc[0..3].
static int features(Ctx *c, const double *b, float *f, float *e)
{
    Mpnn *m = M(c);
    const int N = c->N;
    const int nb = std::min(div_up(N, 256), 1024);
    hipLaunchKernelGGL(k_feat_sums, dim3(nb), dim3(256), 0, c->stream, N, b, c->PA_vals, c->diag_idx, c->vol, c->red);
    launch_finish<3>(c->stream, c->red, nb, EpiFeat1{m->sc, 1.0 / N});
    hipLaunchKernelGGL(k_feat_var, dim3(nb), dim3(256), 0, c->stream, N, b, c->mflux, c->forder, c->fother, m->sc,
                       c->red);
    launch_finish<2>(c->stream, c->red, nb, EpiFeat2{m->sc, 1.0 / N});
    hipLaunchKernelGGL(k_features, dim3(div_up(N, 256)), dim3(256), 0, c->stream, N, b, c->PA_vals, c->diag_idx,
                       m->pmap, c->vol, c->mflux, c->forder, c->fother, m->sc, f, e);
    MC_LAUNCH_CHECK(c);
    return MC_OK;
}

// ------------------------------------------------ library-internal API --
int mpnn_guess(Ctx *c, const double *b, double *g, int *ok)
{
    *ok = 0;
    Mpnn *m = M(c);
    if (!m || !m->ready) return MC_OK;
    MC_TRY(features(c, b, m->f, m->e));
    {
        KTimer kt(c, MC_KC_MPNN, infer_flops(c->N));
        forward_infer(c, m->f, m->e);
    }
    hipLaunchKernelGGL(k_guess_out, dim3(div_up(c->N, 256)), dim3(256), 0, c->stream, c->N, m->y, g);
    MC_LAUNCH_CHECK(c);
    *ok = 1;
    return MC_OK;
}

// after a pressure solve: store (features, target) when collecting
int mpnn_collect_pair(Ctx *c, const double *b, const double *x)
{
    Mpnn *m = M(c);
    if (!m || m->count >= m->cap) return MC_OK;
    const int64_t N = c->N, k = m->count;
    MC_TRY(features(c, b, m->ds_f + k * 4 * N, m->ds_e + k * 6 * N));
    const int nb = std::min(div_up(N, 256), 1024);
    hipLaunchKernelGGL(k_sumsq, dim3(nb), dim3(256), 0, c->stream, (int)N, x, c->red);
    launch_finish<1>(c->stream, c->red, nb, EpiRms{m->sc + 6, 1.0 / (double)N});
    hipLaunchKernelGGL(k_target, dim3(div_up(N, 256)), dim3(256), 0, c->stream, (int)N, x, m->sc, m->ds_t + k * N);
    MC_LAUNCH_CHECK(c);
    m->count++;
    return MC_OK;
}

void mpnn_destroy(Ctx *c)
{
    delete M(c);   // device memory is owned by the context's allocation list
    c->mpnn = nullptr;
}

static int pair_ptrs(Ctx *c, int64_t pair, const float **f, const float **e, const float **t)
{
    Mpnn *m = M(c);
    if (!m || pair < 0 || pair >= m->count) {
        c->err = "mpnn: no such stored pair";
        return MC_ERR_ARG;
    }
    const int64_t N = c->N;
    *f = m->ds_f + pair * 4 * N;
    *e = m->ds_e + pair * 6 * N;
    *t = m->ds_t + pair * N;
    return MC_OK;
}

static int grad_of_pair(Ctx *c, int64_t pair, double *loss_out)
{
    Mpnn *m = M(c);
    const float *f, *e, *t;
    MC_TRY(pair_ptrs(c, pair, &f, &e, &t));
    MC_TRY(ensure_train(c));
    {
        KTimer kt(c, MC_KC_MPNN, forward_flops(c->N));
        forward_train(c, f, e);
    }
    loss(c, t);
    {
        KTimer kt(c, MC_KC_MPNN_BWD, 2.0 * forward_flops(c->N));
        MC_TRY(backward(c, f, e));
    }
    MC_LAUNCH_CHECK(c);
    if (loss_out) {
        MC_HIP(c, hipMemcpyAsync(c->hscal + 700, m->sc + 4, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        MC_HIP(c, hipStreamSynchronize(c->stream));
        *loss_out = c->hscal[700];
    }
    return MC_OK;
}

}  // namespace mc

using namespace mc;

extern "C" {

int64_t mc_mpnn_nparams(void) { return NPARAMS; }

int mc_mpnn_load(mc_ctx *x, const float *weights, int64_t count)
{
    if (!x || !weights) return MC_ERR_ARG;
    Ctx *c = &x->c;
    if (count != NPARAMS) { c->err = "mpnn: wrong parameter count"; return MC_ERR_ARG; }
    MC_TRY(ensure(c));
    Mpnn *m = M(c);
    MC_HIP(c, hipMemcpyAsync(m->w, weights, NPARAMS * sizeof(float), hipMemcpyHostToDevice, c->stream));
    MC_HIP(c, hipMemsetAsync(m->am, 0, NPARAMS * sizeof(float), c->stream));
    MC_HIP(c, hipMemsetAsync(m->av, 0, NPARAMS * sizeof(float), c->stream));
    m->adam_t = 0;
    pack_weights(c);
    MC_LAUNCH_CHECK(c);
    MC_HIP(c, hipStreamSynchronize(c->stream));
    m->ready = true;
    return MC_OK;
}

int mc_mpnn_save(mc_ctx *x, float *weights, int64_t count)
{
    if (!x || !weights) return MC_ERR_ARG;
    Ctx *c = &x->c;
    Mpnn *m = M(c);
    if (!m || !m->ready) { c->err = "mpnn: no weights loaded"; return MC_ERR_STATE; }
    if (count != NPARAMS) { c->err = "mpnn: wrong parameter count"; return MC_ERR_ARG; }
    MC_HIP(c, hipMemcpyAsync(weights, m->w, NPARAMS * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    MC_HIP(c, hipStreamSynchronize(c->stream));
    return MC_OK;
}

int mc_mpnn_collect(mc_ctx *x, int64_t max_pairs)
{
    if (!x || max_pairs < 0) return MC_ERR_ARG;
    Ctx *c = &x->c;
    MC_TRY(ensure(c));
    Mpnn *m = M(c);
    if (max_pairs > m->cap) {
        if (m->cap) { c->err = "mpnn: data set already allocated with a smaller capacity"; return MC_ERR_STATE; }
        const int64_t N = c->N;
        MemScope ms(c, MC_MEM_DATASET);
        MC_TRY(dalloc(c, &m->ds_f, max_pairs * 4 * N));
        MC_TRY(dalloc(c, &m->ds_e, max_pairs * 6 * N));
        MC_TRY(dalloc(c, &m->ds_t, max_pairs * N));
        m->cap = max_pairs;
    } else if (max_pairs < m->cap) {
        m->cap = std::max(max_pairs, m->count);   // stop collecting once full
    }
    return MC_OK;
}

int mc_mpnn_dataset_size(mc_ctx *x, int64_t *pairs)
{
    if (!x || !pairs) return MC_ERR_ARG;
    Ctx *c = &x->c;
    *pairs = M(c) ? M(c)->count : 0;
    return MC_OK;
}

int mc_mpnn_pair_get(mc_ctx *x, int64_t pair, float *feat, float *edge, float *target)
{
    if (!x) return MC_ERR_ARG;
    Ctx *c = &x->c;
    const float *f, *e, *t;
    MC_TRY(pair_ptrs(c, pair, &f, &e, &t));
    const size_t N = c->N;
    if (feat) MC_HIP(c, hipMemcpyAsync(feat, f, 4 * N * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    if (edge) MC_HIP(c, hipMemcpyAsync(edge, e, 6 * N * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    if (target) MC_HIP(c, hipMemcpyAsync(target, t, N * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    MC_HIP(c, hipStreamSynchronize(c->stream));
    return MC_OK;
}

int mc_mpnn_grad(mc_ctx *x, int64_t pair, float *grad_host, double *loss_out)
{
    if (!x) return MC_ERR_ARG;
    Ctx *c = &x->c;
    Mpnn *m = M(c);
    if (!m || !m->ready) { c->err = "mpnn: no weights loaded"; return MC_ERR_STATE; }
    MC_TRY(grad_of_pair(c, pair, loss_out));
    if (grad_host) {
        MC_HIP(c, hipMemcpyAsync(grad_host, m->g, NPARAMS * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        MC_HIP(c, hipStreamSynchronize(c->stream));
    }
    return MC_OK;
}

int mc_mpnn_train_step(mc_ctx *x, int64_t pair, double lr, double *loss_out)
{
    if (!x || !(lr > 0.0)) return MC_ERR_ARG;
    Ctx *c = &x->c;
    Mpnn *m = M(c);
    if (!m || !m->ready) { c->err = "mpnn: no weights loaded"; return MC_ERR_STATE; }
    MC_TRY(grad_of_pair(c, pair, loss_out));
    m->adam_t++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)m->adam_t), c2 = 1.f - powf(b2, (float)m->adam_t);
    hipLaunchKernelGGL(k_adam, dim3(div_up(NPARAMS, 256)), dim3(256), 0, c->stream, NPARAMS, m->w, m->g, m->am, m->av,
                       (float)lr, b1, b2, 1e-8f, c1, c2);
    pack_weights(c);
    MC_LAUNCH_CHECK(c);
    return MC_OK;
}

int mc_mpnn_forward_pair(mc_ctx *x, int64_t pair, float *y_host)
{
    if (!x || !y_host) return MC_ERR_ARG;
    Ctx *c = &x->c;
    Mpnn *m = M(c);
    if (!m || !m->ready) { c->err = "mpnn: no weights loaded"; return MC_ERR_STATE; }
    const float *f, *e, *t;
    MC_TRY(pair_ptrs(c, pair, &f, &e, &t));
    {
        KTimer kt(c, MC_KC_MPNN, infer_flops(c->N));
        forward_infer(c, f, e);
    }
    MC_LAUNCH_CHECK(c);
    MC_HIP(c, hipMemcpyAsync(y_host, m->y, (size_t)c->N * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    MC_HIP(c, hipStreamSynchronize(c->stream));
    return MC_OK;
}

}  // extern "C"
