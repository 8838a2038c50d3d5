// ffh_mpnn.h -- learned preconditioner for the massless CG: a mesh message-passing network
// (embedding 4 -> 32, three message-passing layers of width 32, output 32 -> 3), its inference
// inside flexible CG and its training (forward, backward, Adam) on the GPU.  Included by
// ffh_kernels.hip inside its anonymous namespace (uses SpmvArgs, CgCtl, xcd_block, the
// reduction helpers).  It has no counterpart in the reference, whose only preconditioner is
// Jacobi (src/NoMassCGSolver.cpp:278-292, :325-333).
//
// Network, per linear node i of blob b (pinned and padding rows give z = 0):
//   u_i = D_i^-1 r_i                     Jacobi-scaled residual (3)
//   s_b = sqrt(sum_{i in b} |u_i|^2 / n_b)  per-blob scale (n_b real linear nodes)
//   x_i = (u_i / s_b, c_i),  c_i = 3 zeta_i / tr K_ii     features (4)
//   h0  = relu(We x + be)                                         (32)
//   g_l = sum_j A_ij h_{l-1,j},  A_ij = tr K_ij / sqrt(tr K_ii tr K_jj)   (the matrix's own pattern)
//   h_l = relu(W1_l h_{l-1} + W2_l g_l + b_l),  l = 1..3          (32)
//   o_i = Wo h3 + bo                                              (3)
//   z_i = u_i + s_b o_i                  (Jacobi plus a learned correction; Wo = 0 is Jacobi)
// Storage: activations h, g in bf16 [N][32]; the dense layers run on MFMA
// (v_mfma_f32_32x32x8_bf16, fp32 accumulation) with bf16 weights; embedding and output layers
// in fp32.  Training minimises, per snapshot, mean over blobs of
//   sum_i |o_i - t_i|^2 / sum_i |e_i / s_b|^2,   t_i = (e_i - u_i) / s_b,
// where e = x* - x_k is the exact correction of the residual r_k of a plain CG solve.

typedef unsigned short bf16_t;
typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

constexpr int kHid = 32;
constexpr int kNIn = 4;
constexpr int kNOut = 3;
constexpr int kNLay = 3;
constexpr int kOffWe = 0;                       // [32][4]
constexpr int kOffBe = kOffWe + kHid * kNIn;    // [32]
constexpr int kOffLay = kOffBe + kHid;          // per layer: W1 [32][32], W2 [32][32], b [32]
constexpr int kLayStride = 2 * kHid * kHid + kHid;
constexpr int kOffWo = kOffLay + kNLay * kLayStride;  // [3][32]
constexpr int kOffBo = kOffWo + kNOut * kHid;         // [3]
constexpr int kNumParams = kOffBo + kNOut;            // 6499
constexpr int kWgradRows = 512;                       // rows per weight-gradient block
constexpr int kLayGrad = 2 * kHid * kHid + kHid;      // partial per block: dW1^T, dW2^T, db
constexpr int kOutGrad = kNOut * kHid + kNOut;        // dWo, dbo
constexpr int kEmbGrad = kHid * kNIn + kHid;          // dWe, dbe
static_assert(kNumParams == 6499, "parameter layout");

__device__ __forceinline__ bf16_t f2bf(float f) {
    const uint32_t u = __float_as_uint(f);
    return (bf16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);  // round to nearest even
}
__device__ __forceinline__ float bf2f(bf16_t b) { return __uint_as_float((uint32_t)b << 16); }

static inline uint16_t host_f2bf(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

struct MpnnArgs {
    int N, nslices;
    const int32_t *slice_blob, *blob_slice_off;
    const int32_t *blob_nreal;  // real linear nodes per blob
    const uint8_t *pinned;      // pinned or padding row
    const int64_t *slice_off;
    const int32_t *slice_width, *col;
    const float *aw;    // A_ij per SELL slot
    const float *cf;    // c_i per row
    const float *wf;    // fp32 parameters
    const bf16_t *wT;   // [3][32][64]: forward B operand, wT[l][j][k] = Wcat_l[k][j]
    const bf16_t *wN;   // [3][64][32]: backward B operand, wN[l][k][j] = Wcat_l[k][j]
    const CgCtl *ctl;   // nullptr in training: no blob is skipped
    const double *tot;  // per-blob totals or nullptr
};

__device__ __forceinline__ bool blob_done(const MpnnArgs &a, int slice) {
    return a.ctl && a.ctl[a.slice_blob[slice]].done;
}

// Per-slot coupling A_ij and per-row drag share c_i from the assembled blocks (after
// k_assemble): tr K_ij over sqrt(tr K_ii tr K_jj), with tr K_ii = sum_c 1 / dinv.
template <typename T>
__global__ __launch_bounds__(256) void k_mpnn_edges(SpmvArgs m, int nslices, const double *__restrict__ dinv,
                                                    const double *__restrict__ zeta, float *__restrict__ aw,
                                                    float *__restrict__ cf) {
    const int slice = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (slice >= nslices) return;
    const int row = slice * 64 + lane;
    const T *val = reinterpret_cast<const T *>(m.val);
    const double tri = 1.0 / dinv[3 * row] + 1.0 / dinv[3 * row + 1] + 1.0 / dinv[3 * row + 2];
    cf[row] = (float)(3.0 * zeta[row] / tri);
    const int w = m.slice_width[slice];
    const int64_t base = m.slice_off[slice];
    for (int t = 0; t < w; t++) {
        const int64_t slot = base + (int64_t)t * 64 + lane;
        const int c = m.col[slot];
        const T *vt = val + (base + (int64_t)t * 64) * 9;
        const double tr = (double)vt[tile_entry<T>(0, lane)] + (double)vt[tile_entry<T>(4, lane)] +
                          (double)vt[tile_entry<T>(8, lane)];
        const double trj = 1.0 / dinv[3 * c] + 1.0 / dinv[3 * c + 1] + 1.0 / dinv[3 * c + 2];
        aw[slot] = (float)(tr / sqrt(tri * trj));
    }
}

// Per-blob scale s_b from the per-slice partials of |u|^2 (identical in every block).
__device__ __forceinline__ double blob_scale(const MpnnArgs &a, const double *parts_uu, int b, double *sh) {
    const double uu = blob_total(parts_uu, a.tot, kTotUU, b, a.blob_slice_off[b], a.blob_slice_off[b + 1], sh);
    const int n = a.blob_nreal[b];
    return n > 0 ? sqrt(uu / n) : 0.0;
}

// Embedding.  One block = one slice; thread (row, q) computes hidden units 8q..8q+7 of its row.
// U: u = D^-1 r as double (CG) or float (training).  In training (tgt_e != nullptr) it also
// forms the target t = (e - u) / s_b and the partials of |e / s_b|^2.
template <typename U>
__global__ __launch_bounds__(256) void k_mpnn_embed(MpnnArgs a, const U *__restrict__ u, const double *parts_uu,
                                                    float *__restrict__ feat, bf16_t *__restrict__ H0,
                                                    const float *__restrict__ tgt_e, float *__restrict__ tgt,
                                                    double *parts_den) {
    __shared__ double sh[4];
    const int slice = xcd_block(a.nslices);
    if (slice >= a.nslices || blob_done(a, slice)) return;
    const int b = a.slice_blob[slice];
    const double s = blob_scale(a, parts_uu, b, sh);
    const double inv = s > 0.0 ? 1.0 / s : 0.0;
    const int rloc = threadIdx.x >> 2, q = threadIdx.x & 3;
    const int row = slice * 64 + rloc;
    float x[kNIn];
    for (int c = 0; c < 3; c++) x[c] = (float)((double)u[3 * (size_t)row + c] * inv);
    x[3] = a.cf[row];
    if (q == 0 && feat)
        for (int c = 0; c < kNIn; c++) feat[(size_t)row * kNIn + c] = x[c];
    const float *We = a.wf + kOffWe, *be = a.wf + kOffBe;
    bf16_t hv[8];
    for (int k = 0; k < 8; k++) {
        const int j = 8 * q + k;
        float h = be[j];
        for (int c = 0; c < kNIn; c++) h += We[j * kNIn + c] * x[c];
        hv[k] = f2bf(fmaxf(h, 0.0f));
    }
    uint4 pk;
    pk.x = hv[0] | ((uint32_t)hv[1] << 16);
    pk.y = hv[2] | ((uint32_t)hv[3] << 16);
    pk.z = hv[4] | ((uint32_t)hv[5] << 16);
    pk.w = hv[6] | ((uint32_t)hv[7] << 16);
    *reinterpret_cast<uint4 *>(H0 + (size_t)row * kHid + 8 * q) = pk;
    if (tgt_e) {
        double den = 0.0;
        if (threadIdx.x < 192) {
            const size_t k = (size_t)slice * 192 + threadIdx.x;
            const double e = (double)tgt_e[k] * inv;
            tgt[k] = (float)(e - (double)u[k] * inv);
            den = e * e;
        }
        block_partial_at(den, parts_den, slice, sh);
    }
}

// acc[k] += A_rt h[c_t][k] over row rloc's slots t in order (k = 0..7, the 8 bf16 channels at
// src + c ld): the slots' column indices and couplings are loaded up to eight at a time, then
// their 16-byte gathers, then the sums in slot order -- the order of a one-slot loop, with a
// group's dependent load pairs in flight.
__device__ __forceinline__ void message_rows(const MpnnArgs &a, int64_t base, int w, int rloc,
                                             const bf16_t *__restrict__ src, float acc[8], int ld = kHid) {
    exact_groups(0, w, [&](auto g_, int t0) {
        constexpr int G = decltype(g_)::value;
        int c[G];
        float wg[G];
#pragma unroll
        for (int u = 0; u < G; u++) {
            const int64_t slot = base + (int64_t)(t0 + u) * 64 + rloc;
            c[u] = a.col[slot];
            wg[u] = a.aw[slot];
        }
        uint4 v[G];
#pragma unroll
        for (int u = 0; u < G; u++) v[u] = *reinterpret_cast<const uint4 *>(src + (size_t)c[u] * ld);
#pragma unroll
        for (int u = 0; u < G; u++) {
            const uint32_t vv[4] = {v[u].x, v[u].y, v[u].z, v[u].w};
#pragma unroll
            for (int k = 0; k < 4; k++) {
                acc[2 * k] += wg[u] * bf2f((bf16_t)(vv[k] & 0xffffu));
                acc[2 * k + 1] += wg[u] * bf2f((bf16_t)(vv[k] >> 16));
            }
        }
    });
}

// g = A h (one slice per block, 4 threads per row, 8 channels each).  Input [N][ld] of TI,
// columns off..off+31; output bf16 [N][32].  With mask (backward), the output is
// (in0[:, 0:32] + A in[:, off:off+32]) * (mask > 0) in fp32 instead.
template <typename TI, bool kBwd>
__global__ __launch_bounds__(256) void k_mpnn_spmm(MpnnArgs a, const TI *__restrict__ in, int ld, int off,
                                                   void *__restrict__ out, const bf16_t *__restrict__ mask) {
    const int slice = xcd_block(a.nslices);
    if (slice >= a.nslices || blob_done(a, slice)) return;
    const int rloc = threadIdx.x >> 2, q = threadIdx.x & 3;
    const int row = slice * 64 + rloc;
    const int w = a.slice_width[slice];
    const int64_t base = a.slice_off[slice];
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if constexpr (sizeof(TI) == 2) {
        message_rows(a, base, w, rloc, reinterpret_cast<const bf16_t *>(in) + off + 8 * q, acc, ld);
    } else {
        exact_groups(0, w, [&](auto g_, int t0) {
            constexpr int G = decltype(g_)::value;
            int c[G];
            float wg[G];
#pragma unroll
            for (int u = 0; u < G; u++) {
                const int64_t slot = base + (int64_t)(t0 + u) * 64 + rloc;
                c[u] = a.col[slot];
                wg[u] = a.aw[slot];
            }
            float4 v0[G], v1[G];
#pragma unroll
            for (int u = 0; u < G; u++) {
                const TI *src = in + (size_t)c[u] * ld + off + 8 * q;
                v0[u] = *reinterpret_cast<const float4 *>(src);
                v1[u] = *reinterpret_cast<const float4 *>(src + 4);
            }
#pragma unroll
            for (int u = 0; u < G; u++) {
                acc[0] += wg[u] * v0[u].x;
                acc[1] += wg[u] * v0[u].y;
                acc[2] += wg[u] * v0[u].z;
                acc[3] += wg[u] * v0[u].w;
                acc[4] += wg[u] * v1[u].x;
                acc[5] += wg[u] * v1[u].y;
                acc[6] += wg[u] * v1[u].z;
                acc[7] += wg[u] * v1[u].w;
            }
        });
    }
    if constexpr (kBwd) {
        const float *own = reinterpret_cast<const float *>(in) + (size_t)row * ld + 8 * q;
        float *o = reinterpret_cast<float *>(out) + (size_t)row * kHid + 8 * q;
        const bf16_t *mk = mask + (size_t)row * kHid + 8 * q;
        for (int k = 0; k < 8; k++) o[k] = bf2f(mk[k]) > 0.0f ? own[k] + acc[k] : 0.0f;
    } else {
        bf16_t hv[8];
        for (int k = 0; k < 8; k++) hv[k] = f2bf(acc[k]);
        uint4 pk;
        pk.x = hv[0] | ((uint32_t)hv[1] << 16);
        pk.y = hv[2] | ((uint32_t)hv[3] << 16);
        pk.z = hv[4] | ((uint32_t)hv[5] << 16);
        pk.w = hv[6] | ((uint32_t)hv[7] << 16);
        *reinterpret_cast<uint4 *>(reinterpret_cast<bf16_t *>(out) + (size_t)row * kHid + 8 * q) = pk;
    }
}

// Dense layer on MFMA: h_out = relu([h | g] Wcat + b), one wave per 32 rows, two waves per
// block (one slice).  v_mfma_f32_32x32x8_bf16: lane l supplies A[l%32][4(l/32) + 0..3] and
// B[4(l/32) + 0..3][l%32]; it receives D[8(v/4) + 4(l/32) + v%4][l%32], v = 0..15.
__global__ __launch_bounds__(128) void k_mpnn_dense(MpnnArgs a, int layer, const bf16_t *__restrict__ H,
                                                    const bf16_t *__restrict__ G, bf16_t *__restrict__ Hout) {
    const int slice = blockIdx.x;
    if (slice >= a.nslices || blob_done(a, slice)) return;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row0 = slice * 64 + wv * 32;
    const int r = row0 + (lane & 31), kq = 4 * (lane >> 5);
    const bf16_t *wT = a.wT + (size_t)layer * kHid * 64;
    f32x16_t acc = {0};
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const bf16_t *src = s < 4 ? H : G;
        const bf16x4_t av = *reinterpret_cast<const bf16x4_t *>(src + (size_t)r * kHid + (s & 3) * 8 + kq);
        const bf16x4_t bv = *reinterpret_cast<const bf16x4_t *>(wT + (lane & 31) * 64 + s * 8 + kq);
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, acc, 0, 0, 0);
    }
    const int j = lane & 31;
    const float bj = a.wf[kOffLay + layer * kLayStride + 2 * kHid * kHid + j];
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int i = 8 * (v >> 2) + kq + (v & 3);
        Hout[(size_t)(row0 + i) * kHid + j] = f2bf(fmaxf(acc[v] + bj, 0.0f));
    }
}

// One message-passing layer fused: h_out = relu([h | A h] Wcat + b) for one 64-row slice per
// block.  Threads 4r..4r+3 gather channels 8q..8q+7 of row r's message into LDS (bf16, as
// k_mpnn_spmm stores it), then waves 0 and 1 run the dense layer on MFMA with the h half of the A
// operand from memory and the message half from LDS -- the message never goes to HBM.
__global__ __launch_bounds__(256) void k_mpnn_layer(MpnnArgs a, int layer, const bf16_t *__restrict__ H,
                                                    bf16_t *__restrict__ Hout) {
    __shared__ bf16_t gs[64][kHid + 8];  // +8: rows 80 bytes apart, 8-byte aligned reads
    const int slice = xcd_block(a.nslices);
    if (slice >= a.nslices || blob_done(a, slice)) return;
    {
        const int rloc = threadIdx.x >> 2, q = threadIdx.x & 3;
        const int w = a.slice_width[slice];
        const int64_t base = a.slice_off[slice];
        float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        message_rows(a, base, w, rloc, H + 8 * q, acc);
        for (int k = 0; k < 8; k++) gs[rloc][8 * q + k] = f2bf(acc[k]);
    }
    __syncthreads();
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (wv >= 2) return;
    const int row0 = slice * 64 + wv * 32;
    const int rl = wv * 32 + (lane & 31), kq = 4 * (lane >> 5);
    const int r = slice * 64 + rl;
    const bf16_t *wT = a.wT + (size_t)layer * kHid * 64;
    f32x16_t acc = {0};
#pragma unroll
    for (int s = 0; s < 8; s++) {
        bf16x4_t av;
        if (s < 4) {
            av = *reinterpret_cast<const bf16x4_t *>(H + (size_t)r * kHid + s * 8 + kq);
        } else {
            const bf16_t *g = &gs[rl][(s - 4) * 8 + kq];
            av[0] = (short)g[0];
            av[1] = (short)g[1];
            av[2] = (short)g[2];
            av[3] = (short)g[3];
        }
        const bf16x4_t bv = *reinterpret_cast<const bf16x4_t *>(wT + (lane & 31) * 64 + s * 8 + kq);
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, acc, 0, 0, 0);
    }
    const int j = lane & 31;
    const float bj = a.wf[kOffLay + layer * kLayStride + 2 * kHid * kHid + j];
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int i = 8 * (v >> 2) + kq + (v & 3);
        Hout[(size_t)(row0 + i) * kHid + j] = f2bf(fmaxf(acc[v] + bj, 0.0f));
    }
}

// Output layer inside CG, one wave per slice (one lane per row): z = u + s_b (Wo h3 + bo), 0 on
// pinned rows, with the partials of z.r and z.r_prev.
template <typename V>
__global__ __launch_bounds__(64) void k_mpnn_out_cg1(MpnnArgs a, const bf16_t *__restrict__ H3, const double *parts_uu,
                                                     V *__restrict__ z, const V *__restrict__ r,
                                                     const V *__restrict__ r_prev, double *parts_rz,
                                                     double *parts_zrp) {
    __shared__ double sh[4];
    const int slice = xcd_block(a.nslices);
    if (slice >= a.nslices || blob_done(a, slice)) return;
    const int b = a.slice_blob[slice];
    const double s = blob_scale(a, parts_uu, b, sh);
    const int row = slice * 64 + (int)threadIdx.x;
    float o[kNOut];
    for (int c = 0; c < kNOut; c++) o[c] = a.wf[kOffBo + c];
    const uint4 *hp = reinterpret_cast<const uint4 *>(H3 + (size_t)row * kHid);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const uint4 v = hp[q];
        const uint32_t vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const float h = bf2f((bf16_t)(k & 1 ? vv[k >> 1] >> 16 : vv[k >> 1] & 0xffffu));
            for (int c = 0; c < kNOut; c++) o[c] += a.wf[kOffWo + c * kHid + 8 * q + k] * h;
        }
    }
    const bool zero = a.pinned[row] || s == 0.0;
    double rz = 0.0, zp = 0.0;
    for (int c = 0; c < kNOut; c++) {
        const size_t k = 3 * (size_t)row + c;
        const double zz = (double)(V)(zero ? 0.0 : (double)z[k] + s * (double)o[c]);
        z[k] = (V)zz;
        rz += zz * (double)r[k];
        if (r_prev) zp += zz * (double)r_prev[k];
    }
    block_partial_at(rz, parts_rz, slice, sh);
    if (r_prev) block_partial_at(zp, parts_zrp, slice, sh);
}

// Training: u = D^-1 r for a harvested float residual, and the partials of |u|^2.
__global__ __launch_bounds__(192) void k_mpnn_prep(int nslices, const float *__restrict__ r,
                                                   const double *__restrict__ dinv, float *__restrict__ u,
                                                   double *parts_uu) {
    __shared__ double sh[4];
    const int slice = blockIdx.x;
    if (slice >= nslices) return;
    const size_t k = (size_t)slice * 192 + threadIdx.x;
    const double v = (double)r[k] * dinv[k];
    u[k] = (float)v;
    const double vf = (double)(float)v;
    block_partial_at(vf * vf, parts_uu, slice, sh);
}

// Training: output layer forward + loss + backward to dP3 = dL/dpre3, with per-block partials
// of the loss and of (dWo, dbo).  One block = one slice, one thread per row.
__global__ __launch_bounds__(64) void k_mpnn_out_train(MpnnArgs a, const bf16_t *__restrict__ H3,
                                                       const float *__restrict__ tgt, const double *parts_den,
                                                       float inv_blobs, float *__restrict__ dP3,
                                                       float *__restrict__ part_out, double *part_loss) {
    __shared__ double sh[4];
    __shared__ float red[kOutGrad][65];
    const int slice = blockIdx.x;
    const int b = a.slice_blob[slice];
    const double den = blob_total(parts_den, a.tot, kTotDEN, b, a.blob_slice_off[b], a.blob_slice_off[b + 1], sh);
    const int row = slice * 64 + threadIdx.x;
    const bf16_t *h = H3 + (size_t)row * kHid;
    float hv[kHid];
    for (int j = 0; j < kHid; j++) hv[j] = bf2f(h[j]);
    float dO[kNOut];
    double loss = 0.0;
    const bool live = !a.pinned[row] && den > 0.0;
    for (int c = 0; c < kNOut; c++) {
        float o = a.wf[kOffBo + c];
        for (int j = 0; j < kHid; j++) o += a.wf[kOffWo + c * kHid + j] * hv[j];
        const float d = o - tgt[3 * (size_t)row + c];
        dO[c] = live ? (float)(2.0 * d / den) * inv_blobs : 0.0f;
        if (live) loss += (double)d * d / den;
    }
    for (int j = 0; j < kHid; j++) {
        float g = 0.0f;
        for (int c = 0; c < kNOut; c++) g += a.wf[kOffWo + c * kHid + j] * dO[c];
        dP3[(size_t)row * kHid + j] = hv[j] > 0.0f ? g : 0.0f;
    }
    for (int c = 0; c < kNOut; c++) {
        for (int j = 0; j < kHid; j++) red[c * kHid + j][threadIdx.x] = dO[c] * hv[j];
        red[kNOut * kHid + c][threadIdx.x] = dO[c];
    }
    block_partial_at(loss * inv_blobs, part_loss, slice, sh);  // contains __syncthreads
    for (int p = threadIdx.x; p < kOutGrad; p += 64) {
        float s = 0.0f;
        for (int t = 0; t < 64; t++) s += red[p][t];
        part_out[(size_t)slice * kOutGrad + p] = s;
    }
}

// Training: dX = dP Wcat^T ([N][64] fp32) on MFMA, one wave per 32 rows, two output tiles.
__global__ __launch_bounds__(128) void k_mpnn_dgrad(MpnnArgs a, int layer, const float *__restrict__ dP,
                                                    float *__restrict__ dX) {
    const int slice = blockIdx.x;
    if (slice >= a.nslices) return;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row0 = slice * 64 + wv * 32;
    const int r = row0 + (lane & 31), kq = 4 * (lane >> 5);
    const bf16_t *wN = a.wN + (size_t)layer * 64 * kHid;
    f32x16_t acc0 = {0}, acc1 = {0};
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const float4 d = *reinterpret_cast<const float4 *>(dP + (size_t)r * kHid + s * 8 + kq);
        bf16x4_t av;
        av[0] = (short)f2bf(d.x);
        av[1] = (short)f2bf(d.y);
        av[2] = (short)f2bf(d.z);
        av[3] = (short)f2bf(d.w);
        const bf16x4_t b0 = *reinterpret_cast<const bf16x4_t *>(wN + (lane & 31) * kHid + s * 8 + kq);
        const bf16x4_t b1 = *reinterpret_cast<const bf16x4_t *>(wN + (32 + (lane & 31)) * kHid + s * 8 + kq);
        acc0 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, b0, acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, b1, acc1, 0, 0, 0);
    }
    const int j = lane & 31;
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int i = 8 * (v >> 2) + kq + (v & 3);
        dX[(size_t)(row0 + i) * 64 + j] = acc0[v];
        dX[(size_t)(row0 + i) * 64 + 32 + j] = acc1[v];
    }
}

// Training: per-block partials of dWcat^T = [h | g]^T dP (64 x 32) and db (32) over
// kWgradRows rows, on MFMA: wave 0 the h half, wave 1 the g half (A[m][n] = X[n][m],
// B[n][j] = dP[n][j]); wave 0 also sums db.
__global__ __launch_bounds__(128) void k_mpnn_wgrad(int N, const bf16_t *__restrict__ H, const bf16_t *__restrict__ G,
                                                    const float *__restrict__ dP, float *__restrict__ part) {
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int n0 = blockIdx.x * kWgradRows;
    const bf16_t *X = wv == 0 ? H : G;
    const int m = lane & 31, kq = 4 * (lane >> 5);
    f32x16_t acc = {0};
    float db = 0.0f;
    for (int s = 0; s < kWgradRows / 8; s++) {
        const int nb = n0 + s * 8 + kq;
        bf16x4_t av, bv;
        for (int i = 0; i < 4; i++) {
            const int n = nb + i;
            const bool in = n < N;
            av[i] = in ? (short)X[(size_t)n * kHid + m] : (short)0;
            const float d = in ? dP[(size_t)n * kHid + m] : 0.0f;
            bv[i] = (short)f2bf(d);
            db += d;
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, acc, 0, 0, 0);
    }
    float *out = part + (size_t)blockIdx.x * kLayGrad;
    // D[mrow][j]: mrow = input feature (of h or g), j = hidden output
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const int mrow = 8 * (v >> 2) + kq + (v & 3);
        out[(size_t)(wv * kHid + mrow) * kHid + m] = acc[v];  // layout [k][j]: (W1|W2)^T
    }
    if (wv == 0) {
        db += __shfl_down(db, 32, 64);
        if (lane < 32) out[2 * kHid * kHid + lane] = db;
    }
}

// Training: per-block partials of dWe (32 x 4) and dbe over 256 rows.
__global__ __launch_bounds__(256) void k_mpnn_embed_grad(int N, const float *__restrict__ dP0,
                                                         const float *__restrict__ feat, float *__restrict__ part) {
    __shared__ float red[8][kEmbGrad];
    const int j = threadIdx.x & 31, g = threadIdx.x >> 5;
    float acc[kNIn + 1] = {0, 0, 0, 0, 0};
    for (int rr = g; rr < 256; rr += 8) {
        const int n = blockIdx.x * 256 + rr;
        if (n >= N) break;
        const float d = dP0[(size_t)n * kHid + j];
        for (int c = 0; c < kNIn; c++) acc[c] += d * feat[(size_t)n * kNIn + c];
        acc[kNIn] += d;
    }
    for (int c = 0; c < kNIn; c++) red[g][j * kNIn + c] = acc[c];
    red[g][kHid * kNIn + j] = acc[kNIn];
    __syncthreads();
    for (int p = threadIdx.x; p < kEmbGrad; p += 256) {
        float s = 0.0f;
        for (int t = 0; t < 8; t++) s += red[t][p];
        part[(size_t)blockIdx.x * kEmbGrad + p] = s;
    }
}

// Embedding backward: dP0 = (dX[:, :32] + A dX[:, 32:]) masked -- k_mpnn_spmm<float, true>.

// Fixed-order sum of per-block partials: out[dst[p]] (+)= sum_b part[b * P + p].
// For the layer partials (stored [k][j]) dst maps to the parameter layout W[j][k].
__global__ __launch_bounds__(256) void k_mpnn_reduce(const float *__restrict__ part, int nblk, int P, int kind,
                                                     int layer, float *__restrict__ grad) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= P) return;
    float s = 0.0f;
    for (int b = 0; b < nblk; b++) s += part[(size_t)b * P + p];
    int dst;
    if (kind == 0) {  // layer: [k][j] for k < 64, then db
        if (p < 2 * kHid * kHid) {
            const int k = p / kHid, j = p % kHid;
            dst = kOffLay + layer * kLayStride + (k < kHid ? j * kHid + k : kHid * kHid + j * kHid + (k - kHid));
        } else {
            dst = kOffLay + layer * kLayStride + p;  // db follows W1, W2
        }
    } else if (kind == 1) {  // output: dWo [3][32], dbo [3]
        dst = kOffWo + p;
    } else {  // embedding: dWe [32][4], dbe [32]
        dst = kOffWe + p;
    }
    grad[dst] = s;
}

__global__ void k_mpnn_sum_loss(const double *parts, int n, float *out) {
    __shared__ double sh[4];
    const double s = parts_range(parts, 0, n, sh);
    if (threadIdx.x == 0) *out += (float)s;
}

// Adam (fp32 master weights), then the bf16 operand copies of the dense layers.
__global__ __launch_bounds__(256) void k_mpnn_adam(float *__restrict__ w, float *__restrict__ m1, float *__restrict__ m2,
                                                   const float *__restrict__ g, float lr, float c1, float c2) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= kNumParams) return;
    const float gv = g[p];
    const float a = 0.9f * m1[p] + 0.1f * gv;
    const float b = 0.999f * m2[p] + 0.001f * gv * gv;
    m1[p] = a;
    m2[p] = b;
    w[p] -= lr * (a * c1) / (sqrtf(b * c2) + 1e-8f);
}

__global__ __launch_bounds__(256) void k_mpnn_pack(const float *__restrict__ w, bf16_t *__restrict__ wT,
                                                   bf16_t *__restrict__ wN) {
    const int t = blockIdx.x * 256 + threadIdx.x;  // layer * 2048 + k * 32 + j
    if (t >= kNLay * 64 * kHid) return;
    const int l = t / (64 * kHid), k = (t / kHid) % 64, j = t % kHid;
    const float *L = w + kOffLay + l * kLayStride;
    const float v = k < kHid ? L[j * kHid + k] : L[kHid * kHid + j * kHid + (k - kHid)];
    const bf16_t bv = f2bf(v);
    wT[(size_t)l * kHid * 64 + j * 64 + k] = bv;
    wN[(size_t)l * 64 * kHid + k * kHid + j] = bv;
}

// u = D^-1 r into z with the partials of |u|^2, for a residual given from the host.
__global__ __launch_bounds__(192) void k_mpnn_u(int nslices, const double *__restrict__ dinv, const double *__restrict__ r,
                                                double *__restrict__ z, double *parts_uu) {
    __shared__ double sh[4];
    const int slice = blockIdx.x;
    if (slice >= nslices) return;
    const size_t k = (size_t)slice * 192 + threadIdx.x;
    const double u = dinv[k] * r[k];
    z[k] = u;
    block_partial_at(u * u, parts_uu, slice, sh);
}

// Harvest: snapshot k of a plain solve -- r_k and x_k as float.
template <typename V>
__global__ __launch_bounds__(256) void k_mpnn_harvest(int n3, const V *__restrict__ r, const double *__restrict__ x,
                                                      float *__restrict__ rs, float *__restrict__ xs) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n3) return;
    rs[k] = (float)r[k];
    xs[k] = (float)x[k];
}

// After the solve: e_k = x* - x_k for the iterations each blob really made (snapshots at or
// beyond a blob's final iteration are zeroed: no pair).
__global__ __launch_bounds__(192) void k_mpnn_harvest_finish(int nslices, const int32_t *slice_blob, const CgCtl *ctl,
                                                             int snap, const double *__restrict__ x,
                                                             float *__restrict__ rs, float *__restrict__ xs) {
    const int slice = blockIdx.x;
    if (slice >= nslices) return;
    const size_t k = (size_t)slice * 192 + threadIdx.x;
    if (snap >= ctl[slice_blob[slice]].iter) {
        rs[k] = 0.0f;
        xs[k] = 0.0f;
    } else {
        xs[k] = (float)(x[k] - (double)xs[k]);
    }
}
