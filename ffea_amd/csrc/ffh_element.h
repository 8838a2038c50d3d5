// ffh_element.h -- the element kernel and the assembly of the viscosity matrix.
// Part of ffh_kernels.hip: included inside its anonymous namespace, after the headers it
// lists there; not a stand-alone header.
#pragma once

// ----------------------------------------------------------------------------------------
// Element kernel
// ----------------------------------------------------------------------------------------
struct ElemArgs {
    const double4 *pos;  // (x, y, z, 0) per linear node: one 32-byte load per node
    const int4 *enodes;
    const double *jinv0;  // SoA [9][E]
    const double *vol0, *A, *B, *G, *Em;
    // the same per-element constants rounded to float, SoA [14][E] (jinv0 0-8, vol0, A, B, G,
    // Em): what the FP32 element kernel reads (half the bytes); nullptr on the FP64 path
    const float *elem32;
    int8_t *last_sign;  // [E] sign of the element's determinant at the last evaluation
    void *eforce;  // T SoA [12][E]: node_force[a][i] at (3a+i)*E + e
    int E;
    int calc_noise;
    double kT, dt;
    const uint32_t *rng_state;
    const Mat3u *jump;
    uint32_t elem_base;
    const uint32_t *wave_state;  // [6 per 64 elements]: the stream at each wave's first element
    CgCtl *ctl;
    const double *vel;  // [3N] node velocities (mass-ful solvers: the viscous force K_e u_e)
    double *stress_mag;  // [E] internal_stress_mag per element, or nullptr (not computed)
};

// Entry (R, C) of the element viscosity matrix (create_viscosity_matrix,
// src/tetra_element_linear.cpp:299-328, with the macros of include/tetra_element_linear.h:55-116
// and the diffusion term :599-638): the upper triangle as built, the lower one its mirror.
template <typename R>
__device__ __forceinline__ R visc_entry(const R *dpsi, R A, R B, R vol, int Rr, int Cc) {
    if (Rr > Cc) {
        int t = Rr;
        Rr = Cc;
        Cc = t;
    }
    const int i = Rr >> 2, a = Rr & 3, j = Cc >> 2, b = Cc & 3;
    R v;
    if (i == j) {
        R del2 = dpsi[a] * dpsi[b] + dpsi[4 + a] * dpsi[4 + b] + dpsi[8 + a] * dpsi[8 + b];
        v = dpsi[Rr] * dpsi[Cc] * (A + B);
        v += del2 * A;
    } else if (a == b) {
        v = (B + A) * (dpsi[Rr] * dpsi[Cc]);
    } else {
        v = B * (dpsi[Rr] * dpsi[Cc]) + A * (dpsi[4 * i + b] * dpsi[4 * j + a]);
    }
    return v * vol;
}

// The element geometry from its four vertex positions: see element_geometry.
template <typename R>
__device__ __forceinline__ void element_geometry_p(const double4 p0, const double4 p1, const double4 p2, const double4 p3,
                                                   R J[3][3], R dpsi[12], R &det) {
    const double4 pr[3] = {p1, p2, p3};
#pragma unroll
    for (int r = 0; r < 3; r++) {
        J[r][0] = (R)(pr[r].x - p0.x);
        J[r][1] = (R)(pr[r].y - p0.y);
        J[r][2] = (R)(pr[r].z - p0.z);
    }
    dpsi[1] = J[2][2] * J[1][1] - J[2][1] * J[1][2];
    dpsi[2] = J[2][1] * J[0][2] - J[2][2] * J[0][1];
    dpsi[3] = J[1][2] * J[0][1] - J[1][1] * J[0][2];
    dpsi[5] = J[2][0] * J[1][2] - J[2][2] * J[1][0];
    dpsi[6] = J[2][2] * J[0][0] - J[2][0] * J[0][2];
    dpsi[7] = J[1][0] * J[0][2] - J[1][2] * J[0][0];
    dpsi[9] = J[2][1] * J[1][0] - J[2][0] * J[1][1];
    dpsi[10] = J[2][0] * J[0][1] - J[2][1] * J[0][0];
    dpsi[11] = J[1][1] * J[0][0] - J[1][0] * J[0][1];
    det = J[0][0] * dpsi[1] + J[1][0] * dpsi[2] + J[2][0] * dpsi[3];
    const R dinv = (R)1.0 / det;
    dpsi[1] *= dinv;
    dpsi[2] *= dinv;
    dpsi[3] *= dinv;
    dpsi[5] *= dinv;
    dpsi[6] *= dinv;
    dpsi[7] *= dinv;
    dpsi[9] *= dinv;
    dpsi[10] *= dinv;
    dpsi[11] *= dinv;
    dpsi[0] = -(dpsi[1] + dpsi[2] + dpsi[3]);
    dpsi[4] = -(dpsi[5] + dpsi[6] + dpsi[7]);
    dpsi[8] = -(dpsi[9] + dpsi[10] + dpsi[11]);
}

// calculate_jacobian (src/tetra_element_linear.cpp:193-205), from FP64 positions, and
// calc_shape_function_derivatives_and_volume (:223-293): J, det and the shape gradients
// dpsi[a + 4 dim] (x-major).  Shared by the element kernel and the assembly, which recomputes
// them for every slice that needs the element: the same code rounds the same way.
template <typename R>
__device__ __forceinline__ void element_geometry(const double4 *__restrict__ pos, const int4 n, R J[3][3],
                                                 R dpsi[12], R &det) {
    element_geometry_p<R>(pos[n.x], pos[n.y], pos[n.z], pos[n.w], J, dpsi, det);
}

// The stream state at the first element of every 64-element wave, one thread per wave: the
// element kernel's lanes step on from it with the A^(7 lane) table, so no wave evaluates the
// four-digit jump itself (as uniform scalar code that jump had been a third of the element
// kernel's time at c1: the scalar unit is one per compute unit).
__global__ __launch_bounds__(kBlock) void k_elem_wave_states(const uint32_t *__restrict__ rng_state,
                                                            const Mat3u *__restrict__ jump, uint32_t elem_base,
                                                            int nwaves, uint32_t *__restrict__ out) {
    const int w = blockIdx.x * kBlock + threadIdx.x;
    if (w >= nwaves) return;
    const Stream6 st = rng_jump(rng_state, jump, elem_base + 7u * 64u * (uint32_t)w);
#pragma unroll
    for (int k = 0; k < 6; k++) out[6 * (size_t)w + k] = st.s[k];
}

template <typename R, typename T, bool kMass>
__global__ __launch_bounds__(kBlock) void k_element(ElemArgs g) {
    const int bd = (int)blockDim.x;  // 64, 128 or 256 (tuned); waves stay 64-aligned
    const int e = xcd_block((g.E + bd - 1) / bd) * bd + threadIdx.x;
    if (e >= g.E) return;
    R J[3][3], dpsi[12], det;
    element_geometry<R>(g.pos, g.enodes[e], J, dpsi, det);
    // an inversion is a change of the determinant's sign (last_det * det < 0), so its sign is
    // all that is kept
    const int sg = det > (R)0 ? 1 : (det < (R)0 ? -1 : 0);
    if ((int)g.last_sign[e] * sg < 0) atomicAdd(&g.ctl->inversions, 1);
    g.last_sign[e] = (int8_t)sg;
    const R vol = (R)(1.0 / 6.0) * fabs(det);

    R A, B, G, Em, vol0, Ji0[9];  // Ji0[3 j + k] = jinv0[j][k]
    if constexpr (sizeof(R) == 4) {
        const float *c = g.elem32 + e;
        const size_t E = (size_t)g.E;
#pragma unroll
        for (int q = 0; q < 9; q++) Ji0[q] = c[q * E];
        vol0 = c[9 * E];
        A = c[10 * E];
        B = c[11 * E];
        G = c[12 * E];
        Em = c[13 * E];
    } else {
#pragma unroll
        for (int q = 0; q < 9; q++) Ji0[q] = (R)g.jinv0[(size_t)q * g.E + e];
        vol0 = (R)g.vol0[e];
        A = (R)g.A[e];
        B = (R)g.B[e];
        G = (R)g.G[e];
        Em = (R)g.Em[e];
    }

    // add_shear_elastic_stress (:378-396): F_ij = sum_k J[k][i] Jinv0[j][k]; S = F F^T
    R F[3][3], S[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            R f = (R)0;
#pragma unroll
            for (int k = 0; k < 3; k++) f += J[k][i] * Ji0[3 * j + k];
            F[i][j] = f;
        }
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            R s = (R)0;
#pragma unroll
            for (int k = 0; k < 3; k++) s += F[i][k] * F[j][k];
            S[i][j] = s;
        }
    const R sc = (G * vol0 / vol);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) S[i][j] *= sc;
    S[0][0] -= G;
    S[1][1] -= G;
    S[2][2] -= G;
    // add_bulk_elastic_stress (:401-408)
    {
        const R c_2 = Em - G * (R)2.0 / (R)3.0;
        const R c = G * ((R)1.0 - (vol0 / vol)) + (R)0.5 * c_2 * ((vol / vol0) - (vol0 / vol));
        S[0][0] += c;
        S[1][1] += c;
        S[2][2] += c;
    }
    // add_fluctuating_stress (:417-436): 7 draws at this element's stream offset
    if (g.calc_noise) {
        // the wave's first element's state (k_elem_wave_states), then A^(7 lane) for this lane
        const int lane = threadIdx.x & 63;
        const uint32_t w = __builtin_amdgcn_readfirstlane((uint32_t)(e >> 6));
        Stream6 st;
#pragma unroll
        for (int k = 0; k < 6; k++) st.s[k] = g.wave_state[6 * w + k];
        rng_apply(st, g.jump[kLaneTable + 2 * lane], g.jump[kLaneTable + 2 * lane + 1]);
        const R c = sqrt(((R)24 * (R)g.kT) / (vol * (R)g.dt));
        const R bf = sqrt(B) * (R)rand_half(st);
        S[0][0] += c * (sqrt((R)2 * A) * (R)rand_half(st) + bf);
        S[1][1] += c * (sqrt((R)2 * A) * (R)rand_half(st) + bf);
        S[2][2] += c * (sqrt((R)2 * A) * (R)rand_half(st) + bf);
        S[0][1] += c * sqrt(A) * (R)rand_half(st);
        S[0][2] += c * sqrt(A) * (R)rand_half(st);
        S[1][2] += c * sqrt(A) * (R)rand_half(st);
        S[1][0] = S[0][1];
        S[2][0] = S[0][2];
        S[2][1] = S[1][2];
    }
    // internal_stress_mag (src/Blob.cpp:700): sqrt of mat3_double_contraction_symmetric
    // (src/mat_vec_fns.cpp:131-134) of the stress with its noise
    if (g.stress_mag)
        g.stress_mag[e] = (double)sqrt(S[0][0] * S[0][0] + S[1][1] * S[1][1] + S[2][2] * S[2][2] +
                                       (R)2 * (S[0][1] * S[0][1] + S[0][2] * S[0][2] + S[1][2] * S[1][2]));
    // The mass-ful solvers start du from the viscous force K_e u_e (src/Blob.cpp:703-706):
    // get_element_velocity_vector (:454-469, x-major over the vertices) and mat12_apply
    // (src/mat_vec_fns.cpp:26-34, each row summed in column order from zero); the massless
    // solver starts it from zero.
    R kv[12];
    if constexpr (kMass) {
        const int4 nn = g.enodes[e];
        const int ni[4] = {nn.x, nn.y, nn.z, nn.w};
        R u[12];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int a = 0; a < 4; a++) u[4 * i + a] = (R)g.vel[3 * (size_t)ni[a] + i];
#pragma unroll
        for (int rr = 0; rr < 12; rr++) {
            R acc = (R)0;
#pragma unroll
            for (int cc = 0; cc < 12; cc++) acc += visc_entry<R>(dpsi, A, B, vol, rr, cc) * u[cc];
            kv[rr] = acc;
        }
    } else {
#pragma unroll
        for (int k = 0; k < 12; k++) kv[k] = (R)0;
    }
    // apply_stress_tensor (:441-449) + add_element_force_vector (:474-490)
    T fo[12];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int a = 0; a < 4; a++) {
            R du = kv[4 * i + a] + vol * (dpsi[a] * S[i][0] + dpsi[4 + a] * S[i][1] + dpsi[8 + a] * S[i][2]);
            fo[3 * a + i] = (T)((R)0 - du);
        }
    T *ef = reinterpret_cast<T *>(g.eforce);
#pragma unroll
    for (int k = 0; k < 12; k++) ef[(size_t)k * g.E + e] = fo[k];
}

// ----------------------------------------------------------------------------------------
// Assembly of the block pattern, one workgroup per 64-row slice
// ----------------------------------------------------------------------------------------
struct AsmArgs {
    const double4 *pos;  // positions: the element geometry is recomputed from them
    const int4 *enodes;
    const double *A, *B;
    const int64_t *eoff;      // [slices + 1] into elem
    const int32_t *elem;      // per slice, the distinct elements its rows touch, ascending
    const int32_t *woff;      // [slices + 1] work tiles of each slice (the first: its diagonal tile)
    const int64_t *wcoff;     // [work tiles + 1] first code of each work tile, 128-byte aligned
    const uint16_t *code;     // slice-local element << 4 | a << 2 | b, tile-major [k][64]
    const uint8_t *cnt;       // [work tiles * 64] contributions
    const int32_t *aux;       // [work tiles * 64] diagonal SELL position / mirror slot, -1: none
    const uint8_t *wpos;      // [work tiles * 64] SELL position of an upper block
    const int64_t *slice_off; // [slices + 1] first slot; slot = slice_off[s] + 64 t + lane
    const uint8_t *flags;     // [slots] ffh::kSlotDiag / kSlotIdentity / kSlotPad
    const double *zeta;
    void *val;  // T tiles
    double *dinv;
    int nslices;
};

// SparseMatrixFixedPattern::build (src/SparseMatrixFixedPattern.cpp:85-93): each entry is the
// sum of its sources in registration order -- ascending element, Stokes drag last
// (src/NoMassCGSolver.cpp:93-139, src/SparseMatrixTypes.cpp:52-58); diagonal blocks also give
// calc_inverse_diagonal (:325-333).
//
// Element block (a, b) (row vertex a, column vertex b), entry (i, j) = V[a+4i][b+4j] with
// g_a = grad psi_a, g_b = grad psi_b (dpsi[4i+a] = g_a[i]) and P_ij = g_a[i] g_b[j]; these are
// visc_entry's expressions, every product written in an order that rounds identically:
//   i == j : (P_ii (A+B) + ((P_00 + P_11) + P_22) A) vol
//   i != j : a == b ? ((B+A) P_ij) vol : (B P_ij + A P_ji) vol
// a == b exactly on the diagonal blocks.  Block (b, a) of the same element at entry (j, i)
// multiplies the same pairs of factors (IEEE products and sums of two commute), so a lower block
// (column < row), summed over the same elements in the same order, is its upper mirror's
// transpose bit for bit: only the diagonal and the upper blocks are summed, and each upper block
// is stored twice.
//
// One workgroup per slice.  The slice's distinct elements are evaluated once each, kThreads at
// a time, into LDS (shape gradients, volume, A, B: element_geometry, the element kernel's own
// code, so they round as there; element ids are loaded one chunk ahead and nodes are gathered
// while the previous chunk's geometry is computed).  Wave 0 sums the 64 diagonal blocks (six
// distinct entries each: a diagonal block is symmetric); waves 1.. sum the upper work tiles
// (the upper blocks of SELL tiles whose lanes do not meet share one), kTpt each per pass.  A
// lane adds the contributions of its list that fall in the chunk, reading the next code while it
// sums the current one.  kLdsCodes: the slice's contribution codes (at most kAsmCodeLds) are
// first copied to LDS with 16-byte loads.
constexpr int kAsmCodeLds = 9216;

template <typename T, int kThreads>
__device__ __forceinline__ void asm_add_upper(const T (&lds)[15][kThreads], uint32_t code, int le, T acc[9]) {
    const int a = (code >> 2) & 3, b = code & 3;
    T ga[3], gb[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        ga[i] = lds[a + 4 * i][le];
        gb[i] = lds[b + 4 * i][le];
    }
    const T vol = lds[12][le], A = lds[13][le], B = lds[14][le];
    T P[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) P[i][j] = ga[i] * gb[j];
    const T del2 = P[0][0] + P[1][1] + P[2][2];
    const T AB = A + B, dA = del2 * A;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const T v = i == j ? P[i][i] * AB + dA : B * P[i][j] + A * P[j][i];
            acc[3 * i + j] += v * vol;
        }
}

// Diagonal block (a == b): entries 00, 11, 22, 01, 02, 12 (the block is symmetric).
template <typename T, int kThreads>
__device__ __forceinline__ void asm_add_diag(const T (&lds)[15][kThreads], uint32_t code, int le, T d[6]) {
    const int a = (code >> 2) & 3;
    const T g0 = lds[a][le], g1 = lds[a + 4][le], g2 = lds[a + 8][le];
    const T vol = lds[12][le], A = lds[13][le], B = lds[14][le];
    const T P00 = g0 * g0, P11 = g1 * g1, P22 = g2 * g2;
    const T del2 = P00 + P11 + P22;
    const T AB = A + B, dA = del2 * A;
    d[0] += (P00 * AB + dA) * vol;
    d[1] += (P11 * AB + dA) * vol;
    d[2] += (P22 * AB + dA) * vol;
    d[3] += (AB * (g0 * g1)) * vol;
    d[4] += (AB * (g0 * g2)) * vol;
    d[5] += (AB * (g1 * g2)) * vol;
}

template <typename T, int kWaves, int kTpt, bool kLdsCodes>
__global__ __launch_bounds__(64 * kWaves) void k_assemble_slice(AsmArgs g) {
    static_assert(kWaves >= 2, "wave 0 sums the diagonal blocks, the others the upper tiles");
    constexpr int kThreads = 64 * kWaves;
    constexpr int kUp = (kWaves - 1) * kTpt;  // upper tiles per pass
    __shared__ T lds[15][kThreads];
    __shared__ __align__(16) uint16_t lcode[kLdsCodes ? kAsmCodeLds : 8];
    const int s = xcd_block(g.nslices);
    if (s >= g.nslices) return;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int64_t e0 = g.eoff[s];
    const int ne = (int)(g.eoff[s + 1] - e0);
    const int g0 = g.woff[s], nw = g.woff[s + 1] - g0;
    const int64_t slot0 = g.slice_off[s];
    const int64_t cb = g.wcoff[g0];  // the slice's codes: [cb, wcoff[g0 + nw]), 128-byte aligned
    if constexpr (kLdsCodes) {
        const int n8 = (int)((g.wcoff[g0 + nw] - cb) >> 3);
        const uint4 *src = reinterpret_cast<const uint4 *>(g.code + cb);
        uint4 *dst = reinterpret_cast<uint4 *>(lcode);
        for (int i = tid; i < n8; i += kThreads) dst[i] = src[i];  // visible after the first barrier
    }
    auto code_at = [&](int64_t i) -> uint32_t {
        if constexpr (kLdsCodes)
            return lcode[i];
        else
            return g.code[i];
    };
    T *val = reinterpret_cast<T *>(g.val);
    const int npass = nw > 1 ? (nw - 1 + kUp - 1) / kUp : 1;
    for (int pass = 0; pass < npass; pass++) {
        // this wave's work tiles: wave 0 the diagonal tile (first pass), wave w > 0 the upper
        // tiles 1 + pass kUp + (w - 1) + (kWaves - 1) j
        int gj[kTpt], k[kTpt], cnt[kTpt];
        int64_t co[kTpt];  // this lane's first code: in LDS relative to cb, else in global memory
        T acc[kTpt][9];
#pragma unroll
        for (int j = 0; j < kTpt; j++) {
            const int r = wave == 0 ? (pass == 0 && j == 0 ? 0 : nw) : 1 + pass * kUp + (wave - 1) + (kWaves - 1) * j;
            gj[j] = r < nw ? g0 + r : -1;
            cnt[j] = gj[j] >= 0 ? (int)g.cnt[(int64_t)gj[j] * 64 + lane] : 0;
            co[j] = (gj[j] >= 0 ? g.wcoff[gj[j]] - (kLdsCodes ? cb : 0) : 0) + lane;
            k[j] = 0;
#pragma unroll
            for (int q = 0; q < 9; q++) acc[j][q] = (T)0.0;
        }
        // Every lane loads an element each chunk (past the end, the slice's last one again:
        // loaded, not used), so no load sits in a branch and none is waited for early.
        auto eid = [&](int i) { return g.elem[e0 + min(i, ne - 1)]; };
        int id_c = 0, id_n = 0;
        int4 en_c = make_int4(0, 0, 0, 0);
        if (ne > 0) {
            id_c = eid(tid);
            id_n = eid(kThreads + tid);
            en_c = g.enodes[id_c];
        }
        for (int c0 = 0; c0 < ne; c0 += kThreads) {
            const double4 p0 = g.pos[en_c.x], p1 = g.pos[en_c.y], p2 = g.pos[en_c.z], p3 = g.pos[en_c.w];
            const double Ae = g.A[id_c], Be = g.B[id_c];
            // the next chunk's nodes and the one after's ids, in flight while this one is summed
            const int4 en_n = g.enodes[id_n];
            const int id_nn = eid(c0 + 2 * kThreads + tid);
            if (c0 + tid < ne) {
                T J[3][3], dpsi[12], det;
                element_geometry_p<T>(p0, p1, p2, p3, J, dpsi, det);
#pragma unroll
                for (int q = 0; q < 12; q++) lds[q][tid] = dpsi[q];
                lds[12][tid] = (T)(1.0 / 6.0) * fabs(det);
                lds[13][tid] = (T)Ae;
                lds[14][tid] = (T)Be;
            }
            __syncthreads();
            const uint32_t lim = (uint32_t)(c0 + kThreads) << 4;
#pragma unroll
            for (int j = 0; j < kTpt; j++) {
                if (wave == 0 && j > 0) break;
                // this lane's codes from k[j] on while they lie in the chunk (codes ascend); the
                // next code is read while the current contribution is summed
                int kk = k[j];
                if (kk < cnt[j]) {
                    const int last = cnt[j] - 1;
                    uint32_t code = code_at(co[j] + 64 * kk);
                    if (wave == 0) {
                        while (kk <= last && code < lim) {
                            const uint32_t nxt = code_at(co[j] + 64 * min(kk + 1, last));
                            asm_add_diag<T, kThreads>(lds, code, (int)(code >> 4) - c0, acc[0]);
                            code = nxt;
                            kk++;
                        }
                    } else {
                        while (kk <= last && code < lim) {
                            const uint32_t nxt = code_at(co[j] + 64 * min(kk + 1, last));
                            asm_add_upper<T, kThreads>(lds, code, (int)(code >> 4) - c0, acc[j]);
                            code = nxt;
                            kk++;
                        }
                    }
                }
                k[j] = kk;
            }
            __syncthreads();
            id_c = id_n;
            id_n = id_nn;
            en_c = en_n;
        }
        if (wave == 0) {
            if (gj[0] < 0) continue;
            const int td = g.aux[(int64_t)gj[0] * 64 + lane];
            if (td < 0) continue;  // a padding row past the last node
            const int64_t tile = slot0 + (int64_t)td * 64;
            const uint8_t fl = g.flags[tile + lane];
            if (fl & ffh::kSlotPad) continue;
            const int node = s * 64 + lane;
            const T *d = acc[0];
            T b[9] = {d[0], d[3], d[4], d[3], d[1], d[5], d[4], d[5], d[2]};
            if (fl & ffh::kSlotIdentity) {
#pragma unroll
                for (int q = 0; q < 9; q++) b[q] = (T)0.0;
                b[0] = b[4] = b[8] = (T)1.0;
            } else {
                const T z = (T)g.zeta[node];
                b[0] += z;
                b[4] += z;
                b[8] += z;
            }
            g.dinv[3 * (size_t)node + 0] = 1.0 / (double)b[0];
            g.dinv[3 * (size_t)node + 1] = 1.0 / (double)b[4];
            g.dinv[3 * (size_t)node + 2] = 1.0 / (double)b[8];
            store_block<T>(val + tile * 9, lane, b);
        } else {
#pragma unroll
            for (int j = 0; j < kTpt; j++) {
                if (gj[j] < 0) continue;
                const int m = g.aux[(int64_t)gj[j] * 64 + lane];
                if (m < 0) continue;  // no upper block in this lane
                const int64_t tile = slot0 + (int64_t)g.wpos[(int64_t)gj[j] * 64 + lane] * 64;
                store_block<T>(val + tile * 9, lane, acc[j]);
                const T tr[9] = {acc[j][0], acc[j][3], acc[j][6], acc[j][1], acc[j][4], acc[j][7], acc[j][2], acc[j][5], acc[j][8]};
                store_block<T>(val + (int64_t)(m & ~63) * 9, m & 63, tr);
            }
        }
    }
}
