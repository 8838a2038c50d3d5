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
    double *last_det;
    void *eforce;  // T SoA [12][E]: node_force[a][i] at (3a+i)*E + e
    int E;
    int calc_noise;
    double kT, dt;
    const uint32_t *rng_state;
    const Mat3u *jump;
    uint32_t elem_base;
    CgCtl *ctl;
    const double *vel;  // [3N] node velocities (mass-ful solvers: the viscous force K_e u_e)
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

// calculate_jacobian (src/tetra_element_linear.cpp:193-205), from FP64 positions, and
// calc_shape_function_derivatives_and_volume (:223-293): J, det and the shape gradients
// dpsi[a + 4 dim] (x-major).  Shared by the element kernel and the assembly, which recomputes
// them per contribution: the same code rounds the same way, so the matrix is unchanged.
template <typename R>
__device__ __forceinline__ void element_geometry(const double4 *__restrict__ pos, const int4 n, R J[3][3],
                                                 R dpsi[12], R &det) {
    const double4 p0 = pos[n.x], p1 = pos[n.y], p2 = pos[n.z], p3 = pos[n.w];
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

template <typename R, typename T, bool kMass>
__global__ __launch_bounds__(kBlock) void k_element(ElemArgs g) {
    const int bd = (int)blockDim.x;  // 64, 128 or 256 (tuned); waves stay 64-aligned
    const int e = xcd_block((g.E + bd - 1) / bd) * bd + threadIdx.x;
    if (e >= g.E) return;
    R J[3][3], dpsi[12], det;
    element_geometry<R>(g.pos, g.enodes[e], J, dpsi, det);
    const double last = g.last_det[e];
    if (last * (double)det < 0) atomicAdd(&g.ctl->inversions, 1);
    g.last_det[e] = (double)det;
    const R vol = (R)(1.0 / 6.0) * fabs(det);

    const R A = (R)g.A[e], B = (R)g.B[e], G = (R)g.G[e], Em = (R)g.Em[e], vol0 = (R)g.vol0[e];

    // add_shear_elastic_stress (:378-396): F_ij = sum_k J[k][i] Jinv0[j][k]; S = F F^T
    R F[3][3], S[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            R f = (R)0;
#pragma unroll
            for (int k = 0; k < 3; k++) f += J[k][i] * (R)g.jinv0[(size_t)(3 * j + k) * g.E + e];
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
        // the wave's first element's state (uniform), then A^(7 lane) for this lane
        const int lane = threadIdx.x & 63;
        const uint32_t e0 = __builtin_amdgcn_readfirstlane((uint32_t)(e - lane));
        Stream6 st = rng_jump(g.rng_state, g.jump, g.elem_base + 7u * e0);
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
// Assembly over the upper triangle of the block pattern
// ----------------------------------------------------------------------------------------
struct AsmArgs {
    const double4 *pos;  // positions: the element geometry is recomputed per contribution
    const uint32_t *flat, *eoff;  // contributions entry-major, [eoff[u], eoff[u+1]) for entry u
    const int4 *enodes;
    const double *A, *B;
    const int64_t *toff;
    const uint16_t *twidth, *cnt;
    const uint32_t *ctb;
    const int32_t *slot, *mirror, *node;
    const uint8_t *flags;
    const double *zeta;
    void *val;  // T tiles
    double *dinv;
    int64_t nupper;
};

// SparseMatrixFixedPattern::build (src/SparseMatrixFixedPattern.cpp:85-93): each entry is the
// sum of its sources in registration order -- ascending element, Stokes drag last
// (src/NoMassCGSolver.cpp:93-139, src/SparseMatrixTypes.cpp:52-58).  K is symmetric entry by
// entry (each element block pair is an exact mirror), so the lower block is the transposed
// upper one.  Diagonal blocks also give calc_inverse_diagonal (:325-333).
//
// Block (a,b) of an element, entry (i,j) = V[a+4i][b+4j] with g_a = grad psi_a, g_b = grad psi_b
// (dpsi[4i+a] = g_a[i], recomputed from the positions); these are the expressions of visc_entry with
// every product written in an order that rounds identically:
//   i == j : ((g_a[i] g_b[i]) (A+B) + (g_a.g_b) A) vol
//   i != j : a == b ? ((B+A)(g_a[i] g_a[j])) vol : (B (g_a[i] g_b[j]) + A (g_b[i] g_a[j])) vol
template <typename T, int kUnroll>
__global__ __launch_bounds__(kBlock) void k_assemble(AsmArgs g) {
    const int bd = (int)blockDim.x;  // 64, 128 or 256 (tuned)
    const int64_t u = (int64_t)xcd_block((int)((g.nupper + bd - 1) / bd)) * bd + threadIdx.x;
    if (u >= g.nupper) return;
    const uint8_t fl = g.flags[u];
    if (fl & 4) return;  // tile padding
    const bool diag = (fl & 1) != 0;
    T acc[9];
#pragma unroll
    for (int k = 0; k < 9; k++) acc[k] = (T)0.0;
    const int64_t grp = u >> 6;
    const int lane = (int)(u & 63);
    const int cnt = g.cnt[u];
    const uint32_t *tl = g.ctb + g.toff[grp] + lane;
#pragma unroll kUnroll
    for (int k = 0; k < cnt; k++) {
        const uint32_t code = tl[64 * k];
        const int a = (code >> 2) & 3, b = code & 3;
        const int e = (int)(code >> 4);
        T J[3][3], dpsi[12], det;
        element_geometry<T>(g.pos, g.enodes[e], J, dpsi, det);
        const T vol = (T)(1.0 / 6.0) * fabs(det);
        const T A = (T)g.A[e], B = (T)g.B[e];
        T ga[3], gb[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {  // selects, not a dynamic register index
            const T *d = dpsi + 4 * i;
            ga[i] = a == 0 ? d[0] : a == 1 ? d[1] : a == 2 ? d[2] : d[3];
            gb[i] = b == 0 ? d[0] : b == 1 ? d[1] : b == 2 ? d[2] : d[3];
        }
        const T del2 = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                T v;
                if (i == j) {
                    v = ga[i] * gb[i] * (A + B);
                    v += del2 * A;
                } else if (diag) {
                    v = (B + A) * (ga[i] * gb[j]);
                } else {
                    v = B * (ga[i] * gb[j]) + A * (gb[i] * ga[j]);
                }
                acc[3 * i + j] += v * vol;
            }
    }
    const int node = g.node[u];
    if (diag) {
        if (fl & 2) {
#pragma unroll
            for (int k = 0; k < 9; k++) acc[k] = (T)0.0;
            acc[0] = acc[4] = acc[8] = (T)1.0;
        } else {
            const T z = (T)g.zeta[node];
            acc[0] += z;
            acc[4] += z;
            acc[8] += z;
        }
        g.dinv[3 * (size_t)node + 0] = 1.0 / (double)acc[0];
        g.dinv[3 * (size_t)node + 1] = 1.0 / (double)acc[4];
        g.dinv[3 * (size_t)node + 2] = 1.0 / (double)acc[8];
    }
    T *v = reinterpret_cast<T *>(g.val);
    {
        const int64_t slot = g.slot[u];
        const int64_t tile = slot & ~(int64_t)63;
        const int lane = (int)(slot & 63);
#pragma unroll
        for (int k = 0; k < 9; k++) v[tile * 9 + k * 64 + lane] = acc[k];
    }
    const int32_t mslot = g.mirror[u];
    if (mslot >= 0) {
        const int64_t tile = (int64_t)mslot & ~(int64_t)63;
        const int lane = mslot & 63;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) v[tile * 9 + (3 * i + j) * 64 + lane] = acc[3 * j + i];
    }
}

// One contribution (element block (a,b) of code) of an upper block, entry by entry, premultiplied
// by vol as k_assemble adds it.  a == b exactly for the diagonal blocks.
template <typename T>
__device__ __forceinline__ void contribution(const AsmArgs &g, uint32_t code, T w[9]) {
    const int a = (code >> 2) & 3, b = code & 3;
    const int e = (int)(code >> 4);
    T J[3][3], dpsi[12], det;
    element_geometry<T>(g.pos, g.enodes[e], J, dpsi, det);
    const T vol = (T)(1.0 / 6.0) * fabs(det);
    const T A = (T)g.A[e], B = (T)g.B[e];
    T ga[3], gb[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const T *d = dpsi + 4 * i;
        ga[i] = a == 0 ? d[0] : a == 1 ? d[1] : a == 2 ? d[2] : d[3];
        gb[i] = b == 0 ? d[0] : b == 1 ? d[1] : b == 2 ? d[2] : d[3];
    }
    const T del2 = ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            T v;
            if (i == j) {
                v = ga[i] * gb[i] * (A + B);
                v += del2 * A;
            } else if (a == b) {
                v = (B + A) * (ga[i] * gb[j]);
            } else {
                v = B * (ga[i] * gb[j]) + A * (gb[i] * ga[j]);
            }
            w[3 * i + j] = v * vol;
        }
}

// Two-phase assembly: one workgroup per tile of 64 upper blocks.  Phase 1: all 256 threads
// evaluate the tile's contributions (entry-major, each block's in ascending element order) in
// chunks of 256, nine values each into LDS -- independent gathers in flight instead of one
// dependent chain per lane.  Phase 2: the lane of each block adds its contributions of the
// chunk in order.  The values and the order of additions are k_assemble's: bit-identical.
template <typename T>
__global__ __launch_bounds__(256) void k_assemble2(AsmArgs g) {
    __shared__ T lv[9][256];
    const int ntiles = (int)(g.nupper >> 6);
    const int tile = xcd_block(ntiles);
    if (tile >= ntiles) return;
    const int t = threadIdx.x;
    const int64_t u0 = (int64_t)tile * 64;
    const uint32_t c0 = g.eoff[u0], c1 = g.eoff[u0 + 64];
    uint32_t s = 0, en = 0;
    if (t < 64) {
        s = g.eoff[u0 + t];
        en = g.eoff[u0 + t + 1];
    }
    T acc[9];
#pragma unroll
    for (int k = 0; k < 9; k++) acc[k] = (T)0.0;
    for (uint32_t ch = c0; ch < c1; ch += 256) {
        const uint32_t f = ch + t;
        if (f < c1) {
            T w[9];
            contribution<T>(g, g.flat[f], w);
#pragma unroll
            for (int k = 0; k < 9; k++) lv[k][t] = w[k];
        }
        __syncthreads();
        if (t < 64) {
            const uint32_t lo = s > ch ? s : ch, hi = en < ch + 256 ? en : ch + 256;
            for (uint32_t i = lo; i < hi; i++)
#pragma unroll
                for (int k = 0; k < 9; k++) acc[k] += lv[k][i - ch];
        }
        __syncthreads();
    }
    if (t >= 64) return;
    const int64_t u = u0 + t;
    const uint8_t fl = g.flags[u];
    if (fl & 4) return;  // tile padding
    const bool diag = (fl & 1) != 0;
    const int node = g.node[u];
    if (diag) {
        if (fl & 2) {
#pragma unroll
            for (int k = 0; k < 9; k++) acc[k] = (T)0.0;
            acc[0] = acc[4] = acc[8] = (T)1.0;
        } else {
            const T z = (T)g.zeta[node];
            acc[0] += z;
            acc[4] += z;
            acc[8] += z;
        }
        g.dinv[3 * (size_t)node + 0] = 1.0 / (double)acc[0];
        g.dinv[3 * (size_t)node + 1] = 1.0 / (double)acc[4];
        g.dinv[3 * (size_t)node + 2] = 1.0 / (double)acc[8];
    }
    T *v = reinterpret_cast<T *>(g.val);
    {
        const int64_t slot = g.slot[u];
        const int64_t tl = slot & ~(int64_t)63;
        const int ln = (int)(slot & 63);
#pragma unroll
        for (int k = 0; k < 9; k++) v[tl * 9 + k * 64 + ln] = acc[k];
    }
    const int32_t mslot = g.mirror[u];
    if (mslot >= 0) {
        const int64_t tl = (int64_t)mslot & ~(int64_t)63;
        const int ln = mslot & 63;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) v[tl * 9 + (3 * i + j) * 64 + ln] = acc[3 * j + i];
    }
}
