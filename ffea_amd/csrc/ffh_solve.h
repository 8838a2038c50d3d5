// ffh_solve.h -- the force phase and the solvers: force gather + CG start, the CG
// SpMV and update kernels, integration, matrix apply, strain and kinetic energy.
// Part of ffh_kernels.hip: included inside its anonymous namespace, after the headers it
// lists there; not a stand-alone header.
#pragma once

// ----------------------------------------------------------------------------------------
// RHS: force gather, Stokes noise, pinned rows, CG initialisation.  64 nodes x 3 components
// per block (wave = component).
// ----------------------------------------------------------------------------------------
// The CG work vectors r, p, q, z (and r_prev) hold V values: double on the FP64 path and for
// the mass-ful solvers; float on the mixed path, whose iterate x, right-hand side f, dot
// products and true residual stay FP64.
struct CgArgs {
    int N, nslices;
    double *f, *x, *vel;
    void *r, *p, *q, *z;
    const double *dinv;
    double *parts_rz, *parts_rr, *parts_ff, *parts_pq;  // one partial per slice
    const int32_t *slice_blob, *blob_slice_off;
    double eps2;
    int maxit;
    CgCtl *ctl;  // one per blob
    // learned preconditioner (nullptr otherwise): partials of |D^-1 r|^2, r before the update,
    // partials of z.r_prev for the flexible beta
    double *parts_uu;
    void *r_prev;
    double *parts_zrp;
    const double *tot;  // per-blob totals (large blobs) or nullptr: consumers add the partials
    // mixed path: partials of |f - K x|^2 of a true-residual pass, and its period (0: only when
    // the recurrence residual passes the stopping test)
    double *parts_tr;
    int rr_period;
};

// Vector type and true-residual scheme of the CG kernels for element/matrix type T.
template <typename T, bool kMass>
struct CgVec {
    static constexpr bool kMixed = !kMass && std::is_same<T, float>::value;
    using V = typename std::conditional<kMixed, float, double>::type;
};


struct RhsArgs {
    const void *eforce;  // T SoA [12][E]
    int E;
    const double *f_in;  // solve() RHS given directly ([3N]); nullptr in the time step
    const int64_t *toff;
    const uint16_t *cnt;
    const uint32_t *ctb;
    const double *zeta;
    const uint8_t *pinned;
    const uint32_t *rng_off;
    const uint32_t *rng_state;
    const Mat3u *jump;
    uint32_t node_base;
    int noise;
    double kT, dt;
    const double *ext;  // external forces per linear node [3N], midpoints folded; or nullptr
    const double *vel;     // mass-ful solvers with Stokes drag: node velocities [3N]; or nullptr
    const double *lumped;  // MassLumpedSolver: inverse masses [N] (the solve is x = f / m); or nullptr
    ForceArgs fa;          // springs, constant and surface forces
};

template <typename T, typename V>
__global__ __launch_bounds__(192) void k_rhs(RhsArgs g, CgArgs v) {
    __shared__ double sh[8];
    __shared__ uint32_t rs[6][64];
    const int i = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int blk = xcd_block(v.nslices);
    if (blk >= v.nslices) return;
    const int n = blk * 64 + lane;
    const bool live = !g.pinned[n] || g.f_in;
    // Stokes noise (src/Blob.cpp:3895-3898): draws 3*node + {x,y,z} of the node phase.  Wave 0
    // jumps to the node's first draw; waves 1, 2 step on from the shared state.  Pinned rows
    // are zeroed afterwards in the reference, so their draws are skipped (not evaluated).
    const bool noise = g.noise && !g.f_in;
    if (noise) {
        if (i == 0 && live) {
            Stream6 st = rng_jump(g.rng_state, g.jump, g.node_base + g.rng_off[n]);
#pragma unroll
            for (int k = 0; k < 6; k++) rs[k][lane] = st.s[k];
        }
        __syncthreads();
    }
    double f = 0.0;
    if (g.f_in) {
        f = g.f_in[3 * (size_t)n + i];
    } else if (live) {
        // springs, then constant forces, onto the zeroed force (World::run: apply_springs, then
        // update_internal_forces' apply_ctforces)
        if (g.fa.pre_off) {
            const int t1 = g.fa.pre_off[n + 1];
            for (int t = g.fa.pre_off[n]; t < t1; t++) {
                const int it = g.fa.pre_item[t];
                f += it >= 0 ? spring_component(g.fa, it >> 1, it & 1, i) : const_component(g.fa, ~it, g.fa.pos + n, i);
            }
        }
        const T *ef = reinterpret_cast<const T *>(g.eforce);
        const int cnt = g.cnt[n];
        const uint32_t *tl = g.ctb + g.toff[blk] + lane;
        // terms in groups of up to eight: their codes, then their element forces, then the
        // additions in order
        exact_groups(0, cnt, [&](auto g_, int k0) {
            constexpr int G = decltype(g_)::value;
            uint32_t code[G];
            T fe[G];
#pragma unroll
            for (int u = 0; u < G; u++) code[u] = tl[64 * (k0 + u)];
#pragma unroll
            for (int u = 0; u < G; u++) fe[u] = ef[(size_t)(3 * (code[u] & 3u) + i) * g.E + (code[u] >> 2)];
#pragma unroll
            for (int u = 0; u < G; u++) f += (double)fe[u];
        });
        // surface face forces, face by face, after the element forces (src/Blob.cpp:3842-3851)
        if (g.fa.grp_off) f = face_groups(g.fa, g.fa.grp_off[n], g.fa.grp_off[n + 1], i, f);
        // external forces come after the element forces and before the Stokes noise
        if (g.ext) f += g.ext[3 * (size_t)n + i];
        // mass-ful solvers: Stokes drag on the node's velocity, before its noise (:3854-3877)
        if (g.vel) f -= g.vel[3 * (size_t)n + i] * g.zeta[n];
        if (noise) {
            Stream6 st;
#pragma unroll
            for (int k = 0; k < 6; k++) st.s[k] = rs[k][lane];
            for (int k = 0; k < i; k++) (void)rng_u01(st);
            f -= rand_half(st) * sqrt((24 * g.kT * g.zeta[n]) / (g.dt));
        }
        // forces on edge midpoints, passed on element by element (linearise_force,
        // src/Blob.cpp:1066-1085)
        if (g.fa.fold_off) {
            const int t1 = g.fa.fold_off[n + 1];
            for (int t = g.fa.fold_off[n]; t < t1; t++) {
                const int *mm = g.fa.fold_mid + 3 * (size_t)t;
                const double x = mm[0] >= 0 ? g.fa.mf[3 * (size_t)mm[0] + i] : 0.0;
                const double y = mm[1] >= 0 ? g.fa.mf[3 * (size_t)mm[1] + i] : 0.0;
                const double z = mm[2] >= 0 ? g.fa.mf[3 * (size_t)mm[2] + i] : 0.0;
                f += 0.5 * (x + y + z);
            }
        }
    }
    const size_t k = 3 * (size_t)n + i;
    if (g.lumped) {  // MassLumpedSolver::solve (src/MassLumpedSolver.cpp:82-90)
        v.f[k] = f;
        v.x[k] = f * g.lumped[n];
    } else {
        // stored values (rounded to V on the mixed path) are what the partials sum
        const double r = (double)(V)f, z = (double)(V)(v.dinv[k] * f);
        v.f[k] = f;
        static_cast<V *>(v.r)[k] = (V)r;
        v.x[k] = 0.0;
        static_cast<V *>(v.p)[k] = (V)0.0;
        static_cast<V *>(v.q)[k] = (V)0.0;
        static_cast<V *>(v.z)[k] = (V)z;
        block_partial2_at(r * z, v.parts_rz, f * f, v.parts_ff, blk, sh);
        if (v.parts_uu) block_partial_at(z * z, v.parts_uu, blk, sh);
    }
    const int b = v.slice_blob[blk];
    if (blk == v.blob_slice_off[b] && threadIdx.x == 0) {
        CgCtl *c = &v.ctl[b];
        c->done = 0;
        c->status = 0;
        c->iter = 0;
        c->res = 0.0;
    }
}

// ----------------------------------------------------------------------------------------
// CG iteration `it` (the reference's loop index, src/NoMassCGSolver.cpp:175-194)
// ----------------------------------------------------------------------------------------
struct SpmvArgs {
    const int64_t *slice_off;
    const int32_t *slice_width;
    const int32_t *col;
    const void *val;
};

// Convergence test of update(it-1) (||r||^2/||f||^2 < eps^2, src/NoMassCGSolver.cpp:182-191,
// :233-248); beta = delta_it / delta_(it-1); p = z + beta p (vec3_scale_and_add,
// src/mat_vec_fns.cpp:265-275); q = K p, formed as K z + beta q_old so that only z is gathered
// (K is applied block by block, each scalar row summed in ascending column order like
// SparseMatrixFixedPattern::apply, src/SparseMatrixFixedPattern.cpp:148-191); p.q partials.
// One block = one 64-row slice; wave w computes component w of every row.  Every scalar is
// per blob: the blob's slices are a contiguous range of partials.
// mode 0: Jacobi-PCG (Fletcher-Reeves beta); 1: flexible CG with the learned preconditioner
// (beta = z.(r - r_prev) / delta_prev, Polak-Ribiere form); 2: restart (beta = 0: p = z).
//
// kMass: ConjugateGradientSolver on the mass matrix (src/ConjugateGradientSolver.cpp:205-230),
// whose matrix holds one double per block (the same for the three components) and whose loop
// tests ||r||^2/||f||^2 < eps^2 before each product -- first on r = f, where the ratio is 1 --
// and fails after max_iterations_cg updates without testing the last one.
template <bool kMass>
__device__ __forceinline__ bool cg_stop_first(double f2, double eps2) {
    return f2 == 0.0 || (kMass && 1.0 < eps2);
}
template <bool kMass>
__device__ __forceinline__ bool cg_stop(double res, int it, const CgArgs &v, bool &ok) {
    ok = kMass ? (it < v.maxit && res < v.eps2) : res < v.eps2;
    return ok || it >= v.maxit;
}

// Top of CG iteration `it`, shared by both SpMV variants (tot(parts, slot) = the blob's total
// of a per-slice partial, in the variant's fixed order).  Returns false when the blob has
// stopped or stops here.  Sets beta and, on the mixed path (kMixed), whether this pass also
// forms the true residual f - K x: every rr_period iterations, and whenever the recurrence
// residual passes the stopping test -- the update kernel then stops only if the true residual
// passes it too (the final epsilon check of the FP64 path is on the recurrence residual, as
// NoMassCGSolver::solve's is).
template <bool kMass, bool kMixed, typename Tot, typename Tot2>
__device__ __forceinline__ bool cg_head(const CgArgs &v, CgCtl *ctl, int it, int mode, bool decider, Tot &&tot,
                                        Tot2 &&tot2, double &beta, bool &check) {
    beta = 0.0;
    check = false;
    if (it == 0) {
        double f2, d0;
        tot2(v.parts_ff, kTotFF, v.parts_rz, kTotRZ, f2, d0);
        const bool stop = cg_stop_first<kMass>(f2, v.eps2);
        if (decider) {
            ctl->f2 = f2;
            ctl->delta[0] = d0;
            ctl->check = 0;
            if (stop) {  // zero force: zero solution
                ctl->done = 1;
                ctl->status = 0;
                ctl->iter = 0;
            }
        }
        return !stop;
    }
    const double f2 = ctl->f2;
    double r2, dnew;
    tot2(v.parts_rr, kTotRR, v.parts_rz, kTotRZ, r2, dnew);
    const double res = r2 / f2;
    if constexpr (kMixed) {
        check = res < v.eps2 || (v.rr_period > 0 && it % v.rr_period == 0);
        if (!(res < v.eps2) && it >= v.maxit) {
            if (decider) {
                ctl->done = 1;
                ctl->status = -1;
                ctl->iter = it;
                ctl->res = res;
            }
            return false;
        }
    } else {
        bool ok;
        if (cg_stop<kMass>(res, it, v, ok)) {
            if (decider) {
                ctl->done = 1;
                ctl->status = ok ? 0 : -1;
                ctl->iter = it;
                ctl->res = res;
            }
            return false;
        }
    }
    if (mode == 0)
        beta = dnew / ctl->delta[(it - 1) & 1];
    else if (mode == 1)
        beta = (dnew - tot(v.parts_zrp, kTotZRP)) / ctl->delta[(it - 1) & 1];
    if (decider) {
        ctl->delta[it & 1] = dnew;
        ctl->check = check ? 1 : 0;
    }
    return true;
}

template <typename T, bool kNt, bool kMass>
__global__ __launch_bounds__(192) void k_cg_spmv(SpmvArgs m, CgArgs v, int it, int mode) {
    using V = typename CgVec<T, kMass>::V;
    constexpr bool kMixed = CgVec<T, kMass>::kMixed;
    __shared__ double sh[8];  // two totals at a time (parts_range2)
    const int slice = xcd_block(v.nslices);
    if (slice >= v.nslices) return;
    const int b = v.slice_blob[slice];
    CgCtl *ctl = &v.ctl[b];
    if (ctl->done) return;
    const int s0 = v.blob_slice_off[b], s1 = v.blob_slice_off[b + 1];
    double beta;
    bool check;
    auto tot = [&](const double *parts, int slot) { return blob_total(parts, v.tot, slot, b, s0, s1, sh); };
    auto tot2 = [&](const double *pa, int sa, const double *pb, int sb, double &ta, double &tb) {
        blob_total2(pa, pb, v.tot, sa, sb, b, s0, s1, sh, ta, tb);
    };
    // the slice's shape and this row's z, p, q are loaded before the head, so they arrive while
    // it reduces (f, which only a true-residual pass reads, is loaded there)
    const int ci = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row = slice * 64 + lane;
    const int w = m.slice_width[slice];
    const int64_t base = m.slice_off[slice];
    const size_t k = 3 * (size_t)row + ci;
    V *p = static_cast<V *>(v.p), *q = static_cast<V *>(v.q);
    const double zk = (double)static_cast<const V *>(v.z)[k], pk = (double)p[k], qk = (double)q[k];
    if (!cg_head<kMass, kMixed>(v, ctl, it, mode, slice == s0 && threadIdx.x == 0, tot, tot2, beta, check)) return;
    const T *val = reinterpret_cast<const T *>(m.val);
    const V *__restrict__ z = static_cast<const V *>(v.z);
    const double *__restrict__ x = v.x;
    double acc = 0.0, accx = 0.0;
    if constexpr (kMass) {
        const double *mv = reinterpret_cast<const double *>(m.val);
#pragma unroll 4
        for (int t = 0; t < w; t++) {
            const int64_t tile = base + (int64_t)t * 64;
            const int c = stream_load<kNt>(m.col + tile + lane);
            acc += stream_load<kNt>(mv + tile + lane) * z[3 * c + ci];
        }
    } else if (kMixed && check) {  // K z and K x from one pass over the matrix, tiles in groups
        exact_groups<4>(0, w, [&](auto g_, int t0) {
            constexpr int G = decltype(g_)::value;
            int c[G];
            T vv[G][3];
#pragma unroll
            for (int u = 0; u < G; u++) {
                const int64_t tile = base + (int64_t)(t0 + u) * 64;
                c[u] = stream_load<kNt>(m.col + tile + lane);
                const T *vt = val + tile * 9;
#pragma unroll
                for (int j = 0; j < 3; j++) vv[u][j] = stream_load<kNt>(vt + tile_entry<T>(3 * ci + j, lane));
            }
            V zz[G][3];
            double xx[G][3];
#pragma unroll
            for (int u = 0; u < G; u++)
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    zz[u][j] = z[3 * c[u] + j];
                    xx[u][j] = x[3 * c[u] + j];
                }
#pragma unroll
            for (int u = 0; u < G; u++) {
                const double v0 = vv[u][0], v1 = vv[u][1], v2 = vv[u][2];
                acc += v0 * (double)zz[u][0];
                acc += v1 * (double)zz[u][1];
                acc += v2 * (double)zz[u][2];
                accx += v0 * xx[u][0];
                accx += v1 * xx[u][1];
                accx += v2 * xx[u][2];
            }
        });
    } else {
        // tiles in groups, loads first, sums in tile order (as in k_cg_spmv1)
        auto group = [&](auto g_, int t0) {
            constexpr int G = decltype(g_)::value;
            int c[G];
            T vv[G][3];
#pragma unroll
            for (int u = 0; u < G; u++) {
                const int64_t tile = base + (int64_t)(t0 + u) * 64;
                c[u] = stream_load<kNt>(m.col + tile + lane);
                const T *vt = val + tile * 9;
#pragma unroll
                for (int j = 0; j < 3; j++) vv[u][j] = stream_load<kNt>(vt + tile_entry<T>(3 * ci + j, lane));
            }
            V zz[G][3];
#pragma unroll
            for (int u = 0; u < G; u++)
#pragma unroll
                for (int j = 0; j < 3; j++) zz[u][j] = z[3 * c[u] + j];
#pragma unroll
            for (int u = 0; u < G; u++) {
                acc += (double)vv[u][0] * (double)zz[u][0];
                acc += (double)vv[u][1] * (double)zz[u][1];
                acc += (double)vv[u][2] * (double)zz[u][2];
            }
        };
        exact_groups(0, w, group);
    }
    const double pn = (double)(V)(zk + beta * pk);
    const double qn = (double)(V)(acc + beta * qk);
    p[k] = (V)pn;
    q[k] = (V)qn;
    block_partial_at(pn * qn, v.parts_pq, slice, sh);
    if (kMixed && check) {
        const double rt = v.f[k] - accx;
        static_cast<V *>(v.r)[k] = (V)rt;
        block_partial_at(rt * rt, v.parts_tr, slice, sh);
    }
}

// Sum of parts[i0, i1) in exactly the order parts_range gives with a 192-thread block, computed
// by one 64-thread wave (virtual thread t + 64 w for w = 0, 1, 2).
__device__ __forceinline__ double parts_range_as192(const double *__restrict__ parts, int i0, int i1) {
    const int t = threadIdx.x & 63;
    double tot = 0.0;
#pragma unroll
    for (int w = 0; w < 3; w++) tot += __shfl(wave_sum(strided_sum(parts, i0 + t + 64 * w, i1, 192)), 0, 64);
    return tot;
}

__device__ __forceinline__ void parts_range2_as192(const double *__restrict__ pa, const double *__restrict__ pb, int i0,
                                                   int i1, double &ta, double &tb) {
    const int t = threadIdx.x & 63;
    ta = 0.0;
    tb = 0.0;
#pragma unroll
    for (int w = 0; w < 3; w++) {
        double sa, sb;
        strided_sum2(pa, pb, i0 + t + 64 * w, i1, 192, sa, sb);
        ta += __shfl(wave_sum(sa), 0, 64);
        tb += __shfl(wave_sum(sb), 0, 64);
    }
}

__device__ __forceinline__ double blob_total_as192(const double *parts, const double *tot, int slot, int b, int s0, int s1) {
    if (tot) return tot[b * kTotSlots + slot];
    return parts_range_as192(parts, s0, s1);
}

// k_cg_spmv with one 64-lane wave per slice: each lane computes the 3 components of its row, so
// the column index and the gathered z block are loaded once instead of three times.  Row sums,
// reductions and the partials are formed in k_cg_spmv's exact order, so both variants give
// bit-identical results (the tuner may pick either).
template <typename T, bool kNt, bool kMass>
__global__ __launch_bounds__(64) void k_cg_spmv1(SpmvArgs m, CgArgs v, int it, int mode) {
    using V = typename CgVec<T, kMass>::V;
    constexpr bool kMixed = CgVec<T, kMass>::kMixed;
    const int slice = xcd_block(v.nslices);
    if (slice >= v.nslices) return;
    const int b = v.slice_blob[slice];
    CgCtl *ctl = &v.ctl[b];
    if (ctl->done) return;
    const int s0 = v.blob_slice_off[b], s1 = v.blob_slice_off[b + 1];
    double beta;
    bool check;
    auto tot = [&](const double *parts, int slot) { return blob_total_as192(parts, v.tot, slot, b, s0, s1); };
    auto tot2 = [&](const double *pa, int sa, const double *pb, int sb, double &ta, double &tb) {
        if (v.tot) {
            ta = v.tot[b * kTotSlots + sa];
            tb = v.tot[b * kTotSlots + sb];
        } else {
            parts_range2_as192(pa, pb, s0, s1, ta, tb);
        }
    };
    // the slice's shape and this row's z, p, q are loaded before the head, so they arrive while
    // it reduces (f, which only a true-residual pass reads, is loaded there)
    const int lane = threadIdx.x & 63;
    const int row = slice * 64 + lane;
    const int w = m.slice_width[slice];
    const int64_t base = m.slice_off[slice];
    const size_t k = 3 * (size_t)row;
    V *p = static_cast<V *>(v.p), *q = static_cast<V *>(v.q);
    double zk[3], pk[3], qk[3];
#pragma unroll
    for (int ci = 0; ci < 3; ci++) {
        zk[ci] = (double)static_cast<const V *>(v.z)[k + ci];
        pk[ci] = (double)p[k + ci];
        qk[ci] = (double)q[k + ci];
    }
    if (!cg_head<kMass, kMixed>(v, ctl, it, mode, slice == s0 && threadIdx.x == 0, tot, tot2, beta, check)) return;
    const T *val = reinterpret_cast<const T *>(m.val);
    const V *__restrict__ z = static_cast<const V *>(v.z);
    const double *__restrict__ x = v.x;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, x0 = 0.0, x1 = 0.0, x2 = 0.0;
    if constexpr (kMass) {
        const double *mv = reinterpret_cast<const double *>(m.val);
#pragma unroll 4
        for (int t = 0; t < w; t++) {
            const int64_t tile = base + (int64_t)t * 64;
            const int c = stream_load<kNt>(m.col + tile + lane);
            const double a = stream_load<kNt>(mv + tile + lane);
            a0 += a * z[3 * c + 0];
            a1 += a * z[3 * c + 1];
            a2 += a * z[3 * c + 2];
        }
    } else if (kMixed && check) {  // K z and K x from one pass over the matrix, tiles in groups
        exact_groups<4>(0, w, [&](auto g_, int t0) {
            constexpr int G = decltype(g_)::value;
            int c[G];
            T vt[G][9];
#pragma unroll
            for (int u = 0; u < G; u++) {
                const int64_t tile = base + (int64_t)(t0 + u) * 64;
                c[u] = stream_load<kNt>(m.col + tile + lane);
                load_block<kNt>(val + tile * 9, lane, vt[u]);
            }
            V zz[G][3];
            double xx[G][3];
#pragma unroll
            for (int u = 0; u < G; u++)
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    zz[u][j] = z[3 * c[u] + j];
                    xx[u][j] = x[3 * c[u] + j];
                }
#pragma unroll
            for (int u = 0; u < G; u++) {
                double vv[9];
#pragma unroll
                for (int kk = 0; kk < 9; kk++) vv[kk] = (double)vt[u][kk];
                const double z0 = zz[u][0], z1 = zz[u][1], z2 = zz[u][2];
                const double y0 = xx[u][0], y1 = xx[u][1], y2 = xx[u][2];
                a0 += vv[0] * z0;
                a0 += vv[1] * z1;
                a0 += vv[2] * z2;
                a1 += vv[3] * z0;
                a1 += vv[4] * z1;
                a1 += vv[5] * z2;
                a2 += vv[6] * z0;
                a2 += vv[7] * z1;
                a2 += vv[8] * z2;
                x0 += vv[0] * y0;
                x0 += vv[1] * y1;
                x0 += vv[2] * y2;
                x1 += vv[3] * y0;
                x1 += vv[4] * y1;
                x1 += vv[5] * y2;
                x2 += vv[6] * y0;
                x2 += vv[7] * y1;
                x2 += vv[8] * y2;
            }
        });
    } else {
        // tiles in groups: their column indices and blocks are loaded first, then the gathers of
        // z, then the sums in tile order -- the order of a one-tile loop, with a group's
        // dependent load pairs in flight
        auto group = [&](auto g_, int t0) {
            constexpr int G = decltype(g_)::value;
            int c[G];
            T vv[G][9];
#pragma unroll
            for (int u = 0; u < G; u++) {
                const int64_t tile = base + (int64_t)(t0 + u) * 64;
                c[u] = stream_load<kNt>(m.col + tile + lane);
                load_block<kNt>(val + tile * 9, lane, vv[u]);
            }
            V zz[G][3];
#pragma unroll
            for (int u = 0; u < G; u++)
#pragma unroll
                for (int j = 0; j < 3; j++) zz[u][j] = z[3 * c[u] + j];
#pragma unroll
            for (int u = 0; u < G; u++) {
                const double z0 = zz[u][0], z1 = zz[u][1], z2 = zz[u][2];
                a0 += (double)vv[u][0] * z0;
                a0 += (double)vv[u][1] * z1;
                a0 += (double)vv[u][2] * z2;
                a1 += (double)vv[u][3] * z0;
                a1 += (double)vv[u][4] * z1;
                a1 += (double)vv[u][5] * z2;
                a2 += (double)vv[u][6] * z0;
                a2 += (double)vv[u][7] * z1;
                a2 += (double)vv[u][8] * z2;
            }
        };
        exact_groups(0, w, group);
    }
    const double acc[3] = {a0, a1, a2}, accx[3] = {x0, x1, x2};
    double part = 0.0, tpart = 0.0;
#pragma unroll
    for (int ci = 0; ci < 3; ci++) {
        const double pn = (double)(V)(zk[ci] + beta * pk[ci]);
        const double qn = (double)(V)(acc[ci] + beta * qk[ci]);
        p[k + ci] = (V)pn;
        q[k + ci] = (V)qn;
        part += __shfl(wave_sum(pn * qn), 0, 64);  // ((0 + W0) + W1) + W2, as the 3-wave block
        if (kMixed && check) {
            const double rt = v.f[k + ci] - accx[ci];
            static_cast<V *>(v.r)[k + ci] = (V)rt;
            tpart += __shfl(wave_sum(rt * rt), 0, 64);
        }
    }
    if (lane == 0) {
        v.parts_pq[slice] = part;
        if (kMixed && check) v.parts_tr[slice] = tpart;
    }
}

// alpha = delta / p.q; x += alpha p; r -= alpha q (vec3_add_to_scaled, src/mat_vec_fns.cpp:250-260);
// z = D^-1 r (parallel_apply_preconditioner, src/NoMassCGSolver.cpp:278-292); partials of r.r
// and r.z.  One block = the 3 x 64 scalars of one slice.  On the mixed path (kMixed) a pass whose
// SpMV formed the true residual first decides on it: the blob stops if |f - K x|^2 / |f|^2 <
// eps^2 (or fails at max_iterations_cg); otherwise r, which now holds f - K x, is updated from
// there (residual replacement).
template <typename V, bool kMixed>
__global__ __launch_bounds__(192) void k_cg_update(CgArgs v, int it) {
    __shared__ double sh[8];
    const int slice = xcd_block(v.nslices);
    if (slice >= v.nslices) return;
    const int b = v.slice_blob[slice];
    CgCtl *ctl = &v.ctl[b];
    // This row's x, p, r, q and D^-1 are loaded before the head, so they arrive while it reduces,
    // on the mixed path and where the blob totals are precomputed (a head of scalar loads).  The
    // FP64 path's own re-adding of the partials is a run of vector loads, which complete in issue
    // order and would wait for these too: there they are loaded after it (c1 FP64 update 8.8 ->
    // 8.3 us; c4 FP64 with totals 27.9 -> 26.6 us early; c1 mixed 7.5 -> 7.2 us early).
    const size_t k = (size_t)slice * 192 + threadIdx.x;
    V *p = static_cast<V *>(v.p), *q = static_cast<V *>(v.q), *r = static_cast<V *>(v.r), *z = static_cast<V *>(v.z);
    const bool early = kMixed || v.tot != nullptr;
    double xk = 0.0, pk = 0.0, r0 = 0.0, qk = 0.0, dk = 0.0;
    auto load_row = [&]() {
        xk = v.x[k];
        pk = (double)p[k];
        r0 = (double)r[k];
        qk = (double)q[k];
        dk = v.dinv[k];
    };
    if (early) load_row();
    if (ctl->done) return;
    const int s0 = v.blob_slice_off[b], s1 = v.blob_slice_off[b + 1];
    if constexpr (kMixed) {
        if (ctl->check) {
            const double res = blob_total(v.parts_tr, v.tot, kTotTR, b, s0, s1, sh) / ctl->f2;
            const bool ok = res < v.eps2;
            if (ok || it >= v.maxit) {
                if (slice == s0 && threadIdx.x == 0) {
                    ctl->done = 1;
                    ctl->status = ok ? 0 : -1;
                    ctl->iter = it;
                    ctl->res = res;
                }
                return;
            }
        }
    }
    const double alpha = ctl->delta[it & 1] / blob_total(v.parts_pq, v.tot, kTotPQ, b, s0, s1, sh);
    if (!early) load_row();
    v.x[k] = xk + alpha * pk;
    const double rn = (double)(V)(r0 + (-alpha) * qk);
    r[k] = (V)rn;
    const double zn = (double)(V)(dk * rn);
    z[k] = (V)zn;
    block_partial2_at(rn * rn, v.parts_rr, rn * zn, v.parts_rz, slice, sh);
    if (v.parts_uu) {  // learned preconditioner follows: z holds u = D^-1 r
        static_cast<V *>(v.r_prev)[k] = (V)r0;
        block_partial_at(zn * zn, v.parts_uu, slice, sh);
    }
}

// euler_integrate (src/Blob.cpp:3941-3967): massless, vel = v and pos += v dt; mass-ful
// (massful != 0), vel += a dt and pos += vel dt.  Skipped after a CG failure, as the reference
// skips it.  Thread 0 advances the stream by one step's draws.
__global__ __launch_bounds__(kBlock) void k_integrate(CgArgs v, double4 *pos, double dt,
                                                     uint32_t *rng_state, Mat3u j1, Mat3u j2, int advance,
                                                     int massful, const int *gate) {
    if (gate && *gate == 0) return;  // in a step graph: the host finishes this step itself
    const int n = blockIdx.x * kBlock + threadIdx.x;
    if (n == 0 && advance) {
        uint32_t s[6];
        for (int k = 0; k < 6; k++) s[k] = rng_state[k];
        matvec<ffh::kM1, kC1>(j1, &s[0]);
        matvec<ffh::kM2, kC2>(j2, &s[3]);
        for (int k = 0; k < 6; k++) rng_state[k] = s[k];
    }
    if (n >= v.N || v.ctl[v.slice_blob[n >> 6]].status != 0) return;
    double x0 = v.x[3 * n], x1 = v.x[3 * n + 1], x2 = v.x[3 * n + 2];
    if (massful) {
        x0 = v.vel[3 * n] + x0 * dt;
        x1 = v.vel[3 * n + 1] + x1 * dt;
        x2 = v.vel[3 * n + 2] + x2 * dt;
    }
    v.vel[3 * n] = x0;
    v.vel[3 * n + 1] = x1;
    v.vel[3 * n + 2] = x2;
    double4 p = pos[n];
    p.x += x0 * dt;
    p.y += x1 * dt;
    p.z += x2 * dt;
    pos[n] = p;
}

// The step graph integrates only a step that the host will not continue: every blob's solve has
// stopped (converged or failed), no element has inverted (the reference stops before the
// position phase, src/World.cpp:2025-2028) and the surface-surface interaction raised no error.
// The host polls the same control blocks after the graph and decides the same way.
__global__ void k_step_gate(const CgCtl *ctl, int nblobs, int need_cg, const int *ss_err, int *gate) {
    if (threadIdx.x != 0) return;
    int ok = ctl[0].inversions == 0 && (!ss_err || *ss_err == 0);
    if (need_cg)
        for (int b = 0; b < nblobs; b++) ok = ok && ctl[b].done != 0;
    *gate = ok;
}

__global__ void k_reset_inversions(CgCtl *ctl) {
    if (threadIdx.x == 0) ctl->inversions = 0;
}

// result = K in for Solver::apply_matrix (linear rows; midpoint rows handled on the host).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_apply(SpmvArgs m, int N, const double *__restrict__ in,
                                                  double *__restrict__ out) {
    const int row = blockIdx.x * kBlock + threadIdx.x;
    if (row >= N) return;
    const int s = row >> 6, lane = row & 63;
    const int w = m.slice_width[s];
    const int64_t base = m.slice_off[s];
    const T *val = reinterpret_cast<const T *>(m.val);
    double acc[3] = {0.0, 0.0, 0.0};
    for (int t = 0; t < w; t++) {
        const int64_t tile = base + (int64_t)t * 64;
        const int c = m.col[tile + lane];
        T vt[9];
        load_block<false>(val + tile * 9, lane, vt);
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) acc[i] += (double)vt[3 * i + j] * in[3 * c + j];
    }
    out[3 * row] = acc[0];
    out[3 * row + 1] = acc[1];
    out[3 * row + 2] = acc[2];
}

// Strain energy terms (src/Blob.cpp:1424-1432) recomputed from current positions; one partial
// per block, summed on the host.
__global__ __launch_bounds__(kBlock) void k_strain_energy(ElemArgs g, double *parts) {
    __shared__ double sh[4];
    const int e = blockIdx.x * kBlock + threadIdx.x;
    double part = 0.0;
    if (e < g.E) {
        const int4 n = g.enodes[e];
        const double4 p0 = g.pos[n.x], pr[3] = {g.pos[n.y], g.pos[n.z], g.pos[n.w]};
        double J[3][3];
        for (int r = 0; r < 3; r++) {
            J[r][0] = pr[r].x - p0.x;
            J[r][1] = pr[r].y - p0.y;
            J[r][2] = pr[r].z - p0.z;
        }
        double c1 = J[2][2] * J[1][1] - J[2][1] * J[1][2];
        double c2 = J[2][1] * J[0][2] - J[2][2] * J[0][1];
        double c3 = J[1][2] * J[0][1] - J[1][1] * J[0][2];
        double det = J[0][0] * c1 + J[1][0] * c2 + J[2][0] * c3;
        double vol = (1.0 / 6.0) * fabs(det);
        double F[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double f = 0;
                for (int k = 0; k < 3; k++) f += J[k][i] * g.jinv0[(size_t)(3 * j + k) * g.E + e];
                F[i][j] = f;
            }
        double dc = F[0][0] * F[0][0] + F[1][1] * F[1][1] + F[2][2] * F[2][2] + F[0][1] * F[0][1] +
                    F[1][0] * F[1][0] + F[0][2] * F[0][2] + F[2][0] * F[2][0] + F[1][2] * F[1][2] +
                    F[2][1] * F[2][1];
        double G = g.G[e], C = g.Em[e] - (2.0 / 3.0) * G, vol0 = g.vol0[e];
        double t = vol / vol0;
        part = vol0 * (G * (dc - 3) + 0.5 * C * (t * t - 1) - (C + 2 * G) * log(t));
    }
    block_partial(part, parts, sh);
}

// Kinetic energy of the mass-ful solvers (Blob::make_measurements, src/Blob.cpp:1385-1407):
// per element u_e . (M_e u_e) with apply_element_mass_matrix (src/tetra_element_linear.cpp:
// 534-553); one partial per block.  The caller halves the sum.
__global__ __launch_bounds__(kBlock) void k_kinetic_energy(ElemArgs g, const double *rho, double *parts) {
    __shared__ double sh[4];
    const int e = blockIdx.x * kBlock + threadIdx.x;
    double part = 0.0;
    if (e < g.E) {
        const int4 nn = g.enodes[e];
        const int ni[4] = {nn.x, nn.y, nn.z, nn.w};
        double u[12], mu[12];
        for (int i = 0; i < 3; i++)
            for (int a = 0; a < 4; a++) u[4 * i + a] = g.vel[3 * (size_t)ni[a] + i];
        const double sc = g.vol0[e] * rho[e];
        for (int k = 0; k < 12; k += 4)
            for (int i = 0; i < 4; i++) {
                double t = 0.0;
                for (int j = 0; j < 4; j++) t += (i == j ? .1 : .05) * u[j + k];
                mu[i + k] = t * sc;
            }
        part = u[0] * mu[0] + u[1] * mu[1] + u[2] * mu[2] + u[3] * mu[3] + u[4] * mu[4] + u[5] * mu[5] +
               u[6] * mu[6] + u[7] * mu[7] + u[8] * mu[8] + u[9] * mu[9] + u[10] * mu[10] + u[11] * mu[11];
    }
    block_partial(part, parts, sh);
}
