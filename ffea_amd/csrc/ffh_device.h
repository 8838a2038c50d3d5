// ffh_device.h -- device building blocks: MRG32k3a jumps and draws, fixed-order
// reductions, per-blob totals, streamed loads, XCD-aware block order, the CG control block.
// Part of ffh_kernels.hip: included inside its anonymous namespace, after the headers it
// lists there; not a stand-alone header.
#pragma once

// ----------------------------------------------------------------------------------------
// MRG32k3a on the device: exact integer arithmetic mod m = 2^32 - c.
// ----------------------------------------------------------------------------------------
constexpr uint32_t kC1 = 4294967296ull - ffh::kM1;  // 209
constexpr uint32_t kC2 = 4294967296ull - ffh::kM2;  // 22853
constexpr double kNorm = 1.0 / (4294967087.0 + 1.0);

template <uint32_t M, uint32_t Cc>
__device__ __forceinline__ uint32_t mred(uint64_t x) {
    x = (x >> 32) * Cc + (x & 0xffffffffull);
    x = (x >> 32) * Cc + (x & 0xffffffffull);
    return (uint32_t)(x >= M ? x - M : x);
}

template <uint32_t M, uint32_t Cc>
__device__ __forceinline__ void matvec(const Mat3u &A, uint32_t s[3]) {
    uint32_t v[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        uint64_t acc = (uint64_t)mred<M, Cc>((uint64_t)A.m[i][0] * s[0]) +
                       (uint64_t)mred<M, Cc>((uint64_t)A.m[i][1] * s[1]) +
                       (uint64_t)mred<M, Cc>((uint64_t)A.m[i][2] * s[2]);
        v[i] = mred<M, Cc>(acc);
    }
    s[0] = v[0];
    s[1] = v[1];
    s[2] = v[2];
}

struct Stream6 {
    uint32_t s[6];
};

// State of the stream `off` draws after `base`, via the base-256 digit tables
// tbl[(d*256 + v)*2 + comp] = A_comp^(v 256^d).
__device__ __forceinline__ Stream6 rng_jump(const uint32_t *__restrict__ base, const Mat3u *__restrict__ tbl,
                                            uint32_t off) {
    Stream6 g;
#pragma unroll
    for (int k = 0; k < 6; k++) g.s[k] = base[k];
#pragma unroll
    for (int d = 0; d < 4; d++) {
        uint32_t v = (off >> (8 * d)) & 255u;
        if (v) {
            matvec<ffh::kM1, kC1>(tbl[(d * 256 + v) * 2 + 0], &g.s[0]);
            matvec<ffh::kM2, kC2>(tbl[(d * 256 + v) * 2 + 1], &g.s[3]);
        }
    }
    return g;
}

__device__ __forceinline__ void rng_apply(Stream6 &g, const Mat3u &m1, const Mat3u &m2) {
    matvec<ffh::kM1, kC1>(m1, &g.s[0]);
    matvec<ffh::kM2, kC2>(m2, &g.s[3]);
}

// RngStream::U01 (src/RngStream.cpp:266-289) in exact integers; u rounds as the reference.
__device__ __forceinline__ double rng_u01(Stream6 &g) {
    uint32_t p1 = mred<ffh::kM1, kC1>((uint64_t)1403580u * g.s[1]) +
                  0u;  // a12 * s1
    uint32_t t1 = mred<ffh::kM1, kC1>((uint64_t)(ffh::kM1 - 810728u) * g.s[0]);
    uint64_t q1 = (uint64_t)p1 + t1;
    p1 = (uint32_t)(q1 >= ffh::kM1 ? q1 - ffh::kM1 : q1);
    g.s[0] = g.s[1];
    g.s[1] = g.s[2];
    g.s[2] = p1;
    uint32_t p2 = mred<ffh::kM2, kC2>((uint64_t)527612u * g.s[5]);
    uint32_t t2 = mred<ffh::kM2, kC2>((uint64_t)(ffh::kM2 - 1370589u) * g.s[3]);
    uint64_t q2 = (uint64_t)p2 + t2;
    p2 = (uint32_t)(q2 >= ffh::kM2 ? q2 - ffh::kM2 : q2);
    g.s[3] = g.s[4];
    g.s[4] = g.s[5];
    g.s[5] = p2;
    double d = (double)p1 - (double)p2;
    return (p1 > p2) ? d * kNorm : (d + 4294967087.0) * kNorm;
}

// RAND(-.5,.5) (include/tetra_element_linear.h:118)
__device__ __forceinline__ double rand_half(Stream6 &g) { return (-.5) + ((.5) - (-.5)) * rng_u01(g); }

// ----------------------------------------------------------------------------------------
// Reductions.  A producing kernel writes one partial sum per block (block tree, fixed order);
// every block of the consuming kernel re-adds those partials in the same fixed order, so all
// blocks see bitwise the same scalar and runs are reproducible -- no atomics, no fences.
// ----------------------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Block sum of v; the total is returned to every thread.  sh: >= blockDim.x/64 doubles.
__device__ __forceinline__ double block_sum_all(double v, double *sh) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int i = 0; i < nw; i++) t += sh[i];
    return t;
}

// Writes this block's sum of v to parts[blockIdx.x].
__device__ __forceinline__ void block_partial(double v, double *__restrict__ parts, double *sh) {
    double t = block_sum_all(v, sh);
    if (threadIdx.x == 0) parts[blockIdx.x] = t;
}

// Writes this block's sum of v to parts[idx].
__device__ __forceinline__ void block_partial_at(double v, double *__restrict__ parts, int idx, double *sh) {
    double t = block_sum_all(v, sh);
    if (threadIdx.x == 0) parts[idx] = t;
}

// Writes this block's sums of a and b to pa[idx] and pb[idx], each bitwise block_partial_at's,
// with the two wave trees interleaved and one pair of barriers.  sh: >= 2 * blockDim.x/64 doubles.
__device__ __forceinline__ void block_partial2_at(double a, double *__restrict__ pa, double b,
                                                  double *__restrict__ pb, int idx, double *sh) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_down(a, o, 64);
        b += __shfl_down(b, o, 64);
    }
    __syncthreads();
    if (lane == 0) {
        sh[w] = a;
        sh[nw + w] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ta = 0.0, tb = 0.0;
        for (int i = 0; i < nw; i++) {
            ta += sh[i];
            tb += sh[nw + i];
        }
        pa[idx] = ta;
        pb[idx] = tb;
    }
}

// Visits [t0, t1) in groups: as many of exactly kMax as fit, then one group of the remaining
// 1 .. kMax-1 -- group(std::integral_constant<int, G>(), first).  Every index a group touches is
// in range and used, so the compiler issues a group's loads together ahead of their uses (with a
// clamped last group whose extra members are skipped by a branch, it sinks each load to its use
// and the loads go out one at a time), and a run of 15 -- a row of the tetrahedral meshes --
// takes two dependent rounds of loads, not four.
template <int R, typename F>
__device__ __forceinline__ void rest_group(int r, int t, F &&group) {
    if constexpr (R > 0) {
        if (r == R)
            group(std::integral_constant<int, R>(), t);
        else
            rest_group<R - 1>(r, t, group);
    }
}
template <int kMax = 8, typename F>
__device__ __forceinline__ void exact_groups(int t0, int t1, F &&group) {
    int t = t0;
    for (; t + kMax <= t1; t += kMax) group(std::integral_constant<int, kMax>(), t);
    rest_group<kMax - 1>(t1 - t, t, group);
}

// s = ((0 + p[i]) + p[i + stride]) + ... over [i, i1): a fixed order, with eight loads in
// flight instead of one dependent load per term.
__device__ __forceinline__ double strided_sum(const double *__restrict__ p, int i, int i1, int stride) {
    double s = 0.0;
    for (; i + 7 * stride < i1; i += 8 * stride) {
        double a[8];
#pragma unroll
        for (int k = 0; k < 8; k++) a[k] = p[i + k * stride];
#pragma unroll
        for (int k = 0; k < 8; k++) s += a[k];
    }
    for (; i < i1; i += stride) s += p[i];
    return s;
}

// Fixed-order total of the partials [i0, i1): one blob's slices.
__device__ __forceinline__ double parts_range(const double *__restrict__ parts, int i0, int i1, double *sh) {
    return block_sum_all(strided_sum(parts, i0 + (int)threadIdx.x, i1, (int)blockDim.x), sh);
}

// The totals of two partial arrays over [i0, i1) in one pass and one block tree; each equals
// parts_range's total of that array bit for bit.  sh: >= 2 blockDim.x/64 doubles.
__device__ __forceinline__ void strided_sum2(const double *__restrict__ pa, const double *__restrict__ pb, int i, int i1,
                                             int stride, double &sa, double &sb) {
    sa = 0.0;
    sb = 0.0;
    for (; i + 7 * stride < i1; i += 8 * stride) {
        double a[8], b[8];
#pragma unroll
        for (int k = 0; k < 8; k++) {
            a[k] = pa[i + k * stride];
            b[k] = pb[i + k * stride];
        }
#pragma unroll
        for (int k = 0; k < 8; k++) {
            sa += a[k];
            sb += b[k];
        }
    }
    for (; i < i1; i += stride) {
        sa += pa[i];
        sb += pb[i];
    }
}

__device__ __forceinline__ void parts_range2(const double *pa, const double *pb, int i0, int i1, double *sh, double &ta,
                                             double &tb) {
    double sa, sb;
    strided_sum2(pa, pb, i0 + (int)threadIdx.x, i1, (int)blockDim.x, sa, sb);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    sa = wave_sum(sa);
    sb = wave_sum(sb);
    __syncthreads();
    if (lane == 0) {
        sh[w] = sa;
        sh[nw + w] = sb;
    }
    __syncthreads();
    ta = 0.0;
    tb = 0.0;
    for (int k = 0; k < nw; k++) {
        ta += sh[k];
        tb += sh[nw + k];
    }
}

// Per-blob totals.  A blob of many slices would make every consumer block re-add thousands of
// partials (O(slices^2) work in all); above kTotSlicesMin slices per blob a small kernel adds
// each blob's partials once (same fixed order) into tot[blob][slot] and consumers read that.
constexpr int kTotFF = 0, kTotRZ = 1, kTotRR = 2, kTotPQ = 3, kTotUU = 4, kTotZRP = 5, kTotDEN = 6, kTotTR = 7;
constexpr int kTotSlots = 8;
constexpr int kTotSlicesMin = 3072;

__device__ __forceinline__ double blob_total(const double *__restrict__ parts, const double *__restrict__ tot, int slot,
                                             int b, int s0, int s1, double *sh) {
    if (tot) return tot[b * kTotSlots + slot];
    return parts_range(parts, s0, s1, sh);
}

__device__ __forceinline__ void blob_total2(const double *pa, const double *pb, const double *__restrict__ tot, int slot_a,
                                            int slot_b, int b, int s0, int s1, double *sh, double &ta, double &tb) {
    if (tot) {
        ta = tot[b * kTotSlots + slot_a];
        tb = tot[b * kTotSlots + slot_b];
        return;
    }
    parts_range2(pa, pb, s0, s1, sh, ta, tb);
}

struct TotArgs {
    const double *parts[kTotSlots];
    const int32_t *blob_slice_off;
    double *tot;
    int mask;
};

// One block of 1024 threads per blob; each thread adds every 1024th partial, four loads in
// flight, then a block tree -- a fixed order.
__global__ __launch_bounds__(1024) void k_blob_totals(TotArgs t) {
    __shared__ double sh[16];
    const int b = blockIdx.x;
    const int s0 = t.blob_slice_off[b], s1 = t.blob_slice_off[b + 1];
    for (int k = 0; k < kTotSlots; k++)
        if ((t.mask >> k) & 1) {
            const double *p = t.parts[k];
            double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
            int i = s0 + (int)threadIdx.x;
            for (; i + 3 * 1024 < s1; i += 4 * 1024) {
                a0 += p[i];
                a1 += p[i + 1024];
                a2 += p[i + 2 * 1024];
                a3 += p[i + 3 * 1024];
            }
            for (; i < s1; i += 1024) a0 += p[i];
            const double v = block_sum_all((a0 + a1) + (a2 + a3), sh);
            if (threadIdx.x == 0) t.tot[b * kTotSlots + k] = v;
        }
}

// Streamed-once loads of the SpMV (matrix values, column indices).  Non-temporal (kNt) they do
// not push the gathered vector out of the XCD's L2 -- a gain when the matrix streams from HBM,
// a loss when it fits the Infinity Cache and would be re-read from there; the tuner decides.
template <bool kNt, typename X>
__device__ __forceinline__ X stream_load(const X *p) {
    if constexpr (kNt)
        return __builtin_nontemporal_load(p);
    else
        return *p;
}

// A 64-slot tile of the block matrix holds 9 * 64 values.  FP64 tiles are entry-major (entry
// q = 3 i + j of lane l at q * 64 + l: a wave reads one entry of its 64 blocks with one 512-byte
// load).  FP32 tiles keep a lane's entries 0-3 and 4-7 as two 16-byte groups ([2][64][4]) and
// entry 8 after them ([64]): a lane reads its whole block with two 16-byte loads and one 4-byte
// load, each of them one contiguous run over the wave.
template <typename T>
__host__ __device__ __forceinline__ int tile_entry(int q, int lane) {
    if constexpr (sizeof(T) == 8)
        return q * 64 + lane;
    else
        return q < 8 ? (q >> 2) * 256 + lane * 4 + (q & 3) : 512 + lane;
}

typedef float f32x4 __attribute__((ext_vector_type(4)));

// Lane `lane`'s block (entries 0..8) from the tile at `t` (9 * 64 values, 16-byte aligned).
template <bool kNt, typename T>
__device__ __forceinline__ void load_block(const T *__restrict__ t, int lane, T v[9]) {
    if constexpr (sizeof(T) == 8) {
#pragma unroll
        for (int q = 0; q < 9; q++) v[q] = stream_load<kNt>(t + q * 64 + lane);
    } else {
        const f32x4 a = stream_load<kNt>(reinterpret_cast<const f32x4 *>(t) + lane);
        const f32x4 b = stream_load<kNt>(reinterpret_cast<const f32x4 *>(t) + 64 + lane);
        v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w;
        v[4] = b.x, v[5] = b.y, v[6] = b.z, v[7] = b.w;
        v[8] = stream_load<kNt>(t + 512 + lane);
    }
}

template <typename T>
__device__ __forceinline__ void store_block(T *__restrict__ t, int lane, const T v[9]) {
    if constexpr (sizeof(T) == 8) {
#pragma unroll
        for (int q = 0; q < 9; q++) t[q * 64 + lane] = v[q];
    } else {
        reinterpret_cast<f32x4 *>(t)[lane] = f32x4{v[0], v[1], v[2], v[3]};
        reinterpret_cast<f32x4 *>(t)[64 + lane] = f32x4{v[4], v[5], v[6], v[7]};
        t[512 + lane] = v[8];
    }
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.  The gathering
// kernels remap the dispatch index so that XCD x works on the contiguous logical blocks
// [x per, (x+1) per): neighbouring rows, elements and records then share one L2.  The grid is
// launched with 8 per blocks; logical blocks >= n do no work.
constexpr int kXcds = 8;
__host__ __device__ __forceinline__ int xcd_grid(int n) { return kXcds * ((n + kXcds - 1) / kXcds); }
__device__ __forceinline__ int xcd_block(int n) {
    const int per = (n + kXcds - 1) / kXcds;
    return (int)(blockIdx.x % kXcds) * per + (int)(blockIdx.x / kXcds);
}

// CG control block (device memory), written by one thread of the deciding kernel.
struct CgCtl {
    double res, f2;
    double delta[2];  // delta_k = r_k . D^-1 r_k, slot k & 1
    int iter, done, status, inversions;
    int check;  // mixed path: this iteration's SpMV also formed the true residual f - K x
    int pad_;
};
