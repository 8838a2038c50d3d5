// ffh_box.h -- World::run's periodic box: a blob whose centroid has left the box through a
// periodic wall moves back in by one box length (src/World.cpp:1850-1901).
// Part of ffh_kernels.hip: included inside its anonymous namespace, after the headers it
// lists there; not a stand-alone header.
#pragma once

// The centroid is Blob::get_centroid's (src/Blob.cpp:971-983): the sequential sum over all
// N_t nodes in file order, divided by N_t.  Edge midpoints are .5 (a + b), as
// linearise_element leaves them at the end of every step.  The device sums in a fixed tree
// order and keeps the sum of |x| beside it; when the tree centroid lies within the two sums'
// rounding bound of a wall, the blob's sum is redone sequentially in file order, so the
// decision is always the reference's.
constexpr int kWrapChunk = 4096;  // nodes per partial sum

struct WrapArgs {
    double4 *pos;                 // compact linear nodes
    const int2 *ends;             // [N_t] each node's edge ends (compact ids; a vertex twice)
    const int32_t *blob_item_off; // [nblobs + 1] into ends
    const int32_t *chunk_lo;      // [chunks + 1] item offsets; a blob's chunks are contiguous
    const int32_t *blob_chunk_off;// [nblobs + 1]
    const int32_t *slice_blob;
    double *parts;                // [6 * chunks]: sum x, y, z, then sum |x|, |y|, |z|
    double4 *shift;               // [nblobs]: the move, w = 1 if the blob moves
    double box_x, box_y, box_z;
    int mask;                     // bits 0..5: wall_x_1, wall_x_2, wall_y_1, wall_y_2, wall_z_1, wall_z_2 periodic
    int N;
};

__device__ __forceinline__ double3 wrap_item(const WrapArgs &w, int t) {
    const int2 e = w.ends[t];
    const double4 a = w.pos[e.x], b = w.pos[e.y];
    return make_double3(.5 * (a.x + b.x), .5 * (a.y + b.y), .5 * (a.z + b.z));
}

__global__ __launch_bounds__(256) void k_wrap_partials(WrapArgs w) {
    __shared__ double sh[6][256];
    const int c = blockIdx.x, tid = threadIdx.x;
    const int t0 = w.chunk_lo[c], t1 = w.chunk_lo[c + 1];
    double s[6] = {0, 0, 0, 0, 0, 0};
    for (int t = t0 + tid; t < t1; t += 256) {
        const double3 p = wrap_item(w, t);
        s[0] += p.x;
        s[1] += p.y;
        s[2] += p.z;
        s[3] += fabs(p.x);
        s[4] += fabs(p.y);
        s[5] += fabs(p.z);
    }
    for (int k = 0; k < 6; k++) sh[k][tid] = s[k];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h)
            for (int k = 0; k < 6; k++) sh[k][tid] += sh[k][tid + h];
        __syncthreads();
    }
    if (tid < 6) w.parts[6 * (size_t)c + tid] = sh[tid][0];
}

// One thread per blob: the centroid from the partial sums, per axis the reference's test
// (low wall first, the high wall only when not below zero) and the move.
__global__ __launch_bounds__(64) void k_wrap_decide(WrapArgs w, int nblobs) {
    const int b = blockIdx.x * 64 + threadIdx.x;
    if (b >= nblobs) return;
    const int i0 = w.blob_item_off[b], i1 = w.blob_item_off[b + 1];
    const double n = (double)(i1 - i0);
    const double box[3] = {w.box_x, w.box_y, w.box_z};
    double d[3] = {0, 0, 0};
    bool move = false;
    for (int k = 0; k < 3; k++) {
        double s = 0.0, a = 0.0;
        for (int c = w.blob_chunk_off[b]; c < w.blob_chunk_off[b + 1]; c++) {
            s += w.parts[6 * (size_t)c + k];
            a += w.parts[6 * (size_t)c + 3 + k];
        }
        double com = s / n;
        const double u = 0.5 * __DBL_EPSILON__;
        const double margin = 2.5 * u * a + 4.0 * u * fabs(com) + __DBL_MIN__;
        if (fabs(com) <= margin || fabs(com - box[k]) <= margin) {  // too close to call: file order
            double e = 0.0;
            for (int t = i0; t < i1; t++) {
                const double3 p = wrap_item(w, t);
                e += k == 0 ? p.x : (k == 1 ? p.y : p.z);
            }
            com = e / n;
        }
        if (com < 0) {
            if (w.mask & (1 << (2 * k))) {
                d[k] += box[k];
                move = true;
            }
        } else if (com > box[k]) {
            if (w.mask & (2 << (2 * k))) {
                d[k] -= box[k];
                move = true;
            }
        }
    }
    w.shift[b] = make_double4(d[0], d[1], d[2], move ? 1.0 : 0.0);
}

// Blob::move (src/Blob.cpp:935-949) on the blobs that move.
__global__ __launch_bounds__(256) void k_wrap_move(WrapArgs w) {
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= w.N) return;
    const double4 d = w.shift[w.slice_blob[n >> 6]];
    if (d.w == 0.0) return;
    double4 p = w.pos[n];
    p.x += d.x;
    p.y += d.y;
    p.z += d.z;
    w.pos[n] = p;
}
