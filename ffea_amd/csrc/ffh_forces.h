// ffh_forces.h -- forces from outside the element loop: springs, constant and
// surface forces (their per-step kernels and the helpers the force kernel uses).
// Part of ffh_kernels.hip: included inside its anonymous namespace, after the headers it
// lists there; not a stand-alone header.
#pragma once

// Forces from outside the element loop (ffh::ForcePlan on the device); null lists: none.
struct ForceArgs {
    const int32_t *pre_off, *pre_item;  // springs then constant forces, before the element forces
    const int2 *spring_nodes;
    const double2 *spring_kl;
    const double *cforce;
    const int32_t *grp_off, *grp_ent_off, *grp_ent;  // surface-face groups, after them
    const double *trac;                               // [3 * entries], this step's tractions
    // midpoints carrying a force: their constant items and face groups, this step's totals,
    // and per vertex the fold terms (three midpoint indices each, -1 for none)
    const int32_t *mid_pre_off, *mid_grp_off;
    double *mf;
    const int32_t *fold_off, *fold_mid;
    const double4 *pos;
};

// The sum of a node's face groups for component i: each group is one face's force (its set
// tractions summed in set order from zero), added face by face.
__device__ __forceinline__ double face_groups(const ForceArgs &fa, int q0, int q1, int i, double f) {
    for (int q = q0; q < q1; q++) {
        double sf = 0.0;
        const int t1 = fa.grp_ent_off[q + 1];
        for (int t = fa.grp_ent_off[q]; t < t1; t++) sf += fa.trac[3 * (size_t)fa.grp_ent[t] + i];
        f += sf;
    }
    return f;
}

// This step's force on every midpoint that carries one: constant forces, then face forces
// (the midpoint's element forces are zero, its drag too), before linearise_force moves it.
__global__ __launch_bounds__(256) void k_mid_forces(ForceArgs fa, int M) {
    const int m = blockIdx.x * 256 + threadIdx.x;
    if (m >= M) return;
    for (int i = 0; i < 3; i++) {
        double f = 0.0;
        if (fa.mid_pre_off)
            for (int t = fa.mid_pre_off[m]; t < fa.mid_pre_off[m + 1]; t++) f += fa.cforce[3 * (size_t)(~fa.pre_item[t]) + i];
        if (fa.mid_grp_off) f = face_groups(fa, fa.mid_grp_off[m], fa.mid_grp_off[m + 1], i, f);
        fa.mf[3 * (size_t)m + i] = f;
    }
}

// Component i of spring s's force on its end `end` (World::apply_springs,
// src/World.cpp:3812-3842): direction by arr3Normalise2 (src/mat_vec_fns_II.cpp:164-176, zero
// for coincident ends), magnitude k (|sep| - l) with mag (:229-239), the far end negated.
__device__ __forceinline__ double spring_component(const ForceArgs &fa, int s, int end, int i) {
    const int2 nn = fa.spring_nodes[s];
    const double4 p0 = fa.pos[nn.x], p1 = fa.pos[nn.y];
    const double sx = p1.x - p0.x, sy = p1.y - p0.y, sz = p1.z - p0.z;
    double norm = 0.0;
    norm += sx * sx;
    norm += sy * sy;
    norm += sz * sz;
    const double c = i == 0 ? sx : (i == 1 ? sy : sz);
    const double len = sqrt(norm);
    const double dir = norm == 0.0 ? 0.0 : c / len;
    const double2 kl = fa.spring_kl[s];
    const double fm = kl.x * (len - kl.y);
    return end == 0 ? fm * dir : (-1.0 * fm) * dir;
}

// A face corner's position: a vertex, or an edge midpoint as linearise_element places it
// (.5 (a + b), src/tetra_element_linear.cpp:567-591); a vertex is stored as (v, v), for which
// the same expression is exact.
__device__ __forceinline__ double4 corner(const double4 *pos, const int32_t *ab) {
    const double4 a = pos[ab[0]], b = pos[ab[1]];
    return make_double4(.5 * (a.x + b.x), .5 * (a.y + b.y), .5 * (a.z + b.z), 0.0);
}

// Face::get_area (src/Face.cpp:244-261); ends: the three corners' edge-end pairs.
__device__ __forceinline__ double face_area(const double4 *pos, const int32_t *ends) {
    const double4 a0 = corner(pos, ends), a1 = corner(pos, ends + 2), a2 = corner(pos, ends + 4);
    const double ax = a1.x - a0.x, ay = a1.y - a0.y, az = a1.z - a0.z;
    const double bx = a2.x - a0.x, by = a2.y - a0.y, bz = a2.z - a0.z;
    const double tx = ay * bz - az * by, ty = az * bx - ax * bz, tz = ax * by - ay * bx;
    return .5 * sqrt(tx * tx + ty * ty + tz * tz);
}

// Linear surface forces (Blob::apply_ctforces CTF 3, src/Blob.cpp:2137-2159): one workgroup
// per set; each face's share of the set's force is its area over three times the set's total
// area, summed in the set's face order by one thread.
struct FaceArgs {
    const int32_t *set_off, *ent_ends;
    const double *set_force;
    const double4 *pos;
    double *area, *trac;
};
__global__ __launch_bounds__(256) void k_face_tractions(FaceArgs a) {
    __shared__ double total;
    const int s = blockIdx.x;
    const int e0 = a.set_off[s], e1 = a.set_off[s + 1];
    for (int e = e0 + (int)threadIdx.x; e < e1; e += 256) a.area[e] = face_area(a.pos, a.ent_ends + 6 * (size_t)e);
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int e = e0; e < e1; e++) t += a.area[e];
        total = t;
    }
    __syncthreads();
    const double t3 = 3 * total;
    for (int e = e0 + (int)threadIdx.x; e < e1; e += 256) {
        const double f = a.area[e] / t3;
        for (int k = 0; k < 3; k++) a.trac[3 * (size_t)e + k] = f * a.set_force[3 * (size_t)s + k];
    }
}

// Spring energies (World::make_measurements, src/World.cpp:4322-4328), one per spring.
__global__ __launch_bounds__(256) void k_spring_energy(ForceArgs fa, int n, double *out) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    const int2 nn = fa.spring_nodes[s];
    const double4 a = fa.pos[nn.x], b = fa.pos[nn.y];
    const double cx = a.x - b.x, cy = a.y - b.y, cz = a.z - b.z;
    double m = 0.0;
    m += cx * cx;
    m += cy * cy;
    m += cz * cz;
    m = sqrt(m);
    const double2 kl = fa.spring_kl[s];
    out[s] = 0.5 * kl.x * (m - kl.y) * (m - kl.y);
}
