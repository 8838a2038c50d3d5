// ffh_forces.h -- forces from outside the element loop: springs, constant (linear and
// rotational) and surface forces (their per-step kernels and the helpers the force kernel uses).
// Part of ffh_kernels.hip: included inside its anonymous namespace, after the headers it
// lists there; not a stand-alone header.
#pragma once

// Forces from outside the element loop (ffh::ForcePlan on the device); null lists: none.
struct ForceArgs {
    const int32_t *pre_off, *pre_item;  // springs then constant forces, before the element forces
    const int2 *spring_nodes;
    const double2 *spring_kl;
    const double *cforce;
    int nconst;                         // items ~c with c >= nconst: rotational force c - nconst
    const double *rot;                  // [8 per rotational force]: axis point, direction, F, torque
    const int32_t *grp_off, *grp_ent_off, *grp_ent;  // surface-face groups, after them
    const double *trac;                               // [3 * entries], this step's tractions
    // midpoints carrying a force: their constant items and face groups, this step's totals,
    // and per vertex the fold terms (three midpoint indices each, -1 for none)
    const int32_t *mid_pre_off, *mid_grp_off, *mid_ends;
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

// A face corner's position: a vertex, or an edge midpoint as linearise_element places it
// (.5 (a + b), src/tetra_element_linear.cpp:567-591); a vertex is stored as (v, v), for which
// the same expression is exact.
__device__ __forceinline__ double4 corner(const double4 *pos, const int32_t *ab) {
    const double4 a = pos[ab[0]], b = pos[ab[1]];
    return make_double4(.5 * (a.x + b.x), .5 * (a.y + b.y), .5 * (a.z + b.z), 0.0);
}

// Component i of a rotational constant force on a node at p (Blob::apply_ctforces CTF 2,
// src/Blob.cpp:2104-2136): r from p to its foot on the axis by intersectingPointToLine
// (src/mat_vec_fns_II.cpp:632-680: a point within 3 eps of the axis is its own foot; c by
// detByRows, :261-269), f = r x d with arr3arr3VectorProduct's operand order (:127-134),
// scaled by F, or by F / |r| for a torque.  a: axis point, direction, F, torque flag.
__device__ __forceinline__ double rot_component(const double *a, double4 p, int i) {
    const double dx = a[3], dy = a[4], dz = a[5];
    const double ux = p.x - a[0], uy = p.y - a[1], uz = p.z - a[2];
    const double vx = uy * dz - dy * uz, vy = -ux * dz + dx * uz, vz = ux * dy - dx * uy;
    double s = 0.0;
    s += vx * vx;
    s += vy * vy;
    s += vz * vz;
    double rx = p.x, ry = p.y, rz = p.z;
    if (!(sqrt(s) < 3.0 * __DBL_EPSILON__)) {
        double cx = dy * vz - vy * dz, cy = -dx * vz + vx * dz, cz = dx * vy - vx * dy;
        const double wx = dy * cz - cy * dz, wy = -dx * cz + cx * dz, wz = dx * cy - cx * dy;
        double c = 0;
        c = ux * (dy * wz - dz * wy);
        c += uy * (dz * wx - dx * wz);
        c += uz * (dx * wy - dy * wx);
        c /= (wx * wx + wy * wy + wz * wz);
        cx = c * cx;
        cy = c * cy;
        cz = c * cz;
        rx = p.x + cx;
        ry = p.y + cy;
        rz = p.z + cz;
    }
    rx = rx - p.x;
    ry = ry - p.y;
    rz = rz - p.z;
    const double f = i == 0 ? ry * dz - dy * rz : (i == 1 ? -rx * dz + dx * rz : rx * dy - dx * ry);
    if (a[7] == 0.0) return a[6] * f;
    double m = 0.0;
    m += rx * rx;
    m += ry * ry;
    m += rz * rz;
    return (a[6] / sqrt(m)) * f;
}

// Component i of constant-force item ~c: linear (c < nconst) or rotational, on a node at p.
__device__ __forceinline__ double const_component(const ForceArgs &fa, int c, const double4 *p, int i) {
    return c < fa.nconst ? fa.cforce[3 * (size_t)c + i] : rot_component(fa.rot + 8 * (size_t)(c - fa.nconst), *p, i);
}

// This step's force on every midpoint that carries one: constant forces, then face forces
// (the midpoint's element forces are zero, its drag too), before linearise_force moves it.
__global__ __launch_bounds__(256) void k_mid_forces(ForceArgs fa, int M) {
    const int m = blockIdx.x * 256 + threadIdx.x;
    if (m >= M) return;
    const double4 pm = corner(fa.pos, fa.mid_ends + 2 * (size_t)m);
    for (int i = 0; i < 3; i++) {
        double f = 0.0;
        if (fa.mid_pre_off)
            for (int t = fa.mid_pre_off[m]; t < fa.mid_pre_off[m + 1]; t++) f += const_component(fa, ~fa.pre_item[t], &pm, i);
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
