// ffh_steric.h -- the steric volume exclusion between two surface face elements (the tetrahedron
// of a face's three corners and its opposite element vertex): the overlap test tet_a_tetII
// (src/CheckTetrahedraOverlap.cpp:208-337), the common volume volumeIntersectionII
// (src/VolumeIntersection.cpp:951-1024), the local coordinates getLocalCoordinatesForLinTet
// (src/mat_vec_fns_II.cpp:725-733), and the forces of Face::
// getTetraIntersectionVolumeTotalGradientAndShapeFunctions with VdW_solver::
// do_steric_interaction (src/Face.cpp:473-514, src/VdW_solver.cpp:458-487).  Every sum and
// product rounds as the reference writes it (the library is built without FMA contraction;
// division and sqrt are correctly rounded), so the results are the reference's bit for bit.
// A pair runs on one lane: the overlap test is a few dozen operations; the volume, needed only
// where face elements overlap, is branchy scalar work with the polyhedron's point list
// (at most 56 points) in private memory, evaluated seven times for the gradient.
// Part of ffh_kernels.hip: included inside its anonymous namespace; not a stand-alone header.
#pragma once

namespace st {

// the helpers of src/mat_vec_fns_II.cpp; dot products and squared norms start from 0.0 and
// add the terms in index order, as the reference's loops do
__device__ __forceinline__ double dot(const double *a, const double *b) {
    double r = 0.0;
    r += a[0] * b[0];
    r += a[1] * b[1];
    r += a[2] * b[2];
    return r;
}
__device__ __forceinline__ void sub(const double *a, const double *b, double *r) {
    r[0] = a[0] - b[0];
    r[1] = a[1] - b[1];
    r[2] = a[2] - b[2];
}
__device__ __forceinline__ void add(const double *a, const double *b, double *r) {
    r[0] = a[0] + b[0];
    r[1] = a[1] + b[1];
    r[2] = a[2] + b[2];
}
__device__ __forceinline__ void flip(double *u) {
    u[0] = -1.0 * u[0];
    u[1] = -1.0 * u[1];
    u[2] = -1.0 * u[2];
}
__device__ __forceinline__ void cross(const double *u, const double *v, double *w) {
    w[0] = u[1] * v[2] - v[1] * u[2];
    w[1] = -u[0] * v[2] + v[0] * u[2];
    w[2] = u[0] * v[1] - v[0] * u[1];
}
// getUnitNormal (:311-322)
__device__ __forceinline__ void unit_cross(const double *u, const double *v, double *w) {
    cross(u, v, w);
    const double l = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    w[0] /= l;
    w[1] /= l;
    w[2] /= l;
}
// tangent (:295-306): the unit vector from b to a
__device__ __forceinline__ void tangent(const double *a, const double *b, double *t) {
    double w = 0;
    for (int i = 0; i < 3; i++) {
        t[i] = a[i] - b[i];
        w += t[i] * t[i];
    }
    w = sqrt(w);
    t[0] /= w;
    t[1] /= w;
    t[2] /= w;
}
// arr3Normalise2 (:164-177); a zero vector is where the reference throws
__device__ __forceinline__ void normalise_into(const double *e, double *n, int &degenerate) {
    double s = 0.0;
    s += e[0] * e[0];
    s += e[1] * e[1];
    s += e[2] * e[2];
    if (s == 0.0) degenerate = 1;
    s = sqrt(s);
    n[0] = e[0] / s;
    n[1] = e[1] / s;
    n[2] = e[2] / s;
}
__device__ __forceinline__ double distance(const double *a, const double *b) {
    double d = 0.0;
    for (int i = 0; i < 3; i++) d += (a[i] - b[i]) * (a[i] - b[i]);
    return sqrt(d);
}
__device__ __forceinline__ bool same_sign(double a, double b) { return a * b >= 0; }

// getNormalInwards(f0, f1, f2, p3) (:361-377): unit normal of f0 f1 f2 turned towards p3
__device__ __forceinline__ void normal_inwards(const double *f0, const double *f1, const double *f2, const double *p3,
                                               double *n, int &degenerate) {
    double pl1[3], pl2[3], aux[3];
    sub(f1, f0, pl1);
    sub(f2, f0, pl2);
    cross(pl1, pl2, aux);
    normalise_into(aux, n, degenerate);
    add(aux, f0, aux);
    const double d = -dot(n, f0);
    if (!same_sign(dot(p3, n) + d, dot(aux, n) + d)) flip(n);
}

// sameSidePlane (:383-398), arr3Normalise without the zero test
__device__ __forceinline__ bool same_side_plane(const double *v, const double *test, const double *p1,
                                                const double *p2, const double *p3) {
    double pl1[3], pl2[3], n[3];
    sub(p2, p1, pl1);
    sub(p3, p1, pl2);
    cross(pl1, pl2, n);
    double s = 0.0;
    s += n[0] * n[0];
    s += n[1] * n[1];
    s += n[2] * n[2];
    s = sqrt(s);
    n[0] /= s;
    n[1] /= s;
    n[2] /= s;
    const double d = -dot(n, p1);
    return same_sign(dot(v, n) + d, dot(test, n) + d);
}

// sameSideLine (:405-418)
__device__ __forceinline__ bool same_side_line(const double *e, const double *p1, const double *p2, const double *p3) {
    double pl23[3], pl21[3], pl2e[3], v1[3], ve[3];
    sub(p3, p2, pl23);
    sub(p1, p2, pl21);
    sub(e, p2, pl2e);
    cross(pl21, pl23, v1);
    cross(pl2e, pl23, ve);
    return dot(v1, ve) >= 0;
}

__device__ __forceinline__ bool node_in_tet(const double *v, const double *t0, const double *t1, const double *t2,
                                            const double *t3) {
    return same_side_plane(v, t0, t1, t2, t3) && same_side_plane(v, t1, t2, t3, t0) &&
           same_side_plane(v, t2, t3, t0, t1) && same_side_plane(v, t3, t0, t1, t2);
}

// intersectionPoint (:573-588) with linePlaneIntersectionPoint (:450-470), isPointInFace
// (:531-538)
__device__ bool edge_face_point(double *ip, const double *e1, const double *e2, const double *f1, const double *f2,
                                const double *f3) {
    if (same_side_plane(e1, e2, f1, f2, f3)) return false;
    double v[3], pl1[3], pl2[3], pn[3];
    sub(e2, e1, v);
    sub(f2, f1, pl1);
    sub(f3, f2, pl2);
    unit_cross(pl1, pl2, pn);
    const double d = -dot(pn, f1);
    const double t = -(dot(pn, e1) + d) / dot(pn, v);
    v[0] = t * v[0];
    v[1] = t * v[1];
    v[2] = t * v[2];
    add(e1, v, ip);
    return same_side_line(ip, f1, f2, f3) && same_side_line(ip, f3, f1, f2) && same_side_line(ip, f2, f3, f1);
}

// getBAndN(f0, f1, f2, p3, t) (src/VolumeIntersection.cpp:179-204)
__device__ __forceinline__ void face_frame(const double *f0, const double *f1, const double *f2, const double *p3,
                                           const double *t, double *b, double *n, int &degenerate) {
    double pl2[3], aux[3];
    sub(f2, f0, pl2);
    cross(t, pl2, aux);
    normalise_into(aux, b, degenerate);
    add(aux, f0, aux);
    const double d = -dot(b, f0);
    if (!same_sign(dot(p3, b) + d, dot(aux, b) + d)) flip(b);
    unit_cross(b, t, n);
    add(f0, n, aux);
    if (!same_side_line(aux, f2, f1, f0)) flip(n);
}

// volumeForNode(o0, o1, o2, v) (:259-298): edges v->o0, v->o1, v->o2, each with its two faces
__device__ double vertex_term(const double *o0, const double *o1, const double *o2, const double *v, int &degenerate) {
    const double *o[3] = {o0, o1, o2};
    double vol = 0.0, t[3], b[3], n[3];
#pragma unroll
    for (int e = 0; e < 3; e++) {
        tangent(o[e], v, t);
        const double pt = dot(v, t);
        const double *q1 = o[e == 0 ? 1 : 0], *q2 = o[e == 2 ? 1 : 2];
        face_frame(v, o[e], q1, q2, t, b, n, degenerate);
        vol += pt * dot(v, n) * dot(v, b);
        face_frame(v, o[e], q2, q1, t, b, n, degenerate);
        vol += pt * dot(v, n) * dot(v, b);
    }
    return vol;
}

// volumeForIntPointII (:547-655): ip on edge e1 e2 (tetrahedron e1 e2 e3 e4) and face f1 f2 f3
// (tetrahedron f1 f2 f3 f4)
__device__ double crossing_term(const double *ip, const double *e1, const double *e2, const double *e3,
                                const double *e4, const double *f1, const double *f2, const double *f3,
                                const double *f4, int &degenerate) {
    double B[3][3], T[3][3], aux[3], n[3];
    normal_inwards(f1, f2, f3, f4, B[0], degenerate);
    tangent(e2, e1, T[0]);
    if (same_side_plane(e1, f4, f1, f2, f3)) flip(T[0]);
    normal_inwards(e1, e2, e3, e4, B[1], degenerate);
    unit_cross(B[0], B[1], T[1]);
    add(ip, T[1], aux);
    if (!same_side_plane(aux, e3, e4, e1, e2)) flip(T[1]);
    normal_inwards(e1, e2, e4, e3, B[2], degenerate);
    unit_cross(B[0], B[2], T[2]);
    add(ip, T[2], aux);
    if (!same_side_plane(aux, e4, e3, e1, e2)) flip(T[2]);
    double ipT[3], ipB[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        ipT[i] = dot(ip, T[i]);
        ipB[i] = dot(ip, B[i]);
    }
    double vol = 0;
    unit_cross(B[1], T[0], n);
    add(ip, n, aux);
    if (!same_side_line(aux, e3, e2, e1)) flip(n);
    vol += ipT[0] * dot(ip, n) * ipB[1];
    unit_cross(B[2], T[0], n);
    add(ip, n, aux);
    if (!same_side_line(aux, e4, e2, e1)) flip(n);
    vol += ipT[0] * dot(ip, n) * ipB[2];
    double V0[3], V1[3];
    add(ip, T[1], V0);
    add(ip, T[2], V1);
    // i = 0: the line V0 -> V1 ... (V[i + 1], V[i]) = (V1, V0); i = 1: (V0, V1)
#pragma unroll
    for (int i = 0; i < 2; i++) {
        unit_cross(B[0], T[i + 1], n);
        add(ip, n, aux);
        if (!same_side_line(aux, i == 0 ? V1 : V0, i == 0 ? V0 : V1, ip)) flip(n);
        vol += ipT[i + 1] * dot(ip, n) * ipB[0];
        unit_cross(B[i + 1], T[i + 1], n);
        if (dot(n, B[0]) < 0) flip(n);
        vol += ipT[i + 1] * dot(ip, n) * ipB[i + 1];
    }
    return vol;
}

// The intersection polyhedron's points in the order found (W[56], ips).
struct Points {
    double w[56][3];
    int n;
};

// exists (:32-43)
__device__ __forceinline__ bool listed(const Points &P, const double *p) {
    for (int i = 0; i < P.n; i++)
        if (distance(p, P.w[i]) < 3.0 * 0x1p-52) return true;  // ffea_const::threeErr
    return false;
}
__device__ __forceinline__ void append(Points &P, const double *p) {
    P.w[P.n][0] = p[0];
    P.w[P.n][1] = p[1];
    P.w[P.n][2] = p[2];
    P.n++;
}

// contribVolForIntersections (:929-949): edge ai aj (tetrahedron ai aj a3 a4) against the faces
// (b0 b1 b2), (b1 b2 b3), (b2 b3 b0), (b3 b0 b1)
__device__ void edge_terms(double &vol, Points &P, const double *ai, const double *aj, const double *a3,
                           const double *a4, const double (*b)[3], int &degenerate) {
    double ip[3];
    for (int f = 0; f < 4; f++) {
        const double *q0 = b[f], *q1 = b[(f + 1) & 3], *q2 = b[(f + 2) & 3], *q3 = b[(f + 3) & 3];
        if (edge_face_point(ip, ai, aj, q0, q1, q2) && !listed(P, ip)) {
            append(P, ip);
            vol += crossing_term(ip, ai, aj, a3, a4, q0, q1, q2, q3, degenerate);
        }
    }
}

// volumeIntersectionII with maxVolume (:48-66) and findCM (:95-104): -1/6 of the vertex terms'
// sum (vertices of A inside B, of B inside A, then the edge-face crossings, edge pairs
// (0,1) (0,2) (0,3) (1,2) (1,3) (2,3), A's edge before B's); 0 for three points or fewer;
// replaced by 8 d^3 where negative or larger (d from the points' mean to the last point)
__device__ double intersection(const double (*A)[3], const double (*B)[3], double *cm, int &degenerate) {
    Points P;
    P.n = 0;
    double vol = 0.0;
    for (int side = 0; side < 2; side++) {
        const double(*X)[3] = side ? B : A;
        const double(*Y)[3] = side ? A : B;
        for (int v = 0; v < 4; v++) {
            if (listed(P, X[v]) || !node_in_tet(X[v], Y[0], Y[1], Y[2], Y[3])) continue;
            append(P, X[v]);
            const int r0 = v == 0 ? 1 : 0, r1 = v <= 1 ? 2 : 1, r2 = v <= 2 ? 3 : 2;
            vol += vertex_term(X[r0], X[r1], X[r2], X[v], degenerate);
        }
    }
    constexpr int edge[6][4] = {{0, 1, 2, 3}, {0, 2, 1, 3}, {0, 3, 1, 2}, {1, 2, 0, 3}, {1, 3, 0, 2}, {2, 3, 0, 1}};
    for (int e = 0; e < 6; e++) {
        const int *k = edge[e];
        edge_terms(vol, P, A[k[0]], A[k[1]], A[k[2]], A[k[3]], B, degenerate);
        edge_terms(vol, P, B[k[0]], B[k[1]], B[k[2]], B[k[3]], A, degenerate);
    }
    vol *= -0.166666666666666667;
    if (P.n <= 3) return 0.;
    double c[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < P.n; i++) add(c, P.w[i], c);
    const double s = 1. / P.n;
    c[0] = s * c[0];
    c[1] = s * c[1];
    c[2] = s * c[2];
    double d0 = 0.0, di = 0.0;
    for (int i = 0; i < P.n; i++) {
        di = distance(c, P.w[i]);
        if (di > d0) d0 = di;
    }
    const double w = 8.0 * di * di * di;
    if (fabs(vol) > w || vol < 0) vol = w;
    if (cm) {
        cm[0] = cm[1] = cm[2] = 0.0;
        for (int i = 0; i < P.n; i++) add(cm, P.w[i], cm);
        cm[0] = s * cm[0];
        cm[1] = s * cm[1];
        cm[2] = s * cm[2];
    }
    return vol;
}

// getTetrahedraVolume (src/mat_vec_fns_II.cpp:708-714): six times the signed volume
__device__ __forceinline__ double tet6(const double *p0, const double *p1, const double *p2, const double *p3) {
    return (p1[0] - p0[0]) * ((p1[1] - p2[1]) * (p2[2] - p3[2]) - (p2[1] - p3[1]) * (p1[2] - p2[2])) +
           (p2[0] - p1[0]) * ((p2[1] - p3[1]) * (p0[2] - p1[2]) - (p0[1] - p1[1]) * (p2[2] - p3[2])) +
           (p3[0] - p2[0]) * ((p0[1] - p1[1]) * (p1[2] - p2[2]) - (p1[1] - p2[1]) * (p0[2] - p1[2]));
}

__device__ __forceinline__ void local_coords(const double (*t)[3], const double *p, double *phi) {
    phi[0] = tet6(p, t[1], t[2], t[3]);
    phi[1] = -tet6(p, t[0], t[2], t[3]);
    phi[2] = tet6(p, t[0], t[1], t[3]);
    phi[3] = -tet6(p, t[0], t[1], t[2]);
    const double v = tet6(t[0], t[1], t[2], t[3]);
    for (int i = 0; i < 4; i++) phi[i] /= v;
}

// tet_a_tetII: Ganovelli, Ponchio and Rocchini's separating-axis test (JGT 2002).  C[f][k] is
// vertex k of B against A's face f; a face with all of B outside, or an edge of A whose two
// faces' coordinates admit a separating plane, separates; unless some vertex of B is inside
// all of A's faces, B's faces are tried the same way (faces only).
__device__ __forceinline__ int outside_mask(const double *c) {
    return (c[0] > 0.0 ? 1 : 0) | (c[1] > 0.0 ? 2 : 0) | (c[2] > 0.0 ? 4 : 0) | (c[3] > 0.0 ? 8 : 0);
}
__device__ __forceinline__ bool edge_separates(const double *c0, const double *c1, int m0, int m1) {
    if ((m0 | m1) != 15) return false;
    m0 &= (m0 ^ m1);
    m1 &= (m0 ^ m1);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = i + 1; j < 4; j++) {
            const double det = c0[j] * c1[i] - c0[i] * c1[j];
            if ((m0 & (1 << i)) && (m1 & (1 << j)) && det > 0.0) return false;
            if ((m0 & (1 << j)) && (m1 & (1 << i)) && det < 0.0) return false;
        }
    return true;
}
__device__ __forceinline__ double plane_coord(const double *p, const double *o, const double *n) {
    return (p[0] - o[0]) * n[0] + (p[1] - o[1]) * n[1] + (p[2] - o[2]) * n[2];
}
__device__ bool tets_overlap(const double (*A)[3], const double (*B)[3]) {
    double C[4][4], n[3], e0[3], e1[3], e2[3];
    int mask[4];
    sub(A[1], A[0], e0);
    sub(A[2], A[0], e1);
    sub(A[3], A[0], e2);
#pragma unroll
    for (int f = 0; f < 4; f++) {
        const double *o = A[0];
        if (f == 0) {
            cross(e0, e1, n);
        } else if (f == 1) {
            cross(e2, e0, n);
        } else if (f == 2) {
            cross(e1, e2, n);
        } else {
            double e4[3], e3[3];
            sub(A[3], A[1], e4);
            sub(A[2], A[1], e3);
            cross(e4, e3, n);
            o = A[1];
        }
        // faces 0..2 dot (B_k - A0) formed first, as P_V1; face 3 the difference inside the sum
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (f < 3) {
                double p[3];
                sub(B[k], A[0], p);
                C[f][k] = p[0] * n[0] + p[1] * n[1] + p[2] * n[2];
            } else {
                C[f][k] = plane_coord(B[k], o, n);
            }
        }
        mask[f] = outside_mask(C[f]);
        if (mask[f] == 15) return false;
#pragma unroll
        for (int g = 0; g < f; g++)
            if (edge_separates(C[g], C[f], mask[g], mask[f])) return false;
    }
    if ((mask[0] | mask[1] | mask[2] | mask[3]) != 15) return true;
    double d0[3], d1[3], d2[3];
    sub(B[1], B[0], d0);
    sub(B[2], B[0], d1);
    sub(B[3], B[0], d2);
#pragma unroll
    for (int f = 0; f < 4; f++) {
        if (f == 0) {
            cross(d0, d1, n);
        } else if (f == 1) {
            cross(d2, d0, n);
        } else if (f == 2) {
            cross(d1, d2, n);
        } else {
            double d4[3], d3[3];
            sub(B[3], B[1], d4);
            sub(B[2], B[1], d3);
            cross(d4, d3, n);
        }
        bool all = true;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            double s;
            if (f < 3) {
                double p[3];
                sub(A[k], B[0], p);
                s = p[0] * n[0] + p[1] * n[1] + p[2] * n[2];
            } else {
                s = plane_coord(A[k], B[1], n);
            }
            all = all && s > 0.0;
        }
        if (all) return false;
    }
    return true;
}

// The pair forces: V = volume(A, B) with the points' mean cm; phi1, phi2 the local coordinates
// of cm; dV/dr_d = (V(A, B + dr e_d) - V) / dr - (V(B, A + dr e_d) - V) / dr; V and the gradient
// scaled by the steric factor; energy V, forces phi1_j g on A's node j and -phi2_j g on B's.
// Returns false (nothing applied) where V is 0.
__device__ __noinline__ bool pair_forces(const double (*A)[3], const double (*B)[3], double dr, double factor,
                                         double *energy, double *fA, double *fB, int &degenerate) {
    double cm[3], phi1[4], phi2[4], g[3];
    double vol = intersection(A, B, cm, degenerate);
    if (vol == 0) return false;
    local_coords(A, cm, phi1);
    local_coords(B, cm, phi2);
    for (int dir = 0; dir < 3; dir++) {
        double sA[4][3], sB[4][3];
        for (int i = 0; i < 4; i++)
            for (int j = 0; j < 3; j++) {
                const double dx = j == dir ? dr : 0.0;
                sA[i][j] = A[i][j] + dx;
                sB[i][j] = B[i][j] + dx;
            }
        double vm = intersection(A, sB, nullptr, degenerate);
        g[dir] = (vm - vol) / dr;
        vm = intersection(B, sA, nullptr, degenerate);
        g[dir] -= (vm - vol) / dr;
    }
    vol *= factor;
    g[0] = factor * g[0];
    g[1] = factor * g[1];
    g[2] = factor * g[2];
    *energy = vol;
    for (int j = 0; j < 4; j++)
        for (int x = 0; x < 3; x++) {
            fA[3 * j + x] = phi1[j] * g[x];
            fB[3 * j + x] = (-1.0 * phi2[j]) * g[x];
        }
    return true;
}

}  // namespace st
