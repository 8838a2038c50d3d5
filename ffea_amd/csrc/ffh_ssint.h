// ffh_ssint.h -- surface-surface interactions (VdW_solver, src/VdW_solver.cpp:158-487,
// 663-831): the lennard-jones type, ljsteric's Lennard-Jones with its soft core below Rmin,
// gensoft, and the steric volume exclusion (ffh_steric.h) that steric applies alone and
// ljsteric and gensoft apply instead of their potential where two face elements overlap; face
// geometry and the lookup grid at es_update steps, and every step each active face's
// Face::force[0..3] summed in the reference's order.
// Part of ffh_kernels.hip: included inside its anonymous namespace, after the headers it
// lists there; not a stand-alone header.
#pragma once

// VdW_solver's three-point triangle rule (src/VdW_solver.cpp:62-73): the weight and the
// barycentric coordinates as the decimal constants the reference writes.
constexpr double kSsW = 0.333333333333333;
__constant__ double kSsEta[3][3] = {{0.666666666666667, 0.166666666666667, 0.166666666666667},
                                    {0.166666666666667, 0.666666666666667, 0.166666666666667},
                                    {0.166666666666667, 0.166666666666667, 0.666666666666667}};

struct SsCtl {
    int es_count, es_update, refresh, pad;
};

// The pool is the active faces in the reference's order (blobs in order, faces in surface
// order, src/World.cpp:940-962); pool index = LinkedListNode index.
struct SsArgs {
    const double4 *pos;
    const int32_t *ends;   // [6P] each corner's edge ends (compact ids; a vertex twice)
    const int32_t *blob;   // [P]
    const int32_t *type;   // [P]
    const double *emin, *rmin, *k0;  // [ntypes^2] meso, entry type1 * ntypes + type2
    int ntypes;
    int kind;              // 1 lennard-jones, 2 ljsteric (its soft core below Rmin), 3 gensoft,
                           // 4 steric; 2-4 apply the steric forces where face elements overlap
    const int32_t *nodes4; // [4P] corners and opposite element vertex (N_t ids), for kinds 2-4
    const int32_t *opp;    // [P] the opposite vertex (compact id), for kinds 2-4
    double st_factor, st_dr;  // steric_factor, steric_dr (meso)
    int32_t *deferred;     // [P] kinds 2-4: 1 where some pair of the face overlaps (this step)
    double *geom;          // [8P] centroid, unit normal, area (Face::calc_area_normal_centroid)
    int32_t *cell;         // [P] lookup-grid cell of each face at the last refresh
    int32_t *order;        // [P] 0..P-1, the sort's values
    int32_t *sorted, *sorted_cell;  // [P] pool ids ordered by cell, ascending id within a cell
    int32_t *cstart, *cend;         // [ncell] each cell's range in `sorted`
    int nx, ny, nz;
    double h;
    int inc_self, P, ncell;
    SsCtl *ctl;
    int *err;              // host-visible: 1 a face centroid fell outside the grid, 2 a face
                           // element too degenerate for the steric volume (the reference throws)
    double *out;           // [12P] Face::force[0..3] of every pool face: corners, opposite vertex
};

// LinkedListCube::pbc (src/LinkedListCube.tpp:282-299): one period at most.
__device__ __forceinline__ int ss_wrap(int x, int n) { return x < 0 ? x + n : (x >= n ? x - n : x); }

// adjacent_cell_lookup_table entry c (src/VdW_solver.cpp:31-59): z fastest, each -1..1.
__device__ __forceinline__ int ss_neighbour(const SsArgs &a, int cell, int c) {
    const int cx = cell / (a.ny * a.nz), cy = (cell / a.nz) % a.ny, cz = cell % a.nz;
    const int x = ss_wrap(cx + c / 9 - 1, a.nx), y = ss_wrap(cy + (c / 3) % 3 - 1, a.ny), z = ss_wrap(cz + c % 3 - 1, a.nz);
    return x * a.ny * a.nz + y * a.nz + z;
}

// es_update bookkeeping of World::run (src/World.cpp:1805, 1821-1837): a refresh on the first
// step and every es_update steps after it.
__global__ void k_ss_tick(SsCtl *c) {
    if (threadIdx.x != 0) return;
    const int r = c->es_count == c->es_update;
    c->refresh = r;
    c->es_count = r ? 1 : c->es_count + 1;
}

// Face::calc_area_normal_centroid (src/Face.cpp:197-219) and the face's cell in
// NearestNeighbourLinkedListCube::build_nearest_neighbour_lookup
// (src/NearestNeighbourLinkedListCube.cpp:26-77), at refresh steps.
__global__ __launch_bounds__(256) void k_ss_geom(SsArgs a) {
    const int f = blockIdx.x * 256 + threadIdx.x;
    if (f >= a.P || !a.ctl->refresh) return;
    const double4 p0 = corner(a.pos, a.ends + 6 * (size_t)f), p1 = corner(a.pos, a.ends + 6 * (size_t)f + 2),
                  p2 = corner(a.pos, a.ends + 6 * (size_t)f + 4);
    const double ax = p1.x - p0.x, ay = p1.y - p0.y, az = p1.z - p0.z;
    const double bx = p2.x - p0.x, by = p2.y - p0.y, bz = p2.z - p0.z;
    const double nx = ay * bz - az * by, ny = az * bx - ax * bz, nz = ax * by - ay * bx;
    const double m = sqrt(nx * nx + ny * ny + nz * nz);
    double *g = a.geom + 8 * (size_t)f;
    g[6] = .5 * m;
    g[3] = nx / m;
    g[4] = ny / m;
    g[5] = nz / m;
    g[0] = (1.0 / 3.0) * (p0.x + p1.x + p2.x);
    g[1] = (1.0 / 3.0) * (p0.y + p1.y + p2.y);
    g[2] = (1.0 / 3.0) * (p0.z + p1.z + p2.z);
    int c[3];
    const int n[3] = {a.nx, a.ny, a.nz};
    bool ok = true;
    for (int k = 0; k < 3; k++) {
        const double q = floor(g[k] / a.h);
        ok = ok && q > -4.0e8 && q < 4.0e8;
        c[k] = ok ? ss_wrap((int)q, n[k]) : 0;
        ok = ok && c[k] >= 0 && c[k] < n[k];
    }
    if (!ok) {
        *a.err = 1;
        c[0] = c[1] = c[2] = 0;
    }
    a.cell[f] = c[0] * a.ny * a.nz + c[1] * a.nz + c[2];
}

__global__ __launch_bounds__(256) void k_ss_clear(SsArgs a) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= a.ncell || !a.ctl->refresh) return;
    a.cstart[c] = 0;
    a.cend[c] = 0;
}

__global__ __launch_bounds__(256) void k_ss_bounds(SsArgs a) {
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= a.P || !a.ctl->refresh) return;
    const int c = a.sorted_cell[s];
    if (s == 0 || a.sorted_cell[s - 1] != c) a.cstart[c] = s;
    if (s == a.P - 1 || a.sorted_cell[s + 1] != c) a.cend[c] = s + 1;
}

// consider_interaction (src/VdW_solver.cpp:332-384): other blob unless inc_self_ssint, normals
// whose dot product is not above zero; for the types other than lennard-jones also faces in
// front of each other (sep from centroid i to centroid j, arr3arr3DotProduct: sep.n_i < 0 with
// sep.n_j > 0 rejects) and, within one blob, no shared node among corners and opposite vertex.
__device__ __forceinline__ bool ss_consider(const SsArgs &a, int i, int j) {
    const bool same = a.blob[i] == a.blob[j];
    if (!a.inc_self && same) return false;
    const double *gi = a.geom + 8 * (size_t)i, *gj = a.geom + 8 * (size_t)j;
    if (gi[3] * gj[3] + gi[4] * gj[4] + gi[5] * gj[5] > 0.0) return false;
    if (a.kind == 1) return true;
    const double sx = gj[0] - gi[0], sy = gj[1] - gi[1], sz = gj[2] - gi[2];
    double d1 = 0.0, d2 = 0.0;
    d1 += sx * gi[3];
    d1 += sy * gi[4];
    d1 += sz * gi[5];
    d2 += sx * gj[3];
    d2 += sy * gj[4];
    d2 += sz * gj[5];
    if (d1 < 0 && d2 > 0) return false;
    if (a.inc_self && same) {
        const int32_t *ni = a.nodes4 + 4 * (size_t)i, *nj = a.nodes4 + 4 * (size_t)j;
        if (ni[3] == nj[3]) return false;
        for (int p = 0; p < 4; p++)
            for (int q = 0; q < 4; q++)
                if (nj[q] == ni[p]) return false;
    }
    return true;
}

// do_lj_interaction (src/VdW_solver.cpp:236-324) for faces f1 < f2 with
// calc_lj_force_pair_matrix and calc_lj_factors (:663-731): side 0 adds face 1's corner forces
// to acc, side 1 face 2's; the pair's energy (scaled by both areas) goes to *energy if given.
__device__ void ss_pair(const SsArgs &a, int f1, int f2, int side, double *acc, double *energy) {
    const int t = a.type[f1] * a.ntypes + a.type[f2];
    const double Emin = a.emin[t], Rmin = a.rmin[t], k0 = a.k0[t], Rmini = 1.0 / Rmin;
    double p[3][3], q[3][3];
    {
        const double4 a0 = corner(a.pos, a.ends + 6 * (size_t)f1), a1 = corner(a.pos, a.ends + 6 * (size_t)f1 + 2),
                      a2 = corner(a.pos, a.ends + 6 * (size_t)f1 + 4);
        const double4 b0 = corner(a.pos, a.ends + 6 * (size_t)f2), b1 = corner(a.pos, a.ends + 6 * (size_t)f2 + 2),
                      b2 = corner(a.pos, a.ends + 6 * (size_t)f2 + 4);
        for (int i = 0; i < 3; i++) {
            const double e0 = kSsEta[i][0], e1 = kSsEta[i][1], e2 = kSsEta[i][2];
            p[i][0] = e0 * a0.x + e1 * a1.x + e2 * a2.x;
            p[i][1] = e0 * a0.y + e1 * a1.y + e2 * a2.y;
            p[i][2] = e0 * a0.z + e1 * a1.z + e2 * a2.z;
            q[i][0] = e0 * b0.x + e1 * b1.x + e2 * b2.x;
            q[i][1] = e0 * b0.y + e1 * b1.y + e2 * b2.y;
            q[i][2] = e0 * b0.z + e1 * b1.z + e2 * b2.z;
        }
    }
    const double Rmin_2 = Rmin * Rmin, Rmin_4 = Rmin_2 * Rmin_2, Rmin_6 = Rmin_4 * Rmin_2;
    double fpm[3][3][3];
    double en = 0.0;
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
        for (int l = k; l < 3; l++) {
            const double dx = p[k][0] - q[l][0], dy = p[k][1] - q[l][1], dz = p[k][2] - q[l][2];
            const double mag_r = sqrt(dx * dx + dy * dy + dz * dz);
            double fm, e;
            if (a.kind == 3) {  // calc_gensoft_factors (src/VdW_solver.cpp:754-771)
                const double r2 = mag_r * mag_r, R2 = Rmin * Rmin;
                const double g2 = r2 / R2, g3 = g2 * mag_r / Rmin;
                const double k0rm = k0 * Rmin, k0rm2 = k0rm * Rmin, eps = Emin / Rmin;
                fm = 2 * (k0rm - 6 * eps) * g3 + 3 * (4 * eps - k0rm) * g2 + k0 * mag_r;
                const double emag = (0.5 * k0rm2 - 3 * Emin) * g2 * g2 + (4 * Emin - k0rm2) * g3 + 0.5 * k0 * r2 - Emin;
                e = kSsW * kSsW * emag;
            } else if (a.kind == 2 && mag_r < Rmin) {  // calc_ljinterpolated_factors (:774-784)
                const double x = mag_r * Rmini, x2 = x * x;
                e = kSsW * kSsW * Emin * x2 * (2 * x - 3);
                fm = 6 * Emin * Rmini * x * (1 - x);
            } else {  // calc_lj_factors (:713-731)
                const double ri = 1. / mag_r;
                const double ri2 = ri * ri, ri4 = ri2 * ri2, ri6 = ri4 * ri2, ri7 = ri6 * ri;
                const double f6 = Rmin_6 * ri6;
                fm = 12 * ri7 * Rmin_6 * Emin * (f6 - 1);
                e = kSsW * kSsW * Emin * f6 * (f6 - 2);
            }
            en += l == k ? e : 2 * e;
#pragma unroll
            for (int x = 0; x < 3; x++) {
                fpm[k][l][x] = fm * ((p[k][x] - q[l][x]) / mag_r);
                fpm[l][k][x] = fpm[k][l][x];
            }
        }
    const double ApAq = a.geom[8 * (size_t)f1 + 6] * a.geom[8 * (size_t)f2 + 6];
    if (energy) *energy = en * ApAq;
    if (!acc) return;
#pragma unroll
    for (int j = 0; j < 3; j++) {
        double F[3] = {0, 0, 0};
#pragma unroll
        for (int k = 0; k < 3; k++)
#pragma unroll
            for (int l = 0; l < 3; l++) {
                if (side == 0) {
                    const double c = kSsW * kSsW * kSsEta[l][j];
                    for (int x = 0; x < 3; x++) F[x] += c * fpm[k][l][x];
                } else {
                    const double d = kSsW * kSsW * kSsEta[k][j];
                    for (int x = 0; x < 3; x++) F[x] -= d * fpm[l][k][x];
                }
            }
        for (int x = 0; x < 3; x++) acc[3 * j + x] += F[x] * ApAq;
    }
}

// The tetrahedron of face f: its three corners and the opposite element vertex.
__device__ __forceinline__ void ss_tet(const SsArgs &a, int f, double (*T)[3]) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double4 c = corner(a.pos, a.ends + 6 * (size_t)f + 2 * k);
        T[k][0] = c.x;
        T[k][1] = c.y;
        T[k][2] = c.z;
    }
    const double4 o = a.pos[a.opp[f]];
    T[3][0] = o.x;
    T[3][1] = o.y;
    T[3][2] = o.z;
}

// The interaction of the pair f1 < f2 that passed consider_interaction, as do_interaction of
// each type (src/VdW_solver.cpp:327-330, src/Steric_solver.cpp:28-38,
// src/LJSteric_solver.cpp:34-52, src/GenSoftSSINT_solver.cpp:29-47): for kinds 2-4, where the
// face elements overlap (Face::checkTetraIntersection, src/Face.cpp:399-404) and their common
// volume is not zero, the steric forces on all four nodes of face 1 (side 0) or face 2 (side 1)
// and energy steric_factor * V; otherwise the pair potential (kinds 1-3; steric adds nothing).
// acc [12]: Face::force[0..3] of that face.
__device__ void ss_interact(const SsArgs &a, int f1, int f2, int side, double *acc, double *energy) {
    if (a.kind != 1) {
        double A[4][3], B[4][3];
        ss_tet(a, f1, A);
        ss_tet(a, f2, B);
        if (st::tets_overlap(A, B)) {
            double e, fA[12], fB[12];
            int degenerate = 0;
            const bool applied = st::pair_forces(A, B, a.st_dr, a.st_factor, &e, fA, fB, degenerate);
            if (degenerate) *a.err = 2;
            if (applied) {
                if (energy) *energy = e;
                if (acc)
                    for (int k = 0; k < 12; k++) acc[k] += side == 0 ? fA[k] : fB[k];
                return;
            }
        }
        if (a.kind == 4) {
            if (energy) *energy = 0.0;
            return;
        }
    }
    ss_pair(a, f1, f2, side, acc, energy);
}

// Whether the face elements of f1 and f2 overlap (kinds 2-4 screen every pair with it).
__device__ __forceinline__ bool ss_overlap(const SsArgs &a, int f1, int f2) {
    double A[4][3], B[4][3];
    ss_tet(a, f1, A);
    ss_tet(a, f2, B);
    return st::tets_overlap(A, B);
}

// The force kernel's three modes: kSsLJ, lennard-jones (no steric test); kSsScreen, kinds 2-4
// with the pair potential only, a face with an overlapping pair left to the next mode
// (deferred); kSsFull, the deferred faces with the steric forces.  The volume's private
// point list and its call stay out of the first two, which keep their registers (4 waves
// per SIMD); overlapping face elements are rare, and only their faces pay for it.
enum { kSsLJ = 0, kSsScreen = 1, kSsFull = 2 };

// One pair's contribution in the given mode; returns true (kSsScreen) where the face elements
// overlap and the face must be deferred.
template <int kMode>
__device__ __forceinline__ bool ss_contrib(const SsArgs &a, int f1, int f2, int side, double *acc) {
    if (kMode == kSsFull) {
        ss_interact(a, f1, f2, side, acc, nullptr);
        return false;
    }
    if (kMode == kSsScreen && ss_overlap(a, f1, f2)) return true;
    if (kMode == kSsLJ || a.kind != 4) ss_pair(a, f1, f2, side, acc, nullptr);
    return false;
}

// Face f's pairs as face 2, (i, f) for i < f in ascending i (a cell reached through several
// offsets counts that often), summed into acc by one thread: a merge of the 27 neighbour
// ranges.  The fallback of k_ss_forces for faces with more candidates than its LDS list holds.
// Returns true (kSsScreen) where an overlapping pair defers the face.
template <int kMode>
__device__ bool ss_face2_serial(const SsArgs &a, int f, double *acc) {
    const int cf = a.cell[f];
    int cur[27], lim[27];
    for (int c = 0; c < 27; c++) {
        const int cc = ss_neighbour(a, cf, c);
        cur[c] = a.cstart[cc];
        lim[c] = a.cend[cc];
    }
    for (;;) {
        int best = a.P, bc = -1;
        for (int c = 0; c < 27; c++)
            if (cur[c] < lim[c]) {
                const int v = a.sorted[cur[c]];
                if (v < best) {
                    best = v;
                    bc = c;
                }
            }
        if (bc < 0 || best >= f) break;
        cur[bc]++;
        if (ss_consider(a, best, f) && ss_contrib<kMode>(a, best, f, 1, acc)) return true;
    }
    return false;
}

constexpr int kSsCap = 1024;  // candidates below f a wave sorts (in registers, 16 per lane)

// The cell of position t < pre[27] when the 27 neighbour cells' runs are laid end to end, run
// c starting at pre[c]: the first c with pre[c + 1] > t (a binary search; an empty run has
// pre[c] == pre[c + 1]).
__device__ __forceinline__ int ss_cell_at(const int *pre, int t) {
    int lo = 0, hi = 26;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid + 1] > t)
            hi = mid;
        else
            lo = mid + 1;
    }
    return lo;
}

// The candidate at visiting position t: the runs are the tops of the stacks (cells ending at
// re[c] in the sorted pool), visited top down.
__device__ __forceinline__ int ss_visit(const SsArgs &a, const int *pre, const int *re, int t) {
    const int c = ss_cell_at(pre, t);
    return a.sorted[re[c] - 1 - (t - pre[c])];
}

// Sorts cand[0, na) ascending, 64 * R >= na: a bitonic network over R * 64 entries held R per
// lane (entry r * 64 + lane in v[r], the tail padded with INT_MAX); exchanges across lanes are
// shuffles, those across a lane's own entries plain swaps.
template <int R>
__device__ __forceinline__ void ss_sort_cand(int *cand, int na, int lane) {
    int v[R];
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = r * 64 + lane < na ? cand[r * 64 + lane] : INT_MAX;
#pragma unroll
    for (int k = 2; k <= 64 * R; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= 64) {
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int u = r ^ (j >> 6);
                    if (u > r) {
                        const bool up = ((r * 64 + lane) & k) == 0;
                        const int x = v[r], y = v[u];
                        v[r] = up ? min(x, y) : max(x, y);
                        v[u] = up ? max(x, y) : min(x, y);
                    }
                }
            } else {
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int y = __shfl_xor(v[r], j);
                    const bool up = ((r * 64 + lane) & k) == 0, low = (lane & j) == 0;
                    v[r] = low == up ? min(v[r], y) : max(v[r], y);
                }
            }
        }
#pragma unroll
    for (int r = 0; r < R; r++) cand[r * 64 + lane] = v[r];
}

// acc += column k of the first e staged contributions, in order.  A pair that does not
// interact stages +0.0, and adding +0.0 leaves every sum unchanged: a face force starts at +0.0
// and never becomes -0.0 (in the reference either), so the loop needs no test and its loads
// pipeline.
template <int kN>
__device__ __forceinline__ void ss_ordered_add(const double (*con)[kN], int k, int e, double &acc) {
    int q = 0;
    for (; q + 8 <= e; q += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = con[q + u][k];
#pragma unroll
        for (int u = 0; u < 8; u++) acc += v[u];
    }
    for (; q < e; q++) acc += con[q][k];
}

// Every step: face f's Face::force as VdW_solver::solve leaves it, one wavefront per face.
// The solve visits pool entries i in order, each against the stacks of its 27 neighbour cells
// (top down: descending pool index) with i < j; face f therefore first receives, as face 2,
// the pairs (i, f) for i < f in ascending i, then, as face 1, its own pairs (f, j) in its
// visiting order.  The pairs that pass consider_interaction join a queue in that order (a
// ballot per 64 candidates); the lanes evaluate 64 queued pairs at a time, and the
// contributions then go through LDS and are added one by one in queue order by lanes 0..11 (one
// force component each), so every sum has the reference's order.  Pairs that do not pass add
// nothing in the reference and are left out here.  kMode: kSsLJ, kSsScreen or kSsFull (above).
template <int kMode>
__global__ __launch_bounds__(64) void k_ss_forces(SsArgs a) {
    __shared__ int rs[27], re[27], pa[28], pb[28];
    __shared__ int cand[kSsCap];
    __shared__ int queue[128];  // pending pairs: the other face j for (f, j), ~i for (i, f)
    constexpr int kN = kMode == kSsLJ ? 9 : 12;  // force components: the opposite vertex's only take steric forces
    __shared__ double con[64][kN];
    __shared__ int defer;
    const int f = blockIdx.x, lane = threadIdx.x;
    if (f >= a.P) return;
    if (kMode == kSsFull && !a.deferred[f]) return;  // the whole block: one face
    if (lane < 27) {
        // the cell's stack sorted[s0, s1) ascends by pool index: the entries below f, then f
        // itself (in its own cell), then the entries above f
        const int cc = ss_neighbour(a, a.cell[f], lane);
        const int s0 = a.cstart[cc], s1 = a.cend[cc];
        int lo = s0, hi = s1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (a.sorted[mid] < f)
                lo = mid + 1;
            else
                hi = mid;
        }
        rs[lane] = s0;
        re[lane] = s1;
        pa[lane + 1] = lo - s0;  // counts here, starts after the prefix sums
        pb[lane + 1] = s1 - (lo < s1 && a.sorted[lo] == f ? lo + 1 : lo);
    }
    __syncthreads();
    if (lane == 0) {
        pa[0] = pb[0] = 0;
        for (int c = 0; c < 27; c++) {
            pa[c + 1] += pa[c];
            pb[c + 1] += pb[c];
        }
    }
    __syncthreads();
    const int na = pa[27], nb = pb[27];
    double acc = 0.0;  // lane k < kN: component k of the face force (node k / 3, axis k % 3)
    int nq = 0;        // queued pairs (the same in every lane)
    // (a) as face 2: the candidates below f (the bottoms of the stacks), sorted ascending
    if (na <= kSsCap) {
        for (int t = lane; t < na; t += 64) {
            const int c = ss_cell_at(pa, t);
            cand[t] = a.sorted[rs[c] + (t - pa[c])];
        }
        __syncthreads();
    }
    bool deferred = false;  // the same in every lane
    if (na > kSsCap) {  // too many to sort here: one lane merges
        double acc12[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (lane == 0) {
            defer = ss_face2_serial<kMode>(a, f, acc12);
            for (int k = 0; k < kN; k++) con[0][k] = acc12[k];
        }
        __syncthreads();
        if (lane < kN) acc = con[0][lane];
        deferred = defer;
        __syncthreads();
    } else if (na > 1) {
        int m = 64;
        while (m < na) m <<= 1;
        switch (m) {
            case 64: ss_sort_cand<1>(cand, na, lane); break;
            case 128: ss_sort_cand<2>(cand, na, lane); break;
            case 256: ss_sort_cand<4>(cand, na, lane); break;
            case 512: ss_sort_cand<8>(cand, na, lane); break;
            default: ss_sort_cand<16>(cand, na, lane); break;
        }
        __syncthreads();
    }
    // the queue: part (a) chunks of 64 sorted candidates, then part (b) chunks of 64 visiting
    // positions, then one more pass that evaluates what is left
    const int chunks_a = na > kSsCap ? 0 : (na + 63) / 64, chunks = chunks_a + (nb + 63) / 64;
    for (int ch = 0; ch <= chunks && !deferred; ch++) {
        int i1 = -1, i2 = -1;  // the pair (i1 < i2) this lane queues, if it passes
        if (ch < chunks_a) {
            const int t = ch * 64 + lane;
            if (t < na) {
                i1 = cand[t];
                i2 = f;
            }
        } else if (ch < chunks) {
            const int t = (ch - chunks_a) * 64 + lane;
            if (t < nb) {  // the tops of the stacks, in visiting order
                i1 = f;
                i2 = ss_visit(a, pb, re, t);
            }
        }
        const bool pass = i1 >= 0 && ss_consider(a, i1, i2);
        const unsigned long long m = __ballot(pass);
        if (pass) queue[nq + __popcll(m & ((1ull << lane) - 1))] = i2 == f ? ~i1 : i2;
        nq += __popcll(m);
        __syncthreads();
        const int e = ch == chunks ? nq : (nq >= 64 ? 64 : 0);
        if (e > 0) {  // evaluate the first e queued pairs, one per lane; add them in queue order
            double c12[kN] = {};
            bool ov = false;
            if (lane < e) {
                const int o = queue[lane];
                const bool second = o < 0;
                ov = ss_contrib<kMode>(a, second ? ~o : f, second ? f : o, second ? 1 : 0, c12);
            }
            if (kMode == kSsScreen && __ballot(ov)) {  // wave-uniform: leave the face to kSsFull
                deferred = true;
                break;
            }
            for (int k = 0; k < kN; k++) con[lane][k] = c12[k];
            __syncthreads();
            if (lane < kN) ss_ordered_add<kN>(con, lane, e, acc);
            nq -= e;
            const int carry = lane < nq ? queue[64 + lane] : 0;
            __syncthreads();
            if (lane < nq) queue[lane] = carry;
            __syncthreads();
        }
    }
    if (kMode == kSsScreen && lane == 0) a.deferred[f] = deferred ? 1 : 0;
    if (!deferred && lane < kN) a.out[12 * (size_t)f + lane] = acc;
}

// The energy of the last solve (VdW_solver::get_field_energy(-1, -1),
// src/VdW_solver.cpp:635-661): fieldenergy[b1][b2] is one running sum per blob pair over the
// pairs in the solve's visiting order (i ascending, then its 27 cells, stacks top down), then
// the pairs' sums are added row by row.  k_ss_pair_count counts each face's own pairs (one
// wavefront per face), k_ss_pair_energy writes their energies in visiting order at the face's
// offset, and k_ss_energy_sum adds them up in that order on one thread.
__device__ __forceinline__ void ss_ranges(const SsArgs &a, int f, int lane, int *rs, int *re, int *pre) {
    if (lane < 27) {
        const int cc = ss_neighbour(a, a.cell[f], lane);
        rs[lane] = a.cstart[cc];
        re[lane] = a.cend[cc];
    }
    __syncthreads();
    if (lane == 0) {
        int t = 0;
        for (int c = 0; c < 27; c++) {
            pre[c] = t;
            t += re[c] - rs[c];
        }
        pre[27] = t;
    }
    __syncthreads();
}

__global__ __launch_bounds__(64) void k_ss_pair_count(SsArgs a, int *cnt) {
    __shared__ int rs[27], re[27], pre[28];
    const int f = blockIdx.x, lane = threadIdx.x;
    if (f >= a.P) return;
    ss_ranges(a, f, lane, rs, re, pre);
    const int n = pre[27];
    int k = 0;
    for (int t = lane; t < n; t += 64) {
        const int j = ss_visit(a, pre, re, t);
        k += j > f && ss_consider(a, f, j);
    }
    for (int o = 32; o > 0; o >>= 1) k += __shfl_xor(k, o);
    if (lane == 0) cnt[f] = k;
}

template <bool kSteric>  // kinds 2-4: the steric energy where face elements overlap
__global__ __launch_bounds__(64) void k_ss_pair_energy(SsArgs a, int nblobs, const int *off, double *pe, int *pb) {
    __shared__ int rs[27], re[27], pre[28];
    const int f = blockIdx.x, lane = threadIdx.x;
    if (f >= a.P) return;
    ss_ranges(a, f, lane, rs, re, pre);
    const int n = pre[27];
    int w = off[f];
    for (int base = 0; base < n; base += 64) {
        const int t = base + lane;
        int j = -1;
        bool ok = false;
        if (t < n) {
            j = ss_visit(a, pre, re, t);
            ok = j > f && ss_consider(a, f, j);
        }
        const unsigned long long m = __ballot(ok);
        if (ok) {
            const int r = __popcll(m & ((1ull << lane) - 1ull));
            double e;
            if (kSteric)
                ss_interact(a, f, j, 0, nullptr, &e);
            else
                ss_pair(a, f, j, 0, nullptr, &e);
            pe[w + r] = e;
            pb[w + r] = a.blob[f] * nblobs + a.blob[j];
        }
        w += __popcll(m);
    }
}

// One wavefront: the lanes stage 512 pairs at a time in LDS and all add them in order.  A
// blob pair's running sum stays in a register while consecutive pairs share it and is written
// back when the blob pair changes -- the same additions as fieldenergy[b1][b2] += e.
constexpr int kSsFeLds = 4096;  // blob pairs kept in LDS (64 blobs); more stay in global memory
constexpr int kSsStage = 512;   // pairs staged in LDS per round
__global__ __launch_bounds__(64) void k_ss_energy_sum(int npairs, const double *pe, const int *pb, int nblobs,
                                                       double *fe, double *total) {
    __shared__ double fl[kSsFeLds];
    __shared__ double ce[kSsStage];
    __shared__ int cb[kSsStage];
    const int lane = threadIdx.x, nf = nblobs * nblobs;
    const bool in_lds = nf <= kSsFeLds;
    double *f = in_lds ? fl : fe;
    for (int k = lane; k < nf; k += 64) f[k] = 0.0;
    __syncthreads();
    int key = -1;
    double run = 0.0;
    // each round's pairs are loaded into registers during the previous round's sum
    constexpr int kPer = kSsStage / 64;
    double pv[kPer];
    int pk[kPer];
    auto fetch = [&](int base) {
#pragma unroll
        for (int u = 0; u < kPer; u++) {
            const int t = base + u * 64 + lane;
            pv[u] = t < npairs ? pe[t] : 0.0;
            pk[u] = t < npairs ? pb[t] : 0;
        }
    };
    fetch(0);
    for (int base = 0; base < npairs; base += kSsStage) {
#pragma unroll
        for (int u = 0; u < kPer; u++) {
            ce[u * 64 + lane] = pv[u];
            cb[u * 64 + lane] = pk[u];
        }
        __syncthreads();
        if (base + kSsStage < npairs) fetch(base + kSsStage);
        // every lane adds the same staged pairs in the same order; the blob-pair key is read as
        // a wave-uniform value, so the switches are scalar branches
        const int e = min(kSsStage, npairs - base);
        int q = 0;
        for (; q + 16 <= e; q += 16) {  // sixteen staged pairs at a time into registers
            double v[16];
            int kb[16];
#pragma unroll
            for (int u = 0; u < 16; u++) {
                v[u] = ce[q + u];
                kb[u] = __builtin_amdgcn_readfirstlane(cb[q + u]);
            }
#pragma unroll
            for (int u = 0; u < 16; u++) {
                if (kb[u] != key) {
                    if (key >= 0) f[key] = run;
                    key = kb[u];
                    run = f[key];
                }
                run += v[u];
            }
        }
        for (; q < e; q++) {
            const int kq = __builtin_amdgcn_readfirstlane(cb[q]);
            if (kq != key) {
                if (key >= 0) f[key] = run;
                key = kq;
                run = f[key];
            }
            run += ce[q];
        }
        __syncthreads();
    }
    if (key >= 0) f[key] = run;  // the same value from every lane
    __syncthreads();
    if (lane == 0) {
        double t = 0.0;
        for (int k = 0; k < nf; k++) t += f[k];
        *total = t;
    }
    __syncthreads();
    if (in_lds)
        for (int k = lane; k < nf; k += 64) fe[k] = fl[k];
}
