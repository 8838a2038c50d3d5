// ffh_plan.cpp -- host setup of a blob for the GPU step (compiled with -ffp-contract=off:
// the rest-state arithmetic rounds exactly as the reference's strict build).
//
// Follows: Blob::init order (src/Blob.cpp:335-412), calc_rest_state_info (:3760-3822) with
// mat3_invert (src/mat_vec_fns.cpp:148-177), load_material_params conversions
// (src/Blob.cpp:2861-2866), stokes drag (:408-412), calculate_node_element_connectivity
// (:3974-4019) and the pattern registration of NoMassCGSolver::init
// (src/NoMassCGSolver.cpp:93-139) restricted to linear nodes.
#include "ffh_plan.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>

namespace ffh {

namespace {

// mesoDimensions (include/dimensions.h:30-41)
constexpr double kMesoVolume = 4.913e-30;
constexpr double kMesoMass = 1.994307387553024e-26;
constexpr double kMesoPressure = 8.4305832688784826e8;
constexpr double kMesoTime = 3.7302915560886126e-13;
constexpr double kPi = 3.141592653589793238462643383279502884;

// MRG32k3a one-step transition matrices (reduced mod m: -a13n -> m1 - a13n).
const uint64_t kA[2][3][3] = {
    {{0, 1, 0}, {0, 0, 1}, {kM1 - 810728ull, 1403580ull, 0}},
    {{0, 1, 0}, {0, 0, 1}, {kM2 - 1370589ull, 0, 527612ull}},
};

struct M3 {
    uint64_t a[3][3];
};

M3 mul(const M3 &x, const M3 &y, uint64_t m) {
    M3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            unsigned __int128 s = 0;
            for (int k = 0; k < 3; k++) s += (unsigned __int128)x.a[i][k] * y.a[k][j];
            r.a[i][j] = (uint64_t)(s % m);
        }
    return r;
}

M3 ident() {
    M3 r{};
    for (int i = 0; i < 3; i++) r.a[i][i] = 1;
    return r;
}

M3 base(int comp) {
    M3 r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r.a[i][j] = kA[comp][i][j];
    return r;
}

Mat3u to32(const M3 &x) {
    Mat3u r;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) r.m[i][j] = (uint32_t)x.a[i][j];
    return r;
}

M3 powm(int comp, uint64_t n) {
    uint64_t m = comp ? kM2 : kM1;
    M3 r = ident(), w = base(comp);
    while (n) {
        if (n & 1) r = mul(w, r, m);
        w = mul(w, w, m);
        n >>= 1;
    }
    return r;
}

// Jacobian + cofactors + determinant in the reference's operation order
// (src/tetra_element_linear.cpp:193-205, :223-258).
void jacobian(const double *x, const double *y, const double *z, const int *n, double J[3][3]) {
    const double *p[3] = {x, y, z};
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) J[r][c] = p[c][n[r + 1]] - p[c][n[0]];
}

// Runs f(i, scratch) for i in [0, n) on up to FFH_HOST_THREADS (else OMP_NUM_THREADS, else 16)
// threads; each thread owns one scratch vector.  Iterations must write disjoint data.
template <typename F>
void parallel_for(int64_t n, F &&f) {
    int nt = 16;
    for (const char *v : {"FFH_HOST_THREADS", "OMP_NUM_THREADS"}) {
        const char *x = std::getenv(v);
        if (x && std::atoi(x) > 0) {
            nt = std::atoi(x);
            break;
        }
    }
    nt = std::max(1, std::min<int>({nt, (int)std::max(1u, std::thread::hardware_concurrency()), 64}));
    if (n < 2 * (int64_t)nt || nt == 1) {
        std::vector<int32_t> scratch;
        for (int64_t i = 0; i < n; i++) f(i, scratch);
        return;
    }
    std::atomic<int64_t> next(0);
    const int64_t chunk = std::max<int64_t>(1, n / ((int64_t)nt * 16));
    auto work = [&] {
        std::vector<int32_t> scratch;
        for (;;) {
            const int64_t i0 = next.fetch_add(chunk);
            if (i0 >= n) break;
            const int64_t i1 = std::min(n, i0 + chunk);
            for (int64_t i = i0; i < i1; i++) f(i, scratch);
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nt; t++) th.emplace_back(work);
    work();
    for (auto &t : th) t.join();
}

}  // namespace

Mat3u mrg_pow(int comp, uint64_t n) { return to32(powm(comp, n)); }

std::string make_tiles(const std::vector<uint32_t> &off, const std::vector<uint32_t> &items, Tiles &t) {
    const size_t n = off.size() - 1, groups = (n + kSlice - 1) / kSlice;
    t.off.assign(groups + 1, 0);
    t.width.assign(groups, 0);
    t.cnt.assign(n, 0);
    for (size_t u = 0; u < n; u++) {
        uint32_t c = off[u + 1] - off[u];
        if (c > 65535) return "a contribution list exceeds 65535 entries";
        t.cnt[u] = (uint16_t)c;
        t.width[u / kSlice] = std::max<uint16_t>(t.width[u / kSlice], (uint16_t)c);
    }
    for (size_t g = 0; g < groups; g++) t.off[g + 1] = t.off[g] + (int64_t)t.width[g] * kSlice;
    t.code.assign((size_t)t.off[groups], 0u);
    for (size_t u = 0; u < n; u++) {
        const size_t g = u / kSlice, l = u % kSlice;
        for (uint32_t k = 0; k < t.cnt[u]; k++) t.code[(size_t)t.off[g] + (size_t)k * kSlice + l] = items[off[u] + k];
    }
    return "";
}

std::string build_plan(Plan &P, int num_nodes, int num_elements, const double *pos, const int *topo,
                       const double *material_si, const double *stokes_radius, int num_pinned,
                       const int *pinned, double stokes_visc, int calc_noise, int calc_stokes,
                       int num_blobs, const int *blob_node_off, bool material_meso) {
    if (num_nodes <= 0 || num_elements <= 0) return "mesh must have nodes and elements";
    if ((uint64_t)num_elements >= (1u << 27)) return "more than 2^27 elements per blob is not supported";
    const int NT = num_nodes, E = num_elements;
    P.num_nodes_total = NT;
    P.num_elements = E;

    // --- linear nodes (load_topology marks n[0..3] linear, src/Blob.cpp:2618-2621) ---
    std::vector<uint8_t> is_lin(NT, 0);
    for (int e = 0; e < E; e++)
        for (int j = 0; j < 10; j++) {
            int v = topo[10 * e + j];
            if (v < 0 || v >= NT) return "element " + std::to_string(e) + " references an out of bounds node index";
            if (j < 4) is_lin[v] = 1;
        }
    for (int e = 0; e < E; e++)
        for (int j = 4; j < 10; j++)
            if (is_lin[topo[10 * e + j]])
                return "node " + std::to_string(topo[10 * e + j]) + " is both a vertex and an edge midpoint";
    // blobs and their node ranges
    std::vector<int> boff;
    if (blob_node_off) {
        if (num_blobs < 1) return "need at least one blob";
        boff.assign(blob_node_off, blob_node_off + num_blobs + 1);
        if (boff[0] != 0 || boff[num_blobs] != NT) return "blob node offsets must span all nodes";
        for (int b = 0; b < num_blobs; b++)
            if (boff[b + 1] < boff[b]) return "blob node offsets must be non-decreasing";
    } else {
        num_blobs = 1;
        boff = {0, NT};
    }
    P.num_blobs = num_blobs;
    {
        std::vector<int> node_blob(NT);
        for (int b = 0; b < num_blobs; b++)
            for (int i = boff[b]; i < boff[b + 1]; i++) node_blob[i] = b;
        for (int e = 0; e < E; e++)
            for (int j = 1; j < 10; j++)
                if (node_blob[topo[10 * e + j]] != node_blob[topo[10 * e]])
                    return "element " + std::to_string(e) + " joins nodes of two blobs";
    }
    // compact linear numbering, ascending original index; each blob padded to whole slices
    P.lin_of.assign(NT, -1);
    P.orig_of.clear();
    P.blob_slice_off.assign(num_blobs + 1, 0);
    for (int b = 0; b < num_blobs; b++) {
        P.blob_slice_off[b] = (int)(P.orig_of.size() / kSlice);
        for (int i = boff[b]; i < boff[b + 1]; i++)
            if (is_lin[i]) {
                P.lin_of[i] = (int)P.orig_of.size();
                P.orig_of.push_back(i);
            }
        while (P.orig_of.size() % kSlice) P.orig_of.push_back(-1);
    }
    P.blob_slice_off[num_blobs] = (int)(P.orig_of.size() / kSlice);
    const int N = (int)P.orig_of.size();
    P.num_linear = N;
    P.num_real_linear = 0;
    for (int i = 0; i < N; i++) P.num_real_linear += P.orig_of[i] >= 0;
    if (P.num_real_linear == 0) return "mesh has no linear nodes";
    if (calc_stokes == 1)
        for (int i = 0; i < NT; i++)
            if (!is_lin[i] && stokes_radius[i] != 0.0)
                return "midpoint node " + std::to_string(i) +
                       " has a nonzero Stokes radius; the GPU path requires zero drag on midpoints";

    // --- midpoints for N_t output ---
    {
        std::vector<uint8_t> seen(NT, 0);
        static const int ends[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
        P.mid_index.clear();
        P.mid_a.clear();
        P.mid_b.clear();
        P.emid.resize(6 * (size_t)E);
        for (int e = 0; e < E; e++)
            for (int m = 0; m < 6; m++) {
                int v = topo[10 * e + 4 + m];
                P.emid[6 * (size_t)e + m] = v;
                if (seen[v]) continue;
                seen[v] = 1;
                P.mid_index.push_back(v);
                P.mid_a.push_back(topo[10 * e + ends[m][0]]);
                P.mid_b.push_back(topo[10 * e + ends[m][1]]);
            }
    }

    // --- positions (linear nodes, SoA) ---
    P.pos.assign(3 * (size_t)N, 0.0);
    for (int i = 0; i < N; i++)
        if (P.orig_of[i] >= 0)
            for (int c = 0; c < 3; c++) P.pos[(size_t)c * N + i] = pos[3 * (size_t)P.orig_of[i] + c];
    const double *X = P.pos.data(), *Y = X + N, *Z = Y + N;

    // --- elements: compact connectivity, materials, rest state ---
    P.enodes.resize(4 * (size_t)E);
    P.jinv0.resize(9 * (size_t)E);
    P.vol0.resize(E);
    P.last_det.resize(E);
    P.matA.resize(E);
    P.matB.resize(E);
    P.matG.resize(E);
    P.matE.resize(E);
    P.matRho.resize(E);
    for (int e = 0; e < E; e++) {
        int n[4];
        for (int a = 0; a < 4; a++) n[a] = P.enodes[4 * (size_t)e + a] = P.lin_of[topo[10 * e + a]];
        if (material_meso) {  // rho, A, B, G, E as tetra_element_linear holds them
            const double *m = &material_si[5 * (size_t)e];
            P.matRho[e] = m[0];
            P.matA[e] = m[1];
            P.matB[e] = m[2];
            P.matG[e] = m[3];
            P.matE[e] = m[4];
        } else {
            const double *m = &material_si[6 * (size_t)e];
            P.matA[e] = m[1] / (kMesoPressure * kMesoTime);
            P.matB[e] = m[2] / (kMesoPressure * kMesoTime) - (2.0 / 3.0) * m[1];
            P.matG[e] = m[3] / kMesoPressure;
            P.matE[e] = m[4] / kMesoPressure;
            P.matRho[e] = m[0] * kMesoVolume / kMesoMass;
        }
        double J[3][3];
        jacobian(X, Y, Z, n, J);
        // mat3_invert (src/mat_vec_fns.cpp:148-177)
        double mi[3][3];
        mi[0][0] = J[2][2] * J[1][1] - J[2][1] * J[1][2];
        mi[0][1] = J[2][1] * J[0][2] - J[2][2] * J[0][1];
        mi[0][2] = J[1][2] * J[0][1] - J[1][1] * J[0][2];
        mi[1][0] = J[2][0] * J[1][2] - J[2][2] * J[1][0];
        mi[1][1] = J[2][2] * J[0][0] - J[2][0] * J[0][2];
        mi[1][2] = J[1][0] * J[0][2] - J[1][2] * J[0][0];
        mi[2][0] = J[2][1] * J[1][0] - J[2][0] * J[1][1];
        mi[2][1] = J[2][0] * J[0][1] - J[2][1] * J[0][0];
        mi[2][2] = J[1][1] * J[0][0] - J[1][0] * J[0][1];
        double det = J[0][0] * mi[0][0] + J[1][0] * mi[0][1] + J[2][0] * mi[0][2];
        if (det == 0.0) return "element " + std::to_string(e) + " has zero rest volume";
        double inv = 1.0 / det;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) P.jinv0[(size_t)(3 * i + j) * E + e] = mi[i][j] * inv;
        // calc_shape_function_derivatives_and_volume: the same cofactors give det and vol.
        P.vol0[e] = (1.0 / 6.0) * std::fabs(det);
        P.last_det[e] = det;
    }

    // --- nodes: drag, pins, RNG offsets ---
    P.zeta.assign(N, 0.0);
    P.pinned.assign(N, 0);
    for (int i = 0; i < N; i++)
        if (P.orig_of[i] < 0) P.pinned[i] = 1;  // padding rows: identity, zero force
    for (int i = 0; i < num_pinned; i++) {
        int v = pinned[i];
        if (v < 0 || v >= NT) return "pinned node index out of bounds";
        if (P.lin_of[v] >= 0) P.pinned[P.lin_of[v]] = 1;
        // pinned midpoints carry no force in the reference either (linearise_force zeroes them)
    }
    if (calc_stokes == 1)
        for (int i = 0; i < N; i++)
            if (!P.pinned[i]) P.zeta[i] = 6.0 * kPi * stokes_visc * stokes_radius[P.orig_of[i]];
    P.rng_node_off.resize(N);
    for (int i = 0; i < N; i++) P.rng_node_off[i] = P.orig_of[i] >= 0 ? 3u * (uint32_t)P.orig_of[i] : 0u;

    // --- node force gather (element order), src/Blob.cpp:3974-4019 ---
    P.nctb_off.assign((size_t)N + 1, 0);
    for (int e = 0; e < E; e++)
        for (int a = 0; a < 4; a++) P.nctb_off[P.enodes[4 * (size_t)e + a] + 1]++;
    for (int i = 0; i < N; i++) P.nctb_off[i + 1] += P.nctb_off[i];
    P.nctb.resize(P.nctb_off[N]);
    {
        std::vector<uint32_t> cur(P.nctb_off.begin(), P.nctb_off.end() - 1);
        for (int e = 0; e < E; e++)
            for (int a = 0; a < 4; a++) {
                int n = P.enodes[4 * (size_t)e + a];
                P.nctb[cur[n]++] = ((uint32_t)e << 2) | (uint32_t)a;
            }
    }

    // --- block pattern (linear, unpinned pairs + every diagonal), rows sorted by column ---
    // Row r's columns: r itself and, for an unpinned row, the unpinned vertices of the elements
    // that contain r (the pairs NoMassCGSolver::init registers, src/NoMassCGSolver.cpp:93-139).
    std::vector<int64_t> row_off((size_t)N + 1, 0);
    auto row_columns = [&](int r, std::vector<int32_t> &cols) {
        cols.clear();
        cols.push_back(r);
        if (!P.pinned[r])
            for (uint32_t m = P.nctb_off[r]; m < P.nctb_off[r + 1]; m++) {
                const int e = (int)(P.nctb[m] >> 2);
                for (int b = 0; b < 4; b++) {
                    const int c = P.enodes[4 * (size_t)e + b];
                    if (!P.pinned[c]) cols.push_back(c);
                }
            }
        std::sort(cols.begin(), cols.end());
        cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
    };
    parallel_for(N, [&](int64_t r, std::vector<int32_t> &cols) {
        row_columns((int)r, cols);
        row_off[r + 1] = (int64_t)cols.size();
    });
    for (int i = 0; i < N; i++) row_off[i + 1] += row_off[i];
    const int64_t NB = row_off[N];
    P.num_blocks = NB;
    std::vector<int32_t> bcol(NB);
    parallel_for(N, [&](int64_t r, std::vector<int32_t> &cols) {
        row_columns((int)r, cols);
        std::copy(cols.begin(), cols.end(), bcol.begin() + row_off[r]);
    });

    // --- SELL-64 layout ---
    P.num_slices = (N + kSlice - 1) / kSlice;
    P.slice_off.assign((size_t)P.num_slices + 1, 0);
    P.slice_width.assign(P.num_slices, 0);
    for (int s = 0; s < P.num_slices; s++) {
        int w = 0;
        for (int l = 0; l < kSlice; l++) {
            int r = s * kSlice + l;
            if (r < N) w = std::max<int>(w, (int)(row_off[r + 1] - row_off[r]));
        }
        P.slice_width[s] = w;
        P.slice_off[s + 1] = P.slice_off[s] + (int64_t)w * kSlice;
    }
    P.num_slots = P.slice_off[P.num_slices];
    for (int s = 0; s < P.num_slices; s++)
        if (P.slice_width[s] > 254) return "a node has more than 253 neighbours";
    P.slice_blob.assign(P.num_slices, 0);
    for (int b = 0; b < num_blobs; b++)
        for (int sl = P.blob_slice_off[b]; sl < P.blob_slice_off[b + 1]; sl++) P.slice_blob[sl] = b;
    if (P.num_slots >= (int64_t)1 << 31) return "matrix pattern too large";
    P.slot_col.assign(P.num_slots, 0);
    P.slot_flags.assign(P.num_slots, kSlotPad);
    parallel_for(P.num_slices, [&](int64_t s, std::vector<int32_t> &) {
        for (int t = 0; t < P.slice_width[s]; t++)
            for (int l = 0; l < kSlice; l++) {
                const int64_t slot = P.slice_off[s] + (int64_t)t * kSlice + l;
                const int r = (int)s * kSlice + l;
                if (r < N && t < row_off[r + 1] - row_off[r]) {
                    const int c = bcol[row_off[r] + t];
                    P.slot_col[slot] = c;
                    P.slot_flags[slot] = c == r ? (uint8_t)(kSlotDiag | (P.pinned[r] ? kSlotIdentity : 0)) : 0;
                } else {
                    P.slot_col[slot] = r < N ? r : N - 1;  // zero block on a valid column
                }
            }
    });

    // --- slice assembly lists (SparseMatrixFixedPattern::build's source lists,
    // src/SparseMatrixFixedPattern.cpp:85-93): per slice the distinct elements its rows touch,
    // ascending, and its work tiles.  A lower block (column < row) is the transpose of the upper
    // block (row, column swapped) bit for bit -- the same element blocks, the same products in
    // the same order -- so only the diagonal and the upper blocks are summed; each upper block is
    // also written, transposed, to its lower mirror.  Work tile 0 of a slice holds the 64 rows'
    // diagonal blocks; then one work tile per SELL tile that holds an upper block.  Per lane its
    // contributions (element, local row vertex a, local column vertex b) in ascending element
    // order, as codes local-element << 4 | a << 2 | b, tiled [k][64]; aux: the diagonal's SELL
    // position (tile 0) or the mirror's global slot (upper tiles), -1 where the lane has none ---
    struct SliceLists {
        std::vector<int32_t> elem;
        std::vector<uint16_t> width;  // per work tile: max contributions of its lanes
        std::vector<uint16_t> code;   // the work tiles, each [width][64]
        std::vector<uint8_t> cnt;     // [work tiles * 64]
        std::vector<int32_t> aux;     // [work tiles * 64]
        std::vector<uint8_t> pos;     // [work tiles * 64] SELL position of the lane's block, 255: none
        std::string err;
    };
    std::vector<SliceLists> sl(P.num_slices);
    auto slot_of = [&](int r, int c) -> int64_t {  // global slot of block (r, c), which exists
        const int32_t *cb = bcol.data() + row_off[r], *ce = bcol.data() + row_off[r + 1];
        const int t = (int)(std::lower_bound(cb, ce, c) - cb);
        return P.slice_off[r / kSlice] + (int64_t)t * kSlice + r % kSlice;
    };
    parallel_for(P.num_slices, [&](int64_t s, std::vector<int32_t> &pos_k) {
        SliceLists &o = sl[s];
        const int w = P.slice_width[s];
        // each row's diagonal and upper contributions, element by element (ascending, the node
        // lists' order), then column vertex b: for a fixed block they arrive in ascending
        // element order.  fn(SELL position t, lane, element, a << 2 | b)
        auto each = [&](auto &&fn) {
            for (int l = 0; l < kSlice; l++) {
                const int r = (int)s * kSlice + l;
                if (r >= N || P.pinned[r]) continue;
                const int32_t *cb = bcol.data() + row_off[r], *ce = bcol.data() + row_off[r + 1];
                for (uint32_t m = P.nctb_off[r]; m < P.nctb_off[r + 1]; m++) {
                    const uint32_t e = P.nctb[m] >> 2, a = P.nctb[m] & 3u;
                    for (uint32_t b = 0; b < 4; b++) {
                        const int c = P.enodes[4 * (size_t)e + b];
                        if (P.pinned[c] || c < r) continue;
                        fn((int)(std::lower_bound(cb, ce, c) - cb), l, e, a << 2 | b, c == r);
                    }
                }
            }
        };
        // work tiles: the diagonal tile, then the SELL tiles holding an upper block, merged where
        // their lanes do not meet (densest first, each into the first work tile it fits): a work
        // lane then holds one upper block of its row, at SELL position pos
        std::vector<int> wt_of(w, -1);
        std::vector<uint64_t> wmask(1, 0);
        {
            std::vector<uint64_t> lanes(w, 0);
            each([&](int t, int l, uint32_t, uint32_t, bool diag) {
                if (!diag) lanes[t] |= 1ull << l;
            });
            std::vector<int> order;
            for (int t = 0; t < w; t++)
                if (lanes[t]) order.push_back(t);
            std::stable_sort(order.begin(), order.end(),
                             [&](int x, int y) { return __builtin_popcountll(lanes[x]) > __builtin_popcountll(lanes[y]); });
            for (int t : order) {
                size_t j = 1;
                while (j < wmask.size() && (wmask[j] & lanes[t])) j++;
                if (j == wmask.size()) wmask.push_back(0);
                wmask[j] |= lanes[t];
                wt_of[t] = (int)j;
            }
        }
        const int nw = (int)wmask.size();
        o.cnt.assign((size_t)nw * kSlice, 0);
        o.aux.assign((size_t)nw * kSlice, -1);
        o.pos.assign((size_t)nw * kSlice, 255);
        bool over = false;
        each([&](int t, int l, uint32_t e, uint32_t, bool diag) {
            uint8_t &c = o.cnt[(size_t)(diag ? 0 : wt_of[t]) * kSlice + l];
            if (c == 255) over = true;
            else c++;
            o.elem.push_back((int32_t)e);
        });
        if (over) {
            o.err = "a matrix block has more than 254 element contributions";
            return;
        }
        std::sort(o.elem.begin(), o.elem.end());
        o.elem.erase(std::unique(o.elem.begin(), o.elem.end()), o.elem.end());
        o.elem.shrink_to_fit();
        if (o.elem.size() > 4096) {
            o.err = "a 64-row slice touches more than 4096 elements";
            return;
        }
        // aux: the diagonal's SELL position of every row that has one; an upper block's mirror
        for (int l = 0; l < kSlice; l++) {
            const int r = (int)s * kSlice + l;
            if (r >= N) continue;
            const int32_t *cb = bcol.data() + row_off[r], *ce = bcol.data() + row_off[r + 1];
            o.aux[l] = (int32_t)(std::lower_bound(cb, ce, r) - cb);
            o.pos[l] = (uint8_t)o.aux[l];
            if (P.pinned[r]) continue;
            for (const int32_t *c = cb; c < ce; c++)
                if (*c > r) {
                    o.aux[(size_t)wt_of[c - cb] * kSlice + l] = (int32_t)slot_of(*c, r);
                    o.pos[(size_t)wt_of[c - cb] * kSlice + l] = (uint8_t)(c - cb);
                }
        }
        o.width.assign(nw, 0);
        std::vector<size_t> base(nw + 1, 0);
        for (int j = 0; j < nw; j++) {
            for (int l = 0; l < kSlice; l++) o.width[j] = std::max<uint16_t>(o.width[j], o.cnt[(size_t)j * kSlice + l]);
            base[j + 1] = base[j] + (size_t)o.width[j] * kSlice;
        }
        o.code.assign(base[nw], 0);
        pos_k.assign((size_t)nw * kSlice, 0);
        each([&](int t, int l, uint32_t e, uint32_t ab, bool diag) {
            const uint32_t le = (uint32_t)(std::lower_bound(o.elem.begin(), o.elem.end(), (int32_t)e) - o.elem.begin());
            const int j = diag ? 0 : wt_of[t];
            o.code[base[j] + (size_t)pos_k[(size_t)j * kSlice + l]++ * kSlice + l] = (uint16_t)(le << 4 | ab);
        });
    });
    P.sa_eoff.assign((size_t)P.num_slices + 1, 0);
    P.sa_woff.assign((size_t)P.num_slices + 1, 0);
    for (int s = 0; s < P.num_slices; s++) {
        if (!sl[s].err.empty()) return sl[s].err;
        P.sa_eoff[s + 1] = P.sa_eoff[s] + (int64_t)sl[s].elem.size();
        P.sa_woff[s + 1] = P.sa_woff[s] + (int32_t)sl[s].width.size();
    }
    const int64_t NW = P.sa_woff[P.num_slices];
    P.sa_wcoff.assign((size_t)NW + 1, 0);
    for (int s = 0; s < P.num_slices; s++)
        for (size_t j = 0; j < sl[s].width.size(); j++) {
            const int64_t g = P.sa_woff[s] + (int64_t)j;
            P.sa_wcoff[g + 1] = P.sa_wcoff[g] + (int64_t)sl[s].width[j] * kSlice;
        }
    P.sa_elem.resize(P.sa_eoff[P.num_slices]);
    P.sa_code.resize(P.sa_wcoff.back());
    P.sa_cnt.resize((size_t)NW * kSlice);
    P.sa_aux.resize((size_t)NW * kSlice);
    P.sa_pos.resize((size_t)NW * kSlice);
    parallel_for(P.num_slices, [&](int64_t s, std::vector<int32_t> &) {
        const int64_t g0 = P.sa_woff[s];
        std::copy(sl[s].elem.begin(), sl[s].elem.end(), P.sa_elem.begin() + P.sa_eoff[s]);
        std::copy(sl[s].code.begin(), sl[s].code.end(), P.sa_code.begin() + P.sa_wcoff[g0]);
        std::copy(sl[s].cnt.begin(), sl[s].cnt.end(), P.sa_cnt.begin() + g0 * kSlice);
        std::copy(sl[s].aux.begin(), sl[s].aux.end(), P.sa_aux.begin() + g0 * kSlice);
        std::copy(sl[s].pos.begin(), sl[s].pos.end(), P.sa_pos.begin() + g0 * kSlice);
        SliceLists().elem.swap(sl[s].elem);
        std::vector<uint16_t>().swap(sl[s].code);
    });
    {
        std::string m = make_tiles(P.nctb_off, P.nctb, P.ntile);
        if (!m.empty()) return m;
    }

    // --- RNG: draws per step and digit jump tables ---
    P.elem_draws = calc_noise == 1 ? 7ull * (uint64_t)E : 0;
    uint64_t node_draws = (calc_noise == 1 && calc_stokes == 1) ? 3ull * (uint64_t)NT : 0;
    P.draws_per_step = P.elem_draws + node_draws;
    if (P.draws_per_step >= (1ull << 32)) return "more than 2^32 random draws per step is not supported";
    P.jump.resize(4 * 256 * 2 + 64 * 2);
    for (int comp = 0; comp < 2; comp++) {
        uint64_t m = comp ? kM2 : kM1;
        M3 step = base(comp);  // A^(256^d)
        for (int d = 0; d < 4; d++) {
            M3 acc = ident();
            for (int v = 0; v < 256; v++) {
                P.jump[((size_t)d * 256 + v) * 2 + comp] = to32(acc);
                acc = mul(step, acc, m);
            }
            step = acc;  // A^(256^(d+1))
        }
        P.step_jump[comp] = to32(powm(comp, P.draws_per_step));
        // A^(7 l): the lane offsets of a wavefront of consecutive elements
        for (int l = 0; l < 64; l++) P.jump[4 * 256 * 2 + 2 * l + comp] = to32(powm(comp, 7ull * (uint64_t)l));
    }
    return "";
}

std::string build_mass(const Plan &P, int solver, std::vector<double> &slot_val, std::vector<double> &dinv3,
                       std::vector<double> &inv_mass) {
    const int N = P.num_linear, E = P.num_elements;
    if (solver == 2) {
        // inv_M[ni] += .25 * rho * vol_0 over the elements in order, pinned rows 1, then 1 / inv_M
        inv_mass.assign(N, 0.0);
        for (int e = 0; e < E; e++)
            for (int a = 0; a < 4; a++) inv_mass[P.enodes[4 * (size_t)e + a]] += .25 * P.matRho[e] * P.vol0[e];
        for (int i = 0; i < N; i++) {
            if (P.pinned[i]) inv_mass[i] = 1.0;
            inv_mass[i] = 1.0 / inv_mass[i];
        }
        return "";
    }
    if (solver != 1) return "unknown mass-ful solver";
    // Real slots of a row are its leading slots with strictly increasing columns; the padding
    // slots after them repeat the row's own (diagonal) column.
    std::vector<int> len(N, 0);
    auto slot_at = [&](int r, int t) -> int64_t { return P.slice_off[r / kSlice] + (int64_t)t * kSlice + r % kSlice; };
    for (int r = 0; r < N; r++) {
        const int w = P.slice_width[r / kSlice];
        int prev = -1, t = 0;
        for (; t < w; t++) {
            const int c = P.slot_col[slot_at(r, t)];
            if (c <= prev) break;
            prev = c;
        }
        len[r] = t;
    }
    auto find = [&](int r, int c) -> int64_t {
        int lo = 0, hi = len[r];
        while (lo < hi) {
            const int mid = (lo + hi) / 2;
            if (P.slot_col[slot_at(r, mid)] < c)
                lo = mid + 1;
            else
                hi = mid;
        }
        return (lo < len[r] && P.slot_col[slot_at(r, lo)] == c) ? slot_at(r, lo) : -1;
    };
    slot_val.assign(P.num_slots, 0.0);
    // mass_LU[ni][nj] += (.1 on the diagonal, .05 off it) * rho * vol_0, element by element
    for (int e = 0; e < E; e++)
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) {
                const int r = P.enodes[4 * (size_t)e + a], c = P.enodes[4 * (size_t)e + b];
                if (P.pinned[r] || P.pinned[c]) continue;
                const int64_t k = find(r, c);
                if (k < 0) return "mass matrix entry outside the block pattern";
                slot_val[k] += (a == b ? .1 : .05) * P.matRho[e] * P.vol0[e];
            }
    dinv3.assign(3 * (size_t)N, 0.0);
    for (int r = 0; r < N; r++) {
        const int64_t k = find(r, r);
        if (k < 0) return "mass matrix row without a diagonal";
        if (P.pinned[r]) slot_val[k] = 1.0;
        if (slot_val[k] == 0.0) return "zero mass on a linear node (density 0)";
        const double d = 1.0 / slot_val[k];
        for (int c = 0; c < 3; c++) dinv3[3 * (size_t)r + c] = d;
    }
    return "";
}

std::string build_forces(const Plan &P, const ForceInput &in, ForcePlan &F) {
    const int N = P.num_linear, NT = P.num_nodes_total, E = P.num_elements;
    F = ForcePlan();
    const size_t ns = in.spring_n0.size(), nc = in.const_node.size(), nf = in.faces.size() / 3;
    const size_t nr = in.rot_node.size();
    const size_t nset = in.set_off.empty() ? 0 : in.set_off.size() - 1;
    const size_t nss = in.ssint_faces.size() / 4;
    if (nss && nset)
        return "surface sets (CTF 3) together with surface-surface interactions are not on this path";
    if (in.spring_n1.size() != ns || in.spring_k.size() != ns || in.spring_l.size() != ns ||
        in.const_force.size() != 3 * nc || in.faces.size() != 3 * nf || in.set_force.size() != 3 * nset ||
        in.rot_F.size() != nr || in.rot_axis.size() != 6 * nr || in.rot_torque.size() != nr)
        return "inconsistent force list sizes";
    // edge ends of every midpoint (N_t ids -> compact ids)
    std::vector<int32_t> end_a(NT, -1), end_b(NT, -1);
    for (size_t k = 0; k < P.mid_index.size(); k++) {
        end_a[P.mid_index[k]] = P.lin_of[P.mid_a[k]];
        end_b[P.mid_index[k]] = P.lin_of[P.mid_b[k]];
    }
    std::vector<int32_t> mid_of(NT, -1);  // midpoint N_t id -> carrying-midpoint index
    auto midpoint = [&](int32_t v) -> int32_t {
        if (mid_of[v] < 0) {
            mid_of[v] = (int32_t)(F.mid_ends.size() / 2);
            F.mid_ends.push_back(end_a[v]);
            F.mid_ends.push_back(end_b[v]);
        }
        return mid_of[v];
    };
    std::vector<std::vector<int32_t>> pre(N), mpre;
    F.spring_nodes.resize(2 * ns);
    F.spring_kl.resize(2 * ns);
    F.spring_blob.resize(2 * ns);
    for (size_t s = 0; s < ns; s++) {
        int32_t v[2] = {in.spring_n0[s], in.spring_n1[s]};
        int c[2];
        for (int e = 0; e < 2; e++) {
            if (v[e] < 0 || v[e] >= NT) return "spring " + std::to_string(s) + ": node index out of range";
            c[e] = P.lin_of[v[e]];
            if (c[e] < 0) return "spring " + std::to_string(s) + ": node " + std::to_string(v[e]) + " is an edge midpoint";
        }
        if (c[0] == c[1]) return "spring " + std::to_string(s) + " joins a node to itself";
        for (int e = 0; e < 2; e++) {
            F.spring_nodes[2 * s + e] = c[e];
            F.spring_blob[2 * s + e] = P.slice_blob[c[e] / kSlice];
            pre[c[e]].push_back((int32_t)(2 * s + e));
        }
        F.spring_kl[2 * s] = in.spring_k[s];
        F.spring_kl[2 * s + 1] = in.spring_l[s];
    }
    // linear constant forces, then rotational ones: a node's items in that order
    for (size_t i = 0; i < nc + nr; i++) {
        const bool rot = i >= nc;
        const int32_t v = rot ? in.rot_node[i - nc] : in.const_node[i];
        const std::string what = rot ? "rotational force " + std::to_string(i - nc) : "constant force " + std::to_string(i);
        if (v < 0 || v >= NT) return what + ": node out of range";
        if (rot && in.rot_torque[i - nc] != 0 && in.rot_torque[i - nc] != 1) return what + ": type must be 0 or 1";
        const int32_t item = ~(int32_t)i;
        if (P.lin_of[v] < 0 && end_a[v] < 0) return what + ": node is in no element";
        if (P.lin_of[v] >= 0) {
            pre[P.lin_of[v]].push_back(item);
        } else {
            const int32_t m = midpoint(v);
            if ((size_t)m >= mpre.size()) mpre.resize(m + 1);
            mpre[m].push_back(item);
        }
    }
    F.cforce = in.const_force;
    F.nconst = (int32_t)nc;
    F.rot.resize(8 * nr);
    for (size_t r = 0; r < nr; r++) {
        for (int k = 0; k < 6; k++) F.rot[8 * r + k] = in.rot_axis[6 * r + k];
        F.rot[8 * r + 6] = in.rot_F[r];
        F.rot[8 * r + 7] = in.rot_torque[r];
    }
    // surface sets: entries, and per face the entries that load it (set order)
    std::vector<std::vector<int32_t>> ent_of_face(nf);
    if (nset) {
        F.set_off = in.set_off;
        F.set_force = in.set_force;
        if (F.set_off[0] != 0 || (size_t)F.set_off[nset] != in.set_faces.size()) return "inconsistent surface set offsets";
        for (size_t s = 0; s < nset; s++)
            if (F.set_off[s + 1] < F.set_off[s]) return "inconsistent surface set offsets";
        const size_t nent = in.set_faces.size();
        F.ent_ends.resize(6 * nent);
        for (size_t e = 0; e < nent; e++) {
            const int32_t f = in.set_faces[e];
            if (f < 0 || (size_t)f >= nf) return "surface set face index out of range";
            ent_of_face[f].push_back((int32_t)e);
            for (int k = 0; k < 3; k++) {
                const int32_t v = in.faces[3 * (size_t)f + k];
                if (v < 0 || v >= NT) return "face " + std::to_string(f) + ": node index out of range";
                const int c = P.lin_of[v];
                if (c < 0 && end_a[v] < 0) return "face " + std::to_string(f) + ": node " + std::to_string(v) + " is in no element";
                F.ent_ends[6 * e + 2 * k] = c >= 0 ? c : end_a[v];
                F.ent_ends[6 * e + 2 * k + 1] = c >= 0 ? c : end_b[v];
            }
        }
    }
    // face groups: vertices first, then carrying midpoints; a group key is a set face, or
    // 4 p + k for node k of interaction face p (corners 0-2, opposite vertex 3)
    std::vector<std::vector<int32_t>> grp(N), mgrp;
    auto add_key = [&](int32_t v, int32_t key) {
        if (P.lin_of[v] >= 0) {
            grp[P.lin_of[v]].push_back(key);
        } else {
            const int32_t m = midpoint(v);
            if ((size_t)m >= mgrp.size()) mgrp.resize(m + 1);
            mgrp[m].push_back(key);
        }
    };
    for (size_t f = 0; f < nf; f++) {
        if (ent_of_face[f].empty()) continue;
        for (int k = 0; k < 3; k++) add_key(in.faces[3 * f + k], (int32_t)f);
    }
    F.ssint_row0 = (int32_t)in.set_faces.size();
    for (size_t p = 0; p < nss; p++)
        for (int k = 0; k < 4; k++) {
            const int32_t v = in.ssint_faces[4 * p + k];
            if (k == 3 && v == -1) continue;  // no opposite vertex: no steric forces (lennard-jones)
            if (v < 0 || v >= NT) return "interaction face " + std::to_string(p) + ": node index out of range";
            if (P.lin_of[v] < 0 && end_a[v] < 0)
                return "interaction face " + std::to_string(p) + ": node " + std::to_string(v) + " is in no element";
            add_key(v, (int32_t)(4 * p + k));
        }
    const int M = (int)(F.mid_ends.size() / 2);
    mpre.resize(M);
    mgrp.resize(M);
    bool any_pre = ns > 0;
    for (int i = 0; i < N && !any_pre; i++) any_pre = !pre[i].empty();
    if (any_pre || M > 0) {
        F.pre_off.assign((size_t)N + 1, 0);
        for (int i = 0; i < N; i++) {
            F.pre_item.insert(F.pre_item.end(), pre[i].begin(), pre[i].end());
            F.pre_off[i + 1] = (int32_t)F.pre_item.size();
        }
        F.mid_pre_off.assign((size_t)M + 1, F.pre_off[N]);
        for (int m = 0; m < M; m++) {
            F.pre_item.insert(F.pre_item.end(), mpre[m].begin(), mpre[m].end());
            F.mid_pre_off[m + 1] = (int32_t)F.pre_item.size();
        }
    }
    if (nset || nss) {
        F.grp_off.assign((size_t)N + 1, 0);
        F.grp_ent_off.assign(1, 0);
        auto add_groups = [&](const std::vector<int32_t> &keys) {
            for (int32_t f : keys) {
                if (nset)
                    F.grp_ent.insert(F.grp_ent.end(), ent_of_face[f].begin(), ent_of_face[f].end());
                else
                    F.grp_ent.push_back(F.ssint_row0 + f);
                F.grp_ent_off.push_back((int32_t)F.grp_ent.size());
            }
            return (int32_t)(F.grp_ent_off.size() - 1);
        };
        for (int i = 0; i < N; i++) F.grp_off[i + 1] = add_groups(grp[i]);
        F.mid_grp_off.assign((size_t)M + 1, F.grp_off[N]);
        for (int m = 0; m < M; m++) F.mid_grp_off[m + 1] = add_groups(mgrp[m]);
    }
    // linearise_force: element by element, a vertex takes half the sum of its three adjacent
    // midpoints; the element's midpoints are then zeroed (counted once, in their first element)
    if (M > 0) {
        static const int sets[4][3] = {{0, 1, 2}, {0, 3, 4}, {1, 3, 5}, {2, 4, 5}};
        std::vector<uint8_t> used(NT, 0);
        std::vector<std::vector<int32_t>> terms(N);
        for (int e = 0; e < E; e++) {
            const int32_t *m = &P.emid[6 * (size_t)e];
            for (int a = 0; a < 4; a++) {
                int32_t t[3];
                bool any = false;
                for (int k = 0; k < 3; k++) {
                    const int32_t v = m[sets[a][k]];
                    t[k] = used[v] ? -1 : mid_of[v];
                    any |= t[k] >= 0;
                }
                if (!any) continue;
                auto &tt = terms[P.enodes[4 * (size_t)e + a]];
                tt.insert(tt.end(), t, t + 3);
            }
            for (int k = 0; k < 6; k++) used[m[k]] = 1;
        }
        F.fold_off.assign((size_t)N + 1, 0);
        for (int i = 0; i < N; i++) {
            F.fold_mid.insert(F.fold_mid.end(), terms[i].begin(), terms[i].end());
            F.fold_off[i + 1] = (int32_t)(F.fold_mid.size() / 3);
        }
    }
    return "";
}

}  // namespace ffh
