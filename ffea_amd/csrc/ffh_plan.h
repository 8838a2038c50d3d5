// ffh_plan.h -- host-side setup of a blob for the GPU path: linear-node compaction,
// rest state, the block sparsity pattern in sliced-ELL (SELL-64) order with its
// element-to-slot gather maps, the node force gather maps and the RngStream jump tables.
// Built once per blob (the work NoMassCGSolver::init and Blob::init do on the CPU).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace ffh {

constexpr int kSlice = 64;  // rows per SELL slice = one wavefront
constexpr uint8_t kSlotDiag = 1, kSlotIdentity = 2, kSlotPad = 4;
constexpr uint32_t kM1 = 4294967087u;
constexpr uint32_t kM2 = 4294944443u;

struct Mat3u {
    uint32_t m[3][3];
};

// Per-entry lists laid out in 64-wide tiles: entry u = 64 g + lane finds its k-th item at
// code[off[g] + 64 k + lane] for k < cnt[u]; width[g] = max cnt over the tile.  Lanes of a
// wavefront then read one list position with one coalesced load.
struct Tiles {
    std::vector<int64_t> off;     // [groups + 1]
    std::vector<uint16_t> width;  // [groups]
    std::vector<uint16_t> cnt;    // [entries]
    std::vector<uint32_t> code;   // [off[groups]]
};

struct Plan {
    // sizes
    int num_nodes_total = 0;  // N_t
    int num_linear = 0;       // N, rows on the device (real linear nodes + slice padding)
    int num_real_linear = 0;  // linear nodes of the mesh
    int num_elements = 0;     // E
    // blobs: each blob's rows start on a 64-row slice boundary (padding rows are pinned
    // dummies), so every slice -- and every per-slice partial sum -- belongs to one blob
    int num_blobs = 1;
    std::vector<int32_t> blob_slice_off;  // [num_blobs + 1]
    std::vector<int32_t> slice_blob;      // [num_slices]
    // node maps
    std::vector<int> lin_of;   // [N_t] -> compact index or -1
    std::vector<int> orig_of;  // [N]   -> original index, -1 for padding rows
    // per element (compact node ids), rest data and material in meso units
    std::vector<int32_t> enodes;   // [4*E] AoS (int4 per element)
    std::vector<double> jinv0;     // [9*E] SoA: jinv0[k*E + e]
    std::vector<double> vol0;      // [E]
    std::vector<double> last_det;  // [E]
    std::vector<double> matA, matB, matG, matE;  // [E]
    std::vector<double> matRho;                  // [E] density in meso units
    // per linear node
    std::vector<double> pos;     // [3*N] SoA x|y|z of the current positions
    std::vector<double> zeta;    // [N] Stokes drag (0 when calc_stokes == 0)
    std::vector<uint8_t> pinned; // [N]
    std::vector<uint32_t> rng_node_off;  // [N] 3*original index (offset inside the node draws)
    // midpoint reconstruction for N_t output: for each midpoint, its two edge ends (original ids)
    std::vector<int32_t> mid_index, mid_a, mid_b;
    std::vector<int32_t> emid;     // [6*E] midpoint node ids of each element (original numbering)
    // force gather: per linear node, contributions (e << 2 | a) in ascending element order
    std::vector<uint32_t> nctb_off, nctb;
    Tiles ntile;  // the same lists, tiled per 64 nodes
    // block pattern in SELL-64 order
    int num_slices = 0;
    int64_t num_blocks = 0;  // real 3x3 blocks
    int64_t num_slots = 0;   // padded slots = sum over slices of 64 * width
    std::vector<int64_t> slice_off;  // [num_slices+1] first slot of each slice
    std::vector<int32_t> slice_width;
    std::vector<int32_t> slot_col;    // [num_slots]
    // slot_flags: kSlotDiag (the row's diagonal block: Stokes drag and D^-1), kSlotIdentity (a
    // pinned row's diagonal: identity), kSlotPad (padding: zero for good)
    std::vector<uint8_t> slot_flags;  // [num_slots]
    // slice assembly: per slice its distinct elements (ascending) and its work tiles: the
    // diagonal tile (lane l: row 64 s + l's diagonal block), then the upper blocks (column > row;
    // a lower block is written as its upper mirror's transpose) of SELL tiles whose lanes do not
    // meet, merged into one work tile.  Per work-tile lane its contributions (slice-local element
    // << 4 | local row vertex a << 2 | local column vertex b, ascending element), tiled [k][64]
    // from sa_wcoff[g]; sa_pos: the block's SELL position (255: none); sa_aux: for an upper block
    // its mirror's global slot, -1 where the lane has none
    std::vector<int64_t> sa_eoff;   // [num_slices + 1]
    std::vector<int32_t> sa_elem;   // global element ids
    std::vector<int32_t> sa_woff;   // [num_slices + 1] work tiles of each slice
    std::vector<int64_t> sa_wcoff;  // [work tiles + 1]
    std::vector<uint16_t> sa_code;
    std::vector<uint8_t> sa_cnt;    // [work tiles * 64]
    std::vector<int32_t> sa_aux;    // [work tiles * 64]
    std::vector<uint8_t> sa_pos;    // [work tiles * 64]
    // RNG jump tables: tbl[((d*256 + v)*2 + comp)] = A_comp^(v * 256^d), d=0..3
    std::vector<Mat3u> jump;
    Mat3u step_jump[2];        // A^(draws per step)
    uint64_t draws_per_step = 0;
    uint64_t elem_draws = 0;   // 7*E when calc_noise
};

// Tiles CSR lists (off[n+1], items); returns "" or an error message.
std::string make_tiles(const std::vector<uint32_t> &off, const std::vector<uint32_t> &items, Tiles &t);

// Builds the plan; returns "" on success or an error message.
// blob_node_off: [num_blobs + 1] offsets of each blob's nodes in the concatenated node arrays
// (nullptr: one blob); elements must only reference nodes of their own blob.
// material_meso: material_si holds [5*E] rho, A, B, G, E already in meso units (as
// tetra_element_linear holds them) instead of [6*E] .mat rows in SI.
std::string build_plan(Plan &P, int num_nodes, int num_elements, const double *pos, const int *topo,
                       const double *material_si, const double *stokes_radius, int num_pinned,
                       const int *pinned, double stokes_visc, int calc_noise, int calc_stokes,
                       int num_blobs = 1, const int *blob_node_off = nullptr, bool material_meso = false);

// The mass-ful solvers' matrices on the linear rows; midpoint rows are identity rows with a zero
// right-hand side in the reference (their forces are folded onto the vertices first), so they
// drop out of the system as they do for the viscosity matrix.
//  solver 1 (ConjugateGradientSolver::init, src/ConjugateGradientSolver.cpp:62-203): the
//    consistent mass matrix, one scalar per SELL slot of the block pattern (it acts on the three
//    components alike), each entry summed over the elements in ascending order; identity rows
//    for pinned and padding rows; dinv3 [3N] its inverse diagonal per scalar row.
//  solver 2 (MassLumpedSolver::init, src/MassLumpedSolver.cpp:40-79): inv_mass [N], the inverse
//    of a quarter of the mass of every element at the vertex; 1 on pinned and padding rows.
// Returns "" or an error message.
std::string build_mass(const Plan &P, int solver, std::vector<double> &slot_val, std::vector<double> &dinv3,
                       std::vector<double> &inv_mass);

// Forces from outside the element loop, as the caller gives them (node ids in the N_t
// numbering of the handle: a batch's blobs concatenated; meso units).
struct ForceInput {
    // World::apply_springs (src/World.cpp:3808-3846)
    std::vector<int32_t> spring_n0, spring_n1;
    std::vector<double> spring_k, spring_l;
    // Blob::apply_ctforces CTF 1 (src/Blob.cpp:2096-2103), in the stored order
    std::vector<int32_t> const_node;
    std::vector<double> const_force;  // [3C]
    // CTF 2 (src/Blob.cpp:2104-2136), point + direction axes: per force its node, F, the
    // axis (point, unit direction) and 1 for a torque ('t'), 0 for a circular force ('f')
    std::vector<int32_t> rot_node;
    std::vector<double> rot_F;
    std::vector<double> rot_axis;     // [6R]
    std::vector<int32_t> rot_torque;
    // Blob::apply_ctforces CTF 3 (src/Blob.cpp:2137-2159): surface faces (Face::n[0..2]) and
    // sets of face indices with their total force
    std::vector<int32_t> faces;       // [3F]
    std::vector<int32_t> set_off;     // [S+1]
    std::vector<int32_t> set_faces;
    std::vector<double> set_force;    // [3S]
    // surface-surface interactions (VdW_solver): the active faces' nodes Face::n[0..3] in pool
    // order (the opposite vertex -1 where no steric forces reach it); each node's Face::force
    // is one value row the interaction kernel writes every step
    std::vector<int32_t> ssint_faces; // [4P]
};

// The same forces as per-node lists in the order the reference adds them to a node's force:
//  1. before the element forces: the node's springs (ascending spring index), then its
//     linear constant forces, then its rotational ones (each in stored order) -- World::run
//     applies the springs, then update_internal_forces the constant forces, onto the zeroed
//     force vector;
//  2. after the element forces: one group per surface face holding the node (ascending face
//     index), each the sum of the face's set tractions in set order, or the face node's
//     surface-surface interaction force (Face::force[0..3], the opposite vertex included,
//     gathered in aggregate_forces_and_solve, src/Blob.cpp:3842-3851);
//  3. after the noise: the forces on edge midpoints (constant forces, then face groups, as in
//     1 and 2), passed on to the edge ends element by element as linearise_force does
//     (src/Blob.cpp:1066-1085): each (vertex, element) term is half the sum of the element's
//     three midpoints next to the vertex, a midpoint counting only in the first element that
//     holds it.
// Second-order surfaces are split into triangles that use edge midpoints as corners; such a
// corner's position is the average of its edge ends, as linearise_element leaves it.
struct ForcePlan {
    std::vector<int32_t> pre_off;    // [N+1]
    std::vector<int32_t> pre_item;   // >= 0: 2 * spring + end; < 0: ~c, c < nconst a linear
                                     // constant force, else rotational force c - nconst
    std::vector<int32_t> spring_nodes;  // [2 * springs] compact ids
    std::vector<double> spring_kl;      // [2 * springs]
    std::vector<double> cforce;         // [3 * constants]
    int32_t nconst = 0;
    std::vector<double> rot;            // [8 * rotational]: axis point, direction, F, torque
    std::vector<int32_t> grp_off;       // [N+1] face groups of the vertices
    std::vector<int32_t> grp_ent_off;   // [groups + 1]
    std::vector<int32_t> grp_ent;       // value rows: set entries, then 4 p + k for node k of interaction face p
    int32_t ssint_row0 = 0;             // first interaction row (after the set entries)
    std::vector<int32_t> set_off;       // [S+1] entries of each set
    std::vector<int32_t> ent_ends;      // [6 * entries] per face corner its two edge ends (a vertex twice)
    std::vector<double> set_force;      // [3S]
    // midpoints that carry a force: edge ends, constant-force items, face groups
    std::vector<int32_t> mid_ends;      // [2M] compact ids
    std::vector<int32_t> mid_pre_off;   // [M+1] into pre_item (constant forces only)
    std::vector<int32_t> mid_grp_off;   // [M+1] into the groups
    std::vector<int32_t> fold_off;      // [N+1] fold terms of each vertex
    std::vector<int32_t> fold_mid;      // [3 * terms] midpoint indices, -1 for none
    std::vector<int32_t> spring_blob;   // [2 * springs] blobs of the two ends (energy grouping)
};

// Returns "" or an error message (ids out of range, a spring end on an edge midpoint, a spring
// joining a node to itself, surface sets together with surface-surface interactions).
std::string build_forces(const Plan &P, const ForceInput &in, ForcePlan &F);

// Exact jump matrices of MRG32k3a component c (0: m1, 1: m2) raised to n.
Mat3u mrg_pow(int comp, uint64_t n);

}  // namespace ffh
