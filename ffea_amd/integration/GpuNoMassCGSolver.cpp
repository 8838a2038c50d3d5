// GpuNoMassCGSolver.cpp -- the Solver interface over libffea_hip.so (see the header).
#include "GpuNoMassCGSolver.h"

#include <algorithm>
#include <cstdio>

GpuNoMassCGSolver::~GpuNoMassCGSolver() {
    if (blob_) ffh_blob_destroy(blob_);
}

std::string GpuNoMassCGSolver::marshal(int num_nodes, int num_elements, mesh_node *node, tetra_element_linear *elem,
                                       int num_pinned_nodes, const int *pinned_nodes_list, const set<int> &bsite,
                                       Arrays &out) {
    if (num_nodes <= 0 || num_elements <= 0 || !node || !elem) return "no nodes or no elements";
    out.pos.resize(3 * (size_t)num_nodes);
    out.radius.resize(num_nodes);
    for (int i = 0; i < num_nodes; i++) {
        out.pos[3 * (size_t)i] = node[i].pos.x;
        out.pos[3 * (size_t)i + 1] = node[i].pos.y;
        out.pos[3 * (size_t)i + 2] = node[i].pos.z;
        out.radius[i] = node[i].stokes_radius;
    }
    out.topo.resize(10 * (size_t)num_elements);
    out.material.resize(5 * (size_t)num_elements);
    for (int e = 0; e < num_elements; e++) {
        for (int j = 0; j < 10; j++) {
            const mesh_node *n = elem[e].n[j];
            if (n < node || n >= node + num_nodes)
                return "element " + std::to_string(e) + " points at a node outside the blob's node array";
            out.topo[10 * (size_t)e + j] = (int)(n - node);
        }
        const double m[5] = {elem[e].rho, elem[e].A, elem[e].B, elem[e].G, elem[e].E};
        std::copy(m, m + 5, out.material.begin() + 5 * (size_t)e);
    }
    out.pinned.assign(pinned_nodes_list ? pinned_nodes_list : nullptr,
                      pinned_nodes_list ? pinned_nodes_list + std::max(0, num_pinned_nodes) : nullptr);
    out.pinned.insert(out.pinned.end(), bsite.begin(), bsite.end());
    std::sort(out.pinned.begin(), out.pinned.end());
    out.pinned.erase(std::unique(out.pinned.begin(), out.pinned.end()), out.pinned.end());
    for (int p : out.pinned)
        if (p < 0 || p >= num_nodes) return "pinned node " + std::to_string(p) + " is out of range";
    return "";
}

int GpuNoMassCGSolver::init(int num_nodes, int num_elements, mesh_node *node, tetra_element_linear *elem,
                            SimulationParams *params, int num_pinned_nodes, int *pinned_nodes_list,
                            set<int> bsite_pinned_nodes_list) {
    if (blob_) {  // Blob::reset_solver initialises again (src/Blob.cpp:766-776)
        ffh_blob_destroy(blob_);
        blob_ = nullptr;
    }
    Arrays a;
    const std::string msg =
        marshal(num_nodes, num_elements, node, elem, num_pinned_nodes, pinned_nodes_list, bsite_pinned_nodes_list, a);
    if (!msg.empty()) FFEA_ERROR_MESSG("GpuNoMassCGSolver::init: %s\n", msg.c_str());
    // the solver makes no random draws: any valid RngStream state
    const unsigned long long rng[6] = {12345, 12345, 12345, 12345, 12345, 12345};
    blob_ = ffh_blob_create_meso(num_nodes, num_elements, a.pos.data(), a.topo.data(), a.material.data(),
                                 a.radius.data(), (int)a.pinned.size(), a.pinned.empty() ? nullptr : a.pinned.data(),
                                 params->dt, params->kT, params->epsilon2, params->max_iterations_cg, 0,
                                 params->calc_stokes, params->stokes_visc, rng, FFH_PRECISION_FP64);
    if (!blob_) FFEA_ERROR_MESSG("GpuNoMassCGSolver::init: %s\n", ffh_last_error());
    node_ = node;
    num_nodes_ = num_nodes;
    buf_.assign(3 * (size_t)num_nodes, 0.0);
    return FFEA_OK;
}

int GpuNoMassCGSolver::solve(vector3 *x) {
    if (!blob_) FFEA_ERROR_MESSG("GpuNoMassCGSolver::solve: not initialised\n");
    for (int i = 0; i < num_nodes_; i++) {
        buf_[3 * (size_t)i] = node_[i].pos.x;
        buf_[3 * (size_t)i + 1] = node_[i].pos.y;
        buf_[3 * (size_t)i + 2] = node_[i].pos.z;
    }
    if (ffh_blob_set_positions(blob_, buf_.data()) != FFH_OK || ffh_blob_update_matrix(blob_) != FFH_OK)
        FFEA_ERROR_MESSG("GpuNoMassCGSolver::solve: %s\n", ffh_last_error());
    for (int i = 0; i < num_nodes_; i++) {
        buf_[3 * (size_t)i] = x[i].x;
        buf_[3 * (size_t)i + 1] = x[i].y;
        buf_[3 * (size_t)i + 2] = x[i].z;
    }
    const int rc = ffh_blob_solve(blob_, buf_.data());
    if (rc != FFH_OK && rc != FFH_E_CG) FFEA_ERROR_MESSG("GpuNoMassCGSolver::solve: %s\n", ffh_last_error());
    for (int i = 0; i < num_nodes_; i++) x[i].assign(buf_[3 * (size_t)i], buf_[3 * (size_t)i + 1], buf_[3 * (size_t)i + 2]);
    if (rc == FFH_E_CG)  // as NoMassCGSolver::solve reports it (src/NoMassCGSolver.cpp:196)
        FFEA_ERROR_MESSG("Conjugate gradient solver: %s\n", ffh_last_error());
    return FFEA_OK;
}

void GpuNoMassCGSolver::apply_matrix(scalar *in, scalar *result) {
    if (!blob_ || ffh_blob_apply_matrix(blob_, in, result) != FFH_OK)
        printf("GpuNoMassCGSolver::apply_matrix: %s\n", blob_ ? ffh_last_error() : "not initialised");
}
