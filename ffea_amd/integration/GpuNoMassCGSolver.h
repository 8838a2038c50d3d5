// GpuNoMassCGSolver.h -- FFEA's Solver interface (include/Solver.h:30-55) over the MI355X
// library (include/ffea_hip.h): a drop-in for NoMassCGSolver (src/NoMassCGSolver.cpp), the
// massless solver behind `solver = CG_nomass`.  It is compiled in the simulator's own tree
// (it includes FFEA's headers) and linked with ffea_amd/libffea_hip.so; Blob allocates it where
// it allocates NoMassCGSolver (src/Blob.cpp:369-370).
//
// init marshals the blob's nodes and elements (their meso-unit materials as
// tetra_element_linear holds them) into one GPU handle, which keeps the blob's block sparsity
// pattern and its own copy of the matrix.  solve(x) uploads the current node positions,
// re-assembles the viscosity matrix there on the device (the build step of
// NoMassCGSolver::solve, src/SparseMatrixFixedPattern.cpp:85-93, from the same element
// formulas the reference's element loop evaluates at these positions) and solves in place.
// apply_matrix(in, result) is K in (the reference's NoMassCGSolver leaves it an empty stub,
// include/NoMassCGSolver.h:103-104).
#ifndef GPU_NOMASS_CG_SOLVER_H_INCLUDED
#define GPU_NOMASS_CG_SOLVER_H_INCLUDED

#include <set>
#include <string>
#include <vector>

#include "Solver.h"
#include "ffea_hip.h"

class GpuNoMassCGSolver : public Solver {
public:
    GpuNoMassCGSolver() {}
    ~GpuNoMassCGSolver();

    int init(int num_nodes, int num_elements, mesh_node *node, tetra_element_linear *elem, SimulationParams *params,
             int num_pinned_nodes, int *pinned_nodes_list, set<int> bsite_pinned_nodes_list);
    int solve(vector3 *x);
    void apply_matrix(scalar *in, scalar *result);

    // What init hands to ffh_blob_create_meso: positions [3 N_t], topology [10 E] (node
    // indices), materials [5 E] (rho, A, B, G, E, meso units), Stokes radii [N_t], and the
    // pinned nodes (pinned_nodes_list and the binding-site pins together, as
    // NoMassCGSolver::init marks them, src/NoMassCGSolver.cpp:82-91).
    struct Arrays {
        std::vector<double> pos, material, radius;
        std::vector<int> topo, pinned;
    };
    static std::string marshal(int num_nodes, int num_elements, mesh_node *node, tetra_element_linear *elem,
                               int num_pinned_nodes, const int *pinned_nodes_list, const set<int> &bsite,
                               Arrays &out);

    ffh_blob *handle() const { return blob_; }

private:
    ffh_blob *blob_ = nullptr;
    mesh_node *node_ = nullptr;
    int num_nodes_ = 0;
    std::vector<double> buf_;  // [3 N_t] positions / vectors in the library's layout
};

#endif
