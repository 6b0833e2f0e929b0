// Host-side assembly setup: the counterpart of AssemblerT::PrepareProblem
// (inmost_interface/assembler.inl:194-275) for a GPU assembler.  Produces the
// global DOF numbering, the CSR pattern, the element-to-CSR-slot map and a
// conflict-free colouring, all in "assembly order" (colour-major, caller's tet
// order inside a colour).
#pragma once

#include <cstdint>
#include <functional>
#include <vector>

#include "../../include/anifem_gpu.h"

namespace anifem {

struct AssemblySetup {
    int fem = 0, ncomp = 1, nf = 0, E = 0;
    int64_t n_nodes = 0, n_tets = 0, n_edges = 0, n_faces = 0;
    int64_t ndof = 0, nnz = 0;
    std::vector<int32_t> tets;        // n_tets x 4 after the node reorder (caller's order)
    std::vector<int32_t> dofs;        // n_tets x nf global DOFs (caller's order)
    std::vector<uint8_t> is_dir;      // ndof
    std::vector<int64_t> row_ptr;     // ndof + 1
    std::vector<int32_t> col_idx;     // nnz
    std::vector<int32_t> colour;      // n_tets
    int n_colours = 0;
    std::vector<int64_t> colour_ptr;  // n_colours + 1 (assembly order ranges)
    std::vector<int32_t> perm;        // assembly position -> caller's tet id
    std::vector<int32_t> slot;        // assembly order: n_tets x E (entry ib + nf*ia), -1 dropped
    int64_t n_nonpositive = 0;        // tets with det <= 0 after reorder
};

int build_setup(const anifem_mesh_desc& mesh, const anifem_problem_desc& prob, AssemblySetup& S);

// Parallel loop over [0, n) on the host.
void host_parallel_for(int64_t n, int64_t grain, const std::function<void(int64_t, int64_t)>& fn);

}  // namespace anifem
