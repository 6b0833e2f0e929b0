// Host-side assembly setup: the counterpart of AssemblerT::PrepareProblem
// (inmost_interface/assembler.inl:194-275) for a GPU assembler.  Produces the
// global DOF numbering, the CSR pattern, the element-to-CSR-slot map and the
// row blocks of the owner-computes assembly (blocks.cpp).
#pragma once

#include <cstdint>
#include <functional>
#include <vector>

#include "../../include/anifem_gpu.h"

namespace anifem {

// Row-block budgets of the owner-computes assembly (blocks.cpp).
constexpr int kBlockMaxSlots = 2048;         // CSR slots per block
constexpr int kBlockMaxContrib = 2048;       // (tet, entry) contributions per block
constexpr int kBlockMaxTets = 128;           // default tets per block
constexpr int kBlockHardMaxTets = 1024;      // largest DOF valence supported
constexpr int kBlockKBytes = 24 * 1024;      // default FP64 geometry-tensor bytes per block
constexpr int kBlockMaxKBytes = 96 * 1024;   // ... when raised to the DOF valence

struct AssemblySetup {
    int fem = 0, ncomp = 1, nf = 0, E = 0;
    int64_t n_nodes = 0, n_tets = 0, n_edges = 0, n_faces = 0;
    int64_t ndof = 0, nnz = 0;
    std::vector<int32_t> tets;        // n_tets x 4 after the node reorder (caller's order)
    std::vector<int32_t> dofs;        // n_tets x nf global DOFs (caller's order)
    std::vector<uint8_t> is_dir;      // ndof
    std::vector<int64_t> row_ptr;     // ndof + 1
    std::vector<int32_t> col_idx;     // nnz
    std::vector<int32_t> slot;        // caller's order: n_tets x E (entry ib + nf*ia), -1 dropped
    int64_t n_degenerate = 0;         // tets with a zero or non-finite Jacobian determinant
    int k_per_tet = 1;                // FP64 geometry-tensor values per tet (matrix + load vector)
    // row blocks: consecutive CSR rows [blk_row0[b], blk_row0[b+1])
    int block_tet_cap = kBlockMaxTets;
    bool code16 = true;               // codes fit 16 bits on the device
    int ebits = 1, fbits = 0;         // entry / load-vector index bits of the codes
    int64_t n_blocks = 0;
    std::vector<int32_t> blk_row0;    // n_blocks + 1
    std::vector<int64_t> blk_s0;      // n_blocks + 1: first CSR slot of each block (row_ptr[blk_row0])
    std::vector<int64_t> blk_tet0;    // n_blocks + 1 offsets into blk_tets
    std::vector<int32_t> blk_tets;    // distinct tets of each block, ascending
    std::vector<int64_t> blk_c0;      // n_blocks + 1 offsets into codes
    std::vector<uint32_t> codes;      // per block: local_tet << ebits | entry (E: Dirichlet diagonal, value 1),
                                      // grouped by slot, ascending tet
    std::vector<uint16_t> sstart;     // nnz: first code of each slot, relative to its block
    std::vector<int64_t> blk_g0;      // n_blocks + 1 offsets into fcodes
    std::vector<uint16_t> fcodes;     // per block: local_tet << fbits | i, grouped by row, ascending tet
    std::vector<uint16_t> rstart;     // ndof: first load-vector code of each row, relative to its block
    int64_t max_block_tets = 0, max_block_contrib = 0, max_block_fcontrib = 0, max_block_slots = 0;
};

int build_setup(const anifem_mesh_desc& mesh, const anifem_problem_desc& prob, AssemblySetup& S);
int build_blocks(AssemblySetup& S, const std::vector<int64_t>& aptr, const std::vector<int32_t>& adj);

// Parallel loop over [0, n) on the host.
void host_parallel_for(int64_t n, int64_t grain, const std::function<void(int64_t, int64_t)>& fn);

}  // namespace anifem
