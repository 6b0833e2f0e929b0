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

// Row-block budgets of the owner-computes assembly (blocks.cpp): a block is
// closed before any of them is exceeded.  Pairs are one per lane of the
// row-block workgroup; values and slots are 16-bit positions.
struct BlockCaps {
    int max_slots = 2048;   // CSR slots per block (<= 65535)
    int max_values = 2048;  // element entries per block (<= 65535)
    int max_pairs = 256;    // (row, incident tet) pairs per block (<= workgroup size)
};
// Default budgets by entries per (row, tet) pair: a block of 256 pairs of nf
// entries needs about 256 nf values, so larger element matrices get larger
// value / slot budgets (measured: P2^3 at 8192 values).  A scalar space of
// 11-20 DOFs per cell (P3) takes blocks of at most 128 pairs and 2048 values:
// its pass then runs two lanes per pair and the LDS of four workgroups fits a
// CU (P3 on the 40^3 cube: 0.503 ms against 0.565 at 3072 values and 256
// pairs, three workgroups).
inline BlockCaps default_caps(int nf, bool scalar_space) {
    BlockCaps c;
    if (nf > 20) {
        c.max_slots = 4096;
        c.max_values = 8192;
    } else if (nf > 10) {
        c.max_slots = 4096;
        c.max_values = scalar_space ? 2048 : 3072;
        c.max_pairs = scalar_space ? 128 : 256;
    }
    return c;
}
// Budgets for a mesh one of whose rows exceeds default_caps (a node with
// more incident cells than the default blocks hold: with P3, more than 102).
inline BlockCaps wide_caps(int nf) {
    BlockCaps c;
    c.max_slots = 8192;
    c.max_values = nf > 20 ? 16384 : 8192;
    return c;
}
constexpr uint16_t kPosDropped = 0xFFFF;  // pair_pos of an entry the pattern drops
constexpr uint8_t kPairDirichlet = 0x80;  // pair_ib flag: the pair's row is Dirichlet

// Bounds of one row block, stored in processing order (one 48-byte record).
struct BlockMeta {
    int64_t s0;       // first CSR slot
    int64_t p0;       // first pair
    int64_t q0;       // value positions of pair p0 + i start at pair_pos[q0 + i*nf]
    int32_t r0;       // first row
    uint16_t nr, ns;  // rows, slots
    uint16_t np, nv;  // pairs, values
    uint32_t kind;    // index of the block's element kind (batched assembly), else 0
    uint32_t pad;
};
static_assert(sizeof(BlockMeta) == 48, "block metadata is one 48-byte record");

// Variables of a multi-variable problem (at most kMaxVars).
constexpr int kMaxVars = 4;

struct AssemblySetup {
    int fem = 0, ncomp = 1, nf = 0, E = 0;  // fem / ncomp: those of variable 0
    // the variables (space, components) in NATURAL order VAR, DIM, ELEM_TYPE,
    // ELEM_ID, DOF_ID (global_enumerator.h:396): variable v's DOFs follow
    // variable v-1's; a tet's local DOFs list variable 0's first
    int n_var = 1;
    int var_fem[kMaxVars] = {}, var_ncomp[kMaxVars] = {};
    // a vector space whose operator pair acts component by component (IDEN,
    // GRAD, DUDX..DUDZ on both sides) with a NULL or SCALAR tensor: the
    // entries coupling different components are exactly zero in every cell,
    // so the reference's value rule (|a| > 1e-100, assembler.inl:416) never
    // inserts them and the pattern leaves them out
    bool decoupled = false;
    int64_t n_nodes = 0, n_tets = 0, n_edges = 0, n_faces = 0;
    int64_t ndof = 0, nnz = 0;
    std::vector<int32_t> tets;        // n_tets x 4 after the node reorder (caller's order)
    std::vector<int32_t> dofs;        // n_tets x nf global DOFs (caller's order)
    std::vector<int8_t> sign;         // n_tets x nf orientation signs of oriented DOFs (RT0, ND0), else empty
    std::vector<uint8_t> is_dir;      // ndof
    std::vector<int64_t> row_ptr;     // ndof + 1
    std::vector<int32_t> col_idx;     // nnz
    std::vector<int32_t> slot;        // caller's order: n_tets x E (entry ib + nf*ia), -1 dropped
    int64_t n_degenerate = 0;         // tets with a zero or non-finite Jacobian determinant
    // row blocks: consecutive CSR rows [blk_row0[b], blk_row0[b+1])
    int64_t n_blocks = 0;
    std::vector<int32_t> blk_row0;    // n_blocks + 1
    std::vector<int32_t> blk_order;   // n_blocks: processing order of the blocks
    std::vector<BlockMeta> blk_meta;  // n_blocks: bounds in processing order
    std::vector<int64_t> blk_s0;      // n_blocks + 1: first CSR slot of each block (row_ptr[blk_row0])
    std::vector<int64_t> blk_p0;      // n_blocks + 1 offsets into the pair lists
    std::vector<int64_t> blk_v0;      // n_blocks + 1 offsets of each block's value list
    std::vector<int32_t> pair_tet;    // per pair: tet id
    std::vector<uint8_t> pair_ib;     // per pair: local row index | kPairDirichlet
    std::vector<uint16_t> pair_pos;   // per pair, nf: value-list position of entry (ib, ia), or kPosDropped
    std::vector<uint16_t> sstart;     // nnz: first value of each slot, relative to its block
    std::vector<uint16_t> pstart;     // ndof: first pair of each row, relative to its block
    int64_t max_block_values = 0, max_block_pairs = 0, max_block_slots = 0;
};

int build_setup(const anifem_mesh_desc& mesh, const anifem_problem_desc& prob, AssemblySetup& S,
                const BlockCaps& caps = BlockCaps());
// The phases of build_setup: argument checks and the node reorder; the
// enumeration, Dirichlet flags, CSR pattern and slot map (host or device, with
// identical results); the DOF -> tet adjacency the row blocks are cut from.
int setup_prepare(const anifem_mesh_desc& mesh, const anifem_problem_desc& prob, AssemblySetup& S);
// setup_prepare for the variables vars[0..n_var) (op codes ignored); the
// problem's trial and test are not used.
int setup_prepare_vars(const anifem_mesh_desc& mesh, int n_var, const anifem_op* vars, AssemblySetup& S);
int build_setup_vars(const anifem_mesh_desc& mesh, int n_var, const anifem_op* vars, const anifem_problem_desc& prob,
                     AssemblySetup& S, const BlockCaps& caps = BlockCaps());
int setup_enumerate_host(const anifem_problem_desc& prob, AssemblySetup& S);
int setup_enumerate_device(const anifem_problem_desc& prob, AssemblySetup& S, void* stream);
void dof_adjacency(const AssemblySetup& S, std::vector<int64_t>& aptr, std::vector<int32_t>& adj);
// build_setup with the enumeration phase on the current device (stream-ordered
// on `stream`); the host phase runs instead for meshes beyond its index ranges.
int build_setup_device(const anifem_mesh_desc& mesh, const anifem_problem_desc& prob, AssemblySetup& S,
                       const BlockCaps& caps, void* stream);
int build_blocks(AssemblySetup& S, const std::vector<int64_t>& aptr, const std::vector<int32_t>& adj,
                 const BlockCaps& caps);
// Compact value positions: per pair, the smallest kept position as 16 bits
// and each entry's offset from it as one byte (0xFF: dropped), little endian
// in ceil((2 + nf) / 4) 32-bit words.  Returns the words per pair, or 0 (out
// empty) when some pair's kept positions span 255 or more.
int compact_positions(const AssemblySetup& S, std::vector<uint32_t>& out);

// DOF layout of a space over the entities of a cell (the DofSymmetries of
// spaces/poly_*.h and spaces/*.h): DOFs per node, edge, face and cell, and
// whether the edge / face DOFs are oriented (global basis = sign x local
// basis, the sign following the global node ids: assembler.inl:154-164,
// ordering.inl:152-172).  The local DOF order is nodes, edges (01, 02, 03,
// 12, 13, 23), faces ({k, k+1, k+2 mod 4}), cell.
struct SpaceLayout {
    int node = 0, edge = 0, face = 0, cell = 0;
    bool edge_oriented = false, face_oriented = false;
    bool valid = false;
};
SpaceLayout space_layout(int fem);

// Parallel loop over [0, n) on the host.
void host_parallel_for(int64_t n, int64_t grain, const std::function<void(int64_t, int64_t)>& fn);

}  // namespace anifem
