// anifem_fem3dtet.hpp -- C++ face of the MI355X library, mirroring the
// reference's template API for the element evaluator and the assembler.
//
// The reference calls
//     Ani::fem3Dtet<Ani::Operator<Ani::GRAD, Ani::FemFix<Ani::FEM_P2>>,
//                   Ani::Operator<Ani::GRAD, Ani::FemFix<Ani::FEM_P2>>,
//                   Ani::DfuncTraits<Ani::TENSOR_SYMMETRIC, false>>(XY0, XY1, XY2, XY3, Dfnc, A, order);
// (anifem++/fem/operations/int_tet.h:17-20, int_tet.inl:92-122).  Here the
// same tags select the operator pair and the tensor kind at compile time, and
// the call evaluates a whole batch of tets on the GPU:
//     anifem_gpu::fem3Dtet<anifem_gpu::Operator<anifem_gpu::GRAD, anifem_gpu::FemFix<anifem_gpu::FEM_P2>>,
//                          anifem_gpu::Operator<anifem_gpu::GRAD, anifem_gpu::FemFix<anifem_gpu::FEM_P2>>,
//                          anifem_gpu::DfuncTraits<anifem_gpu::TENSOR_SYMMETRIC, false>>(xyz, n, D, A, order);
// The coefficient is a device view (per tet or per quadrature point) instead
// of a host functor; a functor is evaluated by the caller at the points of
// anifem_quadrature().  Errors come back as the ANIFEM_ERR_* codes of
// anifem_gpu.h (the reference throws).
#ifndef ANIFEM_FEM3DTET_HPP
#define ANIFEM_FEM3DTET_HPP

#include <cstdint>
#include <initializer_list>
#include <type_traits>
#include <utility>
#include <vector>

#include "anifem_gpu.h"

namespace anifem_gpu {

// Finite elements and operators (numbering of operators.h:25-45)
enum FiniteElement {
    FEM_P0 = ANIFEM_FEM_P0,
    FEM_P1 = ANIFEM_FEM_P1,
    FEM_P2 = ANIFEM_FEM_P2,
    FEM_P3 = ANIFEM_FEM_P3,
    // element-level only (operators.h:26-35; spaces/*.h)
    FEM_RT0 = ANIFEM_FEM_RT0,
    FEM_ND0 = ANIFEM_FEM_ND0,
    FEM_CR1 = ANIFEM_FEM_CR1,
    FEM_B2 = ANIFEM_FEM_B2,
    FEM_B4 = ANIFEM_FEM_B4
};
enum OperatorType {
    IDEN = ANIFEM_IDEN,
    GRAD = ANIFEM_GRAD,
    DIV = ANIFEM_DIV,
    CURL = ANIFEM_CURL,
    DUDX = ANIFEM_DUDX,
    DUDY = ANIFEM_DUDY,
    DUDZ = ANIFEM_DUDZ
};

// FemFix<OP> / FemVec<DIM, OP> (operators.h:57-67)
template <int OP>
using FemFix = std::integral_constant<int, OP>;

template <int DIM, int OP>
struct FemVec {
    using Dim = std::integral_constant<int, DIM>;
    using Base = FemFix<OP>;
};

namespace detail {
template <typename T>
struct SpaceOf {
    static constexpr int fem = T::value;
    static constexpr int ncomp = 1;
};
template <int DIM, int OP>
struct SpaceOf<FemVec<DIM, OP>> {
    static constexpr int fem = OP;
    static constexpr int ncomp = DIM;
};
}  // namespace detail

// Operator<OPERATOR, FEMTYPE> (operators.h:50-51,127-155)
template <int OPERATOR, typename FEMTYPE>
struct Operator {
    static_assert(OPERATOR == IDEN || OPERATOR == GRAD || OPERATOR == DIV || OPERATOR == CURL || OPERATOR == DUDX ||
                      OPERATOR == DUDY || OPERATOR == DUDZ,
                  "operators of operators.h:37-45");
    static constexpr anifem_op c() {
        return anifem_op{OPERATOR, detail::SpaceOf<FEMTYPE>::fem, detail::SpaceOf<FEMTYPE>::ncomp};
    }
};

// Tensor kinds (diff_tensor.h:16-27) and DfuncTraits (diff_tensor.h:36-51)
enum TensorType {
    TENSOR_NULL = ANIFEM_TENSOR_NULL,
    TENSOR_SCALAR = ANIFEM_TENSOR_SCALAR,
    TENSOR_SYMMETRIC = ANIFEM_TENSOR_SYMMETRIC,
    TENSOR_GENERAL = ANIFEM_TENSOR_GENERAL,
    TENSOR_LAME = ANIFEM_TENSOR_LAME
};
enum TensorTypeSparsity { PerPoint = 0, PerTetra = -1, PerSelection = -2 };

template <int TensorTypeSparse = PerPoint, bool isConstant = false>
struct DfuncTraits {
    using IsConstant = std::integral_constant<bool, isConstant>;
    using TensorSparsity = std::integral_constant<int, TensorTypeSparse>;
};

// Device view of the coefficient: storage ANIFEM_COEF_CONST / PER_TET /
// PER_POINT, layout as documented in anifem_gpu.h.
struct TensorView {
    int type;
    int storage;
    const double* data;
};

enum Precision { FP64 = ANIFEM_PREC_FP64, FP32 = ANIFEM_PREC_FP32, FP16X3 = ANIFEM_PREC_FP16X3 };

// Checks a view against the traits: a fixed TensorType must match, and a
// constant tensor must use constant storage.
template <typename Traits>
inline int check_view(const TensorView& D) {
    if (Traits::TensorSparsity::value > 0 && D.type != Traits::TensorSparsity::value) return ANIFEM_ERR_ARG;
    if (Traits::IsConstant::value && D.storage != ANIFEM_COEF_CONST) return ANIFEM_ERR_ARG;
    return ANIFEM_OK;
}

// Batched fem3Dtet: element matrices of n_tets tets (xyz: device 12 x n SoA;
// A_out: device, n matrices nfB x nfA column-major, double for FP64, float for
// FP32); stream is a hipStream_t.
template <typename OpA, typename OpB, typename FuncTraits = DfuncTraits<>>
inline int fem3Dtet(const double* xyz, int64_t n_tets, const TensorView& D, void* A_out, int order = 5,
                    Precision precision = FP64, void* stream = nullptr) {
    const int rc = check_view<FuncTraits>(D);
    if (rc) return rc;
    const anifem_tensor t{D.type, D.storage, D.data};
    return anifem_fem3dtet_batch(OpA::c(), OpB::c(), &t, order, n_tets, xyz, precision, A_out, stream);
}

// Batched fem3Dface / fem3DfaceN (int_face.h:15-45): integrals over face
// `face_num` (through vertices face_num, face_num+1, face_num+2 mod 4) of each
// tet; in fem3DfaceN the tensor has 3*dimB rows and is contracted with the
// outward unit normal.  Layouts as fem3Dtet; FP64 or FP32.
template <typename OpA, typename OpB, typename FuncTraits = DfuncTraits<>>
inline int fem3Dface(const double* xyz, int64_t n_tets, int face_num, const TensorView& D, void* A_out,
                     int order = 5, Precision precision = FP64, void* stream = nullptr) {
    const int rc = check_view<FuncTraits>(D);
    if (rc) return rc;
    const anifem_tensor t{D.type, D.storage, D.data};
    return anifem_fem3dface_batch(OpA::c(), OpB::c(), &t, order, face_num, 0, n_tets, xyz, precision, A_out, stream);
}
template <typename OpA, typename OpB, typename FuncTraits = DfuncTraits<>>
inline int fem3DfaceN(const double* xyz, int64_t n_tets, int face_num, const TensorView& D, void* A_out,
                      int order = 5, Precision precision = FP64, void* stream = nullptr) {
    const int rc = check_view<FuncTraits>(D);
    if (rc) return rc;
    const anifem_tensor t{D.type, D.storage, D.data};
    return anifem_fem3dface_batch(OpA::c(), OpB::c(), &t, order, face_num, 1, n_tets, xyz, precision, A_out, stream);
}

// Batched fem3DapplyL / fem3DapplyX (eval.h:13-31): Op(u) at q barycentric
// points common to all tets (xyl, host) or at q physical points per tet
// (X, device, X[(r*q + n)*3 + c]); dofs n x nfa, opU n x q x dim (device).
template <typename Op>
inline int fem3DapplyL(const double* xyz, int64_t n_tets, const double* xyl, int q, const double* dofs, double* opU,
                       void* stream = nullptr) {
    return anifem_fem3dapply_batch(Op::c(), xyl, q, n_tets, xyz, dofs, opU, stream);
}
template <typename Op>
inline int fem3DapplyX(const double* xyz, int64_t n_tets, const double* X, int q, const double* dofs, double* opU,
                       void* stream = nullptr) {
    return anifem_fem3dapplyx_batch(Op::c(), q, n_tets, xyz, X, dofs, opU, stream);
}

// Nfa / Dim of an operator (poly_*.h, operators.h:127-155)
template <typename Op>
inline int nfa() {
    return anifem_op_nfa(Op::c());
}
template <typename Op>
inline int dim() {
    return anifem_op_dim(Op::c());
}

// ---- Operators and spaces chosen at run time ------------------------------
// ApplyOp (operators.h:418-533: ApplyOpBase / ApplyOp / ApplyOpFromTemplate)
// and FemSpace (fem_space.h:404-596) for callers that pick the operator and
// the space at run time (the runtime fem3Dtet of int_tet.h:145-147).  Here an
// ApplyOp is the anifem_op descriptor every entry point takes; its Nfa / Dim
// come from the library, its Order (the polynomial degree of the operator's
// values: Operator<..>::Order of spaces/*.h) from the space's degree.
namespace detail {
inline int space_degree(int fem) {
    switch (fem) {
        case FEM_P0: return 0;
        case FEM_P1: return 1;
        case FEM_P2: return 2;
        case FEM_P3: return 3;
        case FEM_RT0: return 1;
        case FEM_ND0: return 1;
        case FEM_CR1: return 1;
        case FEM_B2: return 2;
        case FEM_B4: return 4;
        default: return -1;
    }
}
// vector-valued reference fields (RT0, ND0): 3 components per basis function
inline int space_dim(int fem) { return fem == FEM_RT0 || fem == FEM_ND0 ? 3 : 1; }
}  // namespace detail

class ApplyOp {
public:
    ApplyOp() = default;
    explicit ApplyOp(anifem_op op) : op_(op), set_(true) {}
    // a descriptor the library does not know has no Nfa / Dim (-1)
    int Nfa() const { return set_ ? anifem_op_nfa(op_) : -1; }
    int Dim() const { return set_ ? anifem_op_dim(op_) : -1; }
    int Order() const {
        const int d = set_ ? detail::space_degree(op_.fem) : -1;
        if (d < 0 || Dim() < 0) return -1;
        return op_.op == IDEN ? d : (d > 0 ? d - 1 : 0);
    }
    bool valid() const { return Nfa() > 0 && Dim() > 0; }
    anifem_op c() const { return op_; }
    bool operator==(const ApplyOp& o) const {
        return set_ == o.set_ && (!set_ || (op_.op == o.op_.op && op_.fem == o.op_.fem && op_.ncomp == o.op_.ncomp));
    }
    bool operator!=(const ApplyOp& o) const { return !(*this == o); }

private:
    anifem_op op_{0, 0, 0};
    bool set_ = false;
};

// ApplyOpFromTemplate<OPERATOR, FEMTYPE> (operators.h:475-533)
template <int OPERATOR, typename FEMTYPE>
inline ApplyOp ApplyOpFromTemplate() {
    return ApplyOp(Operator<OPERATOR, FEMTYPE>::c());
}

// The spaces of spaces/*.h by their reference names (P1Space = GenSpace<
// Polynomial, FEM_P1>, poly_1.h:170; ...), converted to a FemSpace.
template <int FEM>
struct GenSpaceTag {
    static constexpr int fem = FEM;
};
using P0Space = GenSpaceTag<FEM_P0>;
using P1Space = GenSpaceTag<FEM_P1>;
using P2Space = GenSpaceTag<FEM_P2>;
using P3Space = GenSpaceTag<FEM_P3>;
using RT0Space = GenSpaceTag<FEM_RT0>;
using ND0Space = GenSpaceTag<FEM_ND0>;
using CR1Space = GenSpaceTag<FEM_CR1>;
using Bubble2Space = GenSpaceTag<FEM_B2>;
using Bubble4Space = GenSpaceTag<FEM_B4>;
using BubbleSpace = Bubble4Space;

// FemSpace: a space of one element type (UniteFemSpace), its vector power
// (VectorFemSpace, a^k) or a product of such spaces (ComplexFemSpace, a*b):
// the several variables of anifem_assembly_create_vars, local DOFs
// concatenated in factor order.
class FemSpace {
public:
    FemSpace() = default;
    template <int FEM>
    FemSpace(GenSpaceTag<FEM>) : parts_{anifem_op{IDEN, FEM, 1}} {}
    bool empty() const { return parts_.empty(); }
    bool isComplex() const { return parts_.size() > 1; }
    // number of factors and factor i (a unite or vector space)
    int nparts() const { return (int)parts_.size(); }
    FemSpace part(int i) const {
        FemSpace f;
        if (i >= 0 && i < nparts()) f.parts_.push_back(parts_[i]);
        return f;
    }
    // dimension of the space's values (fem_space.h: dim())
    int dim() const {
        int d = 0;
        for (const anifem_op& p : parts_) d += p.ncomp * detail::space_dim(p.fem);
        return d;
    }
    // local DOFs per cell
    int nfa() const {
        int n = 0;
        for (const anifem_op& p : parts_) n += anifem_op_nfa(p);
        return n;
    }
    // the operator op on a unite or vector space (getOP, dimOP, orderOP); a
    // product space has no single operator here: its variables are assembled
    // through anifem_assembly_create_vars
    ApplyOp getOP(int op) const {
        if (parts_.size() != 1) return ApplyOp();
        return ApplyOp(anifem_op{op, parts_[0].fem, parts_[0].ncomp});
    }
    int dimOP(int op) const { return getOP(op).Dim(); }
    int orderOP(int op) const { return getOP(op).Order(); }
    // the descriptor of anifem_assembly_create_vars (at most ANIFEM_MAX_VARS factors)
    int vars(anifem_vars_desc* out) const {
        if (!out || parts_.empty() || parts_.size() > ANIFEM_MAX_VARS) return ANIFEM_ERR_ARG;
        *out = anifem_vars_desc{};
        out->n_var = (int32_t)parts_.size();
        for (size_t i = 0; i < parts_.size(); ++i) out->var[i] = parts_[i];
        return ANIFEM_OK;
    }
    // a^k: the vector space of k copies of a unite space (VectorFemSpace)
    friend FemSpace operator^(const FemSpace& a, int k) {
        FemSpace v;
        if (a.parts_.size() == 1 && a.parts_[0].ncomp == 1 && k >= 1)
            v.parts_.push_back(anifem_op{IDEN, a.parts_[0].fem, k});
        return v;
    }
    // a*b: the product of two spaces (ComplexFemSpace), factors concatenated
    friend FemSpace operator*(const FemSpace& a, const FemSpace& b) {
        FemSpace c;
        if (a.empty() || b.empty()) return c;
        c.parts_ = a.parts_;
        c.parts_.insert(c.parts_.end(), b.parts_.begin(), b.parts_.end());
        return c;
    }
    bool operator==(const FemSpace& o) const {
        if (parts_.size() != o.parts_.size()) return false;
        for (size_t i = 0; i < parts_.size(); ++i)
            if (parts_[i].fem != o.parts_[i].fem || parts_[i].ncomp != o.parts_[i].ncomp) return false;
        return true;
    }

private:
    std::vector<anifem_op> parts_;  // op unused: IDEN
};

// The batched calls with operators chosen at run time (int_tet.h:145-147)
inline int fem3Dtet(const double* xyz, int64_t n_tets, const TensorView& D, const ApplyOp& opA, const ApplyOp& opB,
                    void* A_out, int order = 5, Precision precision = FP64, void* stream = nullptr) {
    if (!opA.valid() || !opB.valid()) return ANIFEM_ERR_ARG;
    const anifem_tensor t{D.type, D.storage, D.data};
    return anifem_fem3dtet_batch(opA.c(), opB.c(), &t, order, n_tets, xyz, precision, A_out, stream);
}
inline int fem3Dface(const double* xyz, int64_t n_tets, int face_num, const TensorView& D, const ApplyOp& opA,
                     const ApplyOp& opB, void* A_out, int order = 5, Precision precision = FP64,
                     void* stream = nullptr) {
    if (!opA.valid() || !opB.valid()) return ANIFEM_ERR_ARG;
    const anifem_tensor t{D.type, D.storage, D.data};
    return anifem_fem3dface_batch(opA.c(), opB.c(), &t, order, face_num, 0, n_tets, xyz, precision, A_out, stream);
}
inline int fem3DfaceN(const double* xyz, int64_t n_tets, int face_num, const TensorView& D, const ApplyOp& opA,
                      const ApplyOp& opB, void* A_out, int order = 5, Precision precision = FP64,
                      void* stream = nullptr) {
    if (!opA.valid() || !opB.valid()) return ANIFEM_ERR_ARG;
    const anifem_tensor t{D.type, D.storage, D.data};
    return anifem_fem3dface_batch(opA.c(), opB.c(), &t, order, face_num, 1, n_tets, xyz, precision, A_out, stream);
}
inline int fem3DapplyL(const double* xyz, int64_t n_tets, const double* xyl, int q, const ApplyOp& op,
                       const double* dofs, double* opU, void* stream = nullptr) {
    if (!op.valid()) return ANIFEM_ERR_ARG;
    return anifem_fem3dapply_batch(op.c(), xyl, q, n_tets, xyz, dofs, opU, stream);
}

// Whole-mesh assembler with operators chosen at run time: PrepareProblem at
// construction (assembler.inl:194-275), Assemble(matrix, rhs) per run
// (assembler.inl:314-488).  The trial and test spaces are the operators'
// spaces; the load vector is fem3Dtet<IDEN P0, IDEN test space> of the
// right-hand-side tensor (examples/tutorials/ex1.cpp:95).  Constructed from a
// FemSpace, it assembles the space's variables (a product space: several)
// from element matrices the caller forms over the concatenated local DOFs.
class RuntimeAssembler {
public:
    RuntimeAssembler(const ApplyOp& trial, const ApplyOp& test, int64_t n_nodes, const double* coords, int64_t n_tets,
                     const int32_t* tets, int order, TensorView D_kind, int rhs_order, TensorView F_kind,
                     int dirichlet, const uint8_t* dirichlet_dof_mask = nullptr, void* stream = nullptr) {
        if (!trial.valid() || !test.valid()) {
            status_ = ANIFEM_ERR_ARG;
            return;
        }
        anifem_mesh_desc mesh{n_nodes, n_tets, coords, tets};
        anifem_problem_desc prob{};
        prob.trial = trial.c();
        prob.test = test.c();
        prob.order = order;
        prob.tensor_type = D_kind.type;
        prob.tensor_storage = D_kind.storage;
        prob.rhs_order = rhs_order;
        prob.rhs_tensor_type = F_kind.type;
        prob.rhs_tensor_storage = F_kind.storage;
        prob.dirichlet = dirichlet;
        prob.dirichlet_dof_mask = dirichlet_dof_mask;
        status_ = anifem_assembly_create(&ctx_, &mesh, &prob, stream);
        if (status_) ctx_ = nullptr;
    }
    // the variables of a space (AssembleElements only); dirichlet as above,
    // on every variable
    RuntimeAssembler(const FemSpace& space, int64_t n_nodes, const double* coords, int64_t n_tets,
                     const int32_t* tets, int dirichlet, const uint8_t* dirichlet_dof_mask = nullptr,
                     void* stream = nullptr) {
        anifem_vars_desc vars{};
        status_ = space.vars(&vars);
        if (status_) return;
        anifem_mesh_desc mesh{n_nodes, n_tets, coords, tets};
        anifem_problem_desc prob{};
        prob.dirichlet = dirichlet;
        prob.dirichlet_dof_mask = dirichlet_dof_mask;
        status_ = anifem_assembly_create_vars(&ctx_, &mesh, &vars, &prob, stream);
        if (status_) ctx_ = nullptr;
    }
    RuntimeAssembler(const RuntimeAssembler&) = delete;
    RuntimeAssembler& operator=(const RuntimeAssembler&) = delete;
    ~RuntimeAssembler() {
        if (ctx_) anifem_assembly_destroy(ctx_);
    }
    int status() const { return status_; }
    int64_t ndof() const { return info().first; }
    int64_t nnz() const { return info().second; }
    // Device CSR pattern views
    int csr(const int64_t** row_ptr, const int32_t** col_idx) const {
        return ctx_ ? anifem_assembly_csr(ctx_, row_ptr, col_idx) : ANIFEM_ERR_ARG;
    }
    // Assemble(A, b): device csr_val (nnz) and rhs (ndof) are overwritten;
    // dir_values (device, ndof, or null = homogeneous) are the Dirichlet
    // values applyDir imposes per cell (dc_on_dof.h:38-45).
    int Assemble(const TensorView& D, const TensorView* F, double* csr_val, double* rhs,
                 anifem_status* st = nullptr, Precision precision = FP64, void* stream = nullptr,
                 const double* dir_values = nullptr) {
        if (!ctx_) return status_ ? status_ : ANIFEM_ERR_ARG;
        const anifem_tensor tD{D.type, D.storage, D.data};
        anifem_tensor tF{};
        if (F) tF = anifem_tensor{F->type, F->storage, F->data};
        return anifem_assembly_run(ctx_, &tD, F ? &tF : nullptr, dir_values, precision, csr_val, rhs, st, stream);
    }
    // Assemble from element matrices computed by the caller (the reference's
    // local-function path): A_elem n_tets x nf x nf column-major, F_elem
    // n_tets x nf or null, FP64 (double) or FP32 (float) values, in the local
    // order of the reordered nodes (Tets()).
    template <typename TE>
    int AssembleElements(const TE* A_elem, const TE* F_elem, double* csr_val, double* rhs,
                         anifem_status* st = nullptr, void* stream = nullptr) {
        static_assert(sizeof(TE) == 8 || sizeof(TE) == 4, "double or float element values");
        if (!ctx_) return status_ ? status_ : ANIFEM_ERR_ARG;
        return anifem_assembly_run_elements(ctx_, A_elem, F_elem, sizeof(TE) == 8 ? FP64 : FP32, csr_val, rhs, st,
                                            stream);
    }
    // AssembleRHS(rhs) (assembler.inl:497-561): rhs from per-cell load vectors
    // F_elem (n_tets x nf), stored (accumulate = false) or added (rhs <- rhs +
    // assembled, the reference's semantics).
    template <typename TE>
    int AssembleRHS(const TE* F_elem, double* rhs, bool accumulate = true, anifem_status* st = nullptr,
                    void* stream = nullptr) {
        static_assert(sizeof(TE) == 8 || sizeof(TE) == 4, "double or float load vectors");
        if (!ctx_) return status_ ? status_ : ANIFEM_ERR_ARG;
        return anifem_assembly_run_rhs(ctx_, F_elem, sizeof(TE) == 8 ? FP64 : FP32, rhs, accumulate ? 1 : 0, st,
                                       stream);
    }
    // Host copy of the tets' nodes after the orientation reorder (n_tets x 4)
    int Tets(int32_t* tets) const { return ctx_ ? anifem_assembly_tets(ctx_, tets) : ANIFEM_ERR_ARG; }
    anifem_assembly* handle() const { return ctx_; }

private:
    std::pair<int64_t, int64_t> info() const {
        int64_t nd = 0, nz = 0, nb = 0;
        int32_t nf = 0;
        if (ctx_) anifem_assembly_info(ctx_, &nd, &nz, &nb, &nf);
        return {nd, nz};
    }
    anifem_assembly* ctx_ = nullptr;
    int status_ = ANIFEM_OK;
};

// Whole-mesh assembler of the reference's template form: the test and trial
// spaces are the Op tags' spaces.
template <typename OpA, typename OpB>
class Assembler : public RuntimeAssembler {
public:
    Assembler(int64_t n_nodes, const double* coords, int64_t n_tets, const int32_t* tets, int order,
              TensorView D_kind, int rhs_order, TensorView F_kind, int dirichlet,
              const uint8_t* dirichlet_dof_mask = nullptr, void* stream = nullptr)
        : RuntimeAssembler(ApplyOp(OpA::c()), ApplyOp(OpB::c()), n_nodes, coords, n_tets, tets, order, D_kind,
                           rhs_order, F_kind, dirichlet, dirichlet_dof_mask, stream) {}
};

}  // namespace anifem_gpu

#endif  // ANIFEM_FEM3DTET_HPP
